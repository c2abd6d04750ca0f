// mhd2d_glm_oracle.cpp -- serial CPU oracle executable (test infrastructure).
//
// Same command line, stdout summary, exit codes and .dat layout as the
// reference CPU driver (mhd2d_glm_cpu.cpp:417-676), so its data rows can be
// md5-compared with the reference build and with the HIP product.  The
// optional 8th argument selects the problem (ot | kh | khb), as the reference
// CUDA driver does (mhd2d_glm_gpu.cu:368-374).
//
// Build: g++ -O2 -std=c++14 [-DORACLE_FLOAT] (no -march: nothing contracts).
// Env: OUTDIR (output directory), MHD_HLLD_TOLERANCE (default 1e-4, the value
// every reference timing and checksum uses).
#include "mhd_oracle.hpp"
#include "problems.hpp"

#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>

#ifdef ORACLE_FLOAT
typedef float Real;
#else
typedef double Real;
#endif

using namespace oracle;

int main(int argc, char* argv[]) {
  int nx = 128, ny = 128;
  Real t_end = (Real)0.5, cfl = (Real)0.25, alpha = (Real)0.18;
  std::string order = "muscl", solver = "hlld", problem = "ot";
  if (argc > 1) nx = std::atoi(argv[1]);
  if (argc > 2) ny = std::atoi(argv[2]);
  if (argc > 3) t_end = (Real)std::atof(argv[3]);
  if (argc > 4) cfl = (Real)std::atof(argv[4]);
  if (argc > 5) order = argv[5];
  if (argc > 6) solver = argv[6];
  if (argc > 7) alpha = (Real)std::atof(argv[7]);
  if (argc > 8) problem = argv[8];
  double tol = 1.0e-4;
  if (const char* s = std::getenv("MHD_HLLD_TOLERANCE")) tol = std::atof(s);

  if (nx <= 0 || ny <= 0 || t_end <= (Real)0.0) {
    std::cerr << "Error: invalid grid or final time." << std::endl;
    return 1;
  }
  if (cfl <= (Real)0.0 || cfl >= (Real)1.0) {
    std::cerr << "Error: CFL must be in (0, 1)." << std::endl;
    return 1;
  }
  if (order != "muscl" && order != "first") {
    std::cerr << "Error: order must be 'muscl' or 'first'." << std::endl;
    return 1;
  }
  if (solver != "hll" && solver != "hlld") {
    std::cerr << "Error: solver must be 'hll' or 'hlld'." << std::endl;
    return 1;
  }
  if (alpha <= (Real)0.0) {
    std::cerr << "Error: alpha must be positive." << std::endl;
    return 1;
  }
  if (problem != "ot" && problem != "kh" && problem != "khb") {
    std::cerr << "Error: problem must be 'ot', 'kh' or 'khb'." << std::endl;
    return 1;
  }
  const bool use_muscl = (order == "muscl");
  const bool use_hlld = (solver == "hlld");
  const Consts<Real> c(tol);
  const Real dx = ((Real)1.0 - (Real)0.0) / (Real)nx;
  const Real dy = ((Real)1.0 - (Real)0.0) / (Real)ny;

  std::cout << "===== 2D ideal MHD Orszag-Tang vortex with mixed GLM =====" << std::endl;
  std::cout << "Solver: " << (use_hlld ? "HLLD" : "HLL") << std::endl;
  std::cout << "Reconstruction/time: "
            << (use_muscl ? "MUSCL + SSPRK2" : "first-order + SSPRK2") << std::endl;
  std::cout << "Precision: " << (int)(sizeof(Real) * 8) << " bit"
            << (sizeof(Real) == 8 ? " (fp64)" : " (fp32)") << std::endl;
  std::cout << "gamma = " << c.gamma << ", alpha = " << alpha << std::endl;
  std::cout << "Nx x Ny = " << nx << " x " << ny << ", CFL = " << cfl
            << ", t_end = " << t_end << std::endl;

  Grid<Real> U(nx, ny), U1(nx, ny);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      Real P[NV];
      initial_primitive<Real>(problem, i, j, dx, dy, P);
      prim_to_state(c, P, U.at(i, j));
    }

  Real t = (Real)0.0, last_ch = (Real)0.0;
  int step = 0;
  FallbackStats stats;
  auto w0 = std::chrono::steady_clock::now();
  while (t < t_end) {
    StepResult r = ssprk2_step(c, U, U1, dx, dy, cfl, alpha, use_muscl,
                               use_hlld, t, t_end, &stats, &last_ch);
    if (r.code == 2) {
      std::cerr << "Error: non-positive maximum signal rate." << std::endl;
      return 2;
    }
    if (r.code == 3 || r.code == 4) {
      std::cerr << (r.code == 3 ? "Error: non-physical RK stage at step "
                                : "Error: non-physical state at step ")
                << step << ", cell (" << r.bad_i << ", " << r.bad_j << ")"
                << std::endl;
      return r.code;
    }
    step++;
  }
  auto w1 = std::chrono::steady_clock::now();
  const double wall = std::chrono::duration<double>(w1 - w0).count();

  Real div_l1 = (Real)0.0, div_linf = (Real)0.0, psi_linf = (Real)0.0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const Real d = div_b(U, i, j, dx, dy);
      div_l1 += std::fabs(d);
      div_linf = smax(div_linf, std::fabs(d));
      psi_linf = smax(psi_linf, (Real)std::fabs(U.at(i, j)[PSI]));
    }
  div_l1 /= (Real)(nx * ny);

  const uint64_t evals = stats.x_evals + stats.y_evals;
  const uint64_t fb = stats.degenerate + stats.nonphysical;
  const double frac = evals > 0 ? (double)fb / (double)evals : 0.0;
  std::cout << "Done. Steps: " << step << ", final t = " << t << std::endl;
  std::cout << "Time loop wall time = " << wall << " s ("
            << ((double)nx * ny * step / wall / 1.0e6) << " Mcell-steps/s)"
            << std::endl;
  std::cout << "c_h(last) = " << last_ch << ", alpha = " << alpha << std::endl;
  std::cout << "divB L1 = " << div_l1 << ", divB Linf = " << div_linf
            << ", psi Linf = " << psi_linf << std::endl;
  if (use_hlld) {
    std::cout << "HLLD flux evaluations = " << evals << " (x=" << stats.x_evals
              << ", y=" << stats.y_evals << ")" << std::endl;
    std::cout << "HLLD relative fallback tolerance = " << std::scientific
              << (double)c.tol << std::defaultfloat << std::endl;
    std::cout << "HLLD->HLL fallbacks = " << fb << " (degenerate="
              << stats.degenerate << ", nonphysical=" << stats.nonphysical
              << "), fraction = " << std::scientific << frac
              << std::defaultfloat << std::endl;
  }

  const char* out_env = std::getenv("OUTDIR");
  const std::string out_dir = out_env ? out_env : "results/mhd/orszag_tang_glm";
  if (std::system(("mkdir -p " + out_dir).c_str()) != 0) { /* ignore */ }
  const std::string base = (problem == "ot") ? "orszag_tang" : problem;
  std::ostringstream fname;
  fname << out_dir << "/" << base << "_glm_" << order << "_" << solver << "_"
        << (sizeof(Real) == 8 ? "fp64" : "fp32") << "_N" << nx << "x" << ny
        << ".dat";
  std::ofstream out(fname.str());
  if (!out) {
    std::cerr << "Error: failed to open output file: " << fname.str() << std::endl;
    return 5;
  }
  out << "# Orszag-Tang ideal MHD with mixed GLM\n";
  out << "# nx " << nx << " ny " << ny << " t " << t << " gamma " << c.gamma
      << " cfl " << cfl << " order " << order << " solver " << solver
      << " alpha " << alpha << " ch_last " << last_ch << "\n";
  if (use_hlld) {
    out << "# hlld_evaluations " << evals << " hlld_tolerance " << c.tol
        << " hlld_x_evaluations " << stats.x_evals << " hlld_y_evaluations "
        << stats.y_evals << " hlld_fallbacks " << fb << " hlld_degenerate "
        << stats.degenerate << " hlld_nonphysical " << stats.nonphysical
        << " hlld_fallback_fraction " << frac << "\n";
  }
  out << "# columns: x y rho vx vy vz p Bx By Bz E divB psi\n";
  out << std::scientific << std::setprecision(std::numeric_limits<Real>::max_digits10);
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      const Real* u = U.at(i, j);
      Real W[NV];
      state_to_prim(c, u, W);
      const Real x = (Real)0.0 + ((Real)i + (Real)0.5) * dx;
      const Real y = (Real)0.0 + ((Real)j + (Real)0.5) * dy;
      out << x << " " << y << " " << W[0] << " " << W[1] << " " << W[2] << " "
          << W[3] << " " << W[4] << " " << W[5] << " " << W[6] << " " << W[7]
          << " " << u[EN] << " " << div_b(U, i, j, dx, dy) << " " << u[PSI]
          << "\n";
    }
  }
  std::cout << "Saved: " << fname.str() << std::endl;
  return 0;
}

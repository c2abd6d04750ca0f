// mhd2d_glm_amd.cpp -- command-line driver, drop-in for the reference
// executables mhd2d_glm_gpu_{fp64,fp32} (mhd2d_glm_gpu.cu:353-653) and
// mhd2d_glm_cpu (mhd2d_glm_cpu.cpp:417-676).
//
//   mhd2d_glm_amd nx ny t_end cfl {muscl|first} {hll|hlld} alpha [ot|kh|khb]
//
// Same argv parsing (atoi/atof, defaults 128 128 0.5 0.25 muscl hlld 0.18),
// OUTDIR convention, stdout lines parsed by the reference scripts
// ("Done. Steps: N", "Time loop wall time = X s", "HLLD->HLL fallbacks = N"),
// .dat name and 13-column layout, and exit codes (1 usage, 2 rate, 3/4
// non-physical state, 5 output file, 9 device).  Precision comes from the
// program name (a name ending in "fp32" runs FP32, as the reference's
// separate -DUSE_FLOAT binary does; a name ending in "_gpu_mixed" runs the
// reference's stage-level mixed form, as mhd2d_glm_gpu_mixed does) or from
// MHD_PRECISION (fp64, fp32, mixed, mixed_kahan, fp64_fma, ref_mixed).  MHD_HLLD_TOLERANCE sets the HLLD tolerance
// (default 1e-4, the value of every reference timing and checksum run);
// MHD_TILE the stage tile shape, MHD_LOWBETA_RATIO the mixed form's FP64
// recovery switch, MHD_FALLBACK its FP64 tile fallback (0 on, 1 off),
// MHD_ROLLBACK its FP64 rollback window in steps (< 0 off) and MHD_FUSE the
// SSPRK2 stages per kernel launch (1 or 2).
//
// This program only talks to the solver through include/mhd_glm.h.
#include "mhd_glm.h"
#include "npz_reader.hpp"
#include "mhd_problems.hpp"

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Args {
  int nx = 128, ny = 128;
  double t_end_raw = 0.5, cfl_raw = 0.25, alpha_raw = 0.18;
  std::string order = "muscl", solver = "hlld", problem = "ot";
};

template <class R>
struct Field {
  int nx, ny;
  std::vector<R> u;  // SoA, plane k at k*nx*ny
  R at(int i, int j, int k) const {
    i = ((i % nx) + nx) % nx;
    j = ((j % ny) + ny) % ny;
    return u[(size_t)k * nx * ny + (size_t)j * nx + i];
  }
};

// Central-difference div B (mhd2d_glm_cpu.cpp:347-358).
template <class R>
R div_b(const Field<R>& f, int i, int j, R dx, R dy) {
  return (f.at(i + 1, j, 5) - f.at(i - 1, j, 5)) / ((R)2.0 * dx) +
         (f.at(i, j + 1, 6) - f.at(i, j - 1, 6)) / ((R)2.0 * dy);
}

template <class R>
void print_failure(const char* label, const Field<R>& f, int bi, int bj, R dx, R dy,
                   R t, R dt, int step, R ch) {
  // printFailureDiagnostics (mhd2d_glm_cpu.cpp:364-415).
  R U[9], P[9];
  for (int k = 0; k < 9; k++) U[k] = f.at(bi, bj, k);
  mhd::host_cons_to_prim<R>(U, P);
  const R rho = U[0];
  const R rs = (rho < (R)1.0e-12) ? (R)1.0e-12 : rho;
  const R vx = U[1] / rs, vy = U[2] / rs, vz = U[3] / rs;
  const R kinetic = (R)0.5 * rs * (vx * vx + vy * vy + vz * vz);
  const R magnetic = (R)0.5 * (U[5] * U[5] + U[6] * U[6] + U[7] * U[7]);
  const R internal = U[4] - kinetic - magnetic;
  const R p = ((R)(5.0 / 3.0) - (R)1.0) * internal;
  std::cerr << std::setprecision(std::numeric_limits<R>::max_digits10);
  std::cerr << "Failure diagnostics (" << label << ")\n";
  std::cerr << "  step=" << step << " t=" << t << " dt=" << dt << " ch=" << ch << "\n";
  std::cerr << "  cell=(" << bi << ", " << bj << ")"
            << " x=" << ((R)bi + (R)0.5) * dx << " y=" << ((R)bj + (R)0.5) * dy << "\n";
  std::cerr << "  rho_raw=" << rho << " rho_prim=" << P[0] << " p=" << p
            << " p_prim=" << P[4] << "\n";
  std::cerr << "  vx=" << vx << " vy=" << vy << " vz=" << vz << "\n";
  std::cerr << "  Bx=" << U[5] << " By=" << U[6] << " Bz=" << U[7] << " psi=" << U[8] << "\n";
  std::cerr << "  E=" << U[4] << " kinetic=" << kinetic << " magnetic=" << magnetic
            << " internal=" << internal << "\n";
  std::cerr << "  divB=" << div_b(f, bi, bj, dx, dy) << "\n";
}

template <class R>
int run(const Args& a, int precision, const char* prec_name, double tol) {
  const R t_end = (R)a.t_end_raw, cfl = (R)a.cfl_raw, alpha = (R)a.alpha_raw;
  if (a.nx <= 0 || a.ny <= 0 || t_end <= (R)0.0) {
    std::cerr << "Error: invalid grid or final time." << std::endl;
    return 1;
  }
  if (cfl <= (R)0.0 || cfl >= (R)1.0) {
    std::cerr << "Error: CFL must be in (0, 1)." << std::endl;
    return 1;
  }
  if (a.order != "muscl" && a.order != "first") {
    std::cerr << "Error: order must be 'muscl' or 'first'." << std::endl;
    return 1;
  }
  if (a.solver != "hll" && a.solver != "hlld") {
    std::cerr << "Error: solver must be 'hll' or 'hlld'." << std::endl;
    return 1;
  }
  if (alpha <= (R)0.0) {
    std::cerr << "Error: alpha must be positive." << std::endl;
    return 1;
  }
  // Problems the command line can name: those without parameters ("otv"
  // needs five, which argv does not carry).
  if (!mhd::problem_known(a.problem.c_str()) || mhd::problem_nparams(a.problem.c_str()) != 0) {
    std::cerr << "Error: problem must be 'ot', 'kh' or 'khb'." << std::endl;
    return 1;
  }
  const bool use_muscl = a.order == "muscl", use_hlld = a.solver == "hlld";
  if (precision == MHD_PREC_REF_MIXED && !use_muscl) {
    std::cerr << "Error: the reference-mixed form needs 'muscl'." << std::endl;
    return 1;
  }
  const int nx = a.nx, ny = a.ny;
  const R dx = ((R)1.0 - (R)0.0) / (R)nx, dy = ((R)1.0 - (R)0.0) / (R)ny;
  const R gamma = (R)(5.0 / 3.0);

  std::cout << "===== 2D ideal MHD Orszag-Tang vortex with mixed GLM (GPU) ====="
            << std::endl;
  std::cout << "Solver: " << (use_hlld ? "HLLD" : "HLL") << std::endl;
  std::cout << "Reconstruction/time: "
            << (use_muscl ? "MUSCL + SSPRK2" : "first-order + SSPRK2") << std::endl;
  if (precision == MHD_PREC_FP64 || precision == MHD_PREC_FP32)
    std::cout << "Precision: " << (int)(sizeof(R) * 8) << " bit"
              << (sizeof(R) == 8 ? " (fp64)" : " (fp32)") << std::endl;
  else
    std::cout << "Precision: " << prec_name << std::endl;
  std::cout << "gamma = " << gamma << ", alpha = " << alpha << std::endl;
  std::cout << "Nx x Ny = " << nx << " x " << ny << ", CFL = " << cfl
            << ", t_end = " << t_end << std::endl;

  mhd_cfg cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  cfg.nx = nx;
  cfg.ny = ny;
  cfg.cfl = (double)cfl;
  cfg.alpha = (double)alpha;
  cfg.muscl = use_muscl;
  cfg.hlld = use_hlld;
  cfg.precision = precision;
  cfg.hlld_tol = tol;
  cfg.tile = -1;
  cfg.batch_steps = 0;
  cfg.lowbeta_ratio = 0.0;
  if (const char* s = std::getenv("MHD_LOWBETA_RATIO")) cfg.lowbeta_ratio = std::atof(s);
  if (const char* s = std::getenv("MHD_TILE")) cfg.tile = std::atoi(s);
  if (const char* s = std::getenv("MHD_FALLBACK")) cfg.fallback = std::atoi(s);
  if (const char* s = std::getenv("MHD_FUSE")) cfg.fuse = std::atoi(s);
  if (const char* s = std::getenv("MHD_ROLLBACK")) cfg.rollback_steps = std::atoi(s);
  // The learned tile selector (mixed form only): $MHD_SELECTOR names a model
  // file as mhdglm.selector writes it (uncompressed .npz of float32
  // arrays: params, mean, std, threshold, audit_budget or budget).
  std::map<std::string, mhd::NpzArray> model;
  const char* sel_path = std::getenv("MHD_SELECTOR");
  const bool selector = sel_path && *sel_path;
  float sel_threshold = 0.f, sel_budget = 0.f;
  if (selector) {
    if (precision != MHD_PREC_MIXED) {
      std::cerr << "Error: MHD_SELECTOR needs MHD_PRECISION=mixed." << std::endl;
      return 1;
    }
    std::string err;
    if (!mhd::read_npz_f32(sel_path, model, err)) {
      std::cerr << "Error: selector model: " << err << std::endl;
      return 1;
    }
    const auto has = [&](const char* k, size_t n) {
      const auto it = model.find(k);
      return it != model.end() && it->second.data.size() == n;
    };
    if (!has("params", (size_t)mhd_mlp_param_count()) || !has("mean", 12) || !has("std", 12) ||
        !has("threshold", 1) || !(has("audit_budget", 1) || has("budget", 1))) {
      std::cerr << "Error: selector model " << sel_path
                << ": needs params, mean, std, threshold and audit_budget or budget" << std::endl;
      return 1;
    }
    sel_threshold = model["threshold"].data[0];
    sel_budget = has("audit_budget", 1) ? model["audit_budget"].data[0]
                                        : 10.0f * model["budget"].data[0];
    std::cout << "Learned tile selector: " << sel_path << " (threshold " << sel_threshold
              << ", 1% audits)" << std::endl;
  }

  mhd_ctx* ctx = nullptr;
  int rc = mhd_create(&cfg, &ctx);
  if (rc != 0) {
    std::cerr << "Error: solver setup failed (code " << rc << ")" << std::endl;
    return rc;
  }
  if ((rc = mhd_init_problem(ctx, a.problem.c_str())) != 0) {
    std::cerr << "Error: " << mhd_last_error(ctx) << std::endl;
    mhd_destroy(ctx);
    return rc;
  }
  if (selector &&
      (rc = mhd_selector_configure(ctx, model["params"].data.data(), model["mean"].data.data(),
                                   model["std"].data.data(), sel_threshold, 0.01f,
                                   sel_budget)) != 0) {
    std::cerr << "Error: " << mhd_last_error(ctx) << std::endl;
    mhd_destroy(ctx);
    return rc;
  }

  mhd_status st;
  auto w0 = std::chrono::steady_clock::now();
  rc = mhd_advance(ctx, (double)t_end, -1, &st);
  auto w1 = std::chrono::steady_clock::now();
  const double wall = std::chrono::duration<double>(w1 - w0).count();
  unsigned long long sel_stats[4] = {0, 0, 0, 0};
  const int sel_tiles = selector ? mhd_selector_tiles(ctx) : 0;
  if (selector && rc != 9) (void)mhd_selector_stats(ctx, sel_stats);
  if (rc == 9) {
    std::cerr << "Error: " << mhd_last_error(ctx) << std::endl;
    mhd_destroy(ctx);
    return 9;
  }

  Field<R> f;
  f.nx = nx;
  f.ny = ny;
  f.u.resize((size_t)9 * nx * ny);
  if (rc == 2) {
    std::cerr << "Error: non-positive maximum signal rate." << std::endl;
    mhd_destroy(ctx);
    return 2;
  }
  if (rc == 3 || rc == 4) {
    mhd_get_state(ctx, rc == 3 ? 1 : 0, f.u.data(), (size_t)nx * ny);
    std::cerr << (rc == 3 ? "Error: non-physical RK stage at step "
                          : "Error: non-physical state at step ")
              << st.fail_step << ", cell (" << st.bad_i << ", " << st.bad_j << ")"
              << std::endl;
    print_failure<R>(rc == 3 ? "RK stage" : "full RK state", f, st.bad_i, st.bad_j, dx, dy,
                     (R)st.t, (R)st.dt, st.fail_step, (R)st.ch_last);
    mhd_destroy(ctx);
    return rc;
  }
  if ((rc = mhd_get_state(ctx, 0, f.u.data(), (size_t)nx * ny)) != 0) {
    std::cerr << "Error: " << mhd_last_error(ctx) << std::endl;
    mhd_destroy(ctx);
    return 9;
  }
  mhd_destroy(ctx);

  const R t = (R)st.t, last_ch = (R)st.ch_last;
  const int step = st.steps;
  std::cout << "Done. Steps: " << step << ", final t = " << t << std::endl;
  std::cout << "c_h(last) = " << last_ch << ", alpha = " << alpha << std::endl;
  std::cout << "Time loop wall time = " << std::fixed << std::setprecision(3) << wall
            << " s (" << std::setprecision(2) << ((double)nx * ny * step / wall / 1.0e6)
            << " Mcell-steps/s)" << std::defaultfloat << std::endl;

  R div_l1 = (R)0.0, div_linf = (R)0.0, psi_linf = (R)0.0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const R d = div_b(f, i, j, dx, dy);
      div_l1 += std::fabs(d);
      div_linf = (div_linf < std::fabs(d)) ? std::fabs(d) : div_linf;
      const R ps = std::fabs(f.at(i, j, 8));
      psi_linf = (psi_linf < ps) ? ps : psi_linf;
    }
  div_l1 /= (R)(nx * ny);
  std::cout << std::setprecision(6);
  std::cout << "divB L1 = " << div_l1 << ", divB Linf = " << div_linf
            << ", psi Linf = " << psi_linf << std::endl;
  const unsigned long long evals = st.x_evaluations + st.y_evaluations;
  const unsigned long long fb = st.degenerate + st.nonphysical;
  const double frac = evals > 0 ? (double)fb / (double)evals : 0.0;
  const R tol_r = (R)tol;
  if (use_hlld) {
    std::cout << "HLLD flux evaluations = " << evals << " (x=" << st.x_evaluations
              << ", y=" << st.y_evaluations << ")" << std::endl;
    std::cout << "HLLD relative fallback tolerance = " << std::scientific << (double)tol_r
              << std::defaultfloat << std::endl;
    std::cout << "HLLD->HLL fallbacks = " << fb << " (degenerate=" << st.degenerate
              << ", nonphysical=" << st.nonphysical << "), fraction = " << std::scientific
              << frac << std::defaultfloat << std::endl;
  }
  if (st.fallback_tiles > 0)
    std::cout << "FP64 fallback stage tiles = " << st.fallback_tiles << std::endl;
  if (st.rollbacks > 0 || st.restarts > 0)
    std::cout << "FP64 rollbacks = " << st.rollbacks << ", restarts = " << st.restarts
              << ", FP64 steps = " << st.exact_steps << std::endl;
  if (selector) {
    const double tile_steps = (double)sel_stats[3] * (double)sel_tiles;
    std::cout << "Selector: FP64 tile-steps = " << sel_stats[2] << " (fraction "
              << std::scientific << (tile_steps > 0 ? (double)sel_stats[2] / tile_steps : 0.0)
              << std::defaultfloat << "), audits = " << sel_stats[0]
              << ", audit misses = " << sel_stats[1] << std::endl;
  }

  const std::string base = a.problem == "ot" ? "orszag_tang" : a.problem;
  const char* out_env = std::getenv("OUTDIR");
  const std::string out_dir = out_env ? std::string(out_env) : "results/mhd/gpu/" + base;
  if (std::system(("mkdir -p " + out_dir).c_str()) != 0) { /* ignore */ }
  std::ostringstream fname;
  fname << out_dir << "/" << base << "_glm_" << a.order << "_" << a.solver << "_"
        << prec_name << "_N" << nx << "x" << ny << ".dat";
  std::ofstream out(fname.str());
  if (!out) {
    std::cerr << "Error: failed to open output file: " << fname.str() << std::endl;
    return 5;
  }
  out << (a.problem == "ot" ? "# Orszag-Tang ideal MHD with mixed GLM (GPU)\n"
                            : "# Kelvin-Helmholtz ideal MHD with mixed GLM (GPU)\n");
  out << "# nx " << nx << " ny " << ny << " t " << t << " gamma " << gamma << " cfl " << cfl
      << " order " << a.order << " solver " << a.solver << " alpha " << alpha
      << " ch_last " << last_ch << "\n";
  if (use_hlld) {
    out << "# hlld_evaluations " << evals << " hlld_tolerance " << tol_r
        << " hlld_x_evaluations " << st.x_evaluations << " hlld_y_evaluations "
        << st.y_evaluations << " hlld_fallbacks " << fb << " hlld_degenerate "
        << st.degenerate << " hlld_nonphysical " << st.nonphysical
        << " hlld_fallback_fraction " << frac << "\n";
  }
  out << "# columns: x y rho vx vy vz p Bx By Bz E divB psi\n";
  out << std::scientific << std::setprecision(std::numeric_limits<R>::max_digits10);
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      R U[9], W[9];
      for (int k = 0; k < 9; k++) U[k] = f.at(i, j, k);
      mhd::host_cons_to_prim<R>(U, W);
      const R x = (R)0.0 + ((R)i + (R)0.5) * dx;
      const R y = (R)0.0 + ((R)j + (R)0.5) * dy;
      out << x << " " << y << " " << W[0] << " " << W[1] << " " << W[2] << " " << W[3]
          << " " << W[4] << " " << W[5] << " " << W[6] << " " << W[7] << " " << U[4] << " "
          << div_b(f, i, j, dx, dy) << " " << U[8] << "\n";
    }
  }
  std::cout << "Saved: " << fname.str() << std::endl;
  return 0;
}

bool ends_with(const std::string& s, const std::string& suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  Args a;
  if (argc > 1) a.nx = std::atoi(argv[1]);
  if (argc > 2) a.ny = std::atoi(argv[2]);
  if (argc > 3) a.t_end_raw = std::atof(argv[3]);
  if (argc > 4) a.cfl_raw = std::atof(argv[4]);
  if (argc > 5) a.order = argv[5];
  if (argc > 6) a.solver = argv[6];
  if (argc > 7) a.alpha_raw = std::atof(argv[7]);
  if (argc > 8) a.problem = argv[8];
  static const struct { const char* name; int prec; } kForms[] = {
      {"fp64", MHD_PREC_FP64}, {"fp32", MHD_PREC_FP32}, {"mixed", MHD_PREC_MIXED},
      {"mixed_kahan", MHD_PREC_MIXED_KAHAN}, {"fp64_fma", MHD_PREC_FP64_FMA},
      {"ref_mixed", MHD_PREC_REF_MIXED}};
  constexpr int kNumForms = 6;
  int form = 0;
  const std::string prog = argv[0];
  for (int q = 0; q < kNumForms; q++)
    if (ends_with(prog, std::string("_") + kForms[q].name)) form = q;
  // A program named as the reference's mixed build (mhd2d_glm_gpu_mixed)
  // runs that build's form and writes its file tag, "mixed".
  const char* tag = nullptr;
  if (ends_with(prog, "_gpu_mixed")) {
    form = 5;
    tag = "mixed";
  }
  if (const char* s = std::getenv("MHD_PRECISION")) {
    form = -1;
    tag = nullptr;
    for (int q = 0; q < kNumForms; q++)
      if (!std::strcmp(s, kForms[q].name)) form = q;
    if (form < 0) {
      std::cerr << "Error: MHD_PRECISION must be fp64, fp32, mixed, mixed_kahan, fp64_fma or "
                   "ref_mixed."
                << std::endl;
      return 1;
    }
  }
  double tol = 1.0e-4;
  if (const char* s = std::getenv("MHD_HLLD_TOLERANCE")) tol = std::atof(s);
  if (kForms[form].prec == MHD_PREC_FP32) return run<float>(a, kForms[form].prec, "fp32", tol);
  return run<double>(a, kForms[form].prec, tag ? tag : kForms[form].name, tol);
}

// mhd_problems.hpp -- initial conditions, evaluated on the host.
//
// Compiled by the host C++ compiler with the reference's flags (-O2, no FMA
// contraction) and glibc's libm, so the initial state is bitwise the
// reference's:
//   ot  : mhd2d_glm_cpu.cpp:486-505 / mhd2d_glm_gpu.cu:443-451
//   kh  : mhd2d_glm_gpu.cu:421-441 (the reference CUDA build's KH)
//   khb : Kelvin-Helmholtz of the 1024^2 benchmark workload (rho=1, p=2.5,
//         shear vx=0.5(tanh((y-.25)/.05)-tanh((y-.75)/.05)-1), Bx=0.1,
//         vy=0.01 sin(4 pi x)(exp(-(y-.25)^2/.005)+exp(-(y-.75)^2/.005)));
//         no reference implementation exists for it.
//   otv : Orszag-Tang with the ensemble workload's variations: params =
//         {B amplitude, phase of vx, vy, Bx, By} added to the sine arguments;
//         with {1, 0, 0, 0, 0} it is bitwise `ot`.
#ifndef MHD_PROBLEMS_HPP
#define MHD_PROBLEMS_HPP

#include <cmath>
#include <cstring>

namespace mhd {

inline bool problem_known(const char* p) {
  return p && (!std::strcmp(p, "ot") || !std::strcmp(p, "kh") || !std::strcmp(p, "khb") ||
               !std::strcmp(p, "otv"));
}

// Number of parameters a problem takes.
inline int problem_nparams(const char* p) { return (p && !std::strcmp(p, "otv")) ? 5 : 0; }

// Primitive state (rho, vx, vy, vz, p, Bx, By, Bz, psi) of cell (i,j).
template <class R>
inline void problem_primitive(const char* problem, int i, int j, R dx, R dy,
                              R P[9], const double* params = nullptr) {
  const R x = (R)0.0 + ((R)i + (R)0.5) * dx;
  const R y = (R)0.0 + ((R)j + (R)0.5) * dy;
  const R pi = std::acos((R)-1.0);
  const R two_pi = (R)2.0 * pi;
  if (!std::strcmp(problem, "kh")) {
    const R width = (R)0.025, sigma = (R)0.05, amp = (R)0.01;
    const R band = (R)0.5 * (std::tanh((y - (R)0.25) / width) -
                             std::tanh((y - (R)0.75) / width));
    const R z1 = (y - (R)0.25) / sigma, z2 = (y - (R)0.75) / sigma;
    P[0] = (R)1.0 + band;
    P[1] = -(R)0.5 + band;
    P[2] = amp * std::sin((R)2.0 * two_pi * x) *
           (std::exp(-(R)0.5 * z1 * z1) + std::exp(-(R)0.5 * z2 * z2));
    P[3] = (R)0.0;
    P[4] = (R)2.5;
    P[5] = (R)0.2;
    P[6] = (R)0.0;
  } else if (!std::strcmp(problem, "khb")) {
    const R w = (R)0.05;
    const R d1 = y - (R)0.25, d2 = y - (R)0.75;
    P[0] = (R)1.0;
    P[1] = (R)0.5 * (std::tanh(d1 / w) - std::tanh(d2 / w) - (R)1.0);
    P[2] = (R)0.01 * std::sin((R)2.0 * two_pi * x) *
           (std::exp(-(d1 * d1) / (R)0.005) + std::exp(-(d2 * d2) / (R)0.005));
    P[3] = (R)0.0;
    P[4] = (R)2.5;
    P[5] = (R)0.1;
    P[6] = (R)0.0;
  } else if (!std::strcmp(problem, "otv")) {
    const R amp = (R)params[0];
    const R inv_sqrt_4pi = (R)1.0 / std::sqrt((R)4.0 * pi);
    P[0] = (R)25.0 / ((R)36.0 * pi);
    P[1] = -std::sin(two_pi * y + (R)params[1]);
    P[2] = std::sin(two_pi * x + (R)params[2]);
    P[3] = (R)0.0;
    P[4] = (R)5.0 / ((R)12.0 * pi);
    P[5] = -amp * std::sin(two_pi * y + (R)params[3]) * inv_sqrt_4pi;
    P[6] = amp * std::sin((R)2.0 * two_pi * x + (R)params[4]) * inv_sqrt_4pi;
  } else {
    const R inv_sqrt_4pi = (R)1.0 / std::sqrt((R)4.0 * pi);
    P[0] = (R)25.0 / ((R)36.0 * pi);
    P[1] = -std::sin(two_pi * y);
    P[2] = std::sin(two_pi * x);
    P[3] = (R)0.0;
    P[4] = (R)5.0 / ((R)12.0 * pi);
    P[5] = -std::sin(two_pi * y) * inv_sqrt_4pi;
    P[6] = std::sin((R)2.0 * two_pi * x) * inv_sqrt_4pi;
  }
  P[7] = (R)0.0;
  P[8] = (R)0.0;
}

// Primitive -> conserved on the host (mhd_common.hpp:149-173), same
// association order as the device version.
template <class R>
inline void host_prim_to_cons(const R W[9], R U[9]) {
  const R gm1 = (R)(5.0 / 3.0) - (R)1.0;
  const R rho = W[0], vx = W[1], vy = W[2], vz = W[3], p = W[4];
  const R bx = W[5], by = W[6], bz = W[7];
  const R v2 = vx * vx + vy * vy + vz * vz;
  const R b2 = bx * bx + by * by + bz * bz;
  U[0] = rho;
  U[1] = rho * vx;
  U[2] = rho * vy;
  U[3] = rho * vz;
  U[4] = p / gm1 + (R)0.5 * rho * v2 + (R)0.5 * b2;
  U[5] = bx;
  U[6] = by;
  U[7] = bz;
  U[8] = W[8];
}

// Host cons -> prim (mhd_common.hpp:178-201) for output and diagnostics.
template <class R>
inline void host_cons_to_prim(const R U[9], R W[9]) {
  const R gm1 = (R)(5.0 / 3.0) - (R)1.0;
  const R rho = (U[0] < (R)1.0e-12) ? (R)1.0e-12 : U[0];
  const R vx = U[1] / rho, vy = U[2] / rho, vz = U[3] / rho;
  const R bx = U[5], by = U[6], bz = U[7];
  const R v2 = vx * vx + vy * vy + vz * vz;
  const R b2 = bx * bx + by * by + bz * bz;
  W[0] = rho;
  W[1] = vx;
  W[2] = vy;
  W[3] = vz;
  W[4] = gm1 * (U[4] - (R)0.5 * rho * v2 - (R)0.5 * b2);
  W[5] = bx;
  W[6] = by;
  W[7] = bz;
  W[8] = U[8];
}

}  // namespace mhd

#endif  // MHD_PROBLEMS_HPP

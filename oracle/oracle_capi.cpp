// oracle_capi.cpp -- C entry points of the CPU oracle, for tests only.
//
// Loaded by tests/ through ctypes to compare the HIP product against the
// oracle on the same inputs.  Arrays are AoS: state n, variable k at n*9+k.
// Each function exists for double (_f64) and float (_f32).
#include "mhd_oracle.hpp"
#include "mhd_refmixed.hpp"
#include "problems.hpp"

#include <cstdint>
#include <string>

using namespace oracle;

#define ORACLE_CAPI(R, SUF)                                                    \
  extern "C" int orc_glm_flux_batch_##SUF(int n, int dir, const R* UL,         \
                                          const R* UR, R ch, int hlld,         \
                                          double tol, R* F, int* status) {     \
    const Consts<R> c(tol);                                                    \
    for (int s = 0; s < n; s++) {                                              \
      const int st = glm_flux(c, UL + 9 * (size_t)s, UR + 9 * (size_t)s, ch,   \
                              hlld != 0, dir, F + 9 * (size_t)s);              \
      if (status) status[s] = st;                                              \
    }                                                                          \
    return 0;                                                                  \
  }                                                                            \
  extern "C" void orc_cons_to_prim_batch_##SUF(int n, const R* U, R* W) {      \
    const Consts<R> c(1.0e-4);                                                 \
    for (int s = 0; s < n; s++) state_to_prim(c, U + 9 * (size_t)s, W + 9 * (size_t)s); \
  }                                                                            \
  extern "C" void orc_prim_to_cons_batch_##SUF(int n, const R* W, R* U) {      \
    const Consts<R> c(1.0e-4);                                                 \
    for (int s = 0; s < n; s++) prim_to_state(c, W + 9 * (size_t)s, U + 9 * (size_t)s); \
  }                                                                            \
  extern "C" void orc_fast_speed_batch_##SUF(int n, const R* W, int dir,       \
                                             R* out) {                         \
    const Consts<R> c(1.0e-4);                                                 \
    for (int s = 0; s < n; s++) out[s] = fast_speed(c, W + 9 * (size_t)s, dir); \
  }                                                                            \
  extern "C" void orc_reconstruct_batch_##SUF(int n, const R* Pm,              \
                                              const R* P0, const R* Pp,        \
                                              R* Pl, R* Pr) {                  \
    const Consts<R> c(1.0e-4);                                                 \
    for (int s = 0; s < n; s++)                                                \
      reconstruct(c, Pm + 9 * (size_t)s, P0 + 9 * (size_t)s,                   \
                  Pp + 9 * (size_t)s, Pl + 9 * (size_t)s, Pr + 9 * (size_t)s); \
  }                                                                            \
  extern "C" void orc_init_params_##SUF(const char* problem,                   \
                                         const double* params, int nx, int ny, \
                                         R* u) {                               \
    const Consts<R> c(1.0e-4);                                                 \
    const R dx = ((R)1.0 - (R)0.0) / (R)nx, dy = ((R)1.0 - (R)0.0) / (R)ny;    \
    const std::string prob(problem);                                           \
    for (int j = 0; j < ny; j++)                                               \
      for (int i = 0; i < nx; i++) {                                           \
        R P[NV];                                                               \
        initial_primitive<R>(prob, i, j, dx, dy, P, params);                   \
        prim_to_state(c, P, u + ((size_t)j * nx + i) * 9);                     \
      }                                                                        \
  }                                                                            \
  extern "C" void orc_init_##SUF(const char* problem, int nx, int ny, R* u) {  \
    orc_init_params_##SUF(problem, nullptr, nx, ny, u);                        \
  }                                                                            \
  extern "C" double orc_max_signal_speed_##SUF(int nx, int ny, const R* u) {   \
    const Consts<R> c(1.0e-4);                                                 \
    Grid<R> U(nx, ny);                                                         \
    for (size_t q = 0; q < U.u.size(); q++) U.u[q] = u[q];                     \
    return (double)max_signal_speed(c, U);                                     \
  }                                                                            \
  extern "C" void orc_compute_rhs_##SUF(int nx, int ny, const R* u, R ch,      \
                                        R alpha, int muscl, int hlld,          \
                                        double tol, R* rhs,                    \
                                        uint64_t* counters4) {                 \
    const Consts<R> c(tol);                                                    \
    Grid<R> U(nx, ny);                                                         \
    for (size_t q = 0; q < U.u.size(); q++) U.u[q] = u[q];                     \
    const R dx = ((R)1.0 - (R)0.0) / (R)nx, dy = ((R)1.0 - (R)0.0) / (R)ny;    \
    FallbackStats st;                                                          \
    std::vector<R> r;                                                          \
    compute_rhs(c, U, dx, dy, muscl != 0, hlld != 0, ch, alpha, &st, r);       \
    for (size_t q = 0; q < r.size(); q++) rhs[q] = r[q];                       \
    if (counters4) {                                                           \
      counters4[0] = st.x_evals; counters4[1] = st.y_evals;                    \
      counters4[2] = st.degenerate; counters4[3] = st.nonphysical;             \
    }                                                                          \
  }                                                                            \
  /* Advances up to max_steps SSPRK2 steps (or to t_end).  Returns the exit   \
     code of the reference time loop: 0, 2, 3 or 4. */                         \
  extern "C" int orc_advance_##SUF(int nx, int ny, R* u, int max_steps,        \
                                   R t_end, R cfl, R alpha, int muscl,         \
                                   int hlld, double tol, R* t_io,              \
                                   int* steps_out, uint64_t* counters4,        \
                                   int* bad_ij, R* last_ch) {                  \
    const Consts<R> c(tol);                                                    \
    Grid<R> U(nx, ny), U1(nx, ny);                                             \
    for (size_t q = 0; q < U.u.size(); q++) U.u[q] = u[q];                     \
    const R dx = ((R)1.0 - (R)0.0) / (R)nx, dy = ((R)1.0 - (R)0.0) / (R)ny;    \
    FallbackStats st;                                                          \
    R t = *t_io;                                                               \
    int steps = 0, code = 0;                                                   \
    while (t < t_end && steps < max_steps) {                                   \
      StepResult r = ssprk2_step(c, U, U1, dx, dy, cfl, alpha, muscl != 0,     \
                                 hlld != 0, t, t_end, &st, last_ch);           \
      if (r.code != 0) {                                                       \
        code = r.code;                                                         \
        if (bad_ij) { bad_ij[0] = r.bad_i; bad_ij[1] = r.bad_j; }              \
        break;                                                                 \
      }                                                                        \
      steps++;                                                                 \
    }                                                                          \
    for (size_t q = 0; q < U.u.size(); q++) u[q] = U.u[q];                     \
    *t_io = t;                                                                 \
    if (steps_out) *steps_out = steps;                                         \
    if (counters4) {                                                           \
      counters4[0] = st.x_evals; counters4[1] = st.y_evals;                    \
      counters4[2] = st.degenerate; counters4[3] = st.nonphysical;             \
    }                                                                          \
    return code;                                                               \
  }

ORACLE_CAPI(double, f64)
ORACLE_CAPI(float, f32)

// FP32 stage functions of the reference-mixed restatement, batched (AoS 9).
extern "C" void orc_refmixed_cons_to_prim_batch(int n, const float* U, float* W) {
  for (int s = 0; s < n; s++) refmixed::cons_to_prim(U + (size_t)s * NV, W + (size_t)s * NV);
}
extern "C" void orc_refmixed_prim_to_cons_batch(int n, const float* W, float* U) {
  for (int s = 0; s < n; s++) refmixed::prim_to_cons(W + (size_t)s * NV, U + (size_t)s * NV);
}
extern "C" void orc_refmixed_fast_speed_batch(int n, const float* W, int dir, float* out) {
  for (int s = 0; s < n; s++) {
    const float* w = W + (size_t)s * NV;
    out[s] = refmixed::fast_speed(w, dir ? w[BY] : w[BX]);
  }
}
extern "C" void orc_refmixed_recon_batch(int n, const float* Pm, const float* P0,
                                         const float* Pp, float* Pl, float* Pr) {
  for (int s = 0; s < n; s++) {
    const size_t o = (size_t)s * NV;
    refmixed::recon(Pm + o, P0 + o, Pp + o, Pl + o, Pr + o);
  }
}

// The reference's stage-level mixed build (mhd_refmixed.hpp), MUSCL, with
// the float regions of `mask` (refmixed::Region; kAllRegions: the mixed
// build, 0: the FP64 step): the same arguments as orc_advance_f64.
extern "C" int orc_refmixed_advance_regions(int nx, int ny, double* u, int max_steps,
                                            double t_end, double cfl, double alpha, int hlld,
                                            double tol, double* t_io, int* steps_out,
                                            uint64_t* counters4, int* bad_ij, double* last_ch,
                                            int mask) {
  const Consts<double> c(tol);
  Grid<double> U(nx, ny), U1(nx, ny);
  for (size_t q = 0; q < U.u.size(); q++) U.u[q] = u[q];
  const double dx = (1.0 - 0.0) / (double)nx, dy = (1.0 - 0.0) / (double)ny;
  FallbackStats st;
  double t = *t_io;
  int steps = 0, code = 0;
  while (t < t_end && steps < max_steps) {
    StepResult r = refmixed::ssprk2_step(c, U, U1, dx, dy, cfl, alpha, hlld != 0, t, t_end, &st,
                                         last_ch, mask);
    if (r.code != 0) {
      code = r.code;
      if (bad_ij) { bad_ij[0] = r.bad_i; bad_ij[1] = r.bad_j; }
      break;
    }
    steps++;
  }
  for (size_t q = 0; q < U.u.size(); q++) u[q] = U.u[q];
  *t_io = t;
  if (steps_out) *steps_out = steps;
  if (counters4) {
    counters4[0] = st.x_evals; counters4[1] = st.y_evals;
    counters4[2] = st.degenerate; counters4[3] = st.nonphysical;
  }
  return code;
}

extern "C" int orc_refmixed_advance(int nx, int ny, double* u, int max_steps, double t_end,
                                    double cfl, double alpha, int hlld, double tol,
                                    double* t_io, int* steps_out, uint64_t* counters4,
                                    int* bad_ij, double* last_ch) {
  return orc_refmixed_advance_regions(nx, ny, u, max_steps, t_end, cfl, alpha, hlld, tol, t_io,
                                      steps_out, counters4, bad_ij, last_ch,
                                      refmixed::kAllRegions);
}

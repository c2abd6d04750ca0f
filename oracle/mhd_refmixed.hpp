// mhd_refmixed.hpp -- TEST INFRASTRUCTURE: a serial CPU restatement of the
// reference's stage-level mixed-precision build, src/mhd/mhd2d_glm_gpu_mixed.cu
// (used only by tests/ to check the device REF_MIXED form).
//
// Allocation (mhd2d_glm_gpu_mixed.cu:3-8): state, storage, the Riemann
// fluxes (GLM-HLLD/HLL, mhd_glm_common.hpp) and the SSPRK2 stage updates in
// double; recovery, the CFL reductions, face reconstruction, the rhs flux
// differences and the psi damping in float, each stage's inputs cast to
// float and its results cast back.  The time accumulator stays double.
// Reductions and the periodic grid follow mhd_oracle.hpp.
//
// Regions.  The five float parts are the regions of the reference's region
// build (mhd2d_glm_raptor.cpp:86-230: recovery, cfl, recon, rhs, glm; its
// flux and update regions are the double parts here).  A region mask puts
// any subset of them in float, the others in double exactly as the FP64
// oracle computes them: mask 0 is the FP64 oracle's step, bit for bit, and
// kAllRegions the mixed build.  The region build truncates each operation
// of a region to float instead of casting the region's inputs, so a
// one-region run is compared with the reference's one-region scan only in
// size (its whole-run distance from FP64, per field).
#ifndef MHD_REFMIXED_HPP
#define MHD_REFMIXED_HPP

#include <cmath>
#include <vector>

#include "mhd_oracle.hpp"

namespace oracle {
namespace refmixed {

enum Region { REG_RECOVERY = 1, REG_CFL = 2, REG_RECON = 4, REG_RHS = 8, REG_GLM = 16 };
constexpr int kAllRegions = REG_RECOVERY | REG_CFL | REG_RECON | REG_RHS | REG_GLM;

constexpr float kRhoFloor = 1.0e-12f;  // F_RHO_FLOOR, mhd2d_glm_gpu_mixed.cu:55
constexpr float kPFloor = 1.0e-12f;    // F_P_FLOOR, :56
const float kGamma = (float)(5.0 / 3.0);  // (float)MHD_GAMMA

// f32Minmod (:57-60)
inline float minmod(float a, float b) {
  if (a * b <= 0.0f) return 0.0f;
  return (std::fabs(a) < std::fabs(b)) ? a : b;
}

// f32ConsToPrim9 (:61-74)
inline void cons_to_prim(const float U[NV], float P[NV]) {
  const float rho = std::fmax(U[RHO], kRhoFloor);
  const float vx = U[MX] / rho, vy = U[MY] / rho, vz = U[MZ] / rho;
  const float bx = U[BX], by = U[BY], bz = U[BZ];
  const float v2 = vx * vx + vy * vy + vz * vz;
  const float b2 = bx * bx + by * by + bz * bz;
  const float p = (kGamma - 1.0f) * (U[EN] - 0.5f * rho * v2 - 0.5f * b2);
  P[0] = rho; P[1] = vx; P[2] = vy; P[3] = vz; P[4] = p;
  P[5] = bx; P[6] = by; P[7] = bz; P[8] = U[PSI];
}

// f32PrimToCons9 (:75-87)
inline void prim_to_cons(const float P[NV], float U[NV]) {
  const float rho = P[0], vx = P[1], vy = P[2], vz = P[3];
  const float bx = P[5], by = P[6], bz = P[7];
  U[RHO] = rho;
  U[MX] = rho * vx; U[MY] = rho * vy; U[MZ] = rho * vz;
  U[EN] = P[4] / (kGamma - 1.0f) + 0.5f * rho * (vx * vx + vy * vy + vz * vz) +
          0.5f * (bx * bx + by * by + bz * bz);
  U[BX] = bx; U[BY] = by; U[BZ] = bz;
  U[PSI] = P[8];
}

// f32FastSpeed (:88-99)
inline float fast_speed(const float P[NV], float bn) {
  const float rho = std::fmax(P[0], kRhoFloor);
  const float p = std::fmax(P[4], kPFloor);
  const float bx = P[5], by = P[6], bz = P[7];
  const float a2 = kGamma * p / rho;
  const float b2 = (bx * bx + by * by + bz * bz) / rho;
  const float can2 = bn * bn / rho;
  const float s = a2 + b2;
  float disc = s * s - 4.0f * a2 * can2;
  disc = std::fmax(disc, 0.0f);
  return std::sqrt(0.5f * (s + std::sqrt(disc)));
}

// f32PrimUsable (:100-106)
inline bool prim_usable(const float P[NV]) {
  if (P[0] <= kRhoFloor || P[4] <= kPFloor) return false;
  for (int k = 0; k < NV; k++)
    if (!std::isfinite((double)P[k])) return false;
  return true;
}

// f32ReconPrim9 (:107-120)
inline void recon(const float Pm[NV], const float P0[NV], const float Pp[NV], float Pl[NV],
                  float Pr[NV]) {
  for (int k = 0; k < NV; k++) {
    const float slope = minmod(P0[k] - Pm[k], Pp[k] - P0[k]);
    Pl[k] = P0[k] - 0.5f * slope;
    Pr[k] = P0[k] + 0.5f * slope;
  }
  if (!prim_usable(Pl) || !prim_usable(Pr))
    for (int k = 0; k < NV; k++) { Pl[k] = P0[k]; Pr[k] = P0[k]; }
}

// f32LoadPrim (:121-125)
inline void load_prim(const double* U9, float P[NV]) {
  float Uf[NV];
  for (int k = 0; k < NV; k++) Uf[k] = (float)U9[k];
  cons_to_prim(Uf, P);
}

// Primitives of a cell: the recovery region.  In float (f32LoadPrim) the
// values are float ones, widened to double.
inline void cell_prim(const Consts<double>& c, const double* U9, int mask, double P[NV]) {
  if (mask & REG_RECOVERY) {
    float Pf[NV];
    load_prim(U9, Pf);
    for (int k = 0; k < NV; k++) P[k] = (double)Pf[k];
  } else {
    state_to_prim(c, U9, P);
  }
}

// Signal speeds of a cell in x and y (the cfl region), from its primitives.
inline void cell_speeds(const Consts<double>& c, const double P[NV], int mask, float& fx,
                        float& fy, double& sx, double& sy) {
  if (mask & REG_CFL) {
    float Pf[NV];
    for (int k = 0; k < NV; k++) Pf[k] = (float)P[k];
    fx = std::fabs(Pf[1]) + fast_speed(Pf, Pf[5]);
    fy = std::fabs(Pf[2]) + fast_speed(Pf, Pf[6]);
  } else {
    sx = std::fabs(P[1]) + oracle::fast_speed(c, P, 0);
    sy = std::fabs(P[2]) + oracle::fast_speed(c, P, 1);
  }
}

// xFaceStatesDev / yFaceStatesDev, MUSCL branch (:178-224): recovery, then
// the recon region (MUSCL and face prim->cons).
inline void face_states(const Consts<double>& c, const Grid<double>& U, int i, int j, int dir,
                        int mask, double UL[NV], double UR[NV]) {
  const int di = (dir == 0) ? 1 : 0, dj = (dir == 0) ? 0 : 1;
  double Pm[NV], P0[NV], Pp[NV], Pq[NV];
  cell_prim(c, U.at(i - di, j - dj), mask, Pm);
  cell_prim(c, U.at(i, j), mask, P0);
  cell_prim(c, U.at(i + di, j + dj), mask, Pp);
  cell_prim(c, U.at(i + 2 * di, j + 2 * dj), mask, Pq);
  if (mask & REG_RECON) {
    float fm[NV], f0[NV], fp[NV], fq[NV];
    for (int k = 0; k < NV; k++) {
      fm[k] = (float)Pm[k]; f0[k] = (float)P0[k]; fp[k] = (float)Pp[k]; fq[k] = (float)Pq[k];
    }
    float Pl0[NV], Pr0[NV], Pl1[NV], Pr1[NV];
    recon(fm, f0, fp, Pl0, Pr0);
    recon(f0, fp, fq, Pl1, Pr1);
    float Ul[NV], Ur[NV];
    prim_to_cons(Pr0, Ul);
    prim_to_cons(Pl1, Ur);
    for (int k = 0; k < NV; k++) { UL[k] = (double)Ul[k]; UR[k] = (double)Ur[k]; }
  } else {
    double Pl0[NV], Pr0[NV], Pl1[NV], Pr1[NV];
    reconstruct(c, Pm, P0, Pp, Pl0, Pr0);
    reconstruct(c, P0, Pp, Pq, Pl1, Pr1);
    prim_to_state(c, Pr0, UL);
    prim_to_state(c, Pl1, UR);
  }
}

// speedKernel (:256-280) + reducePart (:419-423): max over cells of
// (double)fmaxf(sx, sy), then max with 1e-12 (in double without the cfl
// region: maxPhysicalSignalSpeed, mhd2d_glm_cpu.cpp:289-305).
// The maxima below are exact and never take a NaN (smax keeps its first
// argument then), so per-thread partial maxima give the serial loop's value.
inline double max_signal_speed(const Consts<double>& c, const Grid<double>& U, int mask) {
  double m = 0.0;
#pragma omp parallel
  {
    double mine = 0.0;
#pragma omp for schedule(static)
    for (int j = 0; j < U.ny; j++)
      for (int i = 0; i < U.nx; i++) {
        double P[NV], sx = 0, sy = 0;
        float fx = 0, fy = 0;
        cell_prim(c, U.at(i, j), mask, P);
        cell_speeds(c, P, mask, fx, fy, sx, sy);
        mine = (mask & REG_CFL) ? smax(mine, (double)std::fmax(fx, fy)) : smax(mine, smax(sx, sy));
      }
#pragma omp critical
    m = smax(m, mine);
  }
  return m;
}

// rateKernel (:283-308): the rate of every cell in float (in double without
// the cfl region: maxSignalRate, mhd2d_glm_cpu.cpp:307-326).
inline double max_signal_rate(const Consts<double>& c, const Grid<double>& U, double dx,
                              double dy, double ch, int mask) {
  double m = 0.0;
#pragma omp parallel
  {
    double mine = 0.0;
#pragma omp for schedule(static)
    for (int j = 0; j < U.ny; j++)
      for (int i = 0; i < U.nx; i++) {
        double P[NV], sx = 0, sy = 0;
        float fx = 0, fy = 0;
        cell_prim(c, U.at(i, j), mask, P);
        cell_speeds(c, P, mask, fx, fy, sx, sy);
        if (mask & REG_CFL) {
          fx = std::fmax(fx, (float)ch);
          fy = std::fmax(fy, (float)ch);
          mine = smax(mine, (double)(fx / (float)dx + fy / (float)dy));
        } else {
          mine = smax(mine, smax(sx, ch) / dx + smax(sy, ch) / dy);
        }
      }
#pragma omp critical
    m = smax(m, mine);
  }
  return m;
}

// computeRhsGpu (:426-438): fluxX/fluxY kernels (:311-353) in double on the
// face states, rhsKernel (:355-375): the flux differences (rhs region) and
// the psi damping (glm region).  The face fluxes are independent of each
// other, so they are evaluated in parallel where OpenMP is enabled (counters
// summed per thread); every value is the serial loop's.
inline void compute_rhs(const Consts<double>& c, const Grid<double>& U, double dx, double dy,
                        bool use_hlld, double ch, double alpha, FallbackStats* stats,
                        std::vector<double>& rhs, int mask) {
  const int nx = U.nx, ny = U.ny;
  const double damping = ch / smax(alpha, 1.0e-12);
  std::vector<double> Fx((size_t)(nx + 1) * ny * NV), Gy((size_t)nx * (ny + 1) * NV);
  FallbackStats total;
#pragma omp parallel
  {
    FallbackStats mine;
#pragma omp for schedule(static)
    for (int j = 0; j < ny; j++)
      for (int i = -1; i < nx; i++) {
        double UL[NV], UR[NV], F[NV];
        face_states(c, U, i, j, 0, mask, UL, UR);
        const int st = glm_flux(c, UL, UR, ch, use_hlld, 0, F);
        if (use_hlld) mine.record(0, st);
        double* dst = &Fx[((size_t)j * (nx + 1) + (i + 1)) * NV];
        for (int k = 0; k < NV; k++) dst[k] = F[k];
      }
#pragma omp for schedule(static)
    for (int j = -1; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double UL[NV], UR[NV], G[NV];
        face_states(c, U, i, j, 1, mask, UL, UR);
        const int st = glm_flux(c, UL, UR, ch, use_hlld, 1, G);
        if (use_hlld) mine.record(1, st);
        double* dst = &Gy[((size_t)(j + 1) * nx + i) * NV];
        for (int k = 0; k < NV; k++) dst[k] = G[k];
      }
#pragma omp critical
    {
      total.x_evals += mine.x_evals;
      total.y_evals += mine.y_evals;
      total.degenerate += mine.degenerate;
      total.nonphysical += mine.nonphysical;
    }
  }
  if (stats) {
    stats->x_evals += total.x_evals;
    stats->y_evals += total.y_evals;
    stats->degenerate += total.degenerate;
    stats->nonphysical += total.nonphysical;
  }
  const float fdx = (float)dx, fdy = (float)dy;
  rhs.assign((size_t)nx * ny * NV, 0.0);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const double* fr = &Fx[((size_t)j * (nx + 1) + (i + 1)) * NV];
      const double* fl = &Fx[((size_t)j * (nx + 1) + i) * NV];
      const double* gt = &Gy[((size_t)(j + 1) * nx + i) * NV];
      const double* gb = &Gy[((size_t)j * nx + i) * NV];
      double* r = &rhs[((size_t)j * nx + i) * NV];
      for (int k = 0; k < NV; k++) {
        if (mask & REG_RHS) {
          const float d =
              -((float)fr[k] - (float)fl[k]) / fdx - ((float)gt[k] - (float)gb[k]) / fdy;
          r[k] = (double)d;
        } else {
          r[k] = -(fr[k] - fl[k]) / dx - (gt[k] - gb[k]) / dy;
        }
      }
      if (mask & REG_GLM)
        r[PSI] = (double)((float)r[PSI] - (float)damping * (float)U.at(i, j)[PSI]);
      else
        r[PSI] -= damping * U.at(i, j)[PSI];
    }
}

// One step of the host loop (:542-596): speed, rate, dt, two stages with
// the double positivity check of stage1Kernel/stage2Kernel (:377-410).
inline StepResult ssprk2_step(const Consts<double>& c, Grid<double>& U, Grid<double>& U1,
                              double dx, double dy, double cfl, double alpha, bool use_hlld,
                              double& t, double t_end, FallbackStats* stats, double* last_ch,
                              int mask = kAllRegions) {
  StepResult res;
  const double ch = smax(max_signal_speed(c, U, mask), 1.0e-12);
  if (last_ch) *last_ch = ch;
  res.ch = ch;
  const double rate = max_signal_rate(c, U, dx, dy, ch, mask);
  if (rate <= 0.0) { res.code = 2; return res; }
  double dt = cfl / rate;
  if (t + dt > t_end) dt = t_end - t;
  res.dt = dt;
  std::vector<double> rhs;
  compute_rhs(c, U, dx, dy, use_hlld, ch, alpha, stats, rhs, mask);
  U1.nx = U.nx; U1.ny = U.ny; U1.u.resize(U.u.size());
  for (size_t n = 0; n < U.u.size(); n++) U1.u[n] = U.u[n] + dt * rhs[n];
  if (!all_usable(c, U1, res.bad_i, res.bad_j)) { res.code = 3; return res; }
  compute_rhs(c, U1, dx, dy, use_hlld, ch, alpha, stats, rhs, mask);
  for (size_t n = 0; n < U.u.size(); n++)
    U.u[n] = 0.5 * U.u[n] + 0.5 * (U1.u[n] + dt * rhs[n]);
  if (!all_usable(c, U, res.bad_i, res.bad_j)) { res.code = 4; return res; }
  t += dt;
  return res;
}

}  // namespace refmixed
}  // namespace oracle

#endif  // MHD_REFMIXED_HPP

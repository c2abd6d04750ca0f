// mhd_oracle.hpp -- CPU oracle for the 2D ideal-MHD + mixed-GLM step.
//
// TEST INFRASTRUCTURE ONLY.  This header restates, in plain serial C++, the
// arithmetic of the reference solver so that the HIP product path can be
// checked against it.  Only tests/, the smoke check and bench.py's CPU
// baseline leg may build, link or run anything under oracle/; the product
// path never calls it.
//
// Every function names the reference file:line it follows (paths relative to
// the reference repository root).  The association order of every expression
// is the reference's, because bitwise parity depends on it:
//   * a*b*c is (a*b)*c, a+b+c is (a+b)+c, exactly as C++ parses the source;
//   * max/min are std::max/std::min, i.e. (a<b)?b:a and (b<a)?b:a, which
//     differ from fmax/fmin on NaN and on signed zeros;
//   * divisions are true divisions (never a reciprocal multiply);
//   * build with g++ -O2 and no -march so that nothing is contracted to FMA.
//
// Pinning (how this restatement is shown to be the reference's algorithm):
//   * whole-run md5 of the .dat data rows against the reference CPU build
//     (OT 128^2 and 192^2, FP64 and FP32, eps=1e-4) and its own logs
//     (steps, fallback counts, divB, psi);
//   * per-function bitwise comparison with the reference headers compiled
//     from their own sources into oracle/_ref (random and special states);
//   * invariants: HLL/HLLD(U,U) = physical flux, y-flux = rotated x-flux,
//     dt = cfl/(c_h/dx + c_h/dy), prim->cons->prim round trip.
#ifndef MHD_ORACLE_HPP
#define MHD_ORACLE_HPP

#include <cmath>
#include <cstdint>
#include <vector>

namespace oracle {

// Variable slots.  Conserved: rho, mx, my, mz, E, Bx, By, Bz, psi.
// Primitive:  rho, vx, vy, vz, p, Bx, By, Bz, psi.
// (include/mhd/mhd_common.hpp:81-104, include/mhd/mhd_glm_common.hpp:26-27)
enum { RHO = 0, MX = 1, MY = 2, MZ = 3, EN = 4, BX = 5, BY = 6, BZ = 7, PSI = 8 };
enum { NV8 = 8, NV = 9 };

// HLLD status codes (include/mhd/mhd_common.hpp:438-440).
enum { HLLD_OK = 0, HLLD_DEGENERATE = 1, HLLD_NONPHYSICAL = 2 };

template <class R>
struct Consts {
  R gamma;      // (Real)(5.0/3.0): mhd2d_glm_cpu.cpp:25-27, mhd_common.hpp:55
  R rho_floor;  // (Real)1e-12: mhd_common.hpp:56
  R p_floor;    // (Real)1e-12: mhd_common.hpp:57
  R tol;        // (Real)MHD_HLLD_TOLERANCE_VALUE: mhd_common.hpp:77-79,411-413
  explicit Consts(double tol_value = 1.0e-4)
      : gamma((R)(5.0 / 3.0)), rho_floor((R)1.0e-12), p_floor((R)1.0e-12),
        tol((R)tol_value) {}
};

// std::max / std::min semantics (mhd_common.hpp:106-120, host branch).
template <class R> inline R smax(R a, R b) { return (a < b) ? b : a; }
template <class R> inline R smin(R a, R b) { return (b < a) ? b : a; }

// mhd_common.hpp:140-143
template <class R> inline R minmod(R a, R b) {
  if (a * b <= (R)0.0) return (R)0.0;
  return (std::fabs(a) < std::fabs(b)) ? a : b;
}

// mhd_common.hpp:149-173
template <class R>
inline void prim_to_cons8(const Consts<R>& c, const R W[NV8], R U[NV8]) {
  const R rho = W[0], vx = W[1], vy = W[2], vz = W[3], p = W[4];
  const R bx = W[5], by = W[6], bz = W[7];
  const R v2 = vx * vx + vy * vy + vz * vz;
  const R b2 = bx * bx + by * by + bz * bz;
  U[RHO] = rho;
  U[MX] = rho * vx;
  U[MY] = rho * vy;
  U[MZ] = rho * vz;
  U[EN] = p / (c.gamma - (R)1.0) + (R)0.5 * rho * v2 + (R)0.5 * b2;
  U[BX] = bx;
  U[BY] = by;
  U[BZ] = bz;
}

// mhd_common.hpp:178-201
template <class R>
inline void cons_to_prim8(const Consts<R>& c, const R U[NV8], R W[NV8]) {
  const R rho = smax(U[RHO], c.rho_floor);
  const R vx = U[MX] / rho;
  const R vy = U[MY] / rho;
  const R vz = U[MZ] / rho;
  const R bx = U[BX], by = U[BY], bz = U[BZ];
  const R v2 = vx * vx + vy * vy + vz * vz;
  const R b2 = bx * bx + by * by + bz * bz;
  const R p = (c.gamma - (R)1.0) * (U[EN] - (R)0.5 * rho * v2 - (R)0.5 * b2);
  W[0] = rho; W[1] = vx; W[2] = vy; W[3] = vz; W[4] = p;
  W[5] = bx; W[6] = by; W[7] = bz;
}

// mhd_common.hpp:206-238.  dir 0 uses Bx (x speed), dir 1 uses By.
template <class R>
inline R fast_speed(const Consts<R>& c, const R W[NV8], int dir) {
  const R rho = smax(W[0], c.rho_floor);
  const R p = smax(W[4], c.p_floor);
  const R bx = W[5], by = W[6], bz = W[7];
  const R a2 = c.gamma * p / rho;
  const R b2 = (bx * bx + by * by + bz * bz) / rho;
  const R bn = (dir == 0) ? bx : by;
  const R can2 = bn * bn / rho;
  R disc = (a2 + b2) * (a2 + b2) - (R)4.0 * a2 * can2;
  disc = smax(disc, (R)0.0);
  const R cf2 = (R)0.5 * (a2 + b2 + std::sqrt(disc));
  return std::sqrt(smax(cf2, (R)0.0));
}

// mhd_common.hpp:247-270
template <class R>
inline void phys_flux_x(const R W[NV8], R e, R F[NV8]) {
  const R rho = W[0], vx = W[1], vy = W[2], vz = W[3], p = W[4];
  const R bx = W[5], by = W[6], bz = W[7];
  const R b2 = bx * bx + by * by + bz * bz;
  const R ptot = p + (R)0.5 * b2;
  const R vdotb = vx * bx + vy * by + vz * bz;
  F[RHO] = rho * vx;
  F[MX] = rho * vx * vx + ptot - bx * bx;
  F[MY] = rho * vx * vy - bx * by;
  F[MZ] = rho * vx * vz - bx * bz;
  F[EN] = (e + ptot) * vx - bx * vdotb;
  F[BX] = (R)0.0;
  F[BY] = vx * by - bx * vy;
  F[BZ] = vx * bz - bx * vz;
}

// mhd_common.hpp:285-308
template <class R>
inline void phys_flux_y(const R W[NV8], R e, R G[NV8]) {
  const R rho = W[0], vx = W[1], vy = W[2], vz = W[3], p = W[4];
  const R bx = W[5], by = W[6], bz = W[7];
  const R b2 = bx * bx + by * by + bz * bz;
  const R ptot = p + (R)0.5 * b2;
  const R vdotb = vx * bx + vy * by + vz * bz;
  G[RHO] = rho * vy;
  G[MX] = rho * vy * vx - by * bx;
  G[MY] = rho * vy * vy + ptot - by * by;
  G[MZ] = rho * vy * vz - by * bz;
  G[EN] = (e + ptot) * vy - by * vdotb;
  G[BX] = vy * bx - by * vx;
  G[BY] = (R)0.0;
  G[BZ] = vy * bz - by * vz;
}

// mhd_common.hpp:323-377 (dir 0: x, dir 1: y; the y form uses the y flux and
// the y fast speed directly, not a rotation).
template <class R>
inline void hll_flux(const Consts<R>& c, const R UL[NV8], const R UR[NV8],
                     int dir, R F[NV8]) {
  R WL[NV8], WR[NV8], FL[NV8], FR[NV8];
  cons_to_prim8(c, UL, WL);
  cons_to_prim8(c, UR, WR);
  if (dir == 0) {
    phys_flux_x(WL, UL[EN], FL);
    phys_flux_x(WR, UR[EN], FR);
  } else {
    phys_flux_y(WL, UL[EN], FL);
    phys_flux_y(WR, UR[EN], FR);
  }
  const R cfL = fast_speed(c, WL, dir);
  const R cfR = fast_speed(c, WR, dir);
  const int vn = (dir == 0) ? 1 : 2;
  const R sL = smin(WL[vn] - cfL, WR[vn] - cfR);
  const R sR = smax(WL[vn] + cfL, WR[vn] + cfR);
  if (sL >= (R)0.0) {
    for (int k = 0; k < NV8; k++) F[k] = FL[k];
  } else if (sR <= (R)0.0) {
    for (int k = 0; k < NV8; k++) F[k] = FR[k];
  } else {
    const R denom = sR - sL;
    for (int k = 0; k < NV8; k++)
      F[k] = (sR * FL[k] - sL * FR[k] + sL * sR * (UR[k] - UL[k])) / denom;
  }
}

// mhd_common.hpp:381-405 (both rotations are the same swap).
template <class R>
inline void swap_xy(const R A[NV8], R B[NV8]) {
  B[RHO] = A[RHO]; B[MX] = A[MY]; B[MY] = A[MX]; B[MZ] = A[MZ];
  B[EN] = A[EN];   B[BX] = A[BY]; B[BY] = A[BX]; B[BZ] = A[BZ];
}

// mhd_common.hpp:417-419
template <class R>
inline bool hlld_small(const Consts<R>& c, R value, R scale) {
  return std::fabs(value) <= c.tol * scale;
}

// mhd_common.hpp:426-431
template <class R>
inline bool positive_pressure(const Consts<R>& c, const R U[NV8]) {
  if (!(U[RHO] > c.rho_floor)) return false;
  R W[NV8];
  cons_to_prim8(c, U, W);
  return (W[0] > c.rho_floor && W[4] > c.p_floor);
}

// mhd_common.hpp:445-511
template <class R>
inline int hlld_star_x(const Consts<R>& c, const R U[NV8], const R W[NV8], R s,
                       R sm, R pt_star, R bx, R Us[NV8]) {
  const R rho = smax(W[0], c.rho_floor);
  const R vx = W[1], vy = W[2], vz = W[3];
  const R p = smax(W[4], c.p_floor);
  const R by = W[6], bz = W[7];
  const R e = U[EN];
  const R s_minus_vx = s - vx;
  const R s_minus_sm = s - sm;
  const R wave_scale = smax(std::fabs(s), std::fabs(sm));
  if (hlld_small(c, s_minus_sm, wave_scale)) return HLLD_DEGENERATE;
  const R rho_star = rho * s_minus_vx / s_minus_sm;
  if (rho_star <= c.rho_floor) return HLLD_NONPHYSICAL;
  const R denom = rho * s_minus_vx * s_minus_sm - bx * bx;
  const R denom_scale = std::fabs(rho * s_minus_vx * s_minus_sm) + bx * bx;
  if (hlld_small(c, denom, denom_scale)) return HLLD_DEGENERATE;
  const R common = sm - vx;
  const R vy_star = vy - bx * by * common / denom;
  const R vz_star = vz - bx * bz * common / denom;
  const R by_star = by * (rho * s_minus_vx * s_minus_vx - bx * bx) / denom;
  const R bz_star = bz * (rho * s_minus_vx * s_minus_vx - bx * bx) / denom;
  const R b2 = bx * bx + by * by + bz * bz;
  const R pt = p + (R)0.5 * b2;
  const R vdotb = vx * bx + vy * by + vz * bz;
  const R vsdotbs = sm * bx + vy_star * by_star + vz_star * bz_star;
  const R e_star =
      (s_minus_vx * e - pt * vx + pt_star * sm + bx * (vdotb - vsdotbs)) /
      s_minus_sm;
  Us[RHO] = rho_star;
  Us[MX] = rho_star * sm;
  Us[MY] = rho_star * vy_star;
  Us[MZ] = rho_star * vz_star;
  Us[EN] = e_star;
  Us[BX] = bx;
  Us[BY] = by_star;
  Us[BZ] = bz_star;
  return positive_pressure(c, Us) ? HLLD_OK : HLLD_NONPHYSICAL;
}

// mhd_common.hpp:518-654
template <class R>
inline int hlld_flux_x(const Consts<R>& c, const R UL[NV8], const R UR[NV8],
                       R F[NV8]) {
  R WL[NV8], WR[NV8], FL[NV8], FR[NV8];
  cons_to_prim8(c, UL, WL);
  cons_to_prim8(c, UR, WR);
  phys_flux_x(WL, UL[EN], FL);
  phys_flux_x(WR, UR[EN], FR);
  const R cfL = fast_speed(c, WL, 0);
  const R cfR = fast_speed(c, WR, 0);
  const R sL = smin(WL[1] - cfL, WR[1] - cfR);
  const R sR = smax(WL[1] + cfL, WR[1] + cfR);
  if (sL >= (R)0.0) {
    for (int k = 0; k < NV8; k++) F[k] = FL[k];
    return HLLD_OK;
  }
  if (sR <= (R)0.0) {
    for (int k = 0; k < NV8; k++) F[k] = FR[k];
    return HLLD_OK;
  }
  const R rhoL = smax(WL[0], c.rho_floor);
  const R rhoR = smax(WR[0], c.rho_floor);
  const R vxL = WL[1], vxR = WR[1];
  const R bx = (R)0.5 * (WL[5] + WR[5]);
  const R b2L = bx * bx + WL[6] * WL[6] + WL[7] * WL[7];
  const R b2R = bx * bx + WR[6] * WR[6] + WR[7] * WR[7];
  const R ptL = smax(WL[4], c.p_floor) + (R)0.5 * b2L;
  const R ptR = smax(WR[4], c.p_floor) + (R)0.5 * b2R;
  const R denom_sm = rhoR * (sR - vxR) - rhoL * (sL - vxL);
  const R denom_sm_scale =
      std::fabs(rhoR * (sR - vxR)) + std::fabs(rhoL * (sL - vxL));
  const R b_scale = std::sqrt(smax(b2L, b2R));
  if (hlld_small(c, denom_sm, denom_sm_scale) || hlld_small(c, bx, b_scale)) {
    hll_flux(c, UL, UR, 0, F);
    return HLLD_DEGENERATE;
  }
  const R sm =
      (rhoR * vxR * (sR - vxR) - rhoL * vxL * (sL - vxL) + ptL - ptR) / denom_sm;
  const R pt_star_L = ptL + rhoL * (sL - vxL) * (sm - vxL);
  const R pt_star_R = ptR + rhoR * (sR - vxR) * (sm - vxR);
  const R pt_star = (R)0.5 * (pt_star_L + pt_star_R);

  R ULs[NV8], URs[NV8];
  const int ls = hlld_star_x(c, UL, WL, sL, sm, pt_star, bx, ULs);
  const int rs = hlld_star_x(c, UR, WR, sR, sm, pt_star, bx, URs);
  if (ls != HLLD_OK || rs != HLLD_OK) {
    hll_flux(c, UL, UR, 0, F);
    return (ls != HLLD_OK) ? ls : rs;
  }
  const R sqrt_rhoLs = std::sqrt(smax(ULs[RHO], c.rho_floor));
  const R sqrt_rhoRs = std::sqrt(smax(URs[RHO], c.rho_floor));
  const R sgn_bx = (bx < (R)0.0) ? (R)-1.0 : (R)1.0;
  const R sLs = sm - std::fabs(bx) / sqrt_rhoLs;
  const R sRs = sm + std::fabs(bx) / sqrt_rhoRs;
  const R denom = sqrt_rhoLs + sqrt_rhoRs;
  if (denom <= (R)0.0) {
    hll_flux(c, UL, UR, 0, F);
    return HLLD_DEGENERATE;
  }
  const R vyLs = ULs[MY] / ULs[RHO];
  const R vzLs = ULs[MZ] / ULs[RHO];
  const R byLs = ULs[BY];
  const R bzLs = ULs[BZ];
  const R vyRs = URs[MY] / URs[RHO];
  const R vzRs = URs[MZ] / URs[RHO];
  const R byRs = URs[BY];
  const R bzRs = URs[BZ];
  const R vyD =
      (sqrt_rhoLs * vyLs + sqrt_rhoRs * vyRs + sgn_bx * (byRs - byLs)) / denom;
  const R vzD =
      (sqrt_rhoLs * vzLs + sqrt_rhoRs * vzRs + sgn_bx * (bzRs - bzLs)) / denom;
  const R byD = (sqrt_rhoLs * byRs + sqrt_rhoRs * byLs +
                 sgn_bx * sqrt_rhoLs * sqrt_rhoRs * (vyRs - vyLs)) / denom;
  const R bzD = (sqrt_rhoLs * bzRs + sqrt_rhoRs * bzLs +
                 sgn_bx * sqrt_rhoLs * sqrt_rhoRs * (vzRs - vzLs)) / denom;
  R ULD[NV8], URD[NV8];
  for (int k = 0; k < NV8; k++) { ULD[k] = ULs[k]; URD[k] = URs[k]; }
  ULD[MY] = ULD[RHO] * vyD;
  ULD[MZ] = ULD[RHO] * vzD;
  ULD[BY] = byD;
  ULD[BZ] = bzD;
  URD[MY] = URD[RHO] * vyD;
  URD[MZ] = URD[RHO] * vzD;
  URD[BY] = byD;
  URD[BZ] = bzD;
  const R vdotbLD = sm * bx + vyD * byD + vzD * bzD;
  const R vdotbLs = sm * bx + vyLs * byLs + vzLs * bzLs;
  const R vdotbRD = vdotbLD;
  const R vdotbRs = sm * bx + vyRs * byRs + vzRs * bzRs;
  ULD[EN] = ULs[EN] - sgn_bx * sqrt_rhoLs * (vdotbLs - vdotbLD);
  URD[EN] = URs[EN] + sgn_bx * sqrt_rhoRs * (vdotbRs - vdotbRD);
  if (!positive_pressure(c, ULD) || !positive_pressure(c, URD)) {
    hll_flux(c, UL, UR, 0, F);
    return HLLD_NONPHYSICAL;
  }
  if (sLs >= (R)0.0) {
    for (int k = 0; k < NV8; k++) F[k] = FL[k] + sL * (ULs[k] - UL[k]);
  } else if (sm >= (R)0.0) {
    for (int k = 0; k < NV8; k++)
      F[k] = FL[k] + sL * (ULs[k] - UL[k]) + sLs * (ULD[k] - ULs[k]);
  } else if (sRs >= (R)0.0) {
    for (int k = 0; k < NV8; k++)
      F[k] = FR[k] + sR * (URs[k] - UR[k]) + sRs * (URD[k] - URs[k]);
  } else {
    for (int k = 0; k < NV8; k++) F[k] = FR[k] + sR * (URs[k] - UR[k]);
  }
  return HLLD_OK;
}

// mhd_common.hpp:658-667
template <class R>
inline int hlld_flux_y(const Consts<R>& c, const R UL[NV8], const R UR[NV8],
                       R G[NV8]) {
  R ULx[NV8], URx[NV8], Fx[NV8];
  swap_xy(UL, ULx);
  swap_xy(UR, URx);
  const int status = hlld_flux_x(c, ULx, URx, Fx);
  swap_xy(Fx, G);
  return status;
}

// mhd_glm_common.hpp:55-65
template <class R>
inline void glm_normal_state(R bnL, R bnR, R psiL, R psiR, R ch, R& bnStar,
                             R& psiStar) {
  const R ch_safe = smax(ch, (R)1.0e-12);
  bnStar = (R)0.5 * (bnL + bnR) - (R)0.5 * (psiR - psiL) / ch_safe;
  psiStar = (R)0.5 * (psiL + psiR) - (R)0.5 * ch_safe * (bnR - bnL);
}

// mhd_glm_common.hpp:69-126.  dir 0 = x (mhdGlmFluxX), 1 = y (mhdGlmFluxY).
template <class R>
inline int glm_flux(const Consts<R>& c, const R UL[NV], const R UR[NV], R ch,
                    bool use_hlld, int dir, R F[NV]) {
  const int bn = (dir == 0) ? BX : BY;
  R bnStar, psiStar;
  glm_normal_state(UL[bn], UR[bn], UL[PSI], UR[PSI], ch, bnStar, psiStar);
  R UL8[NV8], UR8[NV8], F8[NV8];
  for (int k = 0; k < NV8; k++) { UL8[k] = UL[k]; UR8[k] = UR[k]; }
  UL8[bn] = bnStar;
  UR8[bn] = bnStar;
  int status = HLLD_OK;
  if (use_hlld) {
    status = (dir == 0) ? hlld_flux_x(c, UL8, UR8, F8)
                        : hlld_flux_y(c, UL8, UR8, F8);
  } else {
    hll_flux(c, UL8, UR8, dir, F8);
  }
  for (int k = 0; k < NV8; k++) F[k] = F8[k];
  F[bn] = psiStar;
  F[PSI] = ch * ch * bnStar;
  return status;
}

// 9-variable wrappers: mhd2d_glm_cpu.cpp:90-103 via mhd_glm_common.hpp:34-50.
template <class R>
inline void state_to_prim(const Consts<R>& c, const R U[NV], R P[NV]) {
  cons_to_prim8(c, U, P);
  P[PSI] = U[PSI];
}
template <class R>
inline void prim_to_state(const Consts<R>& c, const R P[NV], R U[NV]) {
  prim_to_cons8(c, P, U);
  U[PSI] = P[PSI];
}

// mhd2d_glm_cpu.cpp:105-111
template <class R>
inline bool prim_usable(const Consts<R>& c, const R P[NV]) {
  if (P[0] <= c.rho_floor || P[4] <= c.p_floor) return false;
  for (int k = 0; k < NV; k++)
    if (!std::isfinite((double)P[k])) return false;
  return true;
}

// mhd2d_glm_cpu.cpp:113-118
template <class R>
inline bool state_usable(const Consts<R>& c, const R U[NV]) {
  if (U[RHO] <= c.rho_floor) return false;
  R P[NV];
  state_to_prim(c, U, P);
  return prim_usable(c, P);
}

// mhd2d_glm_cpu.cpp:144-161
template <class R>
inline void reconstruct(const Consts<R>& c, const R Pm[NV], const R P0[NV],
                        const R Pp[NV], R Pl[NV], R Pr[NV]) {
  for (int k = 0; k < NV; k++) {
    const R slope = minmod(P0[k] - Pm[k], Pp[k] - P0[k]);
    Pl[k] = P0[k] - (R)0.5 * slope;
    Pr[k] = P0[k] + (R)0.5 * slope;
  }
  if (!prim_usable(c, Pl) || !prim_usable(c, Pr)) {
    for (int k = 0; k < NV; k++) { Pl[k] = P0[k]; Pr[k] = P0[k]; }
  }
}

// Counters of HLLD->HLL fallbacks (mhd2d_glm_cpu.cpp:48-74).
struct FallbackStats {
  uint64_t x_evals = 0, y_evals = 0, degenerate = 0, nonphysical = 0;
  void record(int dir, int status) {
    if (dir == 0) x_evals++; else y_evals++;
    if (status == HLLD_DEGENERATE) degenerate++;
    if (status == HLLD_NONPHYSICAL) nonphysical++;
  }
};

// A periodic nx*ny grid of 9-variable conserved states, cell (i,j) at
// j*nx+i.  The reference keeps 2 ghost layers refreshed by
// applyPeriodicBoundary (mhd2d_glm_cpu.cpp:120-142); for nx,ny >= 2 every
// ghost read equals the periodically wrapped interior cell, which is how
// this restatement indexes.
template <class R>
struct Grid {
  int nx = 0, ny = 0;
  std::vector<R> u;  // AoS: (j*nx+i)*9 + k
  Grid() {}
  Grid(int nx_, int ny_) : nx(nx_), ny(ny_), u((size_t)nx_ * ny_ * NV) {}
  int wi(int i) const { return ((i % nx) + nx) % nx; }
  int wj(int j) const { return ((j % ny) + ny) % ny; }
  const R* at(int i, int j) const { return &u[((size_t)wj(j) * nx + wi(i)) * NV]; }
  R* at(int i, int j) { return &u[((size_t)wj(j) * nx + wi(i)) * NV]; }
};

// Face states of the face between cells c0 and c0+1 along dir
// (xFaceStates / yFaceStates, mhd2d_glm_cpu.cpp:163-229).
template <class R>
inline void face_states(const Consts<R>& c, const Grid<R>& U, int i, int j,
                        int dir, bool use_muscl, R UL[NV], R UR[NV]) {
  const int di = (dir == 0) ? 1 : 0, dj = (dir == 0) ? 0 : 1;
  if (!use_muscl) {
    const R* a = U.at(i, j);
    const R* b = U.at(i + di, j + dj);
    for (int k = 0; k < NV; k++) { UL[k] = a[k]; UR[k] = b[k]; }
    return;
  }
  R Pm[NV], P0[NV], Pp[NV], Pq[NV];
  state_to_prim(c, U.at(i - di, j - dj), Pm);
  state_to_prim(c, U.at(i, j), P0);
  state_to_prim(c, U.at(i + di, j + dj), Pp);
  state_to_prim(c, U.at(i + 2 * di, j + 2 * dj), Pq);
  R Pl0[NV], Pr0[NV], Pl1[NV], Pr1[NV];
  reconstruct(c, Pm, P0, Pp, Pl0, Pr0);
  reconstruct(c, P0, Pp, Pq, Pl1, Pr1);
  prim_to_state(c, Pr0, UL);
  prim_to_state(c, Pl1, UR);
}

// computeRhs (mhd2d_glm_cpu.cpp:231-287).  Fx/Gy are evaluated for
// (nx+1) x-faces per row and (ny+1) y-faces per column, exactly as the
// reference loops i in [g-1, nx+g) and j in [g-1, ny+g): the periodic
// boundary face is evaluated (and counted) twice.
template <class R>
inline void compute_rhs(const Consts<R>& c, const Grid<R>& U, R dx, R dy,
                        bool use_muscl, bool use_hlld, R ch, R alpha,
                        FallbackStats* stats, std::vector<R>& rhs) {
  const int nx = U.nx, ny = U.ny;
  const R damping = ch / smax(alpha, (R)1.0e-12);
  std::vector<R> Fx((size_t)(nx + 1) * ny * NV), Gy((size_t)nx * (ny + 1) * NV);
  for (int j = 0; j < ny; j++) {
    for (int i = -1; i < nx; i++) {
      R UL[NV], UR[NV], F[NV];
      face_states(c, U, i, j, 0, use_muscl, UL, UR);
      const int st = glm_flux(c, UL, UR, ch, use_hlld, 0, F);
      if (use_hlld && stats) stats->record(0, st);
      R* dst = &Fx[((size_t)j * (nx + 1) + (i + 1)) * NV];
      for (int k = 0; k < NV; k++) dst[k] = F[k];
    }
  }
  for (int j = -1; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      R UL[NV], UR[NV], G[NV];
      face_states(c, U, i, j, 1, use_muscl, UL, UR);
      const int st = glm_flux(c, UL, UR, ch, use_hlld, 1, G);
      if (use_hlld && stats) stats->record(1, st);
      R* dst = &Gy[((size_t)(j + 1) * nx + i) * NV];
      for (int k = 0; k < NV; k++) dst[k] = G[k];
    }
  }
  rhs.assign((size_t)nx * ny * NV, (R)0);
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      const R* fr = &Fx[((size_t)j * (nx + 1) + (i + 1)) * NV];
      const R* fl = &Fx[((size_t)j * (nx + 1) + i) * NV];
      const R* gt = &Gy[((size_t)(j + 1) * nx + i) * NV];
      const R* gb = &Gy[((size_t)j * nx + i) * NV];
      R* r = &rhs[((size_t)j * nx + i) * NV];
      for (int k = 0; k < NV; k++)
        r[k] = -(fr[k] - fl[k]) / dx - (gt[k] - gb[k]) / dy;
      r[PSI] -= damping * U.at(i, j)[PSI];
    }
  }
}

// maxPhysicalSignalSpeed (mhd2d_glm_cpu.cpp:289-305)
template <class R>
inline R max_signal_speed(const Consts<R>& c, const Grid<R>& U) {
  R m = (R)0.0;
  for (int j = 0; j < U.ny; j++) {
    for (int i = 0; i < U.nx; i++) {
      R P[NV];
      state_to_prim(c, U.at(i, j), P);
      const R sx = std::fabs(P[1]) + fast_speed(c, P, 0);
      const R sy = std::fabs(P[2]) + fast_speed(c, P, 1);
      m = smax(m, smax(sx, sy));
    }
  }
  return m;
}

// maxSignalRate (mhd2d_glm_cpu.cpp:307-326)
template <class R>
inline R max_signal_rate(const Consts<R>& c, const Grid<R>& U, R dx, R dy, R ch) {
  R m = (R)0.0;
  for (int j = 0; j < U.ny; j++) {
    for (int i = 0; i < U.nx; i++) {
      R P[NV];
      state_to_prim(c, U.at(i, j), P);
      const R sx = smax(std::fabs(P[1]) + fast_speed(c, P, 0), ch);
      const R sy = smax(std::fabs(P[2]) + fast_speed(c, P, 1), ch);
      m = smax(m, sx / dx + sy / dy);
    }
  }
  return m;
}

// allInteriorStatesUsable (mhd2d_glm_cpu.cpp:328-345): first bad cell in
// j-major order.
template <class R>
inline bool all_usable(const Consts<R>& c, const Grid<R>& U, int& bi, int& bj) {
  for (int j = 0; j < U.ny; j++)
    for (int i = 0; i < U.nx; i++)
      if (!state_usable(c, U.at(i, j))) { bi = i; bj = j; return false; }
  return true;
}

// divergenceB (mhd2d_glm_cpu.cpp:347-358)
template <class R>
inline R div_b(const Grid<R>& U, int i, int j, R dx, R dy) {
  const R bxp = U.at(i + 1, j)[BX], bxm = U.at(i - 1, j)[BX];
  const R byp = U.at(i, j + 1)[BY], bym = U.at(i, j - 1)[BY];
  return (bxp - bxm) / ((R)2.0 * dx) + (byp - bym) / ((R)2.0 * dy);
}

struct StepResult {
  int code = 0;      // 0 ok, 2 non-positive rate, 3 stage-1 failure, 4 stage-2
  int bad_i = -1, bad_j = -1;
  double ch = 0, dt = 0;
};

// One SSPRK2 step of the time loop (mhd2d_glm_cpu.cpp:513-569).  On success
// U holds U^{n+1} and t is advanced; on failure U/U1 hold the failing state
// and t is unchanged.
template <class R>
inline StepResult ssprk2_step(const Consts<R>& c, Grid<R>& U, Grid<R>& U1,
                              R dx, R dy, R cfl, R alpha, bool use_muscl,
                              bool use_hlld, R& t, R t_end,
                              FallbackStats* stats, R* last_ch) {
  StepResult res;
  const R speed = max_signal_speed(c, U);
  const R ch = smax(speed, (R)1.0e-12);
  if (last_ch) *last_ch = ch;
  res.ch = (double)ch;
  const R rate = max_signal_rate(c, U, dx, dy, ch);
  if (rate <= (R)0.0) { res.code = 2; return res; }
  R dt = cfl / rate;
  if (t + dt > t_end) dt = t_end - t;
  res.dt = (double)dt;
  std::vector<R> rhs;
  compute_rhs(c, U, dx, dy, use_muscl, use_hlld, ch, alpha, stats, rhs);
  U1.nx = U.nx; U1.ny = U.ny; U1.u.resize(U.u.size());
  for (size_t n = 0; n < U.u.size(); n++) U1.u[n] = U.u[n] + dt * rhs[n];
  if (!all_usable(c, U1, res.bad_i, res.bad_j)) { res.code = 3; return res; }
  compute_rhs(c, U1, dx, dy, use_muscl, use_hlld, ch, alpha, stats, rhs);
  for (size_t n = 0; n < U.u.size(); n++)
    U.u[n] = (R)0.5 * U.u[n] + (R)0.5 * (U1.u[n] + dt * rhs[n]);
  if (!all_usable(c, U, res.bad_i, res.bad_j)) { res.code = 4; return res; }
  t += dt;
  return res;
}

}  // namespace oracle

#endif  // MHD_ORACLE_HPP

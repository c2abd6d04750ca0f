// mhd_device.hpp -- pointwise ideal-MHD + mixed-GLM physics for CDNA4.
//
// Device functions on register arrays, templated on the arithmetic type R
// (double or float).  The association order of every expression is the
// reference's (include/mhd/mhd_common.hpp, include/mhd/mhd_glm_common.hpp):
// with FP contraction off the results are bitwise those of the reference
// CPU build.  The GPU-specific restructurings below change no rounding:
//   * y-normal faces are evaluated in the x-normal frame (swap of the
//     (mx,my) and (Bx,By) slots), as the reference's HLLD-y does
//     (mhd_common.hpp:658-667).  The reference's HLL-y (mhd_common.hpp:352-377)
//     uses the y flux directly; it differs from the rotated HLL-x only in the
//     order of the two operands of some additions and multiplications, which
//     IEEE arithmetic makes exact (a+b == b+a, a*b == b*a);
//   * the HLL fallback inside HLLD reuses the primitives, physical fluxes and
//     outer wave speeds already computed in the HLLD prologue, which are the
//     values mhdHllFluxX would recompute from the same inputs;
//   * max/min are written as the host's std::max/std::min ternaries
//     ((a<b)?b:a, (b<a)?b:a), never fmax/fmin, so NaN and signed-zero
//     behaviour matches the reference CPU build.
#ifndef MHD_DEVICE_HPP
#define MHD_DEVICE_HPP

#include <hip/hip_runtime.h>

namespace mhd {

#define MHD_DEV __device__ __forceinline__

enum { RHO = 0, MX = 1, MY = 2, MZ = 3, EN = 4, BX = 5, BY = 6, BZ = 7, PSI = 8 };
enum { HLLD_OK = 0, HLLD_DEGENERATE = 1, HLLD_NONPHYSICAL = 2 };

// Physical constants in the arithmetic type (mhd_common.hpp:52-79; gamma is
// 5/3 as in mhd2d_glm_cpu.cpp:25-27).  gm1 = gamma - 1 evaluated in R, as the
// reference evaluates (MHD_GAMMA - (Real)1.0).
template <class R>
struct Phys {
  R gamma, gm1, rho_floor, p_floor, tol;
};

template <class R> MHD_DEV R smax(R a, R b) { return (a < b) ? b : a; }
template <class R> MHD_DEV R smin(R a, R b) { return (b < a) ? b : a; }
template <class R> MHD_DEV R fabs_(R a) { return __builtin_fabs(a); }
template <> MHD_DEV float fabs_<float>(float a) { return __builtin_fabsf(a); }
template <class R> MHD_DEV R sqrt_(R a) { return __builtin_sqrt(a); }
template <> MHD_DEV float sqrt_<float>(float a) { return __builtin_sqrtf(a); }
template <class R> MHD_DEV bool finite_(R a) { return __builtin_isfinite(a); }

// minmod (mhd_common.hpp:140-143)
template <class R> MHD_DEV R minmod(R a, R b) {
  if (a * b <= (R)0.0) return (R)0.0;
  return (fabs_(a) < fabs_(b)) ? a : b;
}

// Primitive -> conserved for the 8 MHD slots; psi passes through
// (mhd_common.hpp:149-173, mhd_glm_common.hpp:34-41).
template <class R>
MHD_DEV void prim_to_cons(const Phys<R>& c, const R W[9], R U[9]) {
  const R rho = W[0], vx = W[1], vy = W[2], vz = W[3], p = W[4];
  const R bx = W[5], by = W[6], bz = W[7];
  const R v2 = vx * vx + vy * vy + vz * vz;
  const R b2 = bx * bx + by * by + bz * bz;
  U[RHO] = rho;
  U[MX] = rho * vx;
  U[MY] = rho * vy;
  U[MZ] = rho * vz;
  U[EN] = p / c.gm1 + (R)0.5 * rho * v2 + (R)0.5 * b2;
  U[BX] = bx;
  U[BY] = by;
  U[BZ] = bz;
  U[PSI] = W[8];
}

// Conserved -> primitive, pressure recovery (mhd_common.hpp:178-201).
// Writes rho, v, p, B; psi passes through when NPSI is 9.
template <class R, int NPSI = 9>
MHD_DEV void cons_to_prim(const Phys<R>& c, const R U[], R W[]) {
  const R rho = smax(U[RHO], c.rho_floor);
  const R vx = U[MX] / rho;
  const R vy = U[MY] / rho;
  const R vz = U[MZ] / rho;
  const R bx = U[BX], by = U[BY], bz = U[BZ];
  const R v2 = vx * vx + vy * vy + vz * vz;
  const R b2 = bx * bx + by * by + bz * bz;
  W[0] = rho; W[1] = vx; W[2] = vy; W[3] = vz;
  W[4] = c.gm1 * (U[EN] - (R)0.5 * rho * v2 - (R)0.5 * b2);
  W[5] = bx; W[6] = by; W[7] = bz;
  if (NPSI == 9) W[8] = U[PSI];
}

// Fast magnetosonic speed along the axis whose field component is bn
// (mhd_common.hpp:206-238; bn = Bx for x, By for y).
template <class R>
MHD_DEV R fast_speed(const Phys<R>& c, const R W[], R bn) {
  const R rho = smax(W[0], c.rho_floor);
  const R p = smax(W[4], c.p_floor);
  const R bx = W[5], by = W[6], bz = W[7];
  const R a2 = c.gamma * p / rho;
  const R b2 = (bx * bx + by * by + bz * bz) / rho;
  const R can2 = bn * bn / rho;
  R disc = (a2 + b2) * (a2 + b2) - (R)4.0 * a2 * can2;
  disc = smax(disc, (R)0.0);
  const R cf2 = (R)0.5 * (a2 + b2 + sqrt_(disc));
  return sqrt_(smax(cf2, (R)0.0));
}

// x-normal physical flux from primitives and total energy
// (mhd_common.hpp:247-270).
template <class R>
MHD_DEV void phys_flux_x(const R W[], R e, R F[8]) {
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

// Positive density and pressure of a conserved state; NaN fails
// (mhd_common.hpp:426-431).
template <class R>
MHD_DEV bool positive_pressure(const Phys<R>& c, const R U[8]) {
  if (!(U[RHO] > c.rho_floor)) return false;
  R W[8];
  cons_to_prim<R, 8>(c, U, W);
  return (W[0] > c.rho_floor && W[4] > c.p_floor);
}

// Scale-aware degeneracy test (mhd_common.hpp:417-419).
template <class R>
MHD_DEV bool hlld_small(const Phys<R>& c, R value, R scale) {
  return fabs_(value) <= c.tol * scale;
}

// HLLD star state next to an outer fast wave (mhd_common.hpp:445-511).
template <class R>
MHD_DEV int hlld_star(const Phys<R>& c, R e, const R W[8], R s, R sm,
                      R pt_star, R bx, R Us[8]) {
  const R rho = smax(W[0], c.rho_floor);
  const R vx = W[1], vy = W[2], vz = W[3];
  const R p = smax(W[4], c.p_floor);
  const R by = W[6], bz = W[7];
  const R s_minus_vx = s - vx;
  const R s_minus_sm = s - sm;
  const R wave_scale = smax(fabs_(s), fabs_(sm));
  if (hlld_small(c, s_minus_sm, wave_scale)) return HLLD_DEGENERATE;
  const R rho_star = rho * s_minus_vx / s_minus_sm;
  if (rho_star <= c.rho_floor) return HLLD_NONPHYSICAL;
  const R denom = rho * s_minus_vx * s_minus_sm - bx * bx;
  const R denom_scale = fabs_(rho * s_minus_vx * s_minus_sm) + bx * bx;
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

// HLL average of the outer states (mhd_common.hpp:338-347).
template <class R>
MHD_DEV void hll_average(const R UL[], const R UR[], const R FL[8],
                         const R FR[8], R sL, R sR, R F[8]) {
  if (sL >= (R)0.0) {
#pragma unroll
    for (int k = 0; k < 8; k++) F[k] = FL[k];
  } else if (sR <= (R)0.0) {
#pragma unroll
    for (int k = 0; k < 8; k++) F[k] = FR[k];
  } else {
    const R denom = sR - sL;
#pragma unroll
    for (int k = 0; k < 8; k++)
      F[k] = (sR * FL[k] - sL * FR[k] + sL * sR * (UR[k] - UL[k])) / denom;
  }
}

// x-normal HLL or HLLD flux of the 8 MHD slots of UL/UR (UL[BX], UR[BX]
// already replaced by the GLM normal field).  HLLD follows
// mhd_common.hpp:518-654; HLL follows mhd_common.hpp:323-348.
template <class R, bool HLLD>
MHD_DEV int mhd_flux_x(const Phys<R>& c, const R UL[], const R UR[], R F[8]) {
  R WL[8], WR[8], FL[8], FR[8];
  cons_to_prim<R, 8>(c, UL, WL);
  cons_to_prim<R, 8>(c, UR, WR);
  phys_flux_x(WL, UL[EN], FL);
  phys_flux_x(WR, UR[EN], FR);
  const R cfL = fast_speed(c, WL, WL[5]);
  const R cfR = fast_speed(c, WR, WR[5]);
  const R sL = smin(WL[1] - cfL, WR[1] - cfR);
  const R sR = smax(WL[1] + cfL, WR[1] + cfR);
  if (!HLLD || sL >= (R)0.0 || sR <= (R)0.0) {
    hll_average(UL, UR, FL, FR, sL, sR, F);
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
  const R denom_sm_scale = fabs_(rhoR * (sR - vxR)) + fabs_(rhoL * (sL - vxL));
  const R b_scale = sqrt_(smax(b2L, b2R));
  int status = HLLD_OK;
  R ULs[8], URs[8];
  R sLs = (R)0, sRs = (R)0, sm = (R)0, sqrt_rhoLs = (R)0, sqrt_rhoRs = (R)0;
  R sgn_bx = (R)1;
  if (hlld_small(c, denom_sm, denom_sm_scale) || hlld_small(c, bx, b_scale)) {
    status = HLLD_DEGENERATE;
  } else {
    sm = (rhoR * vxR * (sR - vxR) - rhoL * vxL * (sL - vxL) + ptL - ptR) /
         denom_sm;
    const R pt_star_L = ptL + rhoL * (sL - vxL) * (sm - vxL);
    const R pt_star_R = ptR + rhoR * (sR - vxR) * (sm - vxR);
    const R pt_star = (R)0.5 * (pt_star_L + pt_star_R);
    const int ls = hlld_star(c, UL[EN], WL, sL, sm, pt_star, bx, ULs);
    const int rs = hlld_star(c, UR[EN], WR, sR, sm, pt_star, bx, URs);
    if (ls != HLLD_OK || rs != HLLD_OK) {
      status = (ls != HLLD_OK) ? ls : rs;
    } else {
      sqrt_rhoLs = sqrt_(smax(ULs[RHO], c.rho_floor));
      sqrt_rhoRs = sqrt_(smax(URs[RHO], c.rho_floor));
      sgn_bx = (bx < (R)0.0) ? (R)-1.0 : (R)1.0;
      sLs = sm - fabs_(bx) / sqrt_rhoLs;
      sRs = sm + fabs_(bx) / sqrt_rhoRs;
      if (sqrt_rhoLs + sqrt_rhoRs <= (R)0.0) status = HLLD_DEGENERATE;
    }
  }
  if (status != HLLD_OK) {
    hll_average(UL, UR, FL, FR, sL, sR, F);
    return status;
  }
  const R denom = sqrt_rhoLs + sqrt_rhoRs;
  const R vyLs = ULs[MY] / ULs[RHO];
  const R vzLs = ULs[MZ] / ULs[RHO];
  const R byLs = ULs[BY], bzLs = ULs[BZ];
  const R vyRs = URs[MY] / URs[RHO];
  const R vzRs = URs[MZ] / URs[RHO];
  const R byRs = URs[BY], bzRs = URs[BZ];
  const R vyD =
      (sqrt_rhoLs * vyLs + sqrt_rhoRs * vyRs + sgn_bx * (byRs - byLs)) / denom;
  const R vzD =
      (sqrt_rhoLs * vzLs + sqrt_rhoRs * vzRs + sgn_bx * (bzRs - bzLs)) / denom;
  const R byD = (sqrt_rhoLs * byRs + sqrt_rhoRs * byLs +
                 sgn_bx * sqrt_rhoLs * sqrt_rhoRs * (vyRs - vyLs)) / denom;
  const R bzD = (sqrt_rhoLs * bzRs + sqrt_rhoRs * bzLs +
                 sgn_bx * sqrt_rhoLs * sqrt_rhoRs * (vzRs - vzLs)) / denom;
  R ULD[8], URD[8];
#pragma unroll
  for (int k = 0; k < 8; k++) { ULD[k] = ULs[k]; URD[k] = URs[k]; }
  ULD[MY] = ULD[RHO] * vyD;
  ULD[MZ] = ULD[RHO] * vzD;
  ULD[BY] = byD;
  ULD[BZ] = bzD;
  URD[MY] = URD[RHO] * vyD;
  URD[MZ] = URD[RHO] * vzD;
  URD[BY] = byD;
  URD[BZ] = bzD;
  const R vdotbD = sm * bx + vyD * byD + vzD * bzD;
  const R vdotbLs = sm * bx + vyLs * byLs + vzLs * bzLs;
  const R vdotbRs = sm * bx + vyRs * byRs + vzRs * bzRs;
  ULD[EN] = ULs[EN] - sgn_bx * sqrt_rhoLs * (vdotbLs - vdotbD);
  URD[EN] = URs[EN] + sgn_bx * sqrt_rhoRs * (vdotbRs - vdotbD);
  if (!positive_pressure(c, ULD) || !positive_pressure(c, URD)) {
    hll_average(UL, UR, FL, FR, sL, sR, F);
    return HLLD_NONPHYSICAL;
  }
  if (sLs >= (R)0.0) {
#pragma unroll
    for (int k = 0; k < 8; k++) F[k] = FL[k] + sL * (ULs[k] - UL[k]);
  } else if (sm >= (R)0.0) {
#pragma unroll
    for (int k = 0; k < 8; k++)
      F[k] = FL[k] + sL * (ULs[k] - UL[k]) + sLs * (ULD[k] - ULs[k]);
  } else if (sRs >= (R)0.0) {
#pragma unroll
    for (int k = 0; k < 8; k++)
      F[k] = FR[k] + sR * (URs[k] - UR[k]) + sRs * (URD[k] - URs[k]);
  } else {
#pragma unroll
    for (int k = 0; k < 8; k++) F[k] = FR[k] + sR * (URs[k] - UR[k]);
  }
  return HLLD_OK;
}

// GLM flux in the x-normal frame (mhd_glm_common.hpp:55-96): exact upwind
// state of the (Bn, psi) subsystem, Bn replaced on both sides, then the MHD
// flux; F[Bn] = psi*, F[psi] = ch^2 Bn*.  For a y face the caller passes
// swapped states and swaps the flux back (mhd_glm_common.hpp:99-126).
template <class R, bool HLLD>
MHD_DEV int glm_flux_x(const Phys<R>& c, R UL[9], R UR[9], R ch, R F[9]) {
  const R ch_safe = smax(ch, (R)1.0e-12);
  const R bnStar =
      (R)0.5 * (UL[BX] + UR[BX]) - (R)0.5 * (UR[PSI] - UL[PSI]) / ch_safe;
  const R psiStar =
      (R)0.5 * (UL[PSI] + UR[PSI]) - (R)0.5 * ch_safe * (UR[BX] - UL[BX]);
  UL[BX] = bnStar;
  UR[BX] = bnStar;
  const int status = mhd_flux_x<R, HLLD>(c, UL, UR, F);
  F[BX] = psiStar;
  F[PSI] = ch * ch * bnStar;
  return status;
}

// Usability of reconstructed primitives (mhd2d_glm_cpu.cpp:105-111).
template <class R>
MHD_DEV bool prim_usable(const Phys<R>& c, const R P[9]) {
  bool ok = !(P[0] <= c.rho_floor || P[4] <= c.p_floor);
#pragma unroll
  for (int k = 0; k < 9; k++) ok = ok && finite_(P[k]);
  return ok;
}

// Usability of a conserved state (mhd2d_glm_cpu.cpp:113-118).
template <class R>
MHD_DEV bool state_usable(const Phys<R>& c, const R U[9], R P[9]) {
  cons_to_prim<R, 9>(c, U, P);
  return !(U[RHO] <= c.rho_floor) && prim_usable(c, P);
}

// Minmod MUSCL reconstruction of one cell (mhd2d_glm_cpu.cpp:144-161):
// returns the left- or right-face primitive state of the centre cell
// (RIGHT selects P0 + slope/2), first order if either face is unusable.
template <class R, bool RIGHT>
MHD_DEV void reconstruct_face(const Phys<R>& c, const R Pm[9], const R P0[9],
                              const R Pp[9], R out[9]) {
  R Pl[9], Pr[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const R slope = minmod(P0[k] - Pm[k], Pp[k] - P0[k]);
    Pl[k] = P0[k] - (R)0.5 * slope;
    Pr[k] = P0[k] + (R)0.5 * slope;
  }
  const bool ok = prim_usable(c, Pl) && prim_usable(c, Pr);
#pragma unroll
  for (int k = 0; k < 9; k++) out[k] = ok ? (RIGHT ? Pr[k] : Pl[k]) : P0[k];
}

}  // namespace mhd

#endif  // MHD_DEVICE_HPP

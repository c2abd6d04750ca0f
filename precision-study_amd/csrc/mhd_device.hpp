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

// IEEE division n / d with the denominator's reciprocal shared among several
// divisions.  gfx950 divides in FP64 as: ds = div_scale(d; n) (the
// denominator, rescaled only in extreme exponent ranges or when n is 0),
// ns = div_scale(n; d) with a rescale flag, y = rcp(ds) refined by two
// Newton steps, q = ns*y, r = fma(-ds, q, ns), and the quotient
// div_fixup(div_fmas(r, y, q, flag), d, n).  Only y takes the quarter-rate
// rcp and four FMAs, and it depends on ds alone.
//   Div<true, double>::rec  performs that sequence and records (ds, y);
//   Div<true, double>::use  performs it for a later (n, d) with the recorded
//       y, without a branch.  Its quotient is the `/` quotient when its own
//       ds is the recorded one bit for bit, or when n is zero, infinite or
//       NaN (div_fixup then returns a value that does not depend on the
//       quotient it is given).  Otherwise (a denominator that div_scale
//       rescales for this numerator: extreme exponents only) it sets `miss`
//       and the caller evaluates again with `/`: glm_flux_x re-evaluates the
//       whole face flux, for the lanes that missed.
//   Div<false, R> and every FP32 Div use `/`.
// rr() is use() when `reuse`, else rec(), as one division expression for
// FP32: two alike `/` in the arms of a conditional get merged by the
// optimiser, which then drops their accuracy annotation (the mixed forms'
// fast FP32 division would silently become the IEEE one).
template <class R> struct Den {};
template <> struct Den<double> { double ds, y; };

template <bool S, class R>
struct Div {
  static MHD_DEV R rec(R n, R d, Den<R>&) { return n / d; }
  static MHD_DEV R use(R n, R d, const Den<R>&, bool&) { return n / d; }
  static MHD_DEV R rr(R n, R d, Den<R>&, bool, bool&) { return n / d; }
};
template <>
struct Div<true, double> {
  static MHD_DEV double refine(double ds) {
    const double y0 = __builtin_amdgcn_rcp(ds);
    const double e0 = __builtin_fma(-ds, y0, 1.0);
    const double y1 = __builtin_fma(y0, e0, y0);
    const double e1 = __builtin_fma(-ds, y1, 1.0);
    return __builtin_fma(y1, e1, y1);
  }
  static MHD_DEV double finish(double n, double d, double ds, double ns, bool scaled, double y) {
    const double q = ns * y;
    const double r = __builtin_fma(-ds, q, ns);
    return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(r, y, q, scaled), d, n);
  }
  static MHD_DEV double rec(double n, double d, Den<double>& D) {
    bool unused, scaled;
    D.ds = __builtin_amdgcn_div_scale(n, d, false, &unused);
    const double ns = __builtin_amdgcn_div_scale(n, d, true, &scaled);
    D.y = refine(D.ds);
    return finish(n, d, D.ds, ns, scaled, D.y);
  }
  static MHD_DEV double use(double n, double d, const Den<double>& D, bool& miss) {
    bool unused, scaled;
    const double ds = __builtin_amdgcn_div_scale(n, d, false, &unused);
    const double ns = __builtin_amdgcn_div_scale(n, d, true, &scaled);
    const bool counts = __builtin_isfinite(n) && n != 0.0;
    miss = miss || (counts && __double_as_longlong(ds) != __double_as_longlong(D.ds));
    return finish(n, d, ds, ns, scaled, D.y);
  }
  static MHD_DEV double rr(double n, double d, Den<double>& D, bool reuse, bool& miss) {
    return reuse ? use(n, d, D, miss) : rec(n, d, D);
  }
};

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
// Writes rho, v, p, B; psi passes through when NPSI is 9.  With S the
// density's reciprocal is recorded in D (for the wave speeds of the same
// state), or with `reuse` taken from D (Div).  (References and flags rather
// than optional pointers: a private-memory pointer cannot be proven
// non-null, and a select between two of them would put D on the stack.)
template <class R, int NPSI, bool S>
MHD_DEV void cons_to_prim(const Phys<R>& c, const R U[], R W[], Den<R>& D, bool reuse,
                          bool& miss) {
  const R rho = smax(U[RHO], c.rho_floor);
  const R vx = Div<S, R>::rr(U[MX], rho, D, reuse, miss);
  const R vy = Div<S, R>::use(U[MY], rho, D, miss);
  const R vz = Div<S, R>::use(U[MZ], rho, D, miss);
  const R bx = U[BX], by = U[BY], bz = U[BZ];
  const R v2 = vx * vx + vy * vy + vz * vz;
  const R b2 = bx * bx + by * by + bz * bz;
  W[0] = rho; W[1] = vx; W[2] = vy; W[3] = vz;
  W[4] = c.gm1 * (U[EN] - (R)0.5 * rho * v2 - (R)0.5 * b2);
  W[5] = bx; W[6] = by; W[7] = bz;
  if (NPSI == 9) W[8] = U[PSI];
}
template <class R, int NPSI = 9>
MHD_DEV void cons_to_prim(const Phys<R>& c, const R U[], R W[]) {
  Den<R> D;
  bool miss = false;
  cons_to_prim<R, NPSI, false>(c, U, W, D, false, miss);
}

// Fast magnetosonic speed along the axis whose field component is bn
// (mhd_common.hpp:206-238; bn = Bx for x, By for y).
// With S the density's reciprocal is recorded in D, or (reuse) taken from
// it (recorded by the recovery of the same state).
template <bool S, class R>
MHD_DEV R fast_speed(const Phys<R>& c, const R W[], R bn, Den<R>& D, bool reuse, bool& miss) {
  const R rho = smax(W[0], c.rho_floor);
  const R p = smax(W[4], c.p_floor);
  const R bx = W[5], by = W[6], bz = W[7];
  const R a2 = Div<S, R>::rr(c.gamma * p, rho, D, reuse, miss);
  const R b2 = Div<S, R>::use(bx * bx + by * by + bz * bz, rho, D, miss);
  const R can2 = Div<S, R>::use(bn * bn, rho, D, miss);
  R disc = (a2 + b2) * (a2 + b2) - (R)4.0 * a2 * can2;
  disc = smax(disc, (R)0.0);
  const R cf2 = (R)0.5 * (a2 + b2 + sqrt_(disc));
  return sqrt_(smax(cf2, (R)0.0));
}
template <class R>
MHD_DEV R fast_speed(const Phys<R>& c, const R W[], R bn) {
  Den<R> D;
  bool miss = false;
  return fast_speed<false>(c, W, bn, D, false, miss);
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
// (mhd_common.hpp:426-431).  With S the recovery's density reciprocal is
// recorded in D, or (reuse) taken from it.
template <bool S, class R>
MHD_DEV bool positive_pressure(const Phys<R>& c, const R U[8], Den<R>& D, bool reuse,
                               bool& miss) {
  if (!(U[RHO] > c.rho_floor)) return false;
  R W[8];
  cons_to_prim<R, 8, S>(c, U, W, D, reuse, miss);
  return (W[0] > c.rho_floor && W[4] > c.p_floor);
}

// Scale-aware degeneracy test (mhd_common.hpp:417-419).
template <class R>
MHD_DEV bool hlld_small(const Phys<R>& c, R value, R scale) {
  return fabs_(value) <= c.tol * scale;
}

// HLLD star state next to an outer fast wave (mhd_common.hpp:445-511).
// With S, drs receives the reciprocal of the star density (recorded by the
// positivity check's recovery).
template <bool S, class R>
MHD_DEV int hlld_star(const Phys<R>& c, R e, const R W[8], R s, R sm,
                      R pt_star, R bx, R Us[8], Den<R>& drs, bool& miss) {
  const R rho = smax(W[0], c.rho_floor);
  const R vx = W[1], vy = W[2], vz = W[3];
  const R p = smax(W[4], c.p_floor);
  const R by = W[6], bz = W[7];
  const R s_minus_vx = s - vx;
  const R s_minus_sm = s - sm;
  const R wave_scale = smax(fabs_(s), fabs_(sm));
  if (hlld_small(c, s_minus_sm, wave_scale)) return HLLD_DEGENERATE;
  Den<R> dsm, dd;
  const R rho_star = Div<S, R>::rec(rho * s_minus_vx, s_minus_sm, dsm);
  if (rho_star <= c.rho_floor) return HLLD_NONPHYSICAL;
  const R denom = rho * s_minus_vx * s_minus_sm - bx * bx;
  const R denom_scale = fabs_(rho * s_minus_vx * s_minus_sm) + bx * bx;
  if (hlld_small(c, denom, denom_scale)) return HLLD_DEGENERATE;
  const R common = sm - vx;
  const R vy_star = vy - Div<S, R>::rec(bx * by * common, denom, dd);
  const R vz_star = vz - Div<S, R>::use(bx * bz * common, denom, dd, miss);
  const R by_star = Div<S, R>::use(by * (rho * s_minus_vx * s_minus_vx - bx * bx), denom, dd, miss);
  const R bz_star = Div<S, R>::use(bz * (rho * s_minus_vx * s_minus_vx - bx * bx), denom, dd, miss);
  const R b2 = bx * bx + by * by + bz * bz;
  const R pt = p + (R)0.5 * b2;
  const R vdotb = vx * bx + vy * by + vz * bz;
  const R vsdotbs = sm * bx + vy_star * by_star + vz_star * bz_star;
  const R e_star = Div<S, R>::use(
      s_minus_vx * e - pt * vx + pt_star * sm + bx * (vdotb - vsdotbs), s_minus_sm, dsm, miss);
  Us[RHO] = rho_star;
  Us[MX] = rho_star * sm;
  Us[MY] = rho_star * vy_star;
  Us[MZ] = rho_star * vz_star;
  Us[EN] = e_star;
  Us[BX] = bx;
  Us[BY] = by_star;
  Us[BZ] = bz_star;
  return positive_pressure<S>(c, Us, drs, false, miss) ? HLLD_OK : HLLD_NONPHYSICAL;
}

// HLL average of the outer states (mhd_common.hpp:338-347).
template <bool S, class R>
MHD_DEV void hll_average(const R UL[], const R UR[], const R FL[8],
                         const R FR[8], R sL, R sR, R F[8], bool& miss) {
  if (sL >= (R)0.0) {
#pragma unroll
    for (int k = 0; k < 8; k++) F[k] = FL[k];
  } else if (sR <= (R)0.0) {
#pragma unroll
    for (int k = 0; k < 8; k++) F[k] = FR[k];
  } else {
    const R denom = sR - sL;
    Den<R> dd;
    F[0] = Div<S, R>::rec(sR * FL[0] - sL * FR[0] + sL * sR * (UR[0] - UL[0]), denom, dd);
#pragma unroll
    for (int k = 1; k < 8; k++)
      F[k] = Div<S, R>::use(sR * FL[k] - sL * FR[k] + sL * sR * (UR[k] - UL[k]), denom, dd,
                            miss);
  }
}

// x-normal HLL or HLLD flux of the 8 MHD slots of UL/UR (UL[BX], UR[BX]
// already replaced by the GLM normal field).  HLLD follows
// mhd_common.hpp:518-654; HLL follows mhd_common.hpp:323-348.  With S the
// divisions by one denominator share its reciprocal (Div); `miss` marks a
// lane whose result must be evaluated again with S = false.
template <class R, bool HLLD, bool S>
MHD_DEV int mhd_flux_x(const Phys<R>& c, const R UL[], const R UR[], R F[8], bool& miss) {
  R WL[8], WR[8], FL[8], FR[8];
  Den<R> drL, drR;  // reciprocals of the outer densities
  cons_to_prim<R, 8, S>(c, UL, WL, drL, false, miss);
  cons_to_prim<R, 8, S>(c, UR, WR, drR, false, miss);
  phys_flux_x(WL, UL[EN], FL);
  phys_flux_x(WR, UR[EN], FR);
  const R cfL = fast_speed<S>(c, WL, WL[5], drL, true, miss);
  const R cfR = fast_speed<S>(c, WR, WR[5], drR, true, miss);
  const R sL = smin(WL[1] - cfL, WR[1] - cfR);
  const R sR = smax(WL[1] + cfL, WR[1] + cfR);
  if (!HLLD || sL >= (R)0.0 || sR <= (R)0.0) {
    hll_average<S>(UL, UR, FL, FR, sL, sR, F, miss);
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
  Den<R> drLs, drRs;  // reciprocals of the star densities
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
    const int ls = hlld_star<S>(c, UL[EN], WL, sL, sm, pt_star, bx, ULs, drLs, miss);
    const int rs = hlld_star<S>(c, UR[EN], WR, sR, sm, pt_star, bx, URs, drRs, miss);
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
    hll_average<S>(UL, UR, FL, FR, sL, sR, F, miss);
    return status;
  }
  const R denom = sqrt_rhoLs + sqrt_rhoRs;
  const R vyLs = Div<S, R>::use(ULs[MY], ULs[RHO], drLs, miss);
  const R vzLs = Div<S, R>::use(ULs[MZ], ULs[RHO], drLs, miss);
  const R byLs = ULs[BY], bzLs = ULs[BZ];
  const R vyRs = Div<S, R>::use(URs[MY], URs[RHO], drRs, miss);
  const R vzRs = Div<S, R>::use(URs[MZ], URs[RHO], drRs, miss);
  const R byRs = URs[BY], bzRs = URs[BZ];
  Den<R> dd;
  const R vyD =
      Div<S, R>::rec(sqrt_rhoLs * vyLs + sqrt_rhoRs * vyRs + sgn_bx * (byRs - byLs), denom, dd);
  const R vzD =
      Div<S, R>::use(sqrt_rhoLs * vzLs + sqrt_rhoRs * vzRs + sgn_bx * (bzRs - bzLs), denom, dd,
                     miss);
  const R byD = Div<S, R>::use(sqrt_rhoLs * byRs + sqrt_rhoRs * byLs +
                                   sgn_bx * sqrt_rhoLs * sqrt_rhoRs * (vyRs - vyLs),
                               denom, dd, miss);
  const R bzD = Div<S, R>::use(sqrt_rhoLs * bzRs + sqrt_rhoRs * bzLs +
                                   sgn_bx * sqrt_rhoLs * sqrt_rhoRs * (vzRs - vzLs),
                               denom, dd, miss);
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
  if (!positive_pressure<S>(c, ULD, drLs, true, miss) ||
      !positive_pressure<S>(c, URD, drRs, true, miss)) {
    hll_average<S>(UL, UR, FL, FR, sL, sR, F, miss);
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
// With S the divisions share reciprocals and `miss` marks a lane to be
// evaluated again with S = false (Div).
template <class R, bool HLLD, bool S = false>
MHD_DEV int glm_flux_x(const Phys<R>& c, R UL[9], R UR[9], R ch, R F[9], bool& miss) {
  const R ch_safe = smax(ch, (R)1.0e-12);
  const R bnStar =
      (R)0.5 * (UL[BX] + UR[BX]) - (R)0.5 * (UR[PSI] - UL[PSI]) / ch_safe;
  const R psiStar =
      (R)0.5 * (UL[PSI] + UR[PSI]) - (R)0.5 * ch_safe * (UR[BX] - UL[BX]);
  UL[BX] = bnStar;
  UR[BX] = bnStar;
  const int status = mhd_flux_x<R, HLLD, S>(c, UL, UR, F, miss);
  F[BX] = psiStar;
  F[PSI] = ch * ch * bnStar;
  return status;
}
template <class R, bool HLLD>
MHD_DEV int glm_flux_x(const Phys<R>& c, R UL[9], R UR[9], R ch, R F[9]) {
  bool miss = false;
  return glm_flux_x<R, HLLD, false>(c, UL, UR, ch, F, miss);
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

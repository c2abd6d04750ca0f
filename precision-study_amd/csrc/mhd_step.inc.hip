// mhd_step.inc.hip -- CDNA4 (gfx950) kernels of the 2D MHD-GLM SSPRK2 step.
//
// Included by three translation units, each compiled with its own floating
// point flags and its own namespace MHD_NS:
//   k_exact.hip  contraction off, IEEE div/sqrt  -> FP64 and FP32, bitwise
//                the reference CPU build
//   k_fma.hip    FMA contraction on               -> FP64 (relL2 <= 1e-12 bar)
//   k_mixed.hip  contraction on, fast FP32 div/sqrt -> mixed precision forms
//
// One SSPRK2 step (mhd2d_glm_cpu.cpp:513-569) is three launches on one
// stream, with no host synchronisation:
//   k_stage<1>: U1 = U + dt L(U)                        + positivity of U1
//   k_stage<2>: U  = U/2 + (U1 + dt L(U1))/2 (in place) + positivity of U
//               + the max signal speed of the new U (next step's c_h)
//   k_finish  : t += dt, step++; next c_h and dt, or halt with the
//               reference's exit code (3: stage 1, 4: stage 2, 2: rate)
// Each stage kernel is one fused tile pass: a TX x TY cell tile plus a
// 2-cell halo is read once from HBM (SoA planes, coalesced rows), primitives
// are recovered once per cell into LDS, every x and y face of the tile is
// reconstructed (minmod MUSCL) and gets its GLM-HLLD flux from one merged,
// direction-generic face list, fluxes go to LDS, and each cell is updated
// from its four faces.  Only the state itself crosses HBM.
//
// A precision policy fixes the types: R (LDS primitives, reconstruction,
// fluxes), S (stored state element), A (rhs and update arithmetic) and T
// (time, dt, c_h: the control block).
//
// Kernels stop at entry once the device control block says the run is
// done or halted, so a batch of steps (or a replayed graph) can be launched
// without knowing in advance where t_end or a failure falls.
#include "mhd_device.hpp"
#include "mhd_kernels.h"

#include <climits>
#include <type_traits>

#ifndef MHD_NS
#error "MHD_NS must name the translation unit's namespace"
#endif

namespace mhd {
namespace MHD_NS {

// ---------------------------------------------------------------------------
// Wave reductions: DPP within each 16-lane row, readlane across the 4 rows.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned dpp_quad_xor1(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ unsigned dpp_quad_xor2(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
}
__device__ __forceinline__ unsigned dpp_row_ror4(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false);
}
__device__ __forceinline__ unsigned dpp_row_ror8(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
}

// Sum of a 32-bit value over the 64 lanes (all lanes active).
__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
  v += dpp_quad_xor1(v);
  v += dpp_quad_xor2(v);
  v += dpp_row_ror4(v);
  v += dpp_row_ror8(v);
  return (unsigned)__builtin_amdgcn_readlane((int)v, 0) +
         (unsigned)__builtin_amdgcn_readlane((int)v, 16) +
         (unsigned)__builtin_amdgcn_readlane((int)v, 32) +
         (unsigned)__builtin_amdgcn_readlane((int)v, 48);
}

__device__ __forceinline__ unsigned long long u64max(unsigned long long a,
                                                     unsigned long long b) {
  return a < b ? b : a;
}

// Max of a 64-bit unsigned value over the 64 lanes (all lanes active).
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#define MHD_DPP_MAX_STEP(OP)                                               \
  {                                                                        \
    unsigned lo = OP((unsigned)v), hi = OP((unsigned)(v >> 32));           \
    v = u64max(v, ((unsigned long long)hi << 32) | lo);                    \
  }
  MHD_DPP_MAX_STEP(dpp_quad_xor1)
  MHD_DPP_MAX_STEP(dpp_quad_xor2)
  MHD_DPP_MAX_STEP(dpp_row_ror4)
  MHD_DPP_MAX_STEP(dpp_row_ror8)
#undef MHD_DPP_MAX_STEP
  unsigned long long r = 0;
#pragma unroll
  for (int row = 0; row < 4; row++) {
    unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, row * 16);
    unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), row * 16);
    r = u64max(r, ((unsigned long long)hi << 32) | lo);
  }
  return r;
}

// Bits of a non-negative value, ordered like the value.
__device__ __forceinline__ unsigned long long speed_bits(double v) {
  return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ unsigned long long speed_bits(float v) {
  return (unsigned long long)(unsigned)__float_as_uint(v);
}
template <class T> __device__ __forceinline__ T bits_speed(unsigned long long b);
template <> __device__ __forceinline__ double bits_speed<double>(unsigned long long b) {
  return __longlong_as_double((long long)b);
}
template <> __device__ __forceinline__ float bits_speed<float>(unsigned long long b) {
  return __uint_as_float((unsigned)b);
}

// Wave-uniform copies (readfirstlane): values loaded from memory that the
// compiler cannot prove unclobbered (inside the persistent kernels' loops)
// would otherwise occupy vector registers.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned long long uni(unsigned long long v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ size_t uni_size(size_t v) { return (size_t)uni((unsigned long long)v); }
__device__ __forceinline__ double uni(double v) {
  return __longlong_as_double((long long)uni((unsigned long long)__double_as_longlong(v)));
}
__device__ __forceinline__ float uni(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
template <class R>
__device__ __forceinline__ Phys<R> make_phys(double tol) {
  Phys<R> c;
  c.gamma = (R)(5.0 / 3.0);
  c.gm1 = c.gamma - (R)1.0;
  c.rho_floor = (R)1.0e-12;
  c.p_floor = (R)1.0e-12;
  c.tol = (R)tol;
  return c;
}

// Per-cell contribution to the signal-speed maximum of
// maxPhysicalSignalSpeed (mhd2d_glm_cpu.cpp:289-305): max(sx, sy) in the
// std::max form.  A NaN never enters the reference's running maximum and
// the running maximum starts at +0, so NaN and zero both contribute +0
// here, and `valid` records whether any cell was not NaN (the reference
// then reports a non-positive rate).
template <class T>
__device__ __forceinline__ T cell_speed(const Phys<T>& c, const T P[9], bool& valid) {
  const T sx = fabs_(P[1]) + fast_speed(c, P, P[5]);
  const T sy = fabs_(P[2]) + fast_speed(c, P, P[6]);
  const T m = smax(sx, sy);
  valid = (m == m);
  return (m > (T)0.0) ? m : (T)0.0;
}

// FP32 stage functions of the reference's stage-level mixed build
// (mhd2d_glm_gpu_mixed.cu:54-125), with its fmaxf floors.
namespace refmix {
constexpr float kRhoFloor = 1.0e-12f, kPFloor = 1.0e-12f;
constexpr float kGamma = (float)(5.0 / 3.0);

// f32ConsToPrim9 (mhd2d_glm_gpu_mixed.cu:61-74)
__device__ __forceinline__ void cons_to_prim(const float U[9], float P[9]) {
  const float rho = fmaxf(U[RHO], kRhoFloor);
  const float vx = U[MX] / rho, vy = U[MY] / rho, vz = U[MZ] / rho;
  const float bx = U[BX], by = U[BY], bz = U[BZ];
  const float v2 = vx * vx + vy * vy + vz * vz;
  const float b2 = bx * bx + by * by + bz * bz;
  const float p = (kGamma - 1.0f) * (U[EN] - 0.5f * rho * v2 - 0.5f * b2);
  P[0] = rho;
  P[1] = vx;
  P[2] = vy;
  P[3] = vz;
  P[4] = p;
  P[5] = bx;
  P[6] = by;
  P[7] = bz;
  P[8] = U[PSI];
}

// f32PrimToCons9 (mhd2d_glm_gpu_mixed.cu:75-87)
__device__ __forceinline__ void prim_to_cons(const float P[9], float U[9]) {
  const float rho = P[0], vx = P[1], vy = P[2], vz = P[3];
  const float bx = P[5], by = P[6], bz = P[7];
  U[RHO] = rho;
  U[MX] = rho * vx;
  U[MY] = rho * vy;
  U[MZ] = rho * vz;
  U[EN] = P[4] / (kGamma - 1.0f) + 0.5f * rho * (vx * vx + vy * vy + vz * vz) +
          0.5f * (bx * bx + by * by + bz * bz);
  U[BX] = bx;
  U[BY] = by;
  U[BZ] = bz;
  U[PSI] = P[8];
}

// f32FastSpeed (mhd2d_glm_gpu_mixed.cu:88-99)
__device__ __forceinline__ float fast_speed(const float P[9], float bn) {
  const float rho = fmaxf(P[0], kRhoFloor);
  const float p = fmaxf(P[4], kPFloor);
  const float bx = P[5], by = P[6], bz = P[7];
  const float a2 = kGamma * p / rho;
  const float b2 = (bx * bx + by * by + bz * bz) / rho;
  const float can2 = bn * bn / rho;
  const float s = a2 + b2;
  const float disc = fmaxf(s * s - 4.0f * a2 * can2, 0.0f);
  return sqrtf(0.5f * (s + sqrtf(disc)));
}

// f32LoadPrim (mhd2d_glm_gpu_mixed.cu:121-125)
__device__ __forceinline__ void load_prim(const double U9[9], float P[9]) {
  float Uf[9];
#pragma unroll
  for (int k = 0; k < 9; k++) Uf[k] = (float)U9[k];
  cons_to_prim(Uf, P);
}

// Signal speed of one cell as speedKernel takes it
// (mhd2d_glm_gpu_mixed.cu:256-280): (double)fmaxf(sx, sy).
__device__ __forceinline__ double cell_speed(const double U9[9], bool& valid) {
  float P[9];
  load_prim(U9, P);
  const float sx = fabsf(P[1]) + fast_speed(P, P[5]);
  const float sy = fabsf(P[2]) + fast_speed(P, P[6]);
  const float m = fmaxf(sx, sy);
  valid = (m == m);
  return (m > 0.0f) ? (double)m : 0.0;
}
}  // namespace refmix

template <class P>
__device__ __forceinline__ P* uni_ptr(P* p) {
  return (P*)uni((unsigned long long)(uintptr_t)p);
}

__device__ __forceinline__ int wrap(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// Bijective block -> tile remap so that the tiles one XCD receives (blocks
// are dealt round-robin over the 8 XCDs) form one contiguous band of tile
// rows, and halo re-reads between neighbouring tiles hit that XCD's L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7, x = bid & 7;
  return x * q + (x < r ? x : r) + (bid >> 3);
}


// Publishes a block's signal-speed maximum: one slot per block residue, and
// the atomic only when the slot does not already hold a larger value (the
// maximum settles after the first blocks, so most blocks skip the atomic).
template <class T>
__device__ __forceinline__ void publish_speed(Ctrl<T>* ctrl, unsigned long long m, bool valid) {
  unsigned long long* slot = &ctrl->speed_slot[blockIdx.x % Ctrl<T>::kSpeedSlots];
  if (m > __atomic_load_n(slot, __ATOMIC_RELAXED)) atomicMax(slot, m);
  if (valid && __atomic_load_n(&ctrl->valid_acc, __ATOMIC_RELAXED) == 0)
    atomicOr(&ctrl->valid_acc, 1);
}

// Folds the slots into speed_acc and clears them (single thread).
template <class T>
__device__ __forceinline__ void fold_speed_slots(Ctrl<T>* ctrl) {
  unsigned long long m = ctrl->speed_acc;
  for (int q = 0; q < Ctrl<T>::kSpeedSlots; q++) {
    m = u64max(m, ctrl->speed_slot[q]);
    ctrl->speed_slot[q] = 0;
  }
  ctrl->speed_acc = m;
}

// ---------------------------------------------------------------------------
// FP64 checks and signal speed of the mixed forms.  These need FP64-level
// accuracy (c_h sets dt, which every cell sees), not the reference's
// operation sequence: one reciprocal of rho shared by all quotients and
// square roots from v_rsq_f64 with one Newton correction (~1 ulp).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double s0 = x * y;
  const double s = s0 + (0.5 * y) * (x - s0 * s0);
  return (x > 0.0 && x < __builtin_inf()) ? s : x;
}

// Usability (rho > floor, p > floor, all finite) and, when wanted, the
// cell's signal speed max(|vx|+cf_x, |vy|+cf_y), of an FP64 state.
__device__ __forceinline__ bool usable_and_speed_fp64(const double U[9], bool want_speed,
                                                      double& speed, bool& valid) {
  const double gamma = 5.0 / 3.0, gm1 = gamma - 1.0, floor_ = 1.0e-12;
  const double rho = smax(U[RHO], floor_);
  const double inv = 1.0 / rho;
  const double vx = U[MX] * inv, vy = U[MY] * inv, vz = U[MZ] * inv;
  const double bx = U[BX], by = U[BY], bz = U[BZ];
  const double v2 = vx * vx + vy * vy + vz * vz, b2 = bx * bx + by * by + bz * bz;
  const double p = gm1 * (U[EN] - 0.5 * rho * v2 - 0.5 * b2);
  bool ok = U[RHO] > floor_ && p > floor_ && finite_(p) && finite_(v2);
#pragma unroll
  for (int k = 0; k < 9; k++) ok = ok && finite_(U[k]);
  speed = 0.0;
  valid = false;
  if (want_speed) {
    const double a2 = gamma * smax(p, floor_) * inv, bb = b2 * inv, s2 = a2 + bb;
    const double dxn = smax(s2 * s2 - 4.0 * a2 * (bx * bx * inv), 0.0);
    const double dyn = smax(s2 * s2 - 4.0 * a2 * (by * by * inv), 0.0);
    const double cfx = sqrt_nr(smax(0.5 * (s2 + sqrt_nr(dxn)), 0.0));
    const double cfy = sqrt_nr(smax(0.5 * (s2 + sqrt_nr(dyn)), 0.0));
    const double m = smax(fabs_(vx) + cfx, fabs_(vy) + cfy);
    valid = (m == m);
    speed = (m > 0.0) ? m : 0.0;
  }
  return ok;
}

// ---------------------------------------------------------------------------
// Float-float (unevaluated hi + lo) accumulation for the compensated form.
// Only additions: unaffected by FMA contraction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
// (hi, lo) + d
__device__ __forceinline__ void ff_add_f(float hi, float lo, float d, float& rh, float& rl) {
  float s, e;
  two_sum(hi, d, s, e);
  e += lo;
  rh = s + e;
  rl = e - (rh - s);
}
// (ah, al) + (bh, bl)
__device__ __forceinline__ void ff_add(float ah, float al, float bh, float bl, float& rh,
                                       float& rl) {
  float s, e;
  two_sum(ah, bh, s, e);
  e += al + bl;
  rh = s + e;
  rl = e - (rh - s);
}

// ---------------------------------------------------------------------------
// Precision policies.
// ---------------------------------------------------------------------------
// Plain FP64 / FP32: everything in R, state stored in R.
template <class R_>
struct ExactPol {
  using R = R_;   // Riemann flux
  using RP = R_;  // primitives / reconstruction (LDS)
  using RS = R_;  // face fluxes as stored in LDS
  using T = R_;   // time, dt, c_h, update
  using S = R_;   // state in HBM
  static constexpr int kPlanes = 9;
  static constexpr int kMode = sizeof(R_) == 8 ? 0 : 1;
  static constexpr bool kRedo = false;
  static constexpr bool kRefMixed = false;
};
// The reference's stage-level mixed build (mhd2d_glm_gpu_mixed.cu:1-10):
// state, Riemann (GLM-HLLD/HLL) fluxes and the SSPRK2 updates in FP64;
// recovery, CFL reductions, face reconstruction, the rhs flux differences
// and the psi damping in FP32 (each stage's inputs cast to float, results
// cast back).
struct RefMixPol {
  using R = double;
  using RP = float;
  using RS = float;  // the rhs reads (float)F (mhd2d_glm_gpu_mixed.cu:355-375)
  using T = double;
  using S = double;
  static constexpr int kPlanes = 9;
  static constexpr bool kRedo = false;
  static constexpr bool kRefMixed = true;
};
// Mixed: FP32 reconstruction / HLLD / GLM; FP64 state, update accumulation,
// time; FP64 pressure recovery where FP32 would cancel.
struct MixedPol {
  using R = float;
  using RP = float;
  using RS = float;
  using T = double;
  using S = double;
  static constexpr int kPlanes = 9;
  static constexpr bool kRedo = true;  // FP64 fallback of unusable tiles
  static constexpr bool kRefMixed = false;
};
// Allocation variants of the mixed form, for the study of where its error
// comes from: the mixed form with one of its FP32 parts in FP64 (RP: the
// recovered primitives and the reconstruction, FP64 recovery in every cell
// when double; R: the GLM-HLLD/HLL flux; RS: the face fluxes as the rhs
// differences read them).  No FP64 fallback; MUSCL + HLLD, 16x15 tiles.
template <class RP_, class R_, class RS_>
struct MixVarPol {
  using R = R_;
  using RP = RP_;
  using RS = RS_;
  using T = double;
  using S = double;
  static constexpr int kPlanes = 9;
  static constexpr bool kRedo = false;
  static constexpr bool kRefMixed = false;
};
// Compensated: as Mixed, but the state is FP32 hi + FP32 lo (planes 0-8 and
// 9-17), accumulated with float-float additions.
struct KahanPol {
  using R = float;
  using T = double;
  using RP = float;
  using RS = float;
  using S = float;
  static constexpr int kPlanes = 18;
  static constexpr bool kRedo = false;
  static constexpr bool kRefMixed = false;
};

template <class Pol>
struct StageArgs {
  const typename Pol::S* src;   // state L() is evaluated on: U (stage 1) or U1 (stage 2)
  const typename Pol::S* uold;  // stage 2: U^n (same buffer as dst)
  typename Pol::S* dst;         // stage 1: U1; stage 2: U
  size_t plane;
  int nx, ny, tiles_x, ntiles;
  int pow2;                 // dx and dy are powers of two: x/dx == x*(1/dx) exactly
  double inv_dx, inv_dy;
  double tol;
  double alpha;
  float lowbeta_ratio;      // mixed: FP64 recovery where E > ratio * e_internal
  Ctrl<typename Pol::T>* ctrl;
  const int* tiles;            // block b evaluates tiles[b], -1: none (null: every tile)
  // FP64 fallback (mixed policy): a tile with an unusable cell appends
  // redo_base + tile at redo_list[atomicAdd(redo_counter, 1)].
  int* redo_list;              // null: no fallback
  int* redo_counter;
  int redo_base;
  int redo_all;                // queue every tile (check mode)
};

// Buffers that follow a device word (two-stage steps, whose state
// alternates between two buffers): with *cur non-zero the stage reads and
// writes the *_alt buffers instead of StageArgs' own.
template <class Pol>
struct CurBuffers {
  const int* cur;  // null: StageArgs' buffers
  const typename Pol::S* src_alt;
  const typename Pol::S* uold_alt;
  typename Pol::S* dst_alt;
};

template <class Pol>
__device__ __forceinline__ void follow_cur(StageArgs<Pol>& a, const CurBuffers<Pol>& b) {
  if (b.cur && uni(*b.cur) != 0) {
    a.src = b.src_alt;
    a.uold = b.uold_alt;
    a.dst = b.dst_alt;
  }
}

// Cell state as loaded from HBM, widened to double where the policy stores
// more than the flux type.
template <class Pol> struct CellIO;

template <class R>
struct CellIO<ExactPol<R>> {
  R v[9];
  __device__ __forceinline__ void load(const R* p, size_t ps, size_t g) {
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = p[k * ps + g];
  }
  __device__ __forceinline__ R psi() const { return v[PSI]; }
  __device__ __forceinline__ void store(R* p, size_t ps, size_t g) const {
#pragma unroll
    for (int k = 0; k < 9; k++) p[k * ps + g] = v[k];
  }
};

template <>
struct CellIO<RefMixPol> {
  double v[9];
  __device__ __forceinline__ void load(const double* p, size_t ps, size_t g) {
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = p[k * ps + g];
  }
  __device__ __forceinline__ double psi() const { return v[PSI]; }
  __device__ __forceinline__ void store(double* p, size_t ps, size_t g) const {
#pragma unroll
    for (int k = 0; k < 9; k++) p[k * ps + g] = v[k];
  }
};

template <>
struct CellIO<MixedPol> {
  double v[9];
  __device__ __forceinline__ void load(const double* p, size_t ps, size_t g) {
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = p[k * ps + g];
  }
  __device__ __forceinline__ void wide(double out[9]) const {
#pragma unroll
    for (int k = 0; k < 9; k++) out[k] = v[k];
  }
  __device__ __forceinline__ void store(double* p, size_t ps, size_t g) const {
#pragma unroll
    for (int k = 0; k < 9; k++) p[k * ps + g] = v[k];
  }
};

template <class RP, class R, class RS>
struct CellIO<MixVarPol<RP, R, RS>> : CellIO<MixedPol> {};

template <>
struct CellIO<KahanPol> {
  float hi[9], lo[9];
  __device__ __forceinline__ void load(const float* p, size_t ps, size_t g) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
      hi[k] = p[k * ps + g];
      lo[k] = p[(9 + k) * ps + g];
    }
  }
  __device__ __forceinline__ void wide(double out[9]) const {
#pragma unroll
    for (int k = 0; k < 9; k++) out[k] = (double)hi[k] + (double)lo[k];
  }
  __device__ __forceinline__ void store(float* p, size_t ps, size_t g) const {
#pragma unroll
    for (int k = 0; k < 9; k++) {
      p[k * ps + g] = hi[k];
      p[(9 + k) * ps + g] = lo[k];
    }
  }
};

// Phase A recovery: primitives in the flux type.
template <bool MUSCL, class R>
__device__ __forceinline__ void recover(const Phys<R>& c, const CellIO<ExactPol<R>>& u,
                                        float, R P[9]) {
  if (MUSCL) {
    cons_to_prim<R, 9>(c, u.v, P);
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) P[k] = u.v[k];
  }
}

// Mixed recovery: FP32 unless the internal energy is a small remainder of E
// (low plasma beta or high Mach number: E > ratio * e_int, or e_int <= 0),
// where p = (gamma-1)(E - rho v^2/2 - B^2/2) is recomputed in FP64.
template <class Pol, bool MUSCL>
__device__ __forceinline__ void recover_mixed(const Phys<float>& c, const CellIO<Pol>& u,
                                              float ratio, float P[9]) {
  double d[9];
  u.wide(d);
  float f[9];
#pragma unroll
  for (int k = 0; k < 9; k++) f[k] = (float)d[k];
  if (!MUSCL) {
#pragma unroll
    for (int k = 0; k < 9; k++) P[k] = f[k];
    return;
  }
  cons_to_prim<float, 9>(c, f, P);
  if (!(P[4] * ratio > c.gm1 * f[EN])) {
    const Phys<double> cd = make_phys<double>(0.0);
    double W[9];
    cons_to_prim<double, 9>(cd, d, W);
#pragma unroll
    for (int k = 0; k < 9; k++) P[k] = (float)W[k];
  }
}

// Reference-mixed recovery: the FP64 state cast to float, f32ConsToPrim9
// (mhd2d_glm_gpu_mixed.cu:121-125).  First order is not offered by this
// policy (the reference keeps FP64 face states there).
template <bool MUSCL>
__device__ __forceinline__ void recover(const Phys<float>&, const CellIO<RefMixPol>& u, float,
                                        float P[9]) {
  refmix::load_prim(u.v, P);
}

// Face state from a reconstructed primitive: prim->cons in RP, widened to
// the flux type.
template <class Pol, class RP, class R>
__device__ __forceinline__ void face_cons(const Phys<RP>& c, const RP Pf[9], R U[9]) {
  RP Uf[9];
  if constexpr (Pol::kRefMixed)
    refmix::prim_to_cons(Pf, Uf);  // f32PrimToCons9, then (Real) (:194-200)
  else
    prim_to_cons(c, Pf, Uf);
#pragma unroll
  for (int k = 0; k < 9; k++) U[k] = (R)Uf[k];
}

template <bool MUSCL>
__device__ __forceinline__ void recover(const Phys<float>& c, const CellIO<MixedPol>& u,
                                        float ratio, float P[9]) {
  recover_mixed<MixedPol, MUSCL>(c, u, ratio, P);
}
template <bool MUSCL>
__device__ __forceinline__ void recover(const Phys<float>& c, const CellIO<KahanPol>& u,
                                        float ratio, float P[9]) {
  recover_mixed<KahanPol, MUSCL>(c, u, ratio, P);
}
// Allocation variants: the mixed recovery, or FP64 recovery in every cell
// when the primitives are FP64.
template <bool MUSCL, class RP, class R, class RS>
__device__ __forceinline__ void recover(const Phys<RP>& c, const CellIO<MixVarPol<RP, R, RS>>& u,
                                        float ratio, RP P[9]) {
  if constexpr (sizeof(RP) == 4) {
    recover_mixed<MixVarPol<RP, R, RS>, MUSCL>(c, u, ratio, P);
  } else if (MUSCL) {
    cons_to_prim<double, 9>(c, u.v, P);
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) P[k] = u.v[k];
  }
}

// Face-flux differences of one cell: -(F_r - F_l)/dx - (G_t - G_b)/dy
// (mhd2d_glm_cpu.cpp:278-281) in type A.
template <class A, class R>
__device__ __forceinline__ A flux_div(R fr, R fl, R gt, R gb, bool pow2, A dx, A dy, A idx,
                                      A idy) {
  if (pow2) return -((A)fr - (A)fl) * idx - ((A)gt - (A)gb) * idy;
  return -((A)fr - (A)fl) / dx - ((A)gt - (A)gb) / dy;
}

// Phase C update of one cell: the stage updates of mhd2d_glm_cpu.cpp:528-535
// (stage 1: U1 = U + dt rhs) and 549-557 (stage 2: U/2 + (U1 + dt rhs)/2),
// from the stage source `src`, U^n (`old`, stage 2) and the rhs flux
// differences.  `nu` is the new state as stored, `out` the same in T for the
// positivity and speed checks.
//
// Exact policies: everything in R.
template <int STAGE, class R>
__device__ __forceinline__ void advance_cell(const CellIO<ExactPol<R>>& src,
                                             const CellIO<ExactPol<R>>& old, const R rhs_flux[9],
                                             R dt, R damping, CellIO<ExactPol<R>>& nu, R out[9]) {
#pragma unroll
  for (int k = 0; k < 9; k++) {
    R rhs = rhs_flux[k];
    if (k == PSI) rhs -= damping * src.psi();
    if (STAGE == 1)
      out[k] = src.v[k] + dt * rhs;
    else
      out[k] = (R)0.5 * old.v[k] + (R)0.5 * (src.v[k] + dt * rhs);
    nu.v[k] = out[k];
  }
}

// Reference-mixed policy: rhs in FP32 with the psi damping in FP32
// (rhsKernel, mhd2d_glm_gpu_mixed.cu:355-375), the stage update in FP64
// (stage1Kernel/stage2Kernel, :377-410).
template <int STAGE>
__device__ __forceinline__ void advance_cell(const CellIO<RefMixPol>& src,
                                             const CellIO<RefMixPol>& old, const float rhs_flux[9],
                                             double dt, double damping, CellIO<RefMixPol>& nu,
                                             double out[9]) {
#pragma unroll
  for (int k = 0; k < 9; k++) {
    float rhs = rhs_flux[k];
    if (k == PSI) rhs = rhs - (float)damping * (float)src.psi();
    if (STAGE == 1)
      out[k] = src.v[k] + dt * (double)rhs;
    else
      out[k] = 0.5 * old.v[k] + 0.5 * (src.v[k] + dt * (double)rhs);
    nu.v[k] = out[k];
  }
}

// Mixed policy: FP64 rhs accumulation and update.
template <int STAGE>
__device__ __forceinline__ void advance_cell(const CellIO<MixedPol>& src,
                                             const CellIO<MixedPol>& old, const double rhs_flux[9],
                                             double dt, double damping, CellIO<MixedPol>& nu,
                                             double out[9]) {
#pragma unroll
  for (int k = 0; k < 9; k++) {
    double rhs = rhs_flux[k];
    if (k == PSI) rhs -= damping * src.v[PSI];
    if (STAGE == 1)
      out[k] = src.v[k] + dt * rhs;
    else
      out[k] = 0.5 * old.v[k] + 0.5 * (src.v[k] + dt * rhs);
    nu.v[k] = out[k];
  }
}

// Allocation variants: as the mixed policy.
template <int STAGE, class RP, class R, class RS>
__device__ __forceinline__ void advance_cell(const CellIO<MixVarPol<RP, R, RS>>& src,
                                             const CellIO<MixVarPol<RP, R, RS>>& old,
                                             const double rhs_flux[9], double dt, double damping,
                                             CellIO<MixVarPol<RP, R, RS>>& nu, double out[9]) {
#pragma unroll
  for (int k = 0; k < 9; k++) {
    double rhs = rhs_flux[k];
    if (k == PSI) rhs -= damping * src.v[PSI];
    if (STAGE == 1)
      out[k] = src.v[k] + dt * rhs;
    else
      out[k] = 0.5 * old.v[k] + 0.5 * (src.v[k] + dt * rhs);
    nu.v[k] = out[k];
  }
}

// Compensated policy: FP32 rhs, float-float accumulation of the update.
template <int STAGE>
__device__ __forceinline__ void advance_cell(const CellIO<KahanPol>& src,
                                             const CellIO<KahanPol>& old, const float rhs_flux[9],
                                             double dt, double damping, CellIO<KahanPol>& nu,
                                             double out[9]) {
  const float dtf = (float)dt, dampf = (float)damping;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    float rhs = rhs_flux[k];
    if (k == PSI) rhs -= dampf * src.hi[PSI];
    const float d = dtf * rhs;
    float h, l;
    ff_add_f(src.hi[k], src.lo[k], d, h, l);
    if (STAGE == 2) ff_add(0.5f * old.hi[k], 0.5f * old.lo[k], 0.5f * h, 0.5f * l, h, l);
    nu.hi[k] = h;
    nu.lo[k] = l;
    out[k] = (double)h + (double)l;
  }
}

// rhs, stage update and (but for the mixed policy, whose stage_tile stores
// the new state once the tile has passed the positivity check, for the FP64
// fallback) store of one cell.
template <int STAGE, class Pol, class A, class T>
__device__ __forceinline__ void update_cell(const StageArgs<Pol>& a, const CellIO<Pol>& src,
                                            size_t g, const A rhs_flux[9], T dt, T damping,
                                            T out[9]) {
  CellIO<Pol> old, nu;
  if (STAGE == 2) old.load(a.uold, a.plane, g);
  advance_cell<STAGE>(src, old, rhs_flux, dt, damping, nu, out);
  if constexpr (!Pol::kRedo) nu.store(a.dst, a.plane, g);
}

// rhs type of each policy.
template <class Pol> struct RhsType { using type = typename Pol::T; };
template <> struct RhsType<KahanPol> { using type = float; };
template <> struct RhsType<RefMixPol> { using type = float; };

// One face of a tile pass: the 4-cell stencil at base + {0,1,2,3} * stride in
// the LDS primitive planes (PN elements apart), reconstructed (minmod MUSCL)
// or taken first order, its GLM-HLLD flux stored at `slot` of the flux
// planes (FN apart).  A y face is evaluated in the x frame by reading the
// (1,2) and (5,6) slots swapped.  Returns the HLLD status.
template <class Pol, bool MUSCL, bool HLLD, int PN, int FN>
__device__ __forceinline__ int face_flux(const Phys<typename Pol::R>& c,
                                         const Phys<typename Pol::RP>& cp, typename Pol::R ch,
                                         const typename Pol::RP* sP, int base, int stride,
                                         bool ydir, typename Pol::RS* sF, int slot) {
  using R = typename Pol::R;
  using RP = typename Pol::RP;
  using RS = typename Pol::RS;
  const int q1 = (ydir ? 2 : 1) * PN, q2 = (ydir ? 1 : 2) * PN;
  const int q5 = (ydir ? 6 : 5) * PN, q6 = (ydir ? 5 : 6) * PN;
  auto load = [&](int cell, RP out[9]) {
    out[0] = sP[cell];
    out[1] = sP[q1 + cell];
    out[2] = sP[q2 + cell];
    out[3] = sP[3 * PN + cell];
    out[4] = sP[4 * PN + cell];
    out[5] = sP[q5 + cell];
    out[6] = sP[q6 + cell];
    out[7] = sP[7 * PN + cell];
    out[8] = sP[8 * PN + cell];
  };
  // One evaluation of the face: states in RP, widened to the flux type R
  // (a no-op but for the reference-mixed policy, whose face states are
  // FP32 and flux FP64), the flux, its store to LDS.  With sharing (FP64:
  // divisions by one denominator share its reciprocal) a lane may come
  // back with `miss` (an operand range where the sharing does not hold);
  // the face is then evaluated again from LDS with plain division.  Only
  // the face index and the status live across, so the second evaluation
  // adds no register pressure.
  auto evaluate = [&](auto share, bool& miss) -> int {
    R UL[9], UR[9];
    if (MUSCL) {
      RP Pa[9], Pb[9], Pc[9], Pf[9];
      load(base, Pa);
      load(base + stride, Pb);
      load(base + 2 * stride, Pc);
      reconstruct_face<RP, true>(cp, Pa, Pb, Pc, Pf);
      face_cons<Pol>(cp, Pf, UL);
      load(base + 3 * stride, Pa);
      reconstruct_face<RP, false>(cp, Pb, Pc, Pa, Pf);
      face_cons<Pol>(cp, Pf, UR);
    } else {
      RP Ua[9], Ub[9];
      load(base + stride, Ua);
      load(base + 2 * stride, Ub);
#pragma unroll
      for (int k = 0; k < 9; k++) {
        UL[k] = (R)Ua[k];
        UR[k] = (R)Ub[k];
      }
    }
    R F[9];
    const int st = glm_flux_x<R, HLLD, decltype(share)::value>(c, UL, UR, ch, F, miss);
    const int p1 = (ydir ? 2 : 1) * FN, p2 = (ydir ? 1 : 2) * FN;
    const int p5 = (ydir ? 6 : 5) * FN, p6 = (ydir ? 5 : 6) * FN;
    sF[slot] = (RS)F[0];
    sF[p1 + slot] = (RS)F[1];
    sF[p2 + slot] = (RS)F[2];
    sF[3 * FN + slot] = (RS)F[3];
    sF[4 * FN + slot] = (RS)F[4];
    sF[p5 + slot] = (RS)F[5];
    sF[p6 + slot] = (RS)F[6];
    sF[7 * FN + slot] = (RS)F[7];
    sF[8 * FN + slot] = (RS)F[8];
    return st;
  };
  bool miss = false;
  int st = evaluate(std::integral_constant<bool, sizeof(R) == 8>{}, miss);
  if (miss) {
    bool unused = false;
    st = evaluate(std::integral_constant<bool, false>{}, unused);
  }
  return st;
}

// LDS footprint of one tile: primitives (tile + halo) in RP, face fluxes in
// RS.
template <class RP, class RS, int TX, int TY>
struct TileShared {
  static constexpr int NT = TX * TY;
  static constexpr int HX = TX + 4, HY = TY + 4, HN = HX * HY;
  static constexpr int NXF = (TX + 1) * TY, NYF = TX * (TY + 1), NF = NXF + NYF;
  static constexpr int NW = (NT + 63) / 64;
  static constexpr int NTHR = NW * 64;  // threads per block: whole waves
  RP sP[9 * HN];
  RS sF[9 * NF];
  unsigned long long sRed[4 * NW];  // per wave: counters, speed, valid, fallback vote
};
template <class Pol, int TX, int TY>
using TileSharedOf = TileShared<typename Pol::RP, typename Pol::RS, TX, TY>;

// Where phase C finds a thread's own cell state: read again from memory
// (HOLD_NONE), kept in registers since phase A (HOLD_REGS, for kernels with
// the registers to spare), or stored to LDS in phase A (HOLD_LDS, for
// kernels with the LDS to spare: `sU`, 9 planes of NT states).
enum Hold { HOLD_NONE = 0, HOLD_REGS = 1, HOLD_LDS = 2 };

// The single-grid stage kernel's choice: registers for FP32-flux forms,
// LDS for FP64-flux forms where the tile's states fit beside the tile's
// primitives and fluxes within 80 KB (two blocks per CU).
template <class Pol, int TX, int TY>
struct StageHold {
  static constexpr int value =
      sizeof(typename Pol::R) == 4
          ? HOLD_REGS
          : (sizeof(TileSharedOf<Pol, TX, TY>) + 9 * TX * TY * sizeof(typename Pol::S) <= 81920
                 ? HOLD_LDS
                 : HOLD_NONE);
};

// One stage on one tile (tile index `tile` of the instance described by
// `a`), by one workgroup.  The caller has checked that the instance is
// running; every thread of the block must call this.  With a HOLD other
// than HOLD_NONE a thread loads its own tile cell first in phase A, and
// phase C takes that state from registers or from `sU` instead of reading
// it again.
// The control-block values a stage reads, as copied at the start of the
// stage by an ensemble's live-list kernel (so that a block of the flat
// ensemble kernel finds them beside its instance's descriptor instead of
// behind the descriptor's control-block pointer).
template <class T>
struct StepVals {
  T ch, dt, dx, dy;
  int step, exact_until;
};

template <class Pol, bool MUSCL, bool HLLD, int STAGE, int TX, int TY, int HOLD = HOLD_NONE>
__device__ __forceinline__ void stage_tile(const StageArgs<Pol>& a, int tile,
                                           TileSharedOf<Pol, TX, TY>& sh,
                                           typename Pol::S* sU = nullptr,
                                           const StepVals<typename Pol::T>* sv = nullptr) {
  using R = typename Pol::R;
  using RP = typename Pol::RP;
  using RS = typename Pol::RS;
  using T = typename Pol::T;
  using A = typename RhsType<Pol>::type;
  using Sh = TileSharedOf<Pol, TX, TY>;
  constexpr int NT = Sh::NT, HX = Sh::HX, HN = Sh::HN;
  constexpr int NXF = Sh::NXF, NF = Sh::NF, NW = Sh::NW, NTHR = Sh::NTHR;
  RP* sP = sh.sP;
  RS* sF = sh.sF;
  unsigned long long* sRed = sh.sRed;

  // The step's values: from the control block, or (sv) as copied there at
  // the start of the stage.
  Ctrl<T>* ctrl = a.ctrl;
  if constexpr (Pol::kRedo) {
    // Rollback window: the whole tile goes to the FP64 kernel unevaluated.
    const bool window = sv ? uni(sv->step) < uni(sv->exact_until)
                           : uni(ctrl->step) < uni(ctrl->exact_until);
    if (a.redo_list && window) {
      if (threadIdx.x == 0) a.redo_list[atomicAdd(a.redo_counter, 1)] = a.redo_base + tile;
      return;
    }
    // Injected failure (test hook): tile 0 reports cell 0 unusable as the
    // FP64 fallback would, so the step halts with code 3.
    if (STAGE == 1 && tile == 0 && a.redo_list) {
      const int inj = uni(ctrl->fail_inject);
      const int step = sv ? uni(sv->step) : uni(ctrl->step);
      if (inj != 0 && inj == step + 1 && threadIdx.x == 0) atomicMin(&ctrl->bad[0], 0);
    }
  }
  const T ch_t = uni(sv ? sv->ch : ctrl->ch), dt = uni(sv ? sv->dt : ctrl->dt);
  const T dx = uni(sv ? sv->dx : ctrl->dx), dy = uni(sv ? sv->dy : ctrl->dy);
  const R ch = (R)ch_t;
  const Phys<R> c = make_phys<R>(a.tol);
  const Phys<RP> cp = make_phys<RP>(a.tol);  // recovery / reconstruction
  const int nx = a.nx, ny = a.ny;
  const size_t ps = a.plane;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int i0 = tx * TX, j0 = ty * TY;
  const int tid = (int)threadIdx.x;

  // Phase A: tile + halo (no corners) -> primitives in LDS.  A thread's
  // cells (at most kCells) are all loaded before any is recovered, so their
  // loads are in flight together: one memory round trip, not one per cell.
  // Cells in box raster order, or with HOLD the tile's cells first (thread
  // tid < NT takes tile cell tid, the one it updates in phase C), then the
  // halo: 2-cell arms on the tile's rows, then on its columns.
  constexpr int kCells = (HN + NTHR - 1) / NTHR;
  static_assert(kCells <= 2, "phase A holds at most two cells per thread");
  static_assert(HOLD == HOLD_NONE || NT + 4 * TX + 4 * TY <= 2 * NTHR,
                "phase A takes two cells per thread");
  CellIO<Pol> u[kCells];
  bool have[kCells];
  int hbox[kCells];
#pragma unroll
  for (int q = 0; q < kCells; q++) {
    int r, cc;
    if constexpr (HOLD != HOLD_NONE) {
      const int e = tid + q * NTHR;
      if (e < NT) {
        r = e / TX;
        cc = 2 + (e - r * TX);
        r += 2;
      } else if (e < NT + 4 * TY) {
        const int m = e - NT, k = m & 3;
        cc = k < 2 ? k : TX + k;
        r = 2 + (m >> 2);
      } else {
        const int m = e - NT - 4 * TY, k = m / TX;
        cc = 2 + (m - k * TX);
        r = k < 2 ? k : TY + k;
      }
      have[q] = e < NT + 4 * TX + 4 * TY;
    } else {
      const int h = tid + q * NTHR;
      r = h / HX;
      cc = h - r * HX;
      have[q] = h < HN && !((r < 2 || r >= TY + 2) && (cc < 2 || cc >= TX + 2));
    }
    hbox[q] = r * HX + cc;
    if (have[q]) u[q].load(a.src, ps, (size_t)wrap(j0 + r - 2, ny) * nx + wrap(i0 + cc - 2, nx));
  }
  if constexpr (HOLD == HOLD_LDS) {
    if (tid < NT) {
#pragma unroll
      for (int k = 0; k < 9; k++) sU[k * NT + tid] = u[0].v[k];
    }
  }
#pragma unroll
  for (int q = 0; q < kCells; q++) {
    if (!have[q]) continue;
    const int h = hbox[q];
    RP P[9];
    recover<MUSCL>(cp, u[q], a.lowbeta_ratio, P);
#pragma unroll
    for (int k = 0; k < 9; k++) sP[k * HN + h] = P[k];
  }
  __syncthreads();

  // Phase B: faces.  Faces [0, NXF) are x faces (row r, face column fc
  // between local cells fc-1 and fc); [NXF, NF) are y faces (column cc,
  // face row fr between local rows fr-1 and fr).
  unsigned degen = 0, nonphys = 0;
  for (int f = tid; f < NF; f += NTHR) {
    const bool ydir = f >= NXF;
    int base, stride;
    bool owned;
    if (!ydir) {
      const int r = f / (TX + 1), fc = f - r * (TX + 1);
      base = (r + 2) * HX + fc;
      stride = 1;
      owned = (j0 + r < ny) && (fc >= 1 ? (i0 + fc - 1 < nx) : (i0 == 0));
    } else {
      const int g = f - NXF, fr = g / TX, cc = g - fr * TX;
      base = fr * HX + cc + 2;
      stride = HX;
      owned = (i0 + cc < nx) && (fr >= 1 ? (j0 + fr - 1 < ny) : (j0 == 0));
    }
    const int st = face_flux<Pol, MUSCL, HLLD, HN, NF>(c, cp, ch, sP, base, stride, ydir, sF, f);
    if (HLLD && owned) {
      degen += (st == HLLD_DEGENERATE);
      nonphys += (st == HLLD_NONPHYSICAL);
    }
  }
  __syncthreads();

  // Phase C: one cell per thread.  rhs (mhd2d_glm_cpu.cpp:273-284), the
  // stage update (528-535 / 549-557) and the positivity check (328-345).
  const int r = tid / TX, cc = tid - r * TX;
  const int gi = i0 + cc, gj = j0 + r;
  const bool live = tid < NT && gi < nx && gj < ny;
  const size_t g = (size_t)gj * nx + gi;
  unsigned long long sbits = 0;
  unsigned svalid = 0;
  T out[9];
  bool unusable = false;
  if (live) {
    const int xl = r * (TX + 1) + cc, xr = xl + 1;
    const int yb = NXF + r * TX + cc, yt = yb + TX;
    const A adx = (A)dx, ady = (A)dy, aidx = (A)a.inv_dx, aidy = (A)a.inv_dy;
    const bool pow2 = a.pow2 != 0;
    A rhs[9];
#pragma unroll
    for (int k = 0; k < 9; k++)
      rhs[k] = flux_div<A, RS>(sF[k * NF + xr], sF[k * NF + xl], sF[k * NF + yt],
                              sF[k * NF + yb], pow2, adx, ady, aidx, aidy);
    CellIO<Pol> src;
    if constexpr (HOLD == HOLD_REGS) {
      src = u[0];
    } else if constexpr (HOLD == HOLD_LDS) {
#pragma unroll
      for (int k = 0; k < 9; k++) src.v[k] = sU[k * NT + tid];
    } else {
      src.load(a.src, ps, g);
    }
    const T damping = ch_t / smax((T)a.alpha, (T)1.0e-12);
    update_cell<STAGE>(a, src, g, rhs, dt, damping, out);
    if constexpr (sizeof(R) == sizeof(T)) {
      // Exact forms: the reference's check (mhd2d_glm_cpu.cpp:328-345) and
      // wave speeds (289-305), bitwise.  Reference-mixed: the same FP64
      // check (mhd2d_glm_gpu_mixed.cu:377-410), FP32 speeds (256-280).
      const Phys<T> ct = make_phys<T>(a.tol);
      T P[9];
      if (!state_usable(ct, out, P)) atomicMin(&ctrl->bad[STAGE - 1], (int)g);
      if (STAGE == 2) {
        bool valid;
        T s;
        if constexpr (Pol::kRefMixed)
          s = refmix::cell_speed(out, valid);
        else
          s = cell_speed(ct, P, valid);
        sbits = speed_bits(s);
        svalid = valid ? 1u : 0u;
      }
    } else {
      double speed;
      bool valid;
      unusable = !usable_and_speed_fp64(out, STAGE == 2, speed, valid);
      if (unusable && !(Pol::kRedo && a.redo_list)) atomicMin(&ctrl->bad[STAGE - 1], (int)g);
      if (STAGE == 2) {
        sbits = speed_bits(speed);
        svalid = valid ? 1u : 0u;
      }
    }
  }
  // Block reductions (one atomic per block and counter) and, with the FP64
  // fallback, the tile's vote, over one barrier.
  const int lane = tid & 63, wid = tid >> 6;
  const unsigned wd = wave_sum_u32(degen), wn = wave_sum_u32(nonphys);
  unsigned long long wm = 0;
  unsigned wv = 0;
  if (STAGE == 2) {
    wm = wave_max_u64(sbits);
    wv = wave_sum_u32(svalid);
  }
  bool voting = false;
  unsigned long long wvote = 0;
  if constexpr (Pol::kRedo) {
    voting = a.redo_list != nullptr;
    if (voting) wvote = __ballot((unusable || a.redo_all) ? 1 : 0);
  }
  if (lane == 0) {
    sRed[wid] = ((unsigned long long)wn << 32) | wd;
    sRed[NW + wid] = wm;
    sRed[2 * NW + wid] = wv;
    sRed[3 * NW + wid] = wvote;
  }
  __syncthreads();
  if constexpr (Pol::kRedo) {
    // FP64 fallback: a tile with an unusable cell writes nothing (its
    // inputs -- the stage source and, in stage 2, its own cells of U^n --
    // stay intact) and is queued for the FP64 kernel, which also publishes
    // its counters and speeds.
    if (voting) {
      unsigned long long any = 0;
      for (int w = 0; w < NW; w++) any |= sRed[3 * NW + w];
      if (any) {
        if (tid == 0) {
          const int slot = atomicAdd(a.redo_counter, 1);
          a.redo_list[slot] = a.redo_base + tile;
          atomicAdd(&ctrl->redo_tiles, 1ull);
        }
        return;
      }
    }
    if (live) {
#pragma unroll
      for (int k = 0; k < 9; k++) a.dst[k * a.plane + g] = out[k];
    }
  }
  if (tid == 0) {
    unsigned long long d = 0, n = 0, m = 0, v = 0;
    for (int w = 0; w < NW; w++) {
      d += sRed[w] & 0xffffffffull;
      n += sRed[w] >> 32;
      m = u64max(m, sRed[NW + w]);
      v += sRed[2 * NW + w];
    }
    if (d) atomicAdd(&ctrl->degenerate, d);
    if (n) atomicAdd(&ctrl->nonphysical, n);
    if (STAGE == 2) publish_speed(ctrl, m, v != 0);
  }
}

// Waves per SIMD every stage kernel is register-budgeted for: what LDS
// allows (FP64: 64 KB per 4-wave block, 2 blocks per CU; FP32 flux: 32 KB,
// 4 blocks).  Without the target a kernel a register or two over 256 (the
// FP64 list and queue kernels, whose loops keep a little more live) runs at
// half the occupancy.
template <class Pol> struct StageWaves { static constexpr int value = sizeof(typename Pol::R) == 8 ? 2 : 4; };

// One instance, one block per tile (XCD-banded tile order).
template <class Pol, bool MUSCL, bool HLLD, int STAGE, int TX, int TY>
__global__ __launch_bounds__((TX * TY + 63) / 64 * 64)
__attribute__((amdgpu_waves_per_eu(StageWaves<Pol>::value))) void k_stage(StageArgs<Pol> a) {
  constexpr int hold = StageHold<Pol, TX, TY>::value;
  __shared__ TileSharedOf<Pol, TX, TY> sh;
  __shared__ typename Pol::S sU[hold == HOLD_LDS ? 9 * TX * TY : 1];
  Ctrl<typename Pol::T>* ctrl = a.ctrl;
  // The routed tile does not depend on the control block: its load is in
  // flight with the control block's (the grid has one block per slot).
  const int routed = a.tiles ? a.tiles[blockIdx.x] : 0;
  if (ctrl->run != RUN_ACTIVE) return;
  if (STAGE == 2 && ctrl->bad[0] != INT_MAX) return;
  const int tile = a.tiles ? uni(routed) : xcd_tile((int)blockIdx.x, a.ntiles);
  if (tile < 0) return;
  stage_tile<Pol, MUSCL, HLLD, STAGE, TX, TY, hold>(a, tile, sh, sU);
}

// One stage over listed tiles (the mixed form's FP64 fallback queue, the
// selector's FP64 and audit tiles): up to kListSegs segments, each with its
// own arguments (destination, control block) and tile list
// list[0 .. *count); a fixed grid strides over all of them, so a selected
// step's FP64 tiles, fallback queue and audit tiles cost one launch.
constexpr int kListSegs = 3;
template <class Pol>
struct ListSeg {
  StageArgs<Pol> a;
  const int* list;   // null: empty segment
  const int* count;
  CurBuffers<Pol> bufs;
};

template <class Pol, int STAGE>
__device__ __forceinline__ int seg_count(const ListSeg<Pol>& g) {
  if (!g.list) return 0;
  const Ctrl<typename Pol::T>* ctrl = g.a.ctrl;
  if (uni(ctrl->run) != RUN_ACTIVE) return 0;
  if (STAGE == 2 && uni(ctrl->bad[0]) != INT_MAX) return 0;
  return uni(*g.count);
}

template <class Pol>
struct ListSegs {
  ListSeg<Pol> seg[kListSegs];
};

template <class Pol, bool MUSCL, bool HLLD, int STAGE, int TX, int TY>
__global__ __launch_bounds__((TX * TY + 63) / 64 * 64)
__attribute__((amdgpu_waves_per_eu(StageWaves<Pol>::value))) void k_stage_list(ListSegs<Pol> g) {
  __shared__ TileSharedOf<Pol, TX, TY> sh;
  int n[kListSegs], total = 0;
#pragma unroll
  for (int k = 0; k < kListSegs; k++) {
    n[k] = seg_count<Pol, STAGE>(g.seg[k]);
    total += n[k];
  }
  for (int e = (int)blockIdx.x; e < total; e += (int)gridDim.x) {
    int k = 0, x = e;
    while (x >= n[k]) x -= n[k++];
    StageArgs<Pol> a = g.seg[k].a;
    follow_cur(a, g.seg[k].bufs);
    stage_tile<Pol, MUSCL, HLLD, STAGE, TX, TY>(a, uni(g.seg[k].list[x]), sh);
    __syncthreads();
  }
}

// Signal-speed maximum of a state (initial state / after set_state), in T.
template <class Pol>
__device__ __forceinline__ void load_wide(const typename Pol::S* src, size_t ps, size_t g,
                                          typename Pol::T out[9]) {
  CellIO<Pol> u;
  u.load(src, ps, g);
  if constexpr (sizeof(typename Pol::S) == sizeof(typename Pol::T) && Pol::kPlanes == 9) {
#pragma unroll
    for (int k = 0; k < 9; k++) out[k] = u.v[k];
  } else {
    u.wide(out);
  }
}

template <class Pol>
__global__ __launch_bounds__(256) void k_speed(const typename Pol::S* src, size_t plane,
                                               int nx, int ny, double tol,
                                               Ctrl<typename Pol::T>* ctrl) {
  using T = typename Pol::T;
  __shared__ unsigned long long sRed[2 * 4];
  const Phys<T> c = make_phys<T>(tol);
  const long long n = (long long)nx * ny;
  unsigned long long best = 0;
  unsigned valid_any = 0;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n;
       q += (long long)gridDim.x * 256) {
    T U[9], P[9];
    load_wide<Pol>(src, plane, (size_t)q, U);
    bool valid;
    T s;
    if constexpr (Pol::kRefMixed) {
      s = refmix::cell_speed(U, valid);
    } else {
      cons_to_prim<T, 9>(c, U, P);
      s = cell_speed(c, P, valid);
    }
    best = u64max(best, speed_bits(s));
    valid_any |= valid ? 1u : 0u;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const unsigned long long wm = wave_max_u64(best);
  const unsigned wv = wave_sum_u32(valid_any);
  if (lane == 0) { sRed[wid] = wm; sRed[4 + wid] = wv; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = 0, v = 0;
    for (int w = 0; w < 4; w++) { m = u64max(m, sRed[w]); v += sRed[4 + w]; }
    publish_speed(ctrl, m, v != 0);
  }
}

// c_h and dt of the coming step (mhd2d_glm_cpu.cpp:514-524).  The rate
// maximum of maxSignalRate (307-326) is exact in closed form: every cell's
// speed is <= c_h, so max(s, c_h) == c_h bitwise and each cell's rate is
// c_h/dx + c_h/dy; NaN cells never enter the running maximum, so with no
// non-NaN cell the maximum stays 0.
template <class T>
__device__ void prepare_step(Ctrl<T>* ctrl) {
  ctrl->bad[0] = INT_MAX;
  ctrl->bad[1] = INT_MAX;
  ctrl->redo_count[0] = 0;
  ctrl->redo_count[1] = 0;
  if (!(ctrl->t < ctrl->t_end)) {
    ctrl->run = RUN_DONE;
    return;
  }
  const T speed = bits_speed<T>(ctrl->speed_cur);
  const T ch = smax(speed, (T)1.0e-12);
  // The reference-mixed build takes the rate in float (rateKernel,
  // mhd2d_glm_gpu_mixed.cu:283-308); in closed form as for the exact forms.
  const T r = ctrl->rate_f32
                  ? (T)((float)ch / (float)ctrl->dx + (float)ch / (float)ctrl->dy)
                  : ch / ctrl->dx + ch / ctrl->dy;
  const T rate = ctrl->valid_cur ? smax((T)0.0, r) : (T)0.0;
  ctrl->ch = ch;
  if (rate <= (T)0.0) {
    ctrl->run = RUN_HALTED;
    ctrl->code = 2;
    ctrl->fail_step = ctrl->step;
    ctrl->fail_t = ctrl->t;
    ctrl->fail_ch = ch;
    return;
  }
  T dt = ctrl->cfl / rate;
  if (ctrl->t + dt > ctrl->t_end) dt = ctrl->t_end - ctrl->t;
  ctrl->dt = dt;
}

template <class T>
__global__ void k_commit_speed(Ctrl<T>* ctrl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  fold_speed_slots(ctrl);
  ctrl->speed_cur = ctrl->speed_acc;
  ctrl->valid_cur = ctrl->valid_acc;
  ctrl->speed_acc = 0;
  ctrl->valid_acc = 0;
}

template <class T>
__global__ void k_set_target(Ctrl<T>* ctrl, double t_end) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (ctrl->run == RUN_HALTED) return;
  ctrl->t_end = (T)t_end;
  ctrl->run = RUN_ACTIVE;
  prepare_step(ctrl);
}

// End of a step for one instance: verdict of the positivity checks, then
// t += dt, step++ and the next step's c_h and dt (mhd2d_glm_cpu.cpp:538-569).
template <class T>
__device__ __forceinline__ void finish_one(Ctrl<T>* ctrl) {
  if (ctrl->run != RUN_ACTIVE) return;
  if (ctrl->bad[0] != INT_MAX || ctrl->bad[1] != INT_MAX) {
    const bool first = ctrl->bad[0] != INT_MAX;
    ctrl->run = RUN_HALTED;
    ctrl->code = first ? 3 : 4;
    ctrl->fail_cell = first ? ctrl->bad[0] : ctrl->bad[1];
    ctrl->fail_step = ctrl->step;
    ctrl->fail_t = ctrl->t;
    ctrl->fail_dt = ctrl->dt;
    ctrl->fail_ch = ctrl->ch;
    return;
  }
  ctrl->t = ctrl->t + ctrl->dt;
  if (ctrl->step < ctrl->exact_until) ctrl->exact_steps += 1;
  ctrl->step += 1;
  fold_speed_slots(ctrl);
  ctrl->speed_cur = ctrl->speed_acc;
  ctrl->valid_cur = ctrl->valid_acc;
  ctrl->speed_acc = 0;
  ctrl->valid_acc = 0;
  prepare_step(ctrl);
}

template <class T>
__global__ void k_finish(Ctrl<T>* ctrl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  finish_one(ctrl);
}

// ---------------------------------------------------------------------------
// Ensembles: many instances of uneven sizes, each with its own dt and
// control block.  At the start of each stage one block lists the running
// instances' (instance, tile) work items (the live list); the stage kernel
// then takes them one block per item (k_stage_ens), each XCD a contiguous
// run of items so that neighbouring tiles share its L2, or, as a persistent
// grid, pulls them from a work queue split into 8 such runs (k_stage_queue:
// a block that drains its run steals from the others).
// ---------------------------------------------------------------------------
template <class Pol>
struct InstDesc {
  typename Pol::S* u;
  typename Pol::S* u1;
  size_t plane;
  int nx, ny, tiles_x, ntiles;
  int pow2;
  double inv_dx, inv_dy;
  Ctrl<typename Pol::T>* ctrl;
};

template <class Pol>
struct EnsArgs {
  const InstDesc<Pol>* desc;
  const int* prefix;  // prefix[i] = first item of instance i; prefix[n] = total
  int n, total;
  unsigned* queue;    // 2 stages x 8 chunk counters + 2 fallback counts, zeroed
                      // by the finish kernel
  double tol, alpha;
  float lowbeta_ratio;
  int* redo_items[2];   // FP64 fallback queues (null: none)
  int redo_all;
  void* live;           // live list (ens_live_bytes(n) bytes), rebuilt each stage
};

// Live list of an ensemble (layout in mhd_kernels.h): header, the first live
// work item of each live instance, and per live instance its descriptor and
// its first item in the fixed numbering (the fallback queue's).
struct EnsLiveHdr {
  int total;  // live work items
  int n;      // live instances
  int pad[14];
};
static_assert(sizeof(EnsLiveHdr) == kEnsLiveHeaderBytes, "live list header");
template <class Pol>
struct LiveEntry {
  InstDesc<Pol> d;
  int first;
  StepVals<typename Pol::T> v;  // the control block's values for this stage
};

template <class Pol>
__device__ __forceinline__ int* live_first(const EnsArgs<Pol>& e) {
  return (int*)((char*)e.live + kEnsLiveHeaderBytes);
}
template <class Pol>
__device__ __forceinline__ LiveEntry<Pol>* live_entries(const EnsArgs<Pol>& e) {
  static_assert(sizeof(LiveEntry<Pol>) <= kEnsLiveEntryBytes, "live list entry");
  return (LiveEntry<Pol>*)((char*)e.live + kEnsLiveHeaderBytes + ens_live_first_bytes(e.n));
}

// Live slot of live work item `item`: the number of live instances whose
// first item is <= item, less one.  Each lane holds one instance's first
// item, 64 instances per round; `first0` is the lane's value of the first
// round, loaded by the caller (so that it can be in flight with other
// loads).  Wave-uniform.
template <class Pol>
__device__ __forceinline__ int live_slot(const EnsArgs<Pol>& e, int nlive, int first0, int item) {
  const int lane = (int)(threadIdx.x & 63);
  int count = (int)__popcll(__ballot(lane < nlive && first0 <= item));
  for (int c0 = 64; c0 < nlive; c0 += 64) {
    const int p = c0 + lane < nlive ? live_first(e)[c0 + lane] : INT_MAX;
    count += (int)__popcll(__ballot(p <= item));
  }
  return uni(count - 1);
}

// Rebuilds the live list at the start of a stage (one block of 256): the
// instances whose control block says RUN_ACTIVE (and, for stage 2, whose
// stage 1 left no unusable cell), in instance order, their work items
// numbered consecutively from 0, with the control-block values the stage
// reads.
template <class Pol, int STAGE>
__global__ __launch_bounds__(256) void k_ens_live(EnsArgs<Pol> e) {
  __shared__ unsigned long long s_wave[4];
  __shared__ unsigned long long s_base;
  const int tid = (int)threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int* first = live_first(e);
  LiveEntry<Pol>* ent = live_entries(e);
  if (tid == 0) s_base = 0;
  __syncthreads();
  for (int c0 = 0; c0 < e.n; c0 += 256) {
    const int i = c0 + tid;
    // (tiles << 32) | 1 for a live instance: one inclusive scan gives both
    // its live slot (low word) and its first live item (high word).
    unsigned long long v = 0;
    if (i < e.n) {
      const Ctrl<typename Pol::T>* c = e.desc[i].ctrl;
      if (c->run == RUN_ACTIVE && (STAGE == 1 || c->bad[0] == INT_MAX))
        v = ((unsigned long long)(unsigned)e.desc[i].ntiles << 32) | 1ull;
    }
    unsigned long long incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned long long t = __shfl_up(incl, d, 64);
      if (lane >= d) incl += t;
    }
    if (lane == 63) s_wave[wid] = incl;
    __syncthreads();
    unsigned long long off = s_base;
    for (int w = 0; w < wid; w++) off += s_wave[w];
    if (v) {
      const unsigned long long excl = off + incl - v;
      const int slot = (int)(unsigned)(excl & 0xffffffffull);
      first[slot] = (int)(excl >> 32);
      ent[slot].d = e.desc[i];
      ent[slot].first = e.prefix[i];
      const Ctrl<typename Pol::T>* c = e.desc[i].ctrl;
      StepVals<typename Pol::T>& sv = ent[slot].v;
      sv.ch = c->ch;
      sv.dt = c->dt;
      sv.dx = c->dx;
      sv.dy = c->dy;
      sv.step = c->step;
      sv.exact_until = c->exact_until;
    }
    __syncthreads();
    if (tid == 255) s_base = off + incl;
    __syncthreads();
  }
  if (tid == 0) {
    EnsLiveHdr* h = (EnsLiveHdr*)e.live;
    h->total = (int)(s_base >> 32);
    h->n = (int)(unsigned)(s_base & 0xffffffffull);
  }
}

// Stage arguments of one ensemble instance.
template <class Pol, int STAGE>
__device__ __forceinline__ StageArgs<Pol> inst_args(const EnsArgs<Pol>& e, const InstDesc<Pol>& d,
                                                    int first_item) {
  StageArgs<Pol> a;
  a.plane = uni_size(d.plane);
  a.nx = uni(d.nx);
  a.ny = uni(d.ny);
  a.tiles_x = uni(d.tiles_x);
  a.ntiles = uni(d.ntiles);
  a.pow2 = uni(d.pow2);
  a.inv_dx = uni(d.inv_dx);
  a.inv_dy = uni(d.inv_dy);
  a.tol = e.tol;
  a.alpha = e.alpha;
  a.lowbeta_ratio = e.lowbeta_ratio;
  a.ctrl = uni_ptr(d.ctrl);
  a.tiles = nullptr;
  a.redo_list = e.redo_items[STAGE - 1];
  a.redo_counter = (int*)&e.queue[16 + STAGE - 1];
  a.redo_base = first_item;
  a.redo_all = e.redo_all;
  const auto u = uni_ptr(d.u), u1 = uni_ptr(d.u1);
  a.src = STAGE == 1 ? u : u1;
  a.uold = u;
  a.dst = STAGE == 1 ? u1 : u;
  return a;
}

__device__ __forceinline__ int find_instance(const int* prefix, int n, int item) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}


// One stage of every live instance, one block per live work item: the grid
// covers every instance's tiles; block b takes live item (b mod 8) * per +
// b / 8 with per = ceil(live items / 8), so each XCD (blocks are dealt to the
// 8 XCDs round-robin) takes a contiguous run of items, whose halos it then
// shares in its L2, and blocks past the live items exit after one load.
template <class Pol, bool MUSCL, bool HLLD, int STAGE, int TX, int TY>
__global__ __launch_bounds__((TX * TY + 63) / 64 * 64)
__attribute__((amdgpu_waves_per_eu(StageWaves<Pol>::value))) void k_stage_ens(EnsArgs<Pol> e) {
  constexpr int hold = StageHold<Pol, TX, TY>::value;
  __shared__ TileSharedOf<Pol, TX, TY> sh;
  __shared__ typename Pol::S sU[hold == HOLD_LDS ? 9 * TX * TY : 1];
  const EnsLiveHdr* hdr = (const EnsLiveHdr*)e.live;
  const int* first = live_first(e);
  const int lane = (int)(threadIdx.x & 63);
  const int first0 = lane < e.n ? first[lane] : INT_MAX;  // in flight with the header
  const int total = uni(hdr->total), nlive = uni(hdr->n);
  const int b = (int)blockIdx.x, per = (total + 7) >> 3, k = b >> 3;
  if (k >= per) return;
  const int item = (b & 7) * per + k;
  if (item >= total) return;
  // The live list holds only running instances (for stage 2, those whose
  // stage 1 passed) and their step values: no control-block load here.
  const int slot = live_slot(e, nlive, first0, item);
  const LiveEntry<Pol>& le = live_entries(e)[slot];
  stage_tile<Pol, MUSCL, HLLD, STAGE, TX, TY, hold>(inst_args<Pol, STAGE>(e, le.d, uni(le.first)),
                                                    item - uni(first[slot]), sh, sU, &le.v);
}

// The same as a persistent grid pulling live work items from a queue.
template <class Pol, bool MUSCL, bool HLLD, int STAGE, int TX, int TY>
__global__ __launch_bounds__((TX * TY + 63) / 64 * 64)
__attribute__((amdgpu_waves_per_eu(StageWaves<Pol>::value))) void k_stage_queue(EnsArgs<Pol> e) {
  constexpr int hold = StageHold<Pol, TX, TY>::value;
  __shared__ TileSharedOf<Pol, TX, TY> sh;
  __shared__ typename Pol::S sU[hold == HOLD_LDS ? 9 * TX * TY : 1];
  __shared__ int s_item, s_count;
  // Items are claimed kClaim at a time (consecutive tiles, nearly always of
  // one instance): fewer queue atomics per tile at the cost of at most
  // kClaim - 1 tiles of imbalance at the end of a stage.
  constexpr int kClaim = 4;
  const EnsLiveHdr* hdr = (const EnsLiveHdr*)e.live;
  const int* first = live_first(e);
  const int lane = (int)(threadIdx.x & 63);
  const int first0 = lane < e.n ? first[lane] : INT_MAX;
  const int total = uni(hdr->total), nlive = uni(hdr->n);
  unsigned* q = e.queue + (STAGE - 1) * 8;
  const int home = (int)(blockIdx.x & 7);
  int chunk = 0;
  // Item range of the live instance of the previous item (block-uniform),
  // so the slot search runs only when an item leaves it.
  int slot = 0, cur_lo = 0, cur_hi = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      int item = -1, cnt = 0;
      while (chunk < 8) {
        const int x = (home + chunk) & 7;
        const int lo = (int)((long long)total * x / 8);
        const int hi = (int)((long long)total * (x + 1) / 8);
        const unsigned kk = atomicAdd(&q[x], (unsigned)kClaim);
        if ((long long)lo + kk < (long long)hi) {
          item = lo + (int)kk;
          cnt = hi - item < kClaim ? hi - item : kClaim;
          break;
        }
        chunk++;
      }
      s_item = item;
      s_count = cnt;
    }
    __syncthreads();
    // Block-uniform: readfirstlane keeps the item, the slot and the
    // descriptor in scalar registers.
    const int first_item = uni(s_item), count = uni(s_count);
    __syncthreads();
    if (first_item < 0) break;  // every chunk drained: every block gets here
    for (int c = 0; c < count; c++) {
      const int item = first_item + c;
      if (item < cur_lo || item >= cur_hi) {
        slot = live_slot(e, nlive, first0, item);
        cur_lo = uni(first[slot]);
        cur_hi = slot + 1 < nlive ? uni(first[slot + 1]) : total;
      }
      const LiveEntry<Pol>& le = live_entries(e)[slot];
      const InstDesc<Pol>& d = le.d;
      if (uni(d.ctrl->run) != RUN_ACTIVE) continue;
      if (STAGE == 2 && uni(d.ctrl->bad[0]) != INT_MAX) continue;
      stage_tile<Pol, MUSCL, HLLD, STAGE, TX, TY, hold>(inst_args<Pol, STAGE>(e, d, uni(le.first)),
                                                        item - cur_lo, sh, sU);
      __syncthreads();  // LDS is reused by the next item
    }
  }
}

// FP64 fallback of an ensemble: re-evaluates the work items a mixed stage
// queued (count in queue[16 + STAGE - 1]); a fixed grid strides over them.
template <class Pol, bool MUSCL, bool HLLD, int STAGE, int TX, int TY>
__global__ __launch_bounds__((TX * TY + 63) / 64 * 64)
__attribute__((amdgpu_waves_per_eu(StageWaves<Pol>::value))) void k_queue_redo(EnsArgs<Pol> e,
                                                                          const int* list) {
  __shared__ TileSharedOf<Pol, TX, TY> sh;
  const int count = __builtin_amdgcn_readfirstlane((int)e.queue[16 + STAGE - 1]);
  for (int x = (int)blockIdx.x; x < count; x += (int)gridDim.x) {
    const int item = __builtin_amdgcn_readfirstlane(list[x]);
    const int inst = find_instance(e.prefix, e.n, item);
    const InstDesc<Pol>& d = e.desc[inst];
    if (d.ctrl->run != RUN_ACTIVE) continue;
    if (STAGE == 2 && d.ctrl->bad[0] != INT_MAX) continue;
    const int first = e.prefix[inst];
    StageArgs<Pol> a = inst_args<Pol, STAGE>(e, d, first);
    a.redo_list = nullptr;
    stage_tile<Pol, MUSCL, HLLD, STAGE, TX, TY>(a, item - first, sh);
    __syncthreads();
  }
}

template <class T>
__global__ void k_finish_multi(Ctrl<T>* ctrls, int n, unsigned* queue) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) finish_one(&ctrls[i]);
  if (i < kEnsQueueWords) queue[i] = 0;
}

template <class T>
__global__ void k_set_target_multi(Ctrl<T>* ctrls, int n, double t_end, unsigned* queue) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kEnsQueueWords) queue[i] = 0;
  if (i >= n) return;
  Ctrl<T>* ctrl = &ctrls[i];
  if (ctrl->run == RUN_HALTED) return;
  ctrl->t_end = (T)t_end;
  ctrl->run = RUN_ACTIVE;
  prepare_step(ctrl);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
// Device address of a control block's per-stage fallback counter.
template <class T>
static int* redo_counter_of(void* ctrl, int stage) {
  return (int*)((char*)ctrl + offsetof(Ctrl<T>, redo_count) + (stage - 1) * sizeof(int));
}

template <class Pol, int TX, int TY>
static StageArgs<Pol> stage_args(const GridDesc& g, void* ctrl) {
  using T = typename Pol::T;
  const int tiles_x = (g.nx + TX - 1) / TX, tiles_y = (g.ny + TY - 1) / TY;
  StageArgs<Pol> a;
  a.plane = g.plane;
  a.nx = g.nx;
  a.ny = g.ny;
  a.tiles_x = tiles_x;
  a.ntiles = tiles_x * tiles_y;
  a.tol = g.tol;
  a.alpha = g.alpha;
  a.lowbeta_ratio = (float)g.lowbeta_ratio;
  a.ctrl = (Ctrl<T>*)ctrl;
  a.tiles = g.tiles;
  a.redo_list = nullptr;
  a.redo_counter = redo_counter_of<T>(ctrl, 1);
  a.redo_base = 0;
  a.redo_all = g.redo_all;
  // dx = (1-0)/nx in the control type, as the reference computes it.  When
  // nx and ny are powers of two, dx and dy are exact powers of two and
  // dividing by them equals multiplying by their exact inverses.
  const T dx = ((T)1.0 - (T)0.0) / (T)g.nx, dy = ((T)1.0 - (T)0.0) / (T)g.ny;
  a.pow2 = ((g.nx & (g.nx - 1)) == 0 && (g.ny & (g.ny - 1)) == 0) ? 1 : 0;
  a.inv_dx = (double)((T)1.0 / dx);
  a.inv_dy = (double)((T)1.0 / dy);
  return a;
}

template <class Pol, bool MUSCL, bool HLLD, int TX, int TY>
static hipError_t launch_step_t(const GridDesc& g, void* u, void* u1, void* ctrl,
                                hipStream_t s, hipEvent_t* ev) {
  using S = typename Pol::S;
  using T = typename Pol::T;
  StageArgs<Pol> a = stage_args<Pol, TX, TY>(g, ctrl);
  const bool redo = Pol::kRedo && g.redo && g.redo_list[0];
  a.src = (const S*)u;
  a.uold = (const S*)u;
  a.dst = (S*)u1;
  if (redo) a.redo_list = g.redo_list[0];
  a.redo_counter = redo_counter_of<T>(ctrl, 1);
  constexpr int nthr = (TX * TY + 63) / 64 * 64;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL((k_stage<Pol, MUSCL, HLLD, 1, TX, TY>), dim3(a.ntiles), dim3(nthr), 0, s,
                     a);
  if (ev) (void)hipEventRecord(ev[1], s);
  if (redo) (void)g.redo(g, 1, u, u, u1, ctrl, g.redo_list[0], redo_counter_of<T>(ctrl, 1), s);
  a.src = (const S*)u1;
  a.uold = (const S*)u;
  a.dst = (S*)u;
  if (redo) a.redo_list = g.redo_list[1];
  a.redo_counter = redo_counter_of<T>(ctrl, 2);
  hipLaunchKernelGGL((k_stage<Pol, MUSCL, HLLD, 2, TX, TY>), dim3(a.ntiles), dim3(nthr), 0, s,
                     a);
  if (redo) (void)g.redo(g, 2, u1, u, u, ctrl, g.redo_list[1], redo_counter_of<T>(ctrl, 2), s);
  if (ev) (void)hipEventRecord(ev[2], s);
  hipLaunchKernelGGL(k_finish<T>, dim3(1), dim3(64), 0, s, (Ctrl<T>*)ctrl);
  if (ev) (void)hipEventRecord(ev[3], s);
  return hipGetLastError();
}

template <class Pol, bool MUSCL, bool HLLD, int TX, int TY>
static hipError_t launch_stage_t(const GridDesc& g, int stage, const void* src, const void* uold,
                                 void* dst, void* ctrl, hipStream_t s) {
  using S = typename Pol::S;
  StageArgs<Pol> a = stage_args<Pol, TX, TY>(g, ctrl);
  // With a queue and no `redo` launcher the caller re-evaluates the queue
  // itself (the selector folds it into its FP64 list launch).
  const bool queue = Pol::kRedo && g.redo_list[stage - 1];
  const bool redo = queue && g.redo;
  a.src = (const S*)src;
  a.uold = (const S*)uold;
  a.dst = (S*)dst;
  if (queue) a.redo_list = g.redo_list[stage - 1];
  a.redo_counter = redo_counter_of<typename Pol::T>(ctrl, stage);
  constexpr int nthr = (TX * TY + 63) / 64 * 64;
  if (stage == 1)
    hipLaunchKernelGGL((k_stage<Pol, MUSCL, HLLD, 1, TX, TY>), dim3(a.ntiles), dim3(nthr), 0, s,
                       a);
  else
    hipLaunchKernelGGL((k_stage<Pol, MUSCL, HLLD, 2, TX, TY>), dim3(a.ntiles), dim3(nthr), 0, s,
                       a);
  hipError_t e = hipGetLastError();
  if (redo && e == hipSuccess)
    e = g.redo(g, stage, src, uold, dst, ctrl, g.redo_list[stage - 1],
               redo_counter_of<typename Pol::T>(ctrl, stage), s);
  return e;
}

// Grid of the list-driven stage kernel: enough blocks for a few hundred
// listed tiles at once; an empty list costs one short launch.
constexpr int kRedoBlocks = 512;

// With `queue` (mixed form) a listed tile with an unusable cell is queued
// there for the FP64 kernel instead of halting the run.
template <class Pol, bool MUSCL, bool HLLD, int TX, int TY>
static hipError_t launch_list_t(const GridDesc& g, int stage, const ListJob* jobs, int njobs,
                                hipStream_t s, int* queue = nullptr) {
  using S = typename Pol::S;
  auto seg = [&](const ListJob& j) {
    ListSeg<Pol> q;
    q.a = stage_args<Pol, TX, TY>(g, j.ctrl);
    q.a.tiles = nullptr;
    q.a.redo_all = 0;
    if (Pol::kRedo && queue && j.ctrl) {
      q.a.redo_list = queue;
      q.a.redo_counter = redo_counter_of<typename Pol::T>(j.ctrl, stage);
    }
    q.a.src = (const S*)j.src;
    q.a.uold = (const S*)j.uold;
    q.a.dst = (S*)j.dst;
    q.bufs.cur = j.cur;
    q.bufs.src_alt = (const S*)j.src_alt;
    q.bufs.uold_alt = (const S*)j.uold_alt;
    q.bufs.dst_alt = (S*)j.dst_alt;
    q.list = j.list;
    q.count = j.count;
    return q;
  };
  const ListJob none{nullptr, nullptr, nullptr, jobs[0].ctrl, nullptr, nullptr};
  ListSegs<Pol> segs;
  for (int k = 0; k < kListSegs; k++) segs.seg[k] = seg(k < njobs ? jobs[k] : none);
  const int ntiles = segs.seg[0].a.ntiles;
  const int grid = ntiles < kRedoBlocks ? ntiles : kRedoBlocks;
  constexpr int nthr = (TX * TY + 63) / 64 * 64;
  if (stage == 1)
    hipLaunchKernelGGL((k_stage_list<Pol, MUSCL, HLLD, 1, TX, TY>), dim3(grid), dim3(nthr), 0, s,
                       segs);
  else
    hipLaunchKernelGGL((k_stage_list<Pol, MUSCL, HLLD, 2, TX, TY>), dim3(grid), dim3(nthr), 0, s,
                       segs);
  return hipGetLastError();
}

template <class Pol>
static hipError_t launch_listn(const GridDesc& g, int stage, const ListJob* jobs, int njobs,
                               hipStream_t s) {
  if (njobs < 1 || njobs > kListSegs) return hipErrorInvalidValue;
  if (!(g.muscl && g.hlld)) {
    if (g.muscl) return launch_list_t<Pol, true, false, 16, 15>(g, stage, jobs, njobs, s);
    if (g.hlld) return launch_list_t<Pol, false, true, 16, 15>(g, stage, jobs, njobs, s);
    return launch_list_t<Pol, false, false, 16, 15>(g, stage, jobs, njobs, s);
  }
  switch (g.tile) {
    case TILE_16x16: return launch_list_t<Pol, true, true, 16, 16>(g, stage, jobs, njobs, s);
    case TILE_32x8: return launch_list_t<Pol, true, true, 32, 8>(g, stage, jobs, njobs, s);
    case TILE_16x8: return launch_list_t<Pol, true, true, 16, 8>(g, stage, jobs, njobs, s);
    case TILE_32x7: return launch_list_t<Pol, true, true, 32, 7>(g, stage, jobs, njobs, s);
    case TILE_16x7: return launch_list_t<Pol, true, true, 16, 7>(g, stage, jobs, njobs, s);
    default: return launch_list_t<Pol, true, true, 16, 15>(g, stage, jobs, njobs, s);
  }
}

template <class Pol>
static hipError_t launch_list(const GridDesc& g, int stage, const void* src, const void* uold,
                              void* dst, void* ctrl, const int* list, const int* count,
                              hipStream_t s) {
  if (!list || !count) return hipErrorInvalidValue;
  const ListJob j{src, uold, dst, ctrl, list, count};
  return launch_listn<Pol>(g, stage, &j, 1, s);
}

template <class Pol>
static hipError_t launch_stage(const GridDesc& g, int stage, const void* src, const void* uold,
                               void* dst, void* ctrl, hipStream_t s) {
  if (!(g.muscl && g.hlld)) {
    if (g.muscl) return launch_stage_t<Pol, true, false, 16, 15>(g, stage, src, uold, dst, ctrl, s);
    if (g.hlld) return launch_stage_t<Pol, false, true, 16, 15>(g, stage, src, uold, dst, ctrl, s);
    return launch_stage_t<Pol, false, false, 16, 15>(g, stage, src, uold, dst, ctrl, s);
  }
  switch (g.tile) {
    case TILE_16x16: return launch_stage_t<Pol, true, true, 16, 16>(g, stage, src, uold, dst, ctrl, s);
    case TILE_32x8: return launch_stage_t<Pol, true, true, 32, 8>(g, stage, src, uold, dst, ctrl, s);
    case TILE_16x8: return launch_stage_t<Pol, true, true, 16, 8>(g, stage, src, uold, dst, ctrl, s);
    case TILE_32x7: return launch_stage_t<Pol, true, true, 32, 7>(g, stage, src, uold, dst, ctrl, s);
    case TILE_16x7: return launch_stage_t<Pol, true, true, 16, 7>(g, stage, src, uold, dst, ctrl, s);
    default: return launch_stage_t<Pol, true, true, 16, 15>(g, stage, src, uold, dst, ctrl, s);
  }
}

template <class T>
static hipError_t launch_finish(void* ctrl, hipStream_t s) {
  hipLaunchKernelGGL(k_finish<T>, dim3(1), dim3(64), 0, s, (Ctrl<T>*)ctrl);
  return hipGetLastError();
}

template <class Pol, int TX, int TY>
static hipError_t launch_step_tile(const GridDesc& g, void* u, void* u1, void* ctrl,
                                   hipStream_t s, hipEvent_t* ev) {
  if (g.muscl && g.hlld) return launch_step_t<Pol, true, true, TX, TY>(g, u, u1, ctrl, s, ev);
  if (g.muscl) return launch_step_t<Pol, true, false, TX, TY>(g, u, u1, ctrl, s, ev);
  if (g.hlld) return launch_step_t<Pol, false, true, TX, TY>(g, u, u1, ctrl, s, ev);
  return launch_step_t<Pol, false, false, TX, TY>(g, u, u1, ctrl, s, ev);
}

// Every tile shape is instantiated for the production scheme (MUSCL +
// HLLD); the other order/solver combinations use the default tile.
template <class Pol>
static hipError_t launch_step(const GridDesc& g, void* u, void* u1, void* ctrl,
                              hipStream_t s, hipEvent_t* ev) {
  if (!(g.muscl && g.hlld)) return launch_step_tile<Pol, 16, 15>(g, u, u1, ctrl, s, ev);
  switch (g.tile) {
    case TILE_16x16: return launch_step_t<Pol, true, true, 16, 16>(g, u, u1, ctrl, s, ev);
    case TILE_32x8: return launch_step_t<Pol, true, true, 32, 8>(g, u, u1, ctrl, s, ev);
    case TILE_16x8: return launch_step_t<Pol, true, true, 16, 8>(g, u, u1, ctrl, s, ev);
    case TILE_32x7: return launch_step_t<Pol, true, true, 32, 7>(g, u, u1, ctrl, s, ev);
    case TILE_16x7: return launch_step_t<Pol, true, true, 16, 7>(g, u, u1, ctrl, s, ev);
    default: return launch_step_t<Pol, true, true, 16, 15>(g, u, u1, ctrl, s, ev);
  }
}

template <class Pol>
static hipError_t launch_speed(const GridDesc& g, const void* state, void* ctrl,
                               hipStream_t s) {
  const long long n = (long long)g.nx * g.ny;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_speed<Pol>, dim3((unsigned)blocks), dim3(256), 0, s,
                     (const typename Pol::S*)state, g.plane, g.nx, g.ny, g.tol,
                     (Ctrl<typename Pol::T>*)ctrl);
  return hipGetLastError();
}

template <class T>
static hipError_t launch_commit(void* ctrl, hipStream_t s) {
  hipLaunchKernelGGL(k_commit_speed<T>, dim3(1), dim3(64), 0, s, (Ctrl<T>*)ctrl);
  return hipGetLastError();
}

template <class T>
static hipError_t launch_target(void* ctrl, double t_end, hipStream_t s) {
  hipLaunchKernelGGL(k_set_target<T>, dim3(1), dim3(64), 0, s, (Ctrl<T>*)ctrl, t_end);
  return hipGetLastError();
}

template <class Pol>
static int ens_fill(void* host_desc, int i, void* u, void* u1, size_t plane, int nx, int ny,
                    void* ctrl, int tile) {
  using T = typename Pol::T;
  int tx, ty;
  tile_dims(tile, tx, ty);
  InstDesc<Pol>& d = ((InstDesc<Pol>*)host_desc)[i];
  d.u = (typename Pol::S*)u;
  d.u1 = (typename Pol::S*)u1;
  d.plane = plane;
  d.nx = nx;
  d.ny = ny;
  d.tiles_x = (nx + tx - 1) / tx;
  d.ntiles = d.tiles_x * ((ny + ty - 1) / ty);
  const T dx = ((T)1.0 - (T)0.0) / (T)nx, dy = ((T)1.0 - (T)0.0) / (T)ny;
  d.pow2 = ((nx & (nx - 1)) == 0 && (ny & (ny - 1)) == 0) ? 1 : 0;
  d.inv_dx = (double)((T)1.0 / dx);
  d.inv_dy = (double)((T)1.0 / dy);
  d.ctrl = (Ctrl<T>*)ctrl;
  return d.ntiles;
}

template <class Pol>
static EnsArgs<Pol> ens_args(const EnsLaunch& L) {
  EnsArgs<Pol> e;
  e.desc = (const InstDesc<Pol>*)L.desc;
  e.prefix = L.prefix;
  e.n = L.n;
  e.total = L.total;
  e.queue = L.queue;
  e.tol = L.tol;
  e.alpha = L.alpha;
  e.lowbeta_ratio = (float)L.lowbeta_ratio;
  e.redo_items[0] = e.redo_items[1] = nullptr;
  e.redo_all = 0;
  e.live = L.live;
  return e;
}

// FP64 fallback of an ensemble stage.  The descriptors were filled by the
// mixed launchers; InstDesc<MixedPol> and InstDesc<ExactPol<double>> have
// the same layout (double state planes, Ctrl<double>).
template <class Pol, bool MUSCL, bool HLLD, int TX, int TY>
static hipError_t ens_redo_t(const EnsLaunch& L, int stage, hipStream_t s) {
  constexpr int nthr = (TX * TY + 63) / 64 * 64;
  const EnsArgs<Pol> e = ens_args<Pol>(L);
  const int* list = L.redo_items[stage - 1];
  const int grid = L.total < kRedoBlocks ? L.total : kRedoBlocks;
  if (stage == 1)
    hipLaunchKernelGGL((k_queue_redo<Pol, MUSCL, HLLD, 1, TX, TY>), dim3(grid), dim3(nthr), 0, s, e,
                       list);
  else
    hipLaunchKernelGGL((k_queue_redo<Pol, MUSCL, HLLD, 2, TX, TY>), dim3(grid), dim3(nthr), 0, s, e,
                       list);
  return hipGetLastError();
}

template <class Pol>
static hipError_t ens_redo(const EnsLaunch& L, int stage, hipStream_t s) {
  if (!L.redo_items[stage - 1] || L.total < 1) return hipErrorInvalidValue;
  if (!(L.muscl && L.hlld)) {
    if (L.muscl) return ens_redo_t<Pol, true, false, 16, 15>(L, stage, s);
    if (L.hlld) return ens_redo_t<Pol, false, true, 16, 15>(L, stage, s);
    return ens_redo_t<Pol, false, false, 16, 15>(L, stage, s);
  }
  switch (L.tile) {
    case TILE_16x16: return ens_redo_t<Pol, true, true, 16, 16>(L, stage, s);
    case TILE_32x8: return ens_redo_t<Pol, true, true, 32, 8>(L, stage, s);
    case TILE_16x8: return ens_redo_t<Pol, true, true, 16, 8>(L, stage, s);
    case TILE_32x7: return ens_redo_t<Pol, true, true, 32, 7>(L, stage, s);
    case TILE_16x7: return ens_redo_t<Pol, true, true, 16, 7>(L, stage, s);
    default: return ens_redo_t<Pol, true, true, 16, 15>(L, stage, s);
  }
}

template <class Pol, bool MUSCL, bool HLLD, int TX, int TY>
static hipError_t ens_step_t(const EnsLaunch& L, hipStream_t s, hipEvent_t* ev) {
  using T = typename Pol::T;
  constexpr int nthr = (TX * TY + 63) / 64 * 64;
  EnsArgs<Pol> e = ens_args<Pol>(L);
  const bool redo = Pol::kRedo && L.redo && L.redo_items[0] && L.redo_items[1];
  if (redo) {
    e.redo_items[0] = L.redo_items[0];
    e.redo_items[1] = L.redo_items[1];
    e.redo_all = L.redo_all;
  }
  if (!L.live || L.total < 1) return hipErrorInvalidValue;
  const bool queue = L.kernel == ENS_QUEUE;
  // Flat: one block per work item of every instance, rounded up to whole
  // rounds of the 8 XCDs (item x * per + k is block 8k + x).
  int grid = (L.total + 7) / 8 * 8;
  if (queue) {
    grid = L.grid_blocks;
    if (grid <= 0) {
      static int per_cu = 0;
      if (!per_cu) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void*>(&k_stage_queue<Pol, MUSCL, HLLD, 1, TX, TY>),
                nthr, 0) != hipSuccess || per_cu < 1)
          per_cu = 1;
      }
      grid = per_cu * (L.cus > 0 ? L.cus : 256);
    }
  }
  hipLaunchKernelGGL((k_ens_live<Pol, 1>), dim3(1), dim3(256), 0, s, e);
  if (ev) (void)hipEventRecord(ev[0], s);
  if (queue)
    hipLaunchKernelGGL((k_stage_queue<Pol, MUSCL, HLLD, 1, TX, TY>), dim3(grid), dim3(nthr), 0, s, e);
  else
    hipLaunchKernelGGL((k_stage_ens<Pol, MUSCL, HLLD, 1, TX, TY>), dim3(grid), dim3(nthr), 0, s, e);
  if (ev) (void)hipEventRecord(ev[1], s);
  if (redo) (void)L.redo(L, 1, s);
  hipLaunchKernelGGL((k_ens_live<Pol, 2>), dim3(1), dim3(256), 0, s, e);
  if (queue)
    hipLaunchKernelGGL((k_stage_queue<Pol, MUSCL, HLLD, 2, TX, TY>), dim3(grid), dim3(nthr), 0, s, e);
  else
    hipLaunchKernelGGL((k_stage_ens<Pol, MUSCL, HLLD, 2, TX, TY>), dim3(grid), dim3(nthr), 0, s, e);
  if (redo) (void)L.redo(L, 2, s);
  if (ev) (void)hipEventRecord(ev[2], s);
  hipLaunchKernelGGL(k_finish_multi<T>, dim3((L.n + 63) / 64 > 0 ? (L.n + 63) / 64 : 1), dim3(64),
                     0, s, (Ctrl<T>*)L.ctrls, L.n, L.queue);
  if (ev) (void)hipEventRecord(ev[3], s);
  return hipGetLastError();
}

template <class Pol>
static hipError_t ens_step(const EnsLaunch& L, hipStream_t s, hipEvent_t* ev) {
  if (!(L.muscl && L.hlld)) {
    if (L.muscl) return ens_step_t<Pol, true, false, 16, 15>(L, s, ev);
    if (L.hlld) return ens_step_t<Pol, false, true, 16, 15>(L, s, ev);
    return ens_step_t<Pol, false, false, 16, 15>(L, s, ev);
  }
  switch (L.tile) {
    case TILE_16x16: return ens_step_t<Pol, true, true, 16, 16>(L, s, ev);
    case TILE_32x8: return ens_step_t<Pol, true, true, 32, 8>(L, s, ev);
    case TILE_16x8: return ens_step_t<Pol, true, true, 16, 8>(L, s, ev);
    case TILE_32x7: return ens_step_t<Pol, true, true, 32, 7>(L, s, ev);
    case TILE_16x7: return ens_step_t<Pol, true, true, 16, 7>(L, s, ev);
    default: return ens_step_t<Pol, true, true, 16, 15>(L, s, ev);
  }
}

template <class T>
static hipError_t ens_set_target(void* ctrls, int n, double t_end, unsigned* queue, hipStream_t s) {
  hipLaunchKernelGGL(k_set_target_multi<T>, dim3((n + 63) / 64 > 0 ? (n + 63) / 64 : 1), dim3(64), 0,
                     s, (Ctrl<T>*)ctrls, n, t_end, queue);
  return hipGetLastError();
}

// Launchers of an allocation variant: single grids, MUSCL + HLLD, 16x15
// tiles, one stage per launch (other configurations are refused).
template <class Pol>
static hipError_t launch_step_variant(const GridDesc& g, void* u, void* u1, void* ctrl,
                                      hipStream_t s, hipEvent_t* ev) {
  if (!(g.muscl && g.hlld) || g.tile != TILE_16x15) return hipErrorInvalidValue;
  return launch_step_t<Pol, true, true, 16, 15>(g, u, u1, ctrl, s, ev);
}

template <class Pol>
static hipError_t launch_stage_variant(const GridDesc& g, int stage, const void* src,
                                       const void* uold, void* dst, void* ctrl, hipStream_t s) {
  if (!(g.muscl && g.hlld) || g.tile != TILE_16x15) return hipErrorInvalidValue;
  return launch_stage_t<Pol, true, true, 16, 15>(g, stage, src, uold, dst, ctrl, s);
}

template <class Pol>
static Launchers make_variant_launchers(int mode) {
  Launchers L{};
  L.mode = mode;
  L.ctrl_bytes = sizeof(Ctrl<double>);
  L.time_is_float = 0;
  L.state_elem = sizeof(double);
  L.state_planes = 9;
  L.speed = &launch_speed<Pol>;
  L.commit = &launch_commit<double>;
  L.set_target = &launch_target<double>;
  L.step = &launch_step_variant<Pol>;
  L.stage = &launch_stage_variant<Pol>;
  L.finish = &launch_finish<double>;
  return L;
}

#include "mhd_fused.inc.hip"

template <class Pol>
static Launchers make_launchers(int mode) {
  Launchers L;
  L.mode = mode;
  L.ctrl_bytes = sizeof(Ctrl<typename Pol::T>);
  L.time_is_float = sizeof(typename Pol::T) == 4 ? 1 : 0;
  L.state_elem = sizeof(typename Pol::S);
  L.state_planes = Pol::kPlanes;
  L.speed = &launch_speed<Pol>;
  L.commit = &launch_commit<typename Pol::T>;
  L.set_target = &launch_target<typename Pol::T>;
  L.step = &launch_step<Pol>;
  L.step_fused = &launch_step_fused<Pol>;
  L.stage = &launch_stage<Pol>;
  L.finish = &launch_finish<typename Pol::T>;
  if constexpr (std::is_same<Pol, ExactPol<double>>::value) {
    L.stage_list = &launch_list<Pol>;
    L.stage_listn = &launch_listn<Pol>;
    L.ens_redo = &ens_redo<Pol>;
  } else {
    L.stage_list = nullptr;
    L.stage_listn = nullptr;
    L.ens_redo = nullptr;
  }
  L.desc_bytes = sizeof(InstDesc<Pol>);
  L.ens_fill = &ens_fill<Pol>;
  L.ens_step = &ens_step<Pol>;
  L.ens_set_target = &ens_set_target<typename Pol::T>;
  return L;
}

}  // namespace MHD_NS
}  // namespace mhd

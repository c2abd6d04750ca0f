// mhd_kernels.hip -- CDNA4 (gfx950) kernels of the 2D MHD-GLM SSPRK2 step.
//
// One SSPRK2 step (mhd2d_glm_cpu.cpp:513-569) is three launches on one
// stream, with no host synchronisation:
//   k_stage<1>: U1 = U + dt L(U)                      + positivity of U1
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
// Kernels stop at entry once the device control block says the run is
// done or halted, so a batch of steps (or a replayed graph) can be launched
// without knowing in advance where t_end or a failure falls.
#include "mhd_device.hpp"
#include "mhd_kernels.h"

#include <climits>

namespace mhd {

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
template <class R> __device__ __forceinline__ R bits_speed(unsigned long long b);
template <> __device__ __forceinline__ double bits_speed<double>(unsigned long long b) {
  return __longlong_as_double((long long)b);
}
template <> __device__ __forceinline__ float bits_speed<float>(unsigned long long b) {
  return __uint_as_float((unsigned)b);
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
template <class R>
__device__ __forceinline__ R cell_speed(const Phys<R>& c, const R P[9], bool& valid) {
  const R sx = fabs_(P[1]) + fast_speed(c, P, P[5]);
  const R sy = fabs_(P[2]) + fast_speed(c, P, P[6]);
  const R m = smax(sx, sy);
  valid = (m == m);
  return (m > (R)0.0) ? m : (R)0.0;
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

// ---------------------------------------------------------------------------
// Fused stage kernel.
// ---------------------------------------------------------------------------
template <class R>
struct StageArgs {
  const R* src;   // state L() is evaluated on: U (stage 1) or U1 (stage 2)
  const R* uold;  // stage 2: U^n (same buffer as dst)
  R* dst;         // stage 1: U1; stage 2: U
  size_t plane;
  int nx, ny, tiles_x, ntiles;
  double tol;
  R alpha;
  Ctrl<R>* ctrl;
};

template <class R, bool MUSCL, bool HLLD, int STAGE, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_stage(StageArgs<R> a) {
  constexpr int NT = TX * TY;
  constexpr int HX = TX + 4, HY = TY + 4, HN = HX * HY;
  constexpr int NXF = (TX + 1) * TY, NYF = TX * (TY + 1), NF = NXF + NYF;
  constexpr int NW = NT / 64;
  static_assert(NT % 64 == 0, "tile must be a whole number of waves");
  __shared__ R sP[9 * HN];
  __shared__ R sF[9 * NF];
  __shared__ unsigned long long sRed[3 * NW];

  Ctrl<R>* ctrl = a.ctrl;
  if (ctrl->run != RUN_ACTIVE) return;
  if (STAGE == 2 && ctrl->bad[0] != INT_MAX) return;
  const R ch = ctrl->ch, dt = ctrl->dt, dx = ctrl->dx, dy = ctrl->dy;
  const Phys<R> c = make_phys<R>(a.tol);
  const int nx = a.nx, ny = a.ny;
  const size_t ps = a.plane;
  const int tile = xcd_tile((int)blockIdx.x, a.ntiles);
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int i0 = tx * TX, j0 = ty * TY;
  const int tid = (int)threadIdx.x;

  // Phase A: tile + halo (no corners) -> primitives in LDS.
  for (int h = tid; h < HN; h += NT) {
    const int r = h / HX, cc = h - r * HX;
    if ((r < 2 || r >= TY + 2) && (cc < 2 || cc >= TX + 2)) continue;
    const size_t g = (size_t)wrap(j0 + r - 2, ny) * nx + wrap(i0 + cc - 2, nx);
    R U[9];
#pragma unroll
    for (int k = 0; k < 9; k++) U[k] = a.src[k * ps + g];
    if (MUSCL) {
      R P[9];
      cons_to_prim<R, 9>(c, U, P);
#pragma unroll
      for (int k = 0; k < 9; k++) sP[k * HN + h] = P[k];
    } else {
#pragma unroll
      for (int k = 0; k < 9; k++) sP[k * HN + h] = U[k];
    }
  }
  __syncthreads();

  // Phase B: faces.  Faces [0, NXF) are x faces (row r, face column fc
  // between local cells fc-1 and fc); [NXF, NF) are y faces (column cc,
  // face row fr between local rows fr-1 and fr).  A y face is evaluated in
  // the x frame by reading the (1,2) and (5,6) slots swapped.
  unsigned degen = 0, nonphys = 0;
  for (int f = tid; f < NF; f += NT) {
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
    const int q1 = (ydir ? 2 : 1) * HN, q2 = (ydir ? 1 : 2) * HN;
    const int q5 = (ydir ? 6 : 5) * HN, q6 = (ydir ? 5 : 6) * HN;
    auto load = [&](int cell, R out[9]) {
      out[0] = sP[cell];
      out[1] = sP[q1 + cell];
      out[2] = sP[q2 + cell];
      out[3] = sP[3 * HN + cell];
      out[4] = sP[4 * HN + cell];
      out[5] = sP[q5 + cell];
      out[6] = sP[q6 + cell];
      out[7] = sP[7 * HN + cell];
      out[8] = sP[8 * HN + cell];
    };
    R UL[9], UR[9];
    if (MUSCL) {
      R Pa[9], Pb[9], Pc[9], Pf[9];
      load(base, Pa);
      load(base + stride, Pb);
      load(base + 2 * stride, Pc);
      reconstruct_face<R, true>(c, Pa, Pb, Pc, Pf);
      prim_to_cons(c, Pf, UL);
      load(base + 3 * stride, Pa);
      reconstruct_face<R, false>(c, Pb, Pc, Pa, Pf);
      prim_to_cons(c, Pf, UR);
    } else {
      load(base + stride, UL);
      load(base + 2 * stride, UR);
    }
    R F[9];
    const int st = glm_flux_x<R, HLLD>(c, UL, UR, ch, F);
    if (HLLD && owned) {
      degen += (st == HLLD_DEGENERATE);
      nonphys += (st == HLLD_NONPHYSICAL);
    }
    const int p1 = (ydir ? 2 : 1) * NF, p2 = (ydir ? 1 : 2) * NF;
    const int p5 = (ydir ? 6 : 5) * NF, p6 = (ydir ? 5 : 6) * NF;
    sF[f] = F[0];
    sF[p1 + f] = F[1];
    sF[p2 + f] = F[2];
    sF[3 * NF + f] = F[3];
    sF[4 * NF + f] = F[4];
    sF[p5 + f] = F[5];
    sF[p6 + f] = F[6];
    sF[7 * NF + f] = F[7];
    sF[8 * NF + f] = F[8];
  }
  __syncthreads();

  // Phase C: one cell per thread.  rhs (mhd2d_glm_cpu.cpp:273-284), the
  // stage update (528-535 / 549-557) and the positivity check (328-345).
  const int r = tid / TX, cc = tid - r * TX;
  const int gi = i0 + cc, gj = j0 + r;
  unsigned long long sbits = 0;
  unsigned svalid = 0;
  if (gi < nx && gj < ny) {
    const int xl = r * (TX + 1) + cc, xr = xl + 1;
    const int yb = NXF + r * TX + cc, yt = yb + TX;
    const R damping = ch / smax(a.alpha, (R)1.0e-12);
    const size_t g = (size_t)gj * nx + gi;
    R out[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      R rhs = -(sF[k * NF + xr] - sF[k * NF + xl]) / dx -
              (sF[k * NF + yt] - sF[k * NF + yb]) / dy;
      if (k == PSI) rhs -= damping * sP[8 * HN + (r + 2) * HX + cc + 2];
      if (STAGE == 1) {
        out[k] = a.src[k * ps + g] + dt * rhs;
      } else {
        out[k] = (R)0.5 * a.uold[k * ps + g] + (R)0.5 * (a.src[k * ps + g] + dt * rhs);
      }
    }
#pragma unroll
    for (int k = 0; k < 9; k++) a.dst[k * ps + g] = out[k];
    R P[9];
    if (!state_usable(c, out, P)) atomicMin(&ctrl->bad[STAGE - 1], (int)g);
    if (STAGE == 2) {
      bool valid;
      const R s = cell_speed(c, P, valid);
      sbits = speed_bits(s);
      svalid = valid ? 1u : 0u;
    }
  }

  // Block reductions: one atomic per block and counter.
  const int lane = tid & 63, wid = tid >> 6;
  const unsigned wd = wave_sum_u32(degen), wn = wave_sum_u32(nonphys);
  unsigned long long wm = 0;
  unsigned wv = 0;
  if (STAGE == 2) {
    wm = wave_max_u64(sbits);
    wv = wave_sum_u32(svalid);
  }
  if (lane == 0) {
    sRed[wid] = ((unsigned long long)wn << 32) | wd;
    sRed[NW + wid] = wm;
    sRed[2 * NW + wid] = wv;
  }
  __syncthreads();
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
    if (STAGE == 2) {
      if (m) atomicMax(&ctrl->speed_acc, m);
      if (v) atomicOr(&ctrl->valid_acc, 1);
    }
  }
}

// Signal-speed maximum of a state (initial state / after set_state).
template <class R>
__global__ __launch_bounds__(256) void k_speed(const R* src, size_t plane,
                                               int nx, int ny, double tol,
                                               Ctrl<R>* ctrl) {
  __shared__ unsigned long long sRed[2 * 4];
  const Phys<R> c = make_phys<R>(tol);
  const long long n = (long long)nx * ny;
  unsigned long long best = 0;
  unsigned valid_any = 0;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n;
       q += (long long)gridDim.x * 256) {
    R U[9], P[9];
#pragma unroll
    for (int k = 0; k < 9; k++) U[k] = src[k * plane + q];
    cons_to_prim<R, 9>(c, U, P);
    bool valid;
    const R s = cell_speed(c, P, valid);
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
    if (m) atomicMax(&ctrl->speed_acc, m);
    if (v) atomicOr(&ctrl->valid_acc, 1);
  }
}

// c_h and dt of the coming step (mhd2d_glm_cpu.cpp:514-524).  The rate
// maximum of maxSignalRate (307-326) is exact in closed form: every cell's
// speed is <= c_h, so max(s, c_h) == c_h bitwise and each cell's rate is
// c_h/dx + c_h/dy; NaN cells never enter the running maximum, so with no
// non-NaN cell the maximum stays 0.
template <class R>
__device__ void prepare_step(Ctrl<R>* ctrl) {
  ctrl->bad[0] = INT_MAX;
  ctrl->bad[1] = INT_MAX;
  if (!(ctrl->t < ctrl->t_end)) {
    ctrl->run = RUN_DONE;
    return;
  }
  const R speed = bits_speed<R>(ctrl->speed_cur);
  const R ch = smax(speed, (R)1.0e-12);
  const R r = ch / ctrl->dx + ch / ctrl->dy;
  const R rate = ctrl->valid_cur ? smax((R)0.0, r) : (R)0.0;
  ctrl->ch = ch;
  if (rate <= (R)0.0) {
    ctrl->run = RUN_HALTED;
    ctrl->code = 2;
    ctrl->fail_step = ctrl->step;
    ctrl->fail_t = ctrl->t;
    ctrl->fail_ch = ch;
    return;
  }
  R dt = ctrl->cfl / rate;
  if (ctrl->t + dt > ctrl->t_end) dt = ctrl->t_end - ctrl->t;
  ctrl->dt = dt;
}

template <class R>
__global__ void k_commit_speed(Ctrl<R>* ctrl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ctrl->speed_cur = ctrl->speed_acc;
  ctrl->valid_cur = ctrl->valid_acc;
  ctrl->speed_acc = 0;
  ctrl->valid_acc = 0;
}

template <class R>
__global__ void k_set_target(Ctrl<R>* ctrl, double t_end) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (ctrl->run == RUN_HALTED) return;
  ctrl->t_end = (R)t_end;
  ctrl->run = RUN_ACTIVE;
  prepare_step(ctrl);
}

template <class R>
__global__ void k_finish(Ctrl<R>* ctrl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
  ctrl->step += 1;
  ctrl->speed_cur = ctrl->speed_acc;
  ctrl->valid_cur = ctrl->valid_acc;
  ctrl->speed_acc = 0;
  ctrl->valid_acc = 0;
  prepare_step(ctrl);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
template <class R, bool MUSCL, bool HLLD, int TX, int TY>
static hipError_t launch_step_t(const GridDesc& g, R* u, R* u1, Ctrl<R>* ctrl,
                                hipStream_t s, hipEvent_t* ev) {
  const int tiles_x = (g.nx + TX - 1) / TX, tiles_y = (g.ny + TY - 1) / TY;
  StageArgs<R> a;
  a.plane = g.plane;
  a.nx = g.nx;
  a.ny = g.ny;
  a.tiles_x = tiles_x;
  a.ntiles = tiles_x * tiles_y;
  a.tol = g.tol;
  a.alpha = (R)g.alpha;
  a.ctrl = ctrl;
  a.src = u;
  a.uold = u;
  a.dst = u1;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL((k_stage<R, MUSCL, HLLD, 1, TX, TY>), dim3(a.ntiles),
                     dim3(TX * TY), 0, s, a);
  if (ev) (void)hipEventRecord(ev[1], s);
  a.src = u1;
  a.uold = u;
  a.dst = u;
  hipLaunchKernelGGL((k_stage<R, MUSCL, HLLD, 2, TX, TY>), dim3(a.ntiles),
                     dim3(TX * TY), 0, s, a);
  if (ev) (void)hipEventRecord(ev[2], s);
  hipLaunchKernelGGL(k_finish<R>, dim3(1), dim3(64), 0, s, ctrl);
  if (ev) (void)hipEventRecord(ev[3], s);
  return hipGetLastError();
}

template <class R, int TX, int TY>
static hipError_t launch_step_tile(const GridDesc& g, R* u, R* u1,
                                   Ctrl<R>* ctrl, hipStream_t s, hipEvent_t* ev) {
  if (g.muscl && g.hlld) return launch_step_t<R, true, true, TX, TY>(g, u, u1, ctrl, s, ev);
  if (g.muscl) return launch_step_t<R, true, false, TX, TY>(g, u, u1, ctrl, s, ev);
  if (g.hlld) return launch_step_t<R, false, true, TX, TY>(g, u, u1, ctrl, s, ev);
  return launch_step_t<R, false, false, TX, TY>(g, u, u1, ctrl, s, ev);
}

template <class R>
static hipError_t launch_step_any(const GridDesc& g, R* u, R* u1, Ctrl<R>* ctrl,
                                  hipStream_t s, hipEvent_t* ev) {
  switch (g.tile) {
    case TILE_32x8: return launch_step_tile<R, 32, 8>(g, u, u1, ctrl, s, ev);
    case TILE_16x8: return launch_step_tile<R, 16, 8>(g, u, u1, ctrl, s, ev);
    default: return launch_step_tile<R, 16, 16>(g, u, u1, ctrl, s, ev);
  }
}

template <class R>
hipError_t launch_step(const GridDesc& g, R* u, R* u1, Ctrl<R>* ctrl,
                       hipStream_t s) {
  return launch_step_any<R>(g, u, u1, ctrl, s, nullptr);
}

template <class R>
hipError_t launch_step_timed(const GridDesc& g, R* u, R* u1, Ctrl<R>* ctrl,
                             hipStream_t s, hipEvent_t ev[4]) {
  return launch_step_any<R>(g, u, u1, ctrl, s, ev);
}

template <class R>
hipError_t launch_speed(const GridDesc& g, const R* state, Ctrl<R>* ctrl,
                        hipStream_t s) {
  const long long n = (long long)g.nx * g.ny;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_speed<R>, dim3((unsigned)blocks), dim3(256), 0, s, state,
                     g.plane, g.nx, g.ny, g.tol, ctrl);
  return hipGetLastError();
}

template <class R>
hipError_t launch_commit_speed(Ctrl<R>* ctrl, hipStream_t s) {
  hipLaunchKernelGGL(k_commit_speed<R>, dim3(1), dim3(64), 0, s, ctrl);
  return hipGetLastError();
}

template <class R>
hipError_t launch_set_target(Ctrl<R>* ctrl, double t_end, hipStream_t s) {
  hipLaunchKernelGGL(k_set_target<R>, dim3(1), dim3(64), 0, s, ctrl, t_end);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Operator-level batch kernels (one state pair per thread).
// ---------------------------------------------------------------------------
template <class R, bool HLLD>
__global__ void k_glm_flux_batch(int n, int dir, const R* UL, const R* UR, R ch,
                                 double tol, R* F, int* status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const Phys<R> c = make_phys<R>(tol);
  // Read in the x frame: for a y face swap the (1,2) and (5,6) slots.
  const int perm[9] = {0, dir ? 2 : 1, dir ? 1 : 2, 3, 4, dir ? 6 : 5,
                       dir ? 5 : 6, 7, 8};
  R a[9], b[9], f[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    a[k] = UL[(size_t)s * 9 + perm[k]];
    b[k] = UR[(size_t)s * 9 + perm[k]];
  }
  const int st = glm_flux_x<R, HLLD>(c, a, b, ch, f);
#pragma unroll
  for (int k = 0; k < 9; k++) F[(size_t)s * 9 + perm[k]] = f[k];
  if (status) status[s] = st;
}

template <class R>
__global__ void k_cons_to_prim_batch(int n, const R* U, R* W) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const Phys<R> c = make_phys<R>(1.0e-4);
  R u[9], w[9];
#pragma unroll
  for (int k = 0; k < 9; k++) u[k] = U[(size_t)s * 9 + k];
  cons_to_prim<R, 9>(c, u, w);
#pragma unroll
  for (int k = 0; k < 9; k++) W[(size_t)s * 9 + k] = w[k];
}

template <class R>
__global__ void k_fast_speed_batch(int n, const R* W, int dir, R* out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const Phys<R> c = make_phys<R>(1.0e-4);
  R w[9];
#pragma unroll
  for (int k = 0; k < 9; k++) w[k] = W[(size_t)s * 9 + k];
  out[s] = fast_speed(c, w, dir ? w[6] : w[5]);
}

template <class R>
hipError_t launch_glm_flux_batch(int n, int dir, const R* UL, const R* UR,
                                 R ch, int hlld, double tol, R* F, int* status,
                                 hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int blocks = (n + 255) / 256;
  if (hlld)
    hipLaunchKernelGGL((k_glm_flux_batch<R, true>), dim3(blocks), dim3(256), 0,
                       s, n, dir, UL, UR, ch, tol, F, status);
  else
    hipLaunchKernelGGL((k_glm_flux_batch<R, false>), dim3(blocks), dim3(256), 0,
                       s, n, dir, UL, UR, ch, tol, F, status);
  return hipGetLastError();
}

template <class R>
hipError_t launch_cons_to_prim_batch(int n, const R* U, R* W, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cons_to_prim_batch<R>, dim3((n + 255) / 256), dim3(256),
                     0, s, n, U, W);
  return hipGetLastError();
}

template <class R>
hipError_t launch_fast_speed_batch(int n, const R* W, int dir, R* out,
                                   hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fast_speed_batch<R>, dim3((n + 255) / 256), dim3(256), 0,
                     s, n, W, dir, out);
  return hipGetLastError();
}

#define MHD_INSTANTIATE(R)                                                       \
  template hipError_t launch_speed<R>(const GridDesc&, const R*, Ctrl<R>*,      \
                                      hipStream_t);                              \
  template hipError_t launch_commit_speed<R>(Ctrl<R>*, hipStream_t);             \
  template hipError_t launch_set_target<R>(Ctrl<R>*, double, hipStream_t);       \
  template hipError_t launch_step<R>(const GridDesc&, R*, R*, Ctrl<R>*,          \
                                     hipStream_t);                               \
  template hipError_t launch_step_timed<R>(const GridDesc&, R*, R*, Ctrl<R>*,    \
                                           hipStream_t, hipEvent_t*);            \
  template hipError_t launch_glm_flux_batch<R>(int, int, const R*, const R*, R,  \
                                               int, double, R*, int*,            \
                                               hipStream_t);                     \
  template hipError_t launch_cons_to_prim_batch<R>(int, const R*, R*,            \
                                                   hipStream_t);                 \
  template hipError_t launch_fast_speed_batch<R>(int, const R*, int, R*,         \
                                                 hipStream_t);

MHD_INSTANTIATE(double)
MHD_INSTANTIATE(float)

}  // namespace mhd

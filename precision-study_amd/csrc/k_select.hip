// k_select.hip -- kernels of the learned precision selector.
//
// The selector decides, per stage-kernel tile and step, whether the FP32
// flux stage (mixed form) stays inside the error budget or the tile must be
// evaluated by the FP64 kernel.  It is a small MLP 12 -> 64 -> 64 -> 1
// (ReLU, sigmoid) with BF16 operands and FP32 accumulation on MFMA
// (v_mfma_f32_16x16x16_bf16), trained on the GPU by the kernels below on
// labels the solver itself produces: the per-tile difference between an
// FP32-flux and an FP64-flux stage update of the same state.
//
//   k_tile_features  12 features per tile from the FP64 state
//   k_sumsq / k_tile_errors   label source (scaled max difference per tile)
//   k_mlp_infer      probabilities + FP64 flags (with sticky audit flags)
//   k_mlp_grad       forward (BF16 MFMA) + backward (FP32 MFMA tiles) of a
//                    minibatch, one gradient row per block (no atomics)
//   k_sum_partials   the block rows summed in block order (deterministic)
//   k_adam           the same sum, then Adam on the FP32 master weights
//   k_route          FP64 and mixed tile lists of the current flags
//   k_audit_pick     the step's audit sample (and the control-block snapshot)
//   k_tile_errors    (audit mode) audited FP32 tiles against their FP64
//                    recomputation; misses become sticky FP64 tiles
#include <hip/hip_runtime.h>

#include "mhd_kernels.h"
#include "mhd_select.h"

namespace mhd {
namespace sel {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Parameter layout (FP32 masters): W1[16][64], b1[64], W2[64][64], b2[64],
// w3[64], b3.
constexpr int kW1 = 0, kB1 = kW1 + NIN * H, kW2 = kB1 + H, kB2 = kW2 + H * H, kW3 = kB2 + H,
              kB3 = kW3 + H;
static_assert(kB3 + 1 == kParams, "parameter layout");

__device__ __forceinline__ short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(float f) {
  return __uint_as_float(((unsigned)(unsigned short)to_bf16(f)) << 16);
}

// ---------------------------------------------------------------------------
// Block reductions (256 threads = 4 waves).
// ---------------------------------------------------------------------------
template <class Op>
__device__ __forceinline__ double block_reduce(double v, double* red, Op op) {
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  return op(op(red[0], red[1]), op(red[2], red[3]));
}
struct Max { __device__ double operator()(double a, double b) const { return a < b ? b : a; } };
struct Min { __device__ double operator()(double a, double b) const { return b < a ? b : a; } };
struct Sum { __device__ double operator()(double a, double b) const { return a + b; } };

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// What a cell contributes to the features, from its conserved FP64 state:
// density (floored), |m|^2, field, pressure p = (gamma-1)(E - |m|^2/(2 rho) -
// |B|^2/2).  Pressure and the neighbour differences stay FP64 (they cancel);
// the ratios built from them are FP32, the features' own precision.
struct CellFeat {
  double rho, m2, p, bx, by, bz;
};
__device__ __forceinline__ CellFeat cell_at(const double* u, size_t ps, size_t g) {
  const double gm1 = 5.0 / 3.0 - 1.0;
  CellFeat c;
  c.rho = fmax(u[g], 1e-12);
  const double mx = u[ps + g], my = u[2 * ps + g], mz = u[3 * ps + g];
  c.m2 = mx * mx + my * my + mz * mz;
  c.bx = u[5 * ps + g];
  c.by = u[6 * ps + g];
  c.bz = u[7 * ps + g];
  const double b2 = c.bx * c.bx + c.by * c.by + c.bz * c.bz;
  c.p = gm1 * (u[4 * ps + g] - 0.5 * c.m2 / c.rho - 0.5 * b2);
  return c;
}

// Wave reductions of a float: DPP within each 16-lane row (quad xor 1 and
// 2, row rotations by 4 and 8), then the four rows' lanes 0, 16, 32, 48 by
// readlane.  Every lane must be active; the result is wave-uniform.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <class Op>
__device__ __forceinline__ float wave_reduce(float v, Op op) {
  v = op(v, dpp_f<0xB1>(v));   // quad_perm [1,0,3,2]
  v = op(v, dpp_f<0x4E>(v));   // quad_perm [2,3,0,1]
  v = op(v, dpp_f<0x124>(v));  // row_ror:4
  v = op(v, dpp_f<0x128>(v));  // row_ror:8
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
  return op(op(r0, r1), op(r2, r3));
}
// Max and min of non-negative floats (every reduced feature quantity is one)
// as unsigned integers: one VALU instruction, no NaN canonicalisation.
struct FMax {
  __device__ float operator()(float a, float b) const {
    return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
  }
};
struct FMin {
  __device__ float operator()(float a, float b) const {
    return __uint_as_float(min(__float_as_uint(a), __float_as_uint(b)));
  }
};
struct FSum { __device__ float operator()(float a, float b) const { return a + b; } };

// The features' LDS holds a tile plus one column and one row: (tx+1)(ty+1)
// - 1 cells, at most 33 x 9 - 1 over the stage kernels' tile shapes (32x8).
constexpr int kFeatMaxExt = 33 * 9;

// One block (256 threads) per tile.  First every cell of the tile and of
// the column and row beyond it (its cells' right and top neighbours) forms
// its FP64 quantities once, into LDS; then thread t takes cell t of the
// tile and its neighbours from there.  14 wave reductions, one barrier,
// thread 0 combines the waves; the cell count is the tile's live extent.
__global__ __launch_bounds__(256) void k_tile_features(const double* u, size_t ps, int nx, int ny,
                                                       int tx, int ty, int tiles_x, int ntiles,
                                                       const void* ctrl_ch, float* feats) {
  constexpr int NQ = 14;
  __shared__ float red[4][NQ];
  __shared__ double sf[6][kFeatMaxExt];  // rho, m2, p, bx, by, bz
  const int tile = blockIdx.x;
  if (tile >= ntiles) return;
  const float ch = (float)*(const double*)ctrl_ch;
  const int j0 = (tile / tiles_x) * ty, i0 = (tile % tiles_x) * tx;
  const int t = threadIdx.x, li = t % tx, lj = t / tx;
  const int i = i0 + li, j = j0 + lj;
  const bool live = t < tx * ty && i < nx && j < ny;
  const float gamma = 5.0f / 3.0f;
  // The (tx+1) x (ty+1) box at (i0, j0), less its far corner: cells i0+le,
  // j0+je that some live cell of the tile reads (le <= nx - i0 and je <=
  // ny - j0, so one wrap reaches them).
  const int ex = tx + 1, ne = ex * (ty + 1) - 1;
  for (int e = t; e < ne; e += 256) {
    const int je = e / ex, le = e - je * ex;
    if (i0 + le > nx || j0 + je > ny) continue;
    const CellFeat c = cell_at(u, ps, (size_t)wrapi(j0 + je, ny) * nx + wrapi(i0 + le, nx));
    sf[0][e] = c.rho;
    sf[1][e] = c.m2;
    sf[2][e] = c.p;
    sf[3][e] = c.bx;
    sf[4][e] = c.by;
    sf[5][e] = c.bz;
  }
  __syncthreads();
  auto feat = [&](int e) {
    CellFeat c;
    c.rho = sf[0][e];
    c.m2 = sf[1][e];
    c.p = sf[2][e];
    c.bx = sf[3][e];
    c.by = sf[4][e];
    c.bz = sf[5][e];
    return c;
  };
  // q: min beta, max beta, drho, dp, db, mach, psi, |B|, div B, min bn,
  //    v/ch, cf/ch, kinetic, magnetic
  float q[NQ] = {3e38f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 3e38f, 0.f, 0.f, 0.f, 0.f};
  if (live) {
    const size_t g = (size_t)j * nx + i;
    const int e = lj * ex + li;
    const CellFeat c = feat(e);
    const CellFeat r = feat(e + 1);
    const CellFeat tp = feat(e + ex);
    const double bxl = u[5 * ps + (size_t)j * nx + wrapi(i - 1, nx)];
    const double byd = u[6 * ps + (size_t)wrapi(j - 1, ny) * nx + i];
    const double pp = fmax(c.p, 1e-12);
    const float rho = (float)c.rho, p = (float)pp, m2 = (float)c.m2;
    const float b2 = (float)(c.bx * c.bx + c.by * c.by + c.bz * c.bz), bmag = __builtin_amdgcn_sqrtf(b2);
    const float beta = fminf(fmaxf(__fdividef(2.0f * p, fmaxf(b2, 1e-30f)), 1e-6f), 1e6f);
    q[0] = q[1] = beta;
    q[2] = __fdividef((float)fmax(fabs(r.rho - c.rho), fabs(tp.rho - c.rho)), rho);
    q[3] = __fdividef((float)fmax(fabs(fmax(r.p, 1e-12) - pp), fabs(fmax(tp.p, 1e-12) - pp)), p);
    const float dxr = (float)(r.bx - c.bx), dyr = (float)(r.by - c.by), dzr = (float)(r.bz - c.bz);
    const float dxt = (float)(tp.bx - c.bx), dyt = (float)(tp.by - c.by),
                dzt = (float)(tp.bz - c.bz);
    const float dbr = __builtin_amdgcn_sqrtf(dxr * dxr + dyr * dyr + dzr * dzr);
    const float dbt = __builtin_amdgcn_sqrtf(dxt * dxt + dyt * dyt + dzt * dzt);
    q[4] = __fdividef(fmaxf(dbr, dbt), fmaxf(bmag, 1e-12f));
    const float v2 = __fdividef(m2, rho * rho), cs2 = __fdividef(gamma * p, rho);
    q[5] = fminf(__builtin_amdgcn_sqrtf(__fdividef(v2, cs2)), 100.0f);
    q[6] = (float)fabs(u[8 * ps + g]);
    q[7] = bmag;
    const float h = 1.0f / (float)(nx > ny ? nx : ny);
    const float divb = (float)(0.5 * (r.bx - bxl) * nx + 0.5 * (tp.by - byd) * ny);
    q[8] = __fdividef(fabsf(divb) * h, bmag + 1e-12f);
    q[9] = __fdividef(fminf(fabsf((float)c.bx), fabsf((float)c.by)), bmag + 1e-12f);
    q[10] = __fdividef(__builtin_amdgcn_sqrtf(v2), ch);
    q[11] = __fdividef(__builtin_amdgcn_sqrtf(cs2 + __fdividef(b2, rho)), ch);
    q[12] = 0.5f * __fdividef(m2, rho);
    q[13] = 0.5f * b2;
  }
  const FMax mx;
  const FMin mn;
  const FSum sm;
  q[0] = wave_reduce(q[0], mn);
  q[9] = wave_reduce(q[9], mn);
#pragma unroll
  for (int k = 1; k <= 11; k++)
    if (k != 9) q[k] = wave_reduce(q[k], mx);
#pragma unroll
  for (int k = 12; k < NQ; k++) q[k] = wave_reduce(q[k], sm);
  const int lane = t & 63, w = t >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NQ; k++) red[w][k] = q[k];
  __syncthreads();
  if (t == 0) {
    float f[NQ];
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      float v = red[0][k];
      for (int ww = 1; ww < 4; ww++)
        v = (k == 0 || k == 9) ? mn(v, red[ww][k]) : (k >= 12 ? sm(v, red[ww][k]) : mx(v, red[ww][k]));
      f[k] = v;
    }
    const float cnt = (float)(min(tx, nx - i0) * min(ty, ny - j0));
    float* o = feats + (size_t)tile * NIN;
    o[0] = log10f(f[0]);
    o[1] = log10f(f[1]);
    o[2] = log10f(f[2] + 1e-12f);
    o[3] = log10f(f[3] + 1e-12f);
    o[4] = log10f(f[4] + 1e-12f);
    o[5] = f[5];
    o[6] = log10f(f[6] / (f[7] + 1e-12f) + 1e-12f);
    o[7] = log10f(f[8] + 1e-12f);
    o[8] = log10f(f[9] + 1e-12f);
    o[9] = f[10];
    o[10] = f[11];
    o[11] = log10f(f[12] / cnt / (f[13] / cnt + 1e-12f) + 1e-12f);
    for (int k = NFEAT; k < NIN; k++) o[k] = 0.f;
  }
}

// Sum of squares of each of the 9 planes (atomic double adds; out zeroed by
// the caller).
__global__ __launch_bounds__(256) void k_sumsq(const double* u, size_t ps, long long cells,
                                               double* out9) {
  __shared__ double red[4];
  for (int k = 0; k < 9; k++) {
    double s = 0;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < cells;
         q += (long long)gridDim.x * 256) {
      const double v = u[k * ps + q];
      s += v * v;
    }
    const double b = block_reduce(s, red, Sum());
    if (threadIdx.x == 0) atomicAdd(&out9[k], b);
    __syncthreads();
  }
}

// Per-tile max over cells and the 8 MHD variables of |a-b| / (|b| + s_k).
// The scale s_k is the RMS over the grid of the variable's physical
// quantity: rho, |momentum| (k = 1..3), E, |B| (k = 5..7), so that a vector
// component that is small everywhere (v_y and B_y of a shear layer start at
// zero) is measured against its vector.  The GLM potential psi (plane 8) is
// left out: it has no scale of its own (it starts at zero and stays near the
// round-off level of div B).
// With `audit` non-null the listed tiles are audits: a tile whose error
// exceeds the budget becomes sticky and flagged (FP64 from then on: its
// stage 2 of this step already, through the FP64 list audit->fp64_list,
// count audit->fp64_count) and counts a miss; every audit is counted.
struct AuditOut {
  float budget;
  unsigned char* sticky;
  unsigned char* flags;
  int* fp64_list;
  int* fp64_count;
  unsigned long long* counters;
  int* slots;      // the mixed stage launches' slots (-1: none)
  const int* pos;  // slot of each mixed tile
};

__global__ __launch_bounds__(256) void k_tile_errors(const double* a, const double* b, size_t ps,
                                                     int nx, int ny, int tx, int ty, int tiles_x,
                                                     int ntiles, const double* sumsq9,
                                                     const int* list, const int* count,
                                                     float* err, AuditOut audit) {
  __shared__ double red[4];
  const int entries = list ? *count : ntiles;
  const double cells = (double)nx * (double)ny;
  const double s_rho = sqrt(sumsq9[0] / cells) + 1e-30;
  const double s_m = sqrt((sumsq9[1] + sumsq9[2] + sumsq9[3]) / cells) + 1e-30;
  const double s_e = sqrt(sumsq9[4] / cells) + 1e-30;
  const double s_b = sqrt((sumsq9[5] + sumsq9[6] + sumsq9[7]) / cells) + 1e-30;
  for (int x = (int)blockIdx.x; x < entries; x += (int)gridDim.x) {
    const int tile = list ? list[x] : x;
    const int j0 = (tile / tiles_x) * ty, i0 = (tile % tiles_x) * tx;
    const int t = threadIdx.x, i = i0 + t % tx, j = j0 + t / tx;
    double e = 0.0;
    if (t < tx * ty && i < nx && j < ny) {
      const size_t g = (size_t)j * nx + i;
      for (int k = 0; k < 8; k++) {
        const double rms = k == 0 ? s_rho : (k < 4 ? s_m : (k == 4 ? s_e : s_b));
        const double bv = b[k * ps + g];
        e = fmax(e, fabs(a[k * ps + g] - bv) / (fabs(bv) + rms));
      }
    }
    const double m = block_reduce(e, red, Max());
    if (threadIdx.x == 0) {
      err[tile] = (float)m;
      if (audit.counters) {
        atomicAdd(&audit.counters[0], 1ull);
        if ((float)m > audit.budget) {
          audit.sticky[tile] = 1;
          audit.flags[tile] = 1;
          audit.slots[audit.pos[tile]] = -1;
          audit.fp64_list[atomicAdd(audit.fp64_count, 1)] = tile;
          atomicAdd(&audit.counters[1], 1ull);
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// MLP forward on MFMA.  A wave evaluates 16 samples: layer 1 is one K=16
// step over 4 N-blocks, layer 2 four K=16 steps over 4 N-blocks, layer 3 an
// FP32 dot product.  Operand layouts of v_mfma_f32_16x16x16_bf16: lane l
// holds A[l%16][4*(l/16)+0..3], B[4*(l/16)+0..3][l%16], and receives
// D[4*(l/16)+i][l%16], i = 0..3.
// ---------------------------------------------------------------------------
struct WaveScratch {
  float z1[16][H];
  float z2[16][H];
  short a1[16][H];  // bf16 of relu(z1)
};

__device__ __forceinline__ void mlp_forward_wave(const float* x16, int nvalid, const float* P,
                                                 WaveScratch& ws, float* z3_out) {
  const int l = threadIdx.x & 63, row = l & 15, kq = l >> 4;
  // Layer 1: A = x (16 x 16), B = W1 (16 x 64).
  bf16x4 a;
  for (int q = 0; q < 4; q++) {
    const int k = 4 * kq + q;
    a[q] = to_bf16(row < nvalid ? x16[row * NIN + k] : 0.f);
  }
  for (int nb = 0; nb < 4; nb++) {
    bf16x4 b;
    for (int q = 0; q < 4; q++) b[q] = to_bf16(P[kW1 + (4 * kq + q) * H + nb * 16 + row]);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    for (int q = 0; q < 4; q++) {
      const int m = 4 * kq + q, n = nb * 16 + row;
      const float z = acc[q] + P[kB1 + n];
      ws.z1[m][n] = z;
      ws.a1[m][n] = to_bf16(z > 0.f ? z : 0.f);
    }
  }
  __syncthreads();
  // Layer 2: A = a1 (16 x 64) from LDS, B = W2 (64 x 64).
  for (int nb = 0; nb < 4; nb++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 4; ks++) {
      bf16x4 aa, b;
      for (int q = 0; q < 4; q++) {
        const int k = ks * 16 + 4 * kq + q;
        aa[q] = ws.a1[row][k];
        b[q] = to_bf16(P[kW2 + k * H + nb * 16 + row]);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aa, b, acc, 0, 0, 0);
    }
    for (int q = 0; q < 4; q++) {
      const int m = 4 * kq + q, n = nb * 16 + row;
      ws.z2[m][n] = acc[q] + P[kB2 + n];
    }
  }
  __syncthreads();
  // Layer 3 (FP32): lanes 0..15 each finish one sample.
  if (l < 16) {
    float z = P[kB3];
    for (int n = 0; n < H; n++) {
      const float a2 = ws.z2[l][n] > 0.f ? ws.z2[l][n] : 0.f;
      z += a2 * P[kW3 + n];
    }
    z3_out[l] = z;
  }
  __syncthreads();
}

// Standardised features of sample s into x16 (16 floats).
__device__ __forceinline__ void standardise(const float* raw, const float* mean, const float* istd,
                                            float* x16) {
  for (int k = 0; k < NIN; k++) x16[k] = k < NFEAT ? (raw[k] - mean[k]) * istd[k] : 0.f;
}

// Block of 256 threads = 4 waves x 16 samples.  raw: n x NIN features
// (standardised when mean is null); prob[s]; flags[s] = prob > threshold or
// sticky[s].
__global__ __launch_bounds__(256) void k_mlp_infer(const float* raw, int n, const float* P,
                                                   const float* mean, const float* istd,
                                                   float threshold, const unsigned char* sticky,
                                                   float* prob, unsigned char* flags) {
  __shared__ WaveScratch ws[4];
  __shared__ float xs[4][16 * NIN];
  __shared__ float z3[4][16];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int base = blockIdx.x * 64 + w * 16;
  const int nvalid = n - base < 0 ? 0 : (n - base > 16 ? 16 : n - base);
  if (l < 16 && l < nvalid) {
    if (mean)
      standardise(raw + (size_t)(base + l) * NIN, mean, istd, &xs[w][l * NIN]);
    else
      for (int k = 0; k < NIN; k++) xs[w][l * NIN + k] = raw[(size_t)(base + l) * NIN + k];
  }
  __syncthreads();
  mlp_forward_wave(xs[w], nvalid, P, ws[w], z3[w]);
  if (l < nvalid) {
    const float p = 1.f / (1.f + __expf(-z3[w][l]));
    if (prob) prob[base + l] = p;
    if (flags) flags[base + l] = (p > threshold || (sticky && sticky[base + l])) ? 1 : 0;
  }
}

// One 16x16 tile of C = A B over K in steps of 4 on the FP32 matrix cores
// (v_mfma_f32_16x16x4_f32: each lane holds A[l % 16][k0 + l / 16] and
// B[k0 + l / 16][l % 16]; lane l receives C[4 (l / 16) + i][l % 16],
// i = 0..3).  a(m, k) and b(k, n) read the operands.
template <class FA, class FB>
__device__ __forceinline__ f32x4 mfma_tile_f32(int K, FA a, FB b) {
  const int l = threadIdx.x & 63, r = l & 15, kq = l >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 4)
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a(r, k0 + kq), b(k0 + kq, r), acc, 0, 0, 0);
  return acc;
}

// Forward + backward of up to 64 samples per block (4 waves x 16).  The
// gradients of sum_s w_s BCE(p_s, y_s) / wsum and that loss are written to
// this block's row of `partial` (kParams + 1 floats per block), no atomics:
// sel_adam_partials / sel_sum_partials add the rows in block order, so the
// result does not depend on scheduling.  The three weight gradients and the
// back-propagation through W2 are 16x16 tiles on the FP32 matrix cores, with
// the forward's BF16 activations and weights as operands; bias and w3
// gradients are column sums in FP32.  Samples are idx[offset + s].
__global__ __launch_bounds__(256) void k_mlp_grad(const float* X, const float* Y, const float* Wt,
                                                  const int* idx, int offset, int count,
                                                  float wsum, const float* P, float* partial) {
  __shared__ WaveScratch ws[4];
  __shared__ float xs[4][16 * NIN];
  __shared__ float z3[4][16];
  __shared__ float dz3[64];
  __shared__ float dz2[64][H];
  __shared__ float dz1[64][H];
  __shared__ float lred[64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, tid = threadIdx.x;
  const int r = l & 15, kq = l >> 4;
  const int b0 = blockIdx.x * 64;
  const int nblk = count - b0 < 0 ? 0 : (count - b0 > 64 ? 64 : count - b0);
  const int nvalid = nblk - w * 16 < 0 ? 0 : (nblk - w * 16 > 16 ? 16 : nblk - w * 16);
  float* out = partial + (size_t)blockIdx.x * (kParams + 1);
  if (l < 16) {
    const int s = l < nvalid ? idx[offset + b0 + w * 16 + l] : -1;
    for (int k = 0; k < NIN; k++) xs[w][l * NIN + k] = s >= 0 ? X[(size_t)s * NIN + k] : 0.f;
  }
  __syncthreads();
  mlp_forward_wave(xs[w], nvalid, P, ws[w], z3[w]);
  // dL/dz3 and each sample's loss.
  if (tid < 64) {
    float d = 0.f, lp = 0.f;
    if (tid < nblk) {
      const int s = idx[offset + b0 + tid];
      const float z = z3[tid >> 4][tid & 15];
      const float p = 1.f / (1.f + __expf(-z));
      const float y = Y[s], wt = Wt[s];
      d = (p - y) * wt / wsum;
      lp = -wt * (y * __logf(fmaxf(p, 1e-12f)) + (1.f - y) * __logf(fmaxf(1.f - p, 1e-12f))) /
           wsum;
    }
    dz3[tid] = d;
    lred[tid] = lp;
  }
  __syncthreads();
  // dz2 = dz3 w3 (z2 > 0); gw3, gb3 and the loss as column sums.
  for (int e = tid; e < 64 * H; e += 256) {
    const int s = e / H, nn = e % H;
    const float z2 = ws[s >> 4].z2[s & 15][nn];
    dz2[s][nn] = (s < nblk && z2 > 0.f) ? dz3[s] * P[kW3 + nn] : 0.f;
  }
  if (tid < H) {
    float g = 0.f;
    for (int s = 0; s < nblk; s++) {
      const float z2 = ws[s >> 4].z2[s & 15][tid];
      g += (z2 > 0.f ? z2 : 0.f) * dz3[s];
    }
    out[kW3 + tid] = g;
  }
  if (tid == 64) {
    float g = 0.f, lsum = 0.f;
    for (int s = 0; s < nblk; s++) {
      g += dz3[s];
      lsum += lred[s];
    }
    out[kB3] = g;
    out[kParams] = lsum;
  }
  __syncthreads();
  // gW2[k][n] = sum_s a1[s][k] dz2[s][n]: wave w takes rows k of tile w.
  for (int nt = 0; nt < 4; nt++) {
    const f32x4 acc = mfma_tile_f32(
        64,
        [&](int m, int s) {
          return __uint_as_float(((unsigned)(unsigned short)ws[s >> 4].a1[s & 15][w * 16 + m]) << 16);
        },
        [&](int s, int n) { return dz2[s][nt * 16 + n]; });
    for (int i = 0; i < 4; i++) out[kW2 + (w * 16 + 4 * kq + i) * H + nt * 16 + r] = acc[i];
  }
  if (tid < H) {
    float g = 0.f;
    for (int s = 0; s < nblk; s++) g += dz2[s][tid];
    out[kB2 + tid] = g;
  }
  // dz1[s][k] = (sum_n dz2[s][n] W2[k][n]) (z1 > 0): wave w takes its own
  // 16 samples.
  for (int kt = 0; kt < 4; kt++) {
    const f32x4 acc = mfma_tile_f32(
        H, [&](int m, int n) { return dz2[w * 16 + m][n]; },
        [&](int n, int k) { return bf16_val(P[kW2 + (kt * 16 + k) * H + n]); });
    for (int i = 0; i < 4; i++) {
      const int sl = 4 * kq + i, k = kt * 16 + r;
      dz1[w * 16 + sl][k] = ws[w].z1[sl][k] > 0.f ? acc[i] : 0.f;
    }
  }
  __syncthreads();
  // gW1[k][n] = sum_s x[s][k] dz1[s][n] (x as the forward's BF16): wave w
  // takes columns n of tile w.
  {
    const f32x4 acc = mfma_tile_f32(
        64, [&](int k, int s) { return bf16_val(xs[s >> 4][(s & 15) * NIN + k]); },
        [&](int s, int n) { return dz1[s][w * 16 + n]; });
    for (int i = 0; i < 4; i++) out[kW1 + (4 * kq + i) * H + w * 16 + r] = acc[i];
  }
  if (tid < H) {
    float g = 0.f;
    for (int s = 0; s < nblk; s++) g += dz1[s][tid];
    out[kB1 + tid] = g;
  }
}

// Sum of the per-block rows of `partial` in block order: the gradient
// (kParams) and the loss (one more entry, added to *loss).
__global__ void k_sum_partials(const float* partial, int nparts, float* grad, float* loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > kParams) return;
  float g = 0.f;
  for (int b = 0; b < nparts; b++) g += partial[(size_t)b * (kParams + 1) + i];
  if (i < kParams)
    grad[i] = g;
  else if (loss)
    *loss += g;
}

// Adam on the FP32 masters from the per-block gradient rows (summed in block
// order); the minibatch loss is added to *loss.
__global__ void k_adam(float* P, const float* partial, int nparts, float* m, float* v, int step,
                       float lr, float* loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > kParams) return;
  float g = 0.f;
  for (int b = 0; b < nparts; b++) g += partial[(size_t)b * (kParams + 1) + i];
  if (i == kParams) {
    if (loss) *loss += g;
    return;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mi = b1 * m[i] + (1.f - b1) * g, vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - __powf(b1, (float)step)), vh = vi / (1.f - __powf(b2, (float)step));
  float p = P[i] - lr * mh / (sqrtf(vh) + eps);
  if (i >= kW1 + NFEAT * H && i < kB1) p = 0.f;  // padding rows of W1 stay zero
  P[i] = p;
}

// Tile lists of the current flags, by one block of kRouteThreads, each in
// tile order: flagged tiles -> lists[0 .. counts[0]) (FP64 tiles), the
// others -> lists[2 ntiles .. 2 ntiles + counts[2]) (mixed tiles).  The
// flags are taken in rounds of kRouteThreads x 16: a thread reads its 16
// consecutive flags with one 16-byte load (coalesced across the block),
// counts them, a block-wide exclusive scan gives its offsets in both lists,
// and it writes its tiles.  Then it deals the mixed tiles to the mixed
// stage launches' blocks in XCD order (lists + 3 ntiles) and records each
// tile's slot (lists + 4 ntiles).  Runs when the flags are refreshed; an
// audit miss later appends its tile to the FP64 list, sets its flag and
// strikes its slot out (-1), so the mixed stages skip it.
constexpr int kRouteThreads = 1024;
__global__ __launch_bounds__(kRouteThreads) void k_route(const unsigned char* flags, int ntiles,
                                                         int* lists, int* counts) {
  constexpr int kWaves = kRouteThreads / 64, kPer = 16, kRound = kRouteThreads * kPer;
  __shared__ int wsum[kWaves];
  const int tid = (int)threadIdx.x, lane = tid & 63, w = tid >> 6;
  int done_f = 0;  // FP64 tiles in the earlier rounds (block-uniform)
  for (int base = 0; base < ntiles; base += kRound) {
    const int t0 = base + tid * kPer;
    unsigned char f[kPer];
    if (t0 + kPer <= ntiles) {
      // flags is 16-byte aligned (checked by the launcher), t0 a multiple of 16.
      const uint4 v = *(const uint4*)(flags + t0);
      const unsigned wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < kPer; k++) f[k] = (unsigned char)(wv[k >> 2] >> (8 * (k & 3)));
    } else {
#pragma unroll
      for (int k = 0; k < kPer; k++) f[k] = t0 + k < ntiles ? flags[t0 + k] : 0;
    }
    int nf = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) nf += f[k] ? 1 : 0;
    int sf = nf;  // inclusive scan within the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int u = __shfl_up(sf, d);
      if (lane >= d) sf += u;
    }
    if (lane == 63) wsum[w] = sf;
    __syncthreads();
    int bf = 0, tf = 0;
#pragma unroll
    for (int k = 0; k < kWaves; k++) {
      bf += (k < w) ? wsum[k] : 0;
      tf += wsum[k];
    }
    int pf = done_f + bf + sf - nf;  // FP64 tiles before t0
    int pm = t0 - pf;                // mixed tiles before t0
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      if (t0 + k >= ntiles) break;
      if (f[k])
        lists[pf++] = t0 + k;
      else
        lists[2 * ntiles + pm++] = t0 + k;
    }
    done_f += tf;
    __syncthreads();  // wsum is rewritten by the next round
  }
  if (tid == 0) {
    counts[0] = done_f;
    counts[2] = ntiles - done_f;
  }
  // The mixed stage launches' slots: block b of a one-block-per-tile grid
  // takes the mixed tile at xcd_item(b, mixed count), -1 past the list, and
  // pos[tile] is that tile's slot (for an audit miss to strike it out).
  // The mixed list was written by this block above.
  __syncthreads();
  const int nm = ntiles - done_f;
  const int* mixed = lists + 2 * ntiles;
  int* slots = lists + 3 * ntiles;
  int* pos = lists + 4 * ntiles;
  for (int b = tid; b < ntiles; b += kRouteThreads) {
    const int item = xcd_item(b, nm);
    const int t = item < 0 ? -1 : mixed[item];
    slots[b] = t;
    if (t >= 0) pos[t] = b;
  }
}

// Audit sample of a step, by one block: systematic sampling of the mixed
// list, entries floor((k + phi) / frac) for k = 0, 1, ... below its length,
// with the phase phi in [0, 1) a hash of the device step counter (a replayed
// graph samples anew).  Distinct entries (the stride 1/frac is >= 1), a
// fraction `frac` of the mixed tiles -> lists[ntiles .. ntiles + counts[1]).
// Counts the FP64 tile-steps into counters[2] and the steps into
// counters[3].  Also copies the control block (ctrl_words words) into
// ctrl_copy: the audits' snapshot.  A step of a run that has finished or
// halted (*run != 0: the step kernels do nothing, as in the tail of a
// replayed batch) samples and counts nothing, so its audits cannot compare
// stale buffers.
constexpr int kPickThreads = 256;
__global__ __launch_bounds__(kPickThreads) void k_audit_pick(int ntiles, const int* step,
                                                             const int* run, float frac,
                                                             int* lists, int* counts,
                                                             unsigned long long* counters,
                                                             const unsigned* ctrl,
                                                             unsigned* ctrl_copy, int ctrl_words) {
  __shared__ int picked;
  const int tid = (int)threadIdx.x;
  if (tid == 0) picked = 0;
  for (int q = tid; q < ctrl_words; q += kPickThreads) ctrl_copy[q] = ctrl[q];
  __syncthreads();
  const bool active = *run == 0;
  const int n = counts[2];
  int mine = 0;
  if (active && frac > 0.f && n > 0) {
    unsigned h = (unsigned)*step * 2246822519u + 0x9E3779B9u;
    h ^= h >> 15;
    h *= 2654435761u;
    h ^= h >> 13;
    const double phi = (double)(h & 0xFFFFFF) * (1.0 / 16777216.0);
    const double stride = frac >= 1.f ? 1.0 : 1.0 / (double)frac;
    for (int k = tid;; k += kPickThreads) {
      const double x = ((double)k + phi) * stride;
      if (x >= (double)n) break;
      lists[ntiles + k] = lists[2 * ntiles + (int)x];
      mine++;
    }
  }
  if (mine) atomicAdd(&picked, mine);
  __syncthreads();
  if (tid == 0) {
    counts[1] = picked;
    if (active) {
      counters[2] += (unsigned long long)counts[0];
      counters[3] += 1ull;
    }
  }
}

// Word copy (control block snapshots inside captured steps).
__global__ void k_copy_words(const unsigned* src, unsigned* dst, int nwords) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nwords) dst[i] = src[i];
}

}  // namespace sel

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
hipError_t sel_tile_features(const double* u, size_t plane, int nx, int ny, int tx, int ty,
                             const void* ctrl_ch, float* feats, hipStream_t s) {
  if (tx < 1 || ty < 1 || tx * ty > 256 || (tx + 1) * (ty + 1) - 1 > sel::kFeatMaxExt)
    return hipErrorInvalidValue;
  const int tiles_x = (nx + tx - 1) / tx, ntiles = tiles_x * ((ny + ty - 1) / ty);
  hipLaunchKernelGGL(sel::k_tile_features, dim3(ntiles), dim3(256), 0, s, u, plane, nx, ny, tx, ty,
                     tiles_x, ntiles, ctrl_ch, feats);
  return hipGetLastError();
}

hipError_t sel_sumsq(const double* u, size_t plane, long long cells, double* sumsq9, hipStream_t s) {
  hipError_t e = hipMemsetAsync(sumsq9, 0, 9 * sizeof(double), s);
  if (e != hipSuccess) return e;
  long long blocks = (cells + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(sel::k_sumsq, dim3((unsigned)blocks), dim3(256), 0, s, u, plane, cells, sumsq9);
  return hipGetLastError();
}

hipError_t sel_tile_errors(const double* a, const double* b, size_t plane, int nx, int ny, int tx,
                           int ty, const double* sumsq9, const int* list, const int* count,
                           float* err, hipStream_t s) {
  const int tiles_x = (nx + tx - 1) / tx, ntiles = tiles_x * ((ny + ty - 1) / ty);
  const int grid = list ? (ntiles < 1024 ? ntiles : 1024) : ntiles;
  hipLaunchKernelGGL(sel::k_tile_errors, dim3(grid), dim3(256), 0, s, a, b, plane, nx, ny, tx, ty,
                     tiles_x, ntiles, sumsq9, list, count, err, sel::AuditOut{});
  return hipGetLastError();
}

hipError_t sel_audit(const double* a, const double* b, size_t plane, int nx, int ny, int tx, int ty,
                     const double* sumsq9, int* lists, int* counts, float budget,
                     unsigned char* sticky, unsigned char* flags, float* err,
                     unsigned long long* counters, hipStream_t s) {
  const int tiles_x = (nx + tx - 1) / tx, ntiles = tiles_x * ((ny + ty - 1) / ty);
  const int grid = ntiles < 1024 ? ntiles : 1024;
  sel::AuditOut o{budget, sticky, flags, lists, counts, counters, lists + 3 * ntiles,
                  lists + 4 * ntiles};
  hipLaunchKernelGGL(sel::k_tile_errors, dim3(grid), dim3(256), 0, s, a, b, plane, nx, ny, tx, ty,
                     tiles_x, ntiles, sumsq9, lists + ntiles, counts + 1, err, o);
  return hipGetLastError();
}

hipError_t sel_mlp_infer(const float* raw, int n, const float* params, const float* mean,
                         const float* istd, float threshold, const unsigned char* sticky,
                         float* prob, unsigned char* flags, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(sel::k_mlp_infer, dim3((n + 63) / 64), dim3(256), 0, s, raw, n, params, mean,
                     istd, threshold, sticky, prob, flags);
  return hipGetLastError();
}

hipError_t sel_mlp_grad(const float* X, const float* Y, const float* W, const int* idx, int offset,
                        int count, float wsum, const float* params, float* partial,
                        hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(sel::k_mlp_grad, dim3((count + 63) / 64), dim3(256), 0, s, X, Y, W, idx,
                     offset, count, wsum, params, partial);
  return hipGetLastError();
}

hipError_t sel_sum_partials(const float* partial, int nparts, float* grad, float* loss,
                            hipStream_t s) {
  hipLaunchKernelGGL(sel::k_sum_partials, dim3((sel::kParams + 1 + 255) / 256), dim3(256), 0, s,
                     partial, nparts, grad, loss);
  return hipGetLastError();
}

hipError_t sel_adam(float* params, const float* partial, int nparts, float* m, float* v, int step,
                    float lr, float* loss, hipStream_t s) {
  hipLaunchKernelGGL(sel::k_adam, dim3((sel::kParams + 1 + 255) / 256), dim3(256), 0, s, params,
                     partial, nparts, m, v, step, lr, loss);
  return hipGetLastError();
}

hipError_t sel_route(const unsigned char* flags, int ntiles, int* lists, int* counts,
                     hipStream_t s) {
  if ((uintptr_t)flags % 16 != 0) return hipErrorInvalidValue;  // 16-byte flag loads
  hipLaunchKernelGGL(sel::k_route, dim3(1), dim3(sel::kRouteThreads), 0, s, flags, ntiles, lists,
                     counts);
  return hipGetLastError();
}

hipError_t sel_audit_pick(int ntiles, const int* step, const int* run, float frac, int* lists,
                          int* counts, unsigned long long* counters, const void* ctrl,
                          void* ctrl_copy, size_t ctrl_bytes, hipStream_t s) {
  hipLaunchKernelGGL(sel::k_audit_pick, dim3(1), dim3(sel::kPickThreads), 0, s, ntiles, step,
                     run, frac, lists, counts, counters, (const unsigned*)ctrl,
                     (unsigned*)ctrl_copy, (int)((ctrl_bytes + 3) / 4));
  return hipGetLastError();
}

hipError_t sel_copy_words(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const int nwords = (int)((bytes + 3) / 4);
  hipLaunchKernelGGL(sel::k_copy_words, dim3((nwords + 255) / 256), dim3(256), 0, s,
                     (const unsigned*)src, (unsigned*)dst, nwords);
  return hipGetLastError();
}



}  // namespace mhd

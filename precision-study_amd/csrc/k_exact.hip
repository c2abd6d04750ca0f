// k_exact.hip -- FP64 and FP32 step kernels with FP contraction off and IEEE
// division/square root: bitwise the reference CPU build of the same
// precision; the reference's stage-level mixed form under the same flags.
// Also hosts the operator-level batch kernels.
#define MHD_NS exact
#include "mhd_step.inc.hip"

namespace mhd {

const Launchers* launchers_exact(int mode) {
  static const Launchers fp64 = exact::make_launchers<exact::ExactPol<double>>(MODE_FP64);
  static const Launchers fp32 = exact::make_launchers<exact::ExactPol<float>>(MODE_FP32);
  static const Launchers ref_mixed = exact::make_launchers<exact::RefMixPol>(MODE_REF_MIXED);
  if (mode == MODE_FP32) return &fp32;
  if (mode == MODE_REF_MIXED) return &ref_mixed;
  return &fp64;
}

// Operator-level batch kernels (one state pair per thread).
template <class R, bool HLLD>
__global__ void k_glm_flux_batch(int n, int dir, const R* UL, const R* UR, R ch, double tol,
                                 R* F, int* status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const Phys<R> c = exact::make_phys<R>(tol);
  // Read in the x frame: for a y face swap the (1,2) and (5,6) slots.
  const int perm[9] = {0, dir ? 2 : 1, dir ? 1 : 2, 3, 4, dir ? 6 : 5, dir ? 5 : 6, 7, 8};
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
  const Phys<R> c = exact::make_phys<R>(1.0e-4);
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
  const Phys<R> c = exact::make_phys<R>(1.0e-4);
  R w[9];
#pragma unroll
  for (int k = 0; k < 9; k++) w[k] = W[(size_t)s * 9 + k];
  out[s] = fast_speed(c, w, dir ? w[6] : w[5]);
}

template <class R>
hipError_t launch_glm_flux_batch(int n, int dir, const R* UL, const R* UR, R ch, int hlld,
                                 double tol, R* F, int* status, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int blocks = (n + 255) / 256;
  if (hlld)
    hipLaunchKernelGGL((k_glm_flux_batch<R, true>), dim3(blocks), dim3(256), 0, s, n, dir, UL,
                       UR, ch, tol, F, status);
  else
    hipLaunchKernelGGL((k_glm_flux_batch<R, false>), dim3(blocks), dim3(256), 0, s, n, dir, UL,
                       UR, ch, tol, F, status);
  return hipGetLastError();
}

template <class R>
hipError_t launch_cons_to_prim_batch(int n, const R* U, R* W, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cons_to_prim_batch<R>, dim3((n + 255) / 256), dim3(256), 0, s, n, U, W);
  return hipGetLastError();
}

template <class R>
hipError_t launch_fast_speed_batch(int n, const R* W, int dir, R* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fast_speed_batch<R>, dim3((n + 255) / 256), dim3(256), 0, s, n, W, dir,
                     out);
  return hipGetLastError();
}

#define MHD_INSTANTIATE_BATCH(R)                                                               \
  template hipError_t launch_glm_flux_batch<R>(int, int, const R*, const R*, R, int, double,   \
                                               R*, int*, hipStream_t);                         \
  template hipError_t launch_cons_to_prim_batch<R>(int, const R*, R*, hipStream_t);            \
  template hipError_t launch_fast_speed_batch<R>(int, const R*, int, R*, hipStream_t);

MHD_INSTANTIATE_BATCH(double)
MHD_INSTANTIATE_BATCH(float)

}  // namespace mhd

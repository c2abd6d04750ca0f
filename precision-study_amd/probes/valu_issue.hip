// VALU issue-rate probe for the stage kernels' compute bound.
//
// Every thread runs ILP independent dependency chains of one VALU operation
// (FP32 FMA, FP64 FMA, FP32 multiply, 32-bit integer xor-add) for `iters`
// rounds; 8192 blocks of 256 threads keep every SIMD full.  The rate is
// printed as lane-operations per second and as wave64 instructions per CU
// and cycle at the nominal 2.4 GHz over 256 CUs, the unit the stage
// kernels' VALU counters are read in.  Built with the library's flags
// (packed FP32 instructions off), so FP32 here is the non-packed issue the
// mixed kernels get.
//
//   make -C precision-study_amd bin/valu_issue && precision-study_amd/bin/valu_issue
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHECK(x)                                                                  \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "%s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

__device__ inline float fma_t(float x, float y, float z) { return __builtin_fmaf(x, y, z); }
__device__ inline double fma_t(double x, double y, double z) { return __builtin_fma(x, y, z); }

template <class T, int ILP>
__global__ __launch_bounds__(256) void k_fma(T* out, int iters, T a, T b) {
  T acc[ILP];
#pragma unroll
  for (int k = 0; k < ILP; k++) acc[k] = (T)(threadIdx.x + k);
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int k = 0; k < ILP; k++) acc[k] = fma_t(acc[k], a, b);
  }
  T s = 0;
#pragma unroll
  for (int k = 0; k < ILP; k++) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <int ILP>
__global__ __launch_bounds__(256) void k_mul32(float* out, int iters, float a) {
  float acc[ILP];
#pragma unroll
  for (int k = 0; k < ILP; k++) acc[k] = (float)(threadIdx.x + k);
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int k = 0; k < ILP; k++) acc[k] = acc[k] * a;
  }
  float s = 0;
#pragma unroll
  for (int k = 0; k < ILP; k++) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <int ILP>
__global__ __launch_bounds__(256) void k_xad(unsigned* out, int iters, unsigned a) {
  unsigned acc[ILP];
#pragma unroll
  for (int k = 0; k < ILP; k++) acc[k] = threadIdx.x + k;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int k = 0; k < ILP; k++) acc[k] = (acc[k] ^ a) + 7u;  // one v_xad_u32
  }
  unsigned s = 0;
#pragma unroll
  for (int k = 0; k < ILP; k++) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <class F>
static int time_launches(const char* name, F launch, double lane_ops) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  launch();
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  constexpr int kReps = 5;
  CHECK(hipEventRecord(e0));
  for (int r = 0; r < kReps; r++) launch();
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  const double s = ms / kReps * 1e-3;
  std::printf("%-18s %8.3f ms  %7.2f T lane-op/s  %.3f wave64 inst per CU-cycle (256 CUs, 2.4 GHz)\n",
              name, ms / kReps, lane_ops / s * 1e-12, lane_ops / 64.0 / s / 256.0 / 2.4e9);
  std::fflush(stdout);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return 0;
}

int main() {
  constexpr int kBlocks = 256 * 8 * 4, kThreads = 256, kIters = 4096;
  const size_t n = (size_t)kBlocks * kThreads;  // one output per thread
  void* buf = nullptr;
  CHECK(hipMalloc(&buf, n * sizeof(double)));
  const double rounds = (double)n * kIters;
  std::printf("%d blocks x %d threads, %d rounds per chain\n", kBlocks, kThreads, kIters);
  int rc = 0;
  rc |= time_launches("fp32 fma, 8 chains", [&] {
    k_fma<float, 8><<<kBlocks, kThreads>>>((float*)buf, kIters, 0.999f, 1e-3f);
  }, rounds * 8);
  rc |= time_launches("fp64 fma, 8 chains", [&] {
    k_fma<double, 8><<<kBlocks, kThreads>>>((double*)buf, kIters, 0.999, 1e-3);
  }, rounds * 8);
  rc |= time_launches("fp32 mul, 8 chains", [&] {
    k_mul32<8><<<kBlocks, kThreads>>>((float*)buf, kIters, 0.999f);
  }, rounds * 8);
  rc |= time_launches("u32 xad, 8 chains", [&] {
    k_xad<8><<<kBlocks, kThreads>>>((unsigned*)buf, kIters, 3u);
  }, rounds * 8);
  rc |= time_launches("fp32 fma, 16 chains", [&] {
    k_fma<float, 16><<<kBlocks, kThreads>>>((float*)buf, kIters, 0.999f, 1e-3f);
  }, rounds * 16);
  rc |= time_launches("fp64 fma, 16 chains", [&] {
    k_fma<double, 16><<<kBlocks, kThreads>>>((double*)buf, kIters, 0.999, 1e-3);
  }, rounds * 16);
  CHECK(hipFree(buf));
  return rc;
}

// HBM streaming probe: how fast one MI355X reads and copies a large buffer
// with the access widths the solver's kernels use.
//
// A 2 GiB buffer of doubles (far beyond the caches) is summed by every
// thread over a grid-stride loop, with 8-byte (one double per lane, as the
// stage kernels' SoA plane loads) or 16-byte loads and 1 or 4 independent
// loads in flight per thread, and copied into a second buffer with 16-byte
// loads and stores.  Rates are bytes moved over the mean of 5 launches
// (HIP events).
//
//   make -C precision-study_amd bin/hbm_stream && precision-study_amd/bin/hbm_stream
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

// Sum of n doubles, UNROLL loads of one double per lane in flight.
template <int UNROLL>
__global__ __launch_bounds__(256) void k_read8(const double* __restrict__ a, size_t n,
                                               double* __restrict__ out) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s[UNROLL] = {};
  for (; i + (UNROLL - 1) * stride < n; i += UNROLL * stride) {
#pragma unroll
    for (int k = 0; k < UNROLL; k++) s[k] += a[i + k * stride];
  }
  for (; i < n; i += stride) s[0] += a[i];
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < UNROLL; k++) t += s[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = t;
}

// Sum of n doubles read as n/2 double2 (16-byte loads), UNROLL in flight.
template <int UNROLL>
__global__ __launch_bounds__(256) void k_read16(const double2* __restrict__ a, size_t n2,
                                                double* __restrict__ out) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s[UNROLL] = {};
  for (; i + (UNROLL - 1) * stride < n2; i += UNROLL * stride) {
#pragma unroll
    for (int k = 0; k < UNROLL; k++) {
      const double2 v = a[i + k * stride];
      s[k] += v.x + v.y;
    }
  }
  for (; i < n2; i += stride) s[0] += a[i].x + a[i].y;
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < UNROLL; k++) t += s[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = t;
}

// Copy of n2 double2 (16-byte loads and stores), 4 in flight per thread.
__global__ __launch_bounds__(256) void k_copy16(const double2* __restrict__ a,
                                                double2* __restrict__ b, size_t n2) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n2; i += 4 * stride) {
    const double2 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
    b[i] = v0;
    b[i + stride] = v1;
    b[i + 2 * stride] = v2;
    b[i + 3 * stride] = v3;
  }
  for (; i < n2; i += stride) b[i] = a[i];
}

template <class F>
static int time_launches(const char* name, F launch, double bytes) {
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
  std::printf("%-34s %8.3f ms  %6.2f TB/s  %.3f of 8 TB/s\n", name, ms / kReps, bytes / s * 1e-12,
              bytes / s / 8e12);
  std::fflush(stdout);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return 0;
}

int main() {
  const size_t n = (size_t)1 << 28;  // 2 GiB of doubles
  const int threads = 256, blocks = 256 * 16;  // 16 blocks (64 waves) per CU over the grid
  double *a = nullptr, *b = nullptr, *out = nullptr;
  CHECK(hipMalloc(&a, n * sizeof(double)));
  CHECK(hipMalloc(&b, n * sizeof(double)));
  CHECK(hipMalloc(&out, (size_t)blocks * threads * sizeof(double)));
  CHECK(hipMemset(a, 0, n * sizeof(double)));
  const double bytes = (double)n * sizeof(double);
  std::printf("buffer %.2f GiB, %d blocks x %d threads\n", bytes / (1 << 30), blocks, threads);
  int rc = 0;
  rc |= time_launches("read, 8-byte loads, 1 in flight", [&] {
    k_read8<1><<<blocks, threads>>>(a, n, out);
  }, bytes);
  rc |= time_launches("read, 8-byte loads, 4 in flight", [&] {
    k_read8<4><<<blocks, threads>>>(a, n, out);
  }, bytes);
  rc |= time_launches("read, 16-byte loads, 1 in flight", [&] {
    k_read16<1><<<blocks, threads>>>((const double2*)a, n / 2, out);
  }, bytes);
  rc |= time_launches("read, 16-byte loads, 4 in flight", [&] {
    k_read16<4><<<blocks, threads>>>((const double2*)a, n / 2, out);
  }, bytes);
  rc |= time_launches("copy, 16-byte, 4 in flight (r+w)", [&] {
    k_copy16<<<blocks, threads>>>((const double2*)a, (double2*)b, n / 2);
  }, 2.0 * bytes);
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  CHECK(hipFree(out));
  return rc;
}

// k_noalloc.hip -- the fixed device scratch array of the no-allocation run
// (mhd_run_noalloc in mhd_api.cpp).
//
// The array holds, for one run at a time on each device, what a solver
// context would otherwise allocate: the stage-1 state U1, the control block
// and the mixed form's two fallback tile lists.  It is a file-scope device
// array, so the code object's loader reserves it once per device and keeps
// it until the process ends; no call allocates.  Host code never passes the
// array's name to a kernel (on the host that name is a stand-in, not a
// device address): it asks for the device address here, once per device with
// that device current, and uses only that address.
#include "mhd_kernels.h"

#include <hip/hip_runtime.h>

#include <mutex>

namespace mhd {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_addr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(void** ptr, size_t* bytes) {
  if (!ptr || !bytes) return hipErrorInvalidValue;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> hold(g_lock);
  if (!g_addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_addr[dev] = p;
  }
  *ptr = g_addr[dev];
  *bytes = kNoAllocScratchBytes;
  return hipSuccess;
}

}  // namespace mhd

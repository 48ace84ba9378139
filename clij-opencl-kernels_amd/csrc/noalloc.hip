// Device memory of clk_filter_3d_noalloc: one fixed array per GPU that
// holds the float32 (or source-type) temporaries of a 3D filter, so that the
// call allocates nothing.
//
// In host code `scratch` names a host-side stand-in, not the device array:
// every use of the array's memory from the host goes through the device
// address that hipGetSymbolAddress returns for the current device (each
// GPU holds its own copy of the array at its own address).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "clk_internal.h"

namespace clk {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device, resolved once per
// device (with that device current) and kept.
hipError_t noalloc_scratch(void** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = p;
  }
  *ptr = g_ptr[dev];
  return hipSuccess;
}

}  // namespace clk

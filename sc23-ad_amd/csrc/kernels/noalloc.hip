// Device scratch of the allocation-free level operation
// (sc23ad_amg_level_op_noalloc): the two staging vectors that
// sc23ad_amg_level_op allocates per call (the operand and the result in the
// level's smoother numbering) live in one file-scope device array instead.
// Each GPU holds its own copy of the array for the life of the process; host
// code reaches it only through the address hipGetSymbolAddress returns for
// the current device (the array's name is not a device address on the host).
#include <hip/hip_runtime.h>

#include <mutex>

#include "launch.hpp"

namespace sc23ad::dev {

namespace {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch(void** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {  // once per device, with that device current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = p;
  }
  *ptr = g_ptr[dev];
  return hipSuccess;
}

}  // namespace sc23ad::dev

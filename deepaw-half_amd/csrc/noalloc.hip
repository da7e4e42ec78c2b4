// The workspace of half_amd_operator_apply_noalloc: one file-scope device
// array.  The code object holds it on every device the library runs on,
// from its load until the process ends, so the entry point that uses it
// never calls hipMalloc / hipFree.  In host code the array's name is only a
// stand-in for the symbol: kernels get the device address that
// hipGetSymbolAddress returns for the current device, never the name.
#include "noalloc.hpp"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace half {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mutex;
std::vector<void *> g_scratch_ptr;  // per device ordinal
}  // namespace

void *noalloc_scratch() {
  int dev = 0;
  HALF_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (dev >= static_cast<int>(g_scratch_ptr.size())) g_scratch_ptr.resize(size_t(dev) + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(kInternal, std::string("hipGetSymbolAddress of the workspace array failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace half

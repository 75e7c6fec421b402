// Device scratch of the allocation-free evaluation entry point
// (lkgpu_ll_grad_batch_noalloc): one static array per device, held until
// the process ends.  It holds a batch's slot tables (slot descriptors,
// parameters, results) followed by every slot's workspace (R / L, L^-1, the
// vector block, the reduction partials), laid out exactly as the pooled
// workspace of the allocating entry point.  Host code never passes `scratch`
// itself to a kernel or a copy: its device address comes from
// hipGetSymbolAddress, once per device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lk_kernels.h"

namespace lk {

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

// device address of `scratch` on `device` (made current by the caller),
// looked up the first time and kept per device
hipError_t noalloc_scratch(int device, char** ptr) {
  static std::mutex mu;
  static char* addr[64] = {};
  if (device < 0 || device >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (addr[device] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = static_cast<char*>(p);
  }
  *ptr = addr[device];
  return hipSuccess;
}

}  // namespace lk

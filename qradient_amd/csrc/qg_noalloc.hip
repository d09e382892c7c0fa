// Device scratch of the allocation-free entry point qg_mcclean_grad_run_noalloc.
//
// A file-scope __device__ array, one copy per GPU, held until the process
// ends.  Its device address is taken once per device with
// hipGetSymbolAddress when a context is created on that device (the
// library's init call, with the device current); host code uses only that
// address, never the array's name (which is a host-side stand-in: passing it
// to a kernel would fault).  Every call on a GPU shares the GPU's array, so
// calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "qg_internal.h"

static __device__ __align__(256) unsigned char scratch[qg::kNoAllocScratchBytes];

namespace {
std::mutex g_mu;
void* g_ptr[qg::kNoAllocMaxDevices] = {};
}  // namespace

namespace qg {

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return hipSuccess;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
  void* ptr = nullptr;
  e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_ptr[device] = ptr;
  return hipSuccess;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

}  // namespace qg

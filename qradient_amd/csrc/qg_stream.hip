// Streaming copy: the achievable-bandwidth reference for the pass kernels.
//
// Each lane moves 16 bytes per access, four independent accesses in flight
// per iteration of a grid-stride loop; the grid is a multiple of the CU
// count.  Optionally non-temporal (streaming) loads and stores, which bypass
// the memory-side cache's retention.  Used by the benchmark to measure the
// copy ceiling on the device it runs on (HBM with buffers larger than the
// 256 MB memory-side cache; that cache when a buffer pair fits it).
#include <hip/hip_runtime.h>

#include "qg_internal.h"

namespace qg {

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

template <bool NT>
__global__ __launch_bounds__(256) void k_stream_copy(const v4u* __restrict__ src, v4u* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    v4u a0, a1, a2, a3;
    if (NT) {
      a0 = __builtin_nontemporal_load(src + i);
      a1 = __builtin_nontemporal_load(src + i + stride);
      a2 = __builtin_nontemporal_load(src + i + 2 * stride);
      a3 = __builtin_nontemporal_load(src + i + 3 * stride);
      __builtin_nontemporal_store(a0, dst + i);
      __builtin_nontemporal_store(a1, dst + i + stride);
      __builtin_nontemporal_store(a2, dst + i + 2 * stride);
      __builtin_nontemporal_store(a3, dst + i + 3 * stride);
    } else {
      a0 = src[i];
      a1 = src[i + stride];
      a2 = src[i + 2 * stride];
      a3 = src[i + 3 * stride];
      dst[i] = a0;
      dst[i + stride] = a1;
      dst[i + 2 * stride] = a2;
      dst[i + 3 * stride] = a3;
    }
  }
  for (; i < n; i += stride) {
    if (NT) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
    else dst[i] = src[i];
  }
}

hipError_t launch_stream_copy(void* dst, const void* src, size_t bytes, bool nontemporal, int grid,
                              hipStream_t st) {
  const size_t n = bytes / 16;
  if (nontemporal)
    hipLaunchKernelGGL(k_stream_copy<true>, dim3(grid), dim3(256), 0, st, (const v4u*)src, (v4u*)dst, n);
  else
    hipLaunchKernelGGL(k_stream_copy<false>, dim3(grid), dim3(256), 0, st, (const v4u*)src, (v4u*)dst, n);
  return hipGetLastError();
}

}  // namespace qg

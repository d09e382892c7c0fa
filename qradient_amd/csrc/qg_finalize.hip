// Deterministic final reduction of the partial rows written by the passes,
// and the pass-kernel dispatch used by the host planner.
#include <hip/hip_runtime.h>

#include "qg_internal.h"

namespace qg {

hipError_t launch_pass_c64(int R, bool two, int phase, bool post, const PassDesc* d, int K, int grid, int batch,
                           int threads, bool big, hipStream_t st, int* occ, int tab);
hipError_t launch_pass_c128(int R, bool two, int phase, bool post, const PassDesc* d, int K, int grid, int batch,
                            int threads, bool big, hipStream_t st, int* occ, int tab);
hipError_t launch_pass_c64x(int R, bool two, int phase, bool post, const PassDesc* d, int K, int grid, int batch,
                            int threads, bool big, hipStream_t st, int* occ, int tab);
hipError_t set_attributes_c64();
hipError_t set_attributes_c128();
hipError_t set_attributes_c64x();

// out[i] = scale[i] * sum over the rows r of output i of sum_{w < cnt[r]} part[r*stride + w].
// One workgroup per output; each thread sums a fixed strided subset, then a
// fixed-shape tree: the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_finalize(int nout, const int* __restrict__ ptr, const int* __restrict__ rows,
                                                  const int* __restrict__ cnt, const double* __restrict__ scale,
                                                  const double* __restrict__ part, int stride, double* __restrict__ out) {
  __shared__ double red[256];
  const int i = blockIdx.x;
  if (i >= nout) return;
  double s = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const int r = rows[k];
    for (int w = threadIdx.x; w < cnt[r]; w += blockDim.x) s += part[(size_t)r * stride + w];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = scale[i] * red[0];
}

hipError_t launch_finalize(int nout, const int* ptr, const int* rows, const int* cnt, const double* scale,
                           const double* part, int stride, double* out, hipStream_t st) {
  if (nout <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3(nout), dim3(256), 0, st, nout, ptr, rows, cnt, scale, part, stride, out);
  return hipGetLastError();
}

hipError_t set_pass_attributes() {
  hipError_t e = set_attributes_c64();
  if (e == hipSuccess) e = set_attributes_c128();
  if (e == hipSuccess) e = set_attributes_c64x();
  return e;
}

// dtype 1: complex64 storage, 0: complex128, 2: complex64 storage with FP64
// LDS tiles (phase-free passes only).  post: an adjoint pass of the general
// form.  With occ != NULL the pass is not launched: *occ
// receives the workgroups per CU.  tab: bytes of the table the pass stages in
// LDS.
hipError_t launch_pass(int dtype, int R, bool two, int phase, bool post, const PassDesc* d, int K, int grid,
                       int batch, int threads, int n, hipStream_t st, int* occ, int tab) {
  // byte offsets of a 2^n state: 32-bit addressing below 4 GiB
  const bool big = (n + (dtype == 0 ? 4 : 3)) > 32;
  if (dtype == 1) return launch_pass_c64(R, two, phase, post, d, K, grid, batch, threads, big, st, occ, tab);
  if (dtype == 2) return launch_pass_c64x(R, two, phase, post, d, K, grid, batch, threads, big, st, occ, tab);
  return launch_pass_c128(R, two, phase, post, d, K, grid, batch, threads, big, st, occ, tab);
}

}  // namespace qg

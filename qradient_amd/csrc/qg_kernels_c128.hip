// complex128-storage instantiations of the tile-pass kernel (the plain
// method: FP64 state, FP64 LDS tiles, FP64 arithmetic).
#include "qg_pass.hpp"

namespace qg {

template <int R>
static hipError_t launch_c128_r(bool two, int phase, const PassDesc* d, int K, int grid, int batch, int threads,
                                bool big, hipStream_t st, int* occ) {
  if (!two) {
    if (phase == 0) return launch_pass_t<double, R, false, 0, double>(d, K, grid, batch, threads, big, st, occ);
    if (phase == 1) return launch_pass_t<double, R, false, 1, double>(d, K, grid, batch, threads, big, st, occ);
    return launch_pass_t<double, R, false, 2, double>(d, K, grid, batch, threads, big, st, occ);
  }
  if (phase == 0) return launch_pass_t<double, R, true, 0, double>(d, K, grid, batch, threads, big, st, occ);
  if (phase == 1) return launch_pass_t<double, R, true, 1, double>(d, K, grid, batch, threads, big, st, occ);
  return launch_pass_t<double, R, true, 2, double>(d, K, grid, batch, threads, big, st, occ);
}

hipError_t launch_pass_c128(int R, bool two, int phase, const PassDesc* d, int K, int grid, int batch, int threads,
                            bool big, hipStream_t st, int* occ) {
  switch (R) {
    case 4: return launch_c128_r<4>(two, phase, d, K, grid, batch, threads, big, st, occ);
    case 3: return launch_c128_r<3>(two, phase, d, K, grid, batch, threads, big, st, occ);
    case 2: return launch_c128_r<2>(two, phase, d, K, grid, batch, threads, big, st, occ);
    case 1: return launch_c128_r<1>(two, phase, d, K, grid, batch, threads, big, st, occ);
  }
  return hipErrorInvalidValue;
}

hipError_t set_attributes_c128() {
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = set_attr_t<double, 4, false, 0, double>();
  if (e == hipSuccess) e = set_attr_t<double, 4, false, 1, double>();
  if (e == hipSuccess) e = set_attr_t<double, 4, false, 2, double>();
  if (e == hipSuccess) e = set_attr_t<double, 4, true, 0, double>();
  if (e == hipSuccess) e = set_attr_t<double, 4, true, 1, double>();
  if (e == hipSuccess) e = set_attr_t<double, 4, true, 2, double>();
  return e;
}

}  // namespace qg

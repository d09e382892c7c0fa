// complex64-storage instantiations of the tile-pass kernel (FP32 LDS tiles,
// FP64 arithmetic for the propagated state).
#include "qg_pass.hpp"

namespace qg {

template <int R>
static hipError_t launch_c64_r(bool two, int phase, bool ta_float, const PassDesc* d, int K, int grid, int batch,
                               int threads, bool big, hipStream_t st, int* occ) {
  if (!two) {
    if (phase == 0) return launch_pass_t<float, R, false, 0, double>(d, K, grid, batch, threads, big, st, occ);
    if (phase == 1) return launch_pass_t<float, R, false, 1, double>(d, K, grid, batch, threads, big, st, occ);
    return launch_pass_t<float, R, false, 2, double>(d, K, grid, batch, threads, big, st, occ);
  }
  if constexpr (R == 5) {
    if (ta_float) {
      if (phase == 0) return launch_pass_t<float, R, true, 0, float>(d, K, grid, batch, threads, big, st, occ);
      if (phase == 1) return launch_pass_t<float, R, true, 1, float>(d, K, grid, batch, threads, big, st, occ);
      return launch_pass_t<float, R, true, 2, float>(d, K, grid, batch, threads, big, st, occ);
    }
    return hipErrorInvalidValue;  // FP64 companion with 32 amplitudes per thread: not instantiated
  } else {
    if (ta_float) {
      if constexpr (R == 4) {
        if (phase == 0) return launch_pass_t<float, R, true, 0, float>(d, K, grid, batch, threads, big, st, occ);
        if (phase == 1) return launch_pass_t<float, R, true, 1, float>(d, K, grid, batch, threads, big, st, occ);
        return launch_pass_t<float, R, true, 2, float>(d, K, grid, batch, threads, big, st, occ);
      }
      return hipErrorInvalidValue;
    }
    if (phase == 0) return launch_pass_t<float, R, true, 0, double>(d, K, grid, batch, threads, big, st, occ);
    if (phase == 1) return launch_pass_t<float, R, true, 1, double>(d, K, grid, batch, threads, big, st, occ);
    return launch_pass_t<float, R, true, 2, double>(d, K, grid, batch, threads, big, st, occ);
  }
}

hipError_t launch_pass_c64(int R, bool two, int phase, bool ta_float, const PassDesc* d, int K, int grid, int batch,
                           int threads, bool big, hipStream_t st, int* occ) {
  switch (R) {
    case 5: return launch_c64_r<5>(two, phase, ta_float, d, K, grid, batch, threads, big, st, occ);
    case 4: return launch_c64_r<4>(two, phase, ta_float, d, K, grid, batch, threads, big, st, occ);
    case 3: return launch_c64_r<3>(two, phase, ta_float, d, K, grid, batch, threads, big, st, occ);
    case 2: return launch_c64_r<2>(two, phase, ta_float, d, K, grid, batch, threads, big, st, occ);
    case 1: return launch_c64_r<1>(two, phase, ta_float, d, K, grid, batch, threads, big, st, occ);
  }
  return hipErrorInvalidValue;
}

hipError_t set_attributes_c64() {
  hipError_t e = hipSuccess;
#define QG_A(R, TWO, PH, TA) \
  if (e == hipSuccess) e = set_attr_t<float, R, TWO, PH, TA>();
#define QG_R(R) QG_A(R, false, 0, double) QG_A(R, false, 1, double) QG_A(R, false, 2, double)
  QG_R(5) QG_R(4) QG_R(3) QG_R(2) QG_R(1)
  QG_A(5, true, 0, float) QG_A(5, true, 1, float) QG_A(5, true, 2, float)
  QG_A(4, true, 0, double) QG_A(4, true, 1, double) QG_A(4, true, 2, double)
  QG_A(4, true, 0, float) QG_A(4, true, 1, float) QG_A(4, true, 2, float)
#undef QG_R
#undef QG_A
  return e;
}

}  // namespace qg

// complex128-storage instantiations of the tile-pass kernel (the plain method:
// FP64 state, FP64 LDS tiles, FP64 arithmetic).
#include "qg_pass.hpp"

namespace qg {

// post: an adjoint pass of the general form (both states rotated); forward
// passes take post = false (their phase code runs before or after the gates).
template <int R>
static hipError_t launch_c128_r(bool two, int phase, bool post, const PassDesc* d, int K, int grid, int batch,
                               int threads, bool big, hipStream_t st, int* occ, int tab) {
  if (!two) {
    if (post) return hipErrorInvalidValue;
    if (phase == 0) return launch_pass_t<double, R, false, 0, false>(d, K, grid, batch, threads, big, st, occ, tab);
    if (phase == 1) return launch_pass_t<double, R, false, 1, false>(d, K, grid, batch, threads, big, st, occ, tab);
    return launch_pass_t<double, R, false, 2, false>(d, K, grid, batch, threads, big, st, occ, tab);
  }
  if (phase == 0) {
    if (post) return launch_pass_t<double, R, true, 0, true>(d, K, grid, batch, threads, big, st, occ, tab);
    return launch_pass_t<double, R, true, 0, false>(d, K, grid, batch, threads, big, st, occ, tab);
  }
  if (phase == 1) {
    if (post) return launch_pass_t<double, R, true, 1, true>(d, K, grid, batch, threads, big, st, occ, tab);
    return launch_pass_t<double, R, true, 1, false>(d, K, grid, batch, threads, big, st, occ, tab);
  }
  if (post) return launch_pass_t<double, R, true, 2, true>(d, K, grid, batch, threads, big, st, occ, tab);
  return launch_pass_t<double, R, true, 2, false>(d, K, grid, batch, threads, big, st, occ, tab);
}

hipError_t launch_pass_c128(int R, bool two, int phase, bool post, const PassDesc* d, int K, int grid, int batch,
                           int threads, bool big, hipStream_t st, int* occ, int tab) {
  switch (R) {
    case 4: return launch_c128_r<4>(two, phase, post, d, K, grid, batch, threads, big, st, occ, tab);
    case 3: return launch_c128_r<3>(two, phase, post, d, K, grid, batch, threads, big, st, occ, tab);
    case 2: return launch_c128_r<2>(two, phase, post, d, K, grid, batch, threads, big, st, occ, tab);
    case 1: return launch_c128_r<1>(two, phase, post, d, K, grid, batch, threads, big, st, occ, tab);
  }
  return hipErrorInvalidValue;
}

hipError_t set_attributes_c128() {
  hipError_t e = hipSuccess;
#define QG_A(R, TWO, PH, POST) \
  if (e == hipSuccess) e = set_attr_t<double, R, TWO, PH, POST>();
#define QG_R(R)                                                                               \
  QG_A(R, false, 0, false) QG_A(R, false, 1, false) QG_A(R, false, 2, false) QG_A(R, true, 0, false) \
  QG_A(R, true, 0, true) QG_A(R, true, 1, true) QG_A(R, true, 2, true) QG_A(R, true, 1, false) QG_A(R, true, 2, false)
  QG_R(4) QG_R(3) QG_R(2) QG_R(1)
#undef QG_R
#undef QG_A
  return e;
}

}  // namespace qg

// complex64-storage instantiations of the tile-pass kernel with FP64 LDS
// tiles: the state is rounded to complex64 only when a pass stores it.  Used
// for long McClean circuits, where the roundings at the transposes would add
// up past the complex64 error bound.  Phase-free passes, 16 amplitudes per
// thread.
#include "qg_pass.hpp"

namespace qg {

hipError_t launch_pass_c64x(int R, bool two, int phase, bool post, const PassDesc* d, int K, int grid, int batch,
                            int threads, bool big, hipStream_t st, int* occ, int tab) {
  if (R != 4 || phase != 0) return hipErrorInvalidValue;
  if (!two) {
    if (post) return hipErrorInvalidValue;
    return launch_pass_t<F32L64, 4, false, 0, false>(d, K, grid, batch, threads, big, st, occ, tab);
  }
  if (post) return launch_pass_t<F32L64, 4, true, 0, true>(d, K, grid, batch, threads, big, st, occ, tab);
  return launch_pass_t<F32L64, 4, true, 0, false>(d, K, grid, batch, threads, big, st, occ, tab);
}

hipError_t set_attributes_c64x() {
  hipError_t e = set_attr_t<F32L64, 4, false, 0, false>();
  if (e == hipSuccess) e = set_attr_t<F32L64, 4, true, 0, false>();
  if (e == hipSuccess) e = set_attr_t<F32L64, 4, true, 0, true>();
  return e;
}

}  // namespace qg

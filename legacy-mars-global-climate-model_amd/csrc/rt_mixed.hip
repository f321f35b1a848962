// rt_mixed.hip -- the mixed-precision path: FP32 k-interpolation, optical
// depths and (below the top FP64 layers) layer constants and transmissions;
// FP64 elimination, substitution, flux sums and heating.
#include "rt_column.h"

namespace mgcm {
hipError_t launch_mixed(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  return launch_column<true>(dt, a, st);
}
hipError_t launch_batch_mixed(int cls, int kind, const DevTables &dt, const RunArgs *grids, const uint2 *work,
                          int nwork, int nblocks, size_t lds, hipStream_t st) {
  return launch_batch_class<true>(cls, kind, dt, grids, work, nwork, nblocks, lds, st);
}
}  // namespace mgcm

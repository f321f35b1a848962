// rt_mixed.hip -- the mixed-precision path: the k-tables are read in FP32 (half
// the gather bytes and corner registers); every operation, the interpolation
// included, is FP64, as on the plain path.
#include "rt_column.h"

namespace mgcm {
hipError_t launch_mixed(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  return launch_column<P_TAB32>(dt, a, st);
}
hipError_t launch_batch_mixed(int cls, int kind, int waves, const DevTables &dt, const RunArgs *grids,
                              const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st) {
  return launch_batch_class<P_TAB32>(cls, kind, waves, dt, grids, work, nwork, nblocks, lds, st);
}
}  // namespace mgcm

// rt_mixed.hip -- the mixed-precision path: the k-tables are read in FP32 (half
// the gather bytes and corner registers); every operation, the interpolation
// included, is FP64, as on the plain path.  The one-slot tile here; the 2- and
// 4-slot tiles in rt_tile_mixed.hip.
#include "rt_column.h"
#include "rt_ctx.h"

namespace mgcm {
hipError_t launch_mixed(const DevTables &dt, const RunArgs &a, int tile, hipStream_t st) {
  return tile == 1 ? launch_column<P_TAB32, 1>(dt, a, st) : launch_mixed_tile(dt, a, tile, st);
}
hipError_t launch_batch_mixed(int cls, int kind, int waves, int tile, const DevTables &dt, const RunArgs *grids,
                              const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st) {
  return tile == 1 ? launch_batch_class<P_TAB32, 1>(cls, kind, waves, dt, grids, work, nwork, nblocks, lds, st)
                   : launch_batch_mixed_tile(cls, kind, waves, tile, dt, grids, work, nwork, nblocks, lds, st);
}
}  // namespace mgcm

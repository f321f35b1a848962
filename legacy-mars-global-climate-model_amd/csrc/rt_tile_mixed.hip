// rt_tile_mixed.hip -- the mixed-precision path's column kernels with 2 and 4
// band slots per column; bit-identical to its one-slot tile.
#include "rt_column.h"
#include "rt_ctx.h"

namespace mgcm {
hipError_t launch_mixed_tile(const DevTables &dt, const RunArgs &a, int tile, hipStream_t st) {
  switch (tile) {
    case 2: return launch_column<P_TAB32, 2>(dt, a, st);
    case 4: return launch_column<P_TAB32, 4>(dt, a, st);
    case TILE_BANDS: return launch_bands<P_TAB32>(dt, a, st);
    default: return hipErrorInvalidValue;
  }
}
hipError_t launch_batch_mixed_tile(int cls, int kind, int waves, int tile, const DevTables &dt, const RunArgs *grids,
                                   const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st) {
  switch (tile) {
    case 2: return launch_batch_class<P_TAB32, 2>(cls, kind, waves, dt, grids, work, nwork, nblocks, lds, st);
    case 4: return launch_batch_class<P_TAB32, 4>(cls, kind, waves, dt, grids, work, nwork, nblocks, lds, st);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace mgcm

// rt_tile_plain.hip -- the plain FP64 path's column kernels with 2 and 4 band
// slots per column (2 and 1 columns per wave), for grids too small to fill the
// device at 4 columns per wave; bit-identical to the one-slot tile.
#include "rt_column.h"
#include "rt_ctx.h"

namespace mgcm {
hipError_t launch_plain_tile(const DevTables &dt, const RunArgs &a, int tile, hipStream_t st) {
  switch (tile) {
    case 2: return launch_column<0, 2>(dt, a, st);
    case 4: return launch_column<0, 4>(dt, a, st);
    case TILE_BANDS: return launch_bands<0>(dt, a, st);
    default: return hipErrorInvalidValue;
  }
}
hipError_t launch_batch_plain_tile(int cls, int kind, int waves, int tile, const DevTables &dt, const RunArgs *grids,
                                   const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st) {
  switch (tile) {
    case 2: return launch_batch_class<0, 2>(cls, kind, waves, dt, grids, work, nwork, nblocks, lds, st);
    case 4: return launch_batch_class<0, 4>(cls, kind, waves, dt, grids, work, nwork, nblocks, lds, st);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace mgcm

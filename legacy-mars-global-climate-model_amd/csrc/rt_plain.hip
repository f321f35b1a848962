// rt_plain.hip -- the plain FP64 path: every statement in the reference's
// FP64 evaluation order (compiled with -ffp-contract=off).  The one-slot tile
// (4 columns per wave) here; the 2- and 4-slot tiles in rt_tile_plain.hip.
#include "rt_column.h"
#include "rt_ctx.h"

namespace mgcm {
size_t clr_workspace_bytes(int ncol, int nlay) { return clr_bytes(ncol, nlay); }
size_t grid_workspace_bytes(int ncol, int nlay, int tile) {
  return clr_bytes(ncol, nlay) + (tile == TILE_BANDS ? part_bytes(ncol, nlay) : 0);
}
int grid_tile(const mgcm_rt_ctx *c, int ncol, int nlay, size_t max_ws) {
  if (c->tile) return c->tile;
  return auto_tile(ncol, nlay, c->ncu, clr_bytes(ncol, nlay) + part_bytes(ncol, nlay) <= max_ws);
}

hipError_t launch_plain(const DevTables &dt, const RunArgs &a, int tile, hipStream_t st) {
  return tile == 1 ? launch_column<0, 1>(dt, a, st) : launch_plain_tile(dt, a, tile, st);
}
hipError_t launch_batch_plain(int cls, int kind, int waves, int tile, const DevTables &dt, const RunArgs *grids,
                              const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st) {
  return tile == 1 ? launch_batch_class<0, 1>(cls, kind, waves, dt, grids, work, nwork, nblocks, lds, st)
                   : launch_batch_plain_tile(cls, kind, waves, tile, dt, grids, work, nwork, nblocks, lds, st);
}
}  // namespace mgcm

// rt_plain.hip -- the plain FP64 path: every statement in the reference's
// FP64 evaluation order (compiled with -ffp-contract=off).
#include "rt_column.h"

namespace mgcm {
size_t clr_workspace_bytes(int ncol, int nlay) { return clr_bytes(ncol, nlay); }

hipError_t launch_plain(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  return launch_column<0>(dt, a, st);
}
hipError_t launch_batch_plain(int cls, int kind, int waves, const DevTables &dt, const RunArgs *grids,
                              const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st) {
  return launch_batch_class<0>(cls, kind, waves, dt, grids, work, nwork, nblocks, lds, st);
}
}  // namespace mgcm

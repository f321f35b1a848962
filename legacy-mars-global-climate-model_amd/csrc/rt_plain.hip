// rt_plain.hip -- the plain FP64 path: every statement in the reference's
// FP64 evaluation order (compiled with -ffp-contract=off).
#include "rt_column.h"

namespace mgcm {
hipError_t launch_plain(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  return launch_column<false>(dt, a, st);
}
}  // namespace mgcm

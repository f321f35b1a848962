// rt_mixed.hip -- mixed-precision path (placeholder until the FP32 optics land).
#include "rt_internal.h"
namespace mgcm {
hipError_t launch_mixed(const DevTables &, const RunArgs &, hipStream_t) { return hipErrorNotSupported; }
}  // namespace mgcm

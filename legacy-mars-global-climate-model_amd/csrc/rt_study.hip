// rt_study.hip -- the precision study: the column kernel under reduced-
// precision policies (one component of the work moved to cheaper arithmetic
// at a time, and the combination the mixed-precision configuration names),
// at the 25..36-layer class only.  Reached through mgcm_rt_run with
// MGCM_RT_STUDY | policy bits; the continuum point g = 17 stays FP64.
#include "rt_column.h"
#include "rt_ctx.h"

namespace mgcm {
namespace {
constexpr int S_INTERP = P_TAB32 | P_INTERP32;
constexpr int S_ALL = P_TAB32 | P_INTERP32 | P_TAU32 | P_TRANS32;
constexpr int S_FAST_T32 = P_FASTDIV | P_TAB32;
}  // namespace

hipError_t launch_study(int policy, const DevTables &dt, const RunArgs &a, hipStream_t st) {
  const int nl = a.nlay + 1;
  if (nl <= 25 || nl > 37) return hipErrorInvalidValue;
  switch (policy) {
    case S_INTERP: return launch_nlm<37, S_INTERP, 1>(dt, a, st);
    case P_TAU32: return launch_nlm<37, P_TAU32, 1>(dt, a, st);
    case P_TRANS32: return launch_nlm<37, P_TRANS32, 1>(dt, a, st);
    case S_ALL: return launch_nlm<37, S_ALL, 1>(dt, a, st);
    case P_FASTDIV: return launch_nlm<37, P_FASTDIV, 1>(dt, a, st);
    case S_FAST_T32: return launch_nlm<37, S_FAST_T32, 1>(dt, a, st);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace mgcm
static_assert(mgcm::P_TAB32 << 8 == MGCM_RT_P_TAB32 && mgcm::P_INTERP32 << 8 == MGCM_RT_P_INTERP32 &&
                  mgcm::P_TAU32 << 8 == MGCM_RT_P_TAU32 && mgcm::P_TRANS32 << 8 == MGCM_RT_P_TRANS32 &&
                  mgcm::P_FASTDIV << 8 == MGCM_RT_P_FASTDIV,
              "precision policy bits of the header");

// rt_batched.hip -- many independent grids (ragged layer counts) per call.
#include "rt_ctx.h"

namespace mgcm {
void batch_free(void *) {}

int run_batched_impl(mgcm_rt_ctx *c, int ngrids, const mgcm_rt_columns *grids, const mgcm_rt_fluxes *outs,
                     int precision, hipStream_t st) {
  for (int i = 0; i < ngrids; i++) {
    if (grids[i].ncol == 0) continue;
    RunArgs a;
    fill_args(a, &grids[i], &outs[i], c->d_flag);
    const hipError_t e = precision == MGCM_RT_FP64 ? launch_plain(c->dt, a, st) : launch_mixed(c->dt, a, st);
    if (e != hipSuccess) return MGCM_RT_ERR_DEVICE;
  }
  return MGCM_RT_OK;
}
}  // namespace mgcm

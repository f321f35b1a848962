// rt_batched.hip -- many independent grids (ragged layer counts) per call.
#include "../../include/mgcm_rt.h"
#include "rt_internal.h"

namespace mgcm {
hipError_t launch_plain(const DevTables &dt, const RunArgs &a, hipStream_t st);
hipError_t launch_mixed(const DevTables &dt, const RunArgs &a, hipStream_t st);
size_t workspace_bytes(int ncol, int nlay);
void bind_workspace(RunArgs &a, void *ws);
void fill_args(RunArgs &a, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out, int *flag);
}  // namespace mgcm

using namespace mgcm;

struct mgcm_rt_ctx;
struct BatchState {
  void *ws = nullptr;
  size_t bytes = 0;
};

void batch_free(void *p) {
  BatchState *b = (BatchState *)p;
  if (b->ws) (void)hipFree(b->ws);
  delete b;
}

namespace mgcm {
struct CtxView {
  int device;
  RtConst *d_const;
  double *d_kv, *d_ki, *d_planck;
  float *d_kv32, *d_ki32;
  DevTables dt;
  double *d_ws;
  size_t ws_bytes;
  int *d_flag;
  size_t host_bytes;
  double *d_host;
  void *batch;
};

int run_batched_impl(::mgcm_rt_ctx *ctx, int ngrids, const mgcm_rt_columns *grids,
                     const mgcm_rt_fluxes *outs, int precision, hipStream_t st) {
  CtxView *c = (CtxView *)ctx;
  if (!c->batch) c->batch = new BatchState();
  BatchState *b = (BatchState *)c->batch;
  size_t total = 0;
  for (int i = 0; i < ngrids; i++) total += (workspace_bytes(grids[i].ncol, grids[i].nlay) + 255) & ~(size_t)255;
  if (total > b->bytes) {
    if (b->ws) {
      if (hipDeviceSynchronize() != hipSuccess) return MGCM_RT_ERR_DEVICE;
      (void)hipFree(b->ws);
    }
    b->ws = nullptr;
    b->bytes = 0;
    if (hipMalloc(&b->ws, total) != hipSuccess) return MGCM_RT_ERR_NOMEM;
    b->bytes = total;
  }
  char *w = (char *)b->ws;
  for (int i = 0; i < ngrids; i++) {
    if (grids[i].ncol == 0) continue;
    RunArgs a;
    fill_args(a, &grids[i], &outs[i], c->d_flag);
    bind_workspace(a, w);
    w += (workspace_bytes(grids[i].ncol, grids[i].nlay) + 255) & ~(size_t)255;
    const hipError_t e = precision == MGCM_RT_FP64 ? launch_plain(c->dt, a, st) : launch_mixed(c->dt, a, st);
    if (e != hipSuccess) return MGCM_RT_ERR_DEVICE;
  }
  return MGCM_RT_OK;
}
}  // namespace mgcm

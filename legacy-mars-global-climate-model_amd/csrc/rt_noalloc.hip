// rt_noalloc.hip -- mgcm_rt_run_noalloc: the device-pointer radiation call for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops).  It runs the same kernels as mgcm_rt_run; the one
// scratch buffer of that path -- the continuum-point records the g = 17
// kernel writes and the column kernel reads -- is the file-scope array below
// instead of the context's grow-only workspace.
//
// Every device holds its own copy of the array for the life of the process.
// Host code never passes the array's name to a kernel (on the host it is only a
// stand-in for the symbol): mgcm_rt_init looks up the device address once per
// device (noalloc_setup) and the entry point passes that address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rt_ctx.h"

namespace mgcm {

// 128 MiB: the records of 26,886 columns at 24 layers (4,992 B each), 3,461 at
// the maximum of 99 layers.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)128 << 20;
static_assert(NOALLOC_SCRATCH_BYTES == MGCM_RT_NOALLOC_SCRATCH_BYTES, "scratch size of the header");

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_addr[MAX_DEVICES] = {nullptr};
}  // namespace

// Called by mgcm_rt_init with the context's device current: the device
// address of `scratch` on that device, looked up on the first call per device.
int noalloc_setup(mgcm_rt_ctx *c) {
  c->d_scratch = nullptr;
  if (c->device < 0 || c->device >= MAX_DEVICES) return MGCM_RT_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_addr[c->device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MGCM_RT_ERR_DEVICE;
    g_addr[c->device] = p;
  }
  c->d_scratch = (double *)g_addr[c->device];
  return MGCM_RT_OK;
}

}  // namespace mgcm

extern "C" int mgcm_rt_run_noalloc(mgcm_rt_ctx *c, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out,
                                   int precision, void *stream) {
  using namespace mgcm;
  if (!c) return MGCM_RT_ERR_ARG;
  int s = check_columns(in, out);
  if (s != MGCM_RT_OK) return s;
  int path, bands;
  s = decode_mode(precision, &path, &bands);
  if (s != MGCM_RT_OK) return s;
  if (in->ncol == 0) return MGCM_RT_OK;
  const int tile = grid_tile(c, in->ncol, in->nlay, NOALLOC_SCRATCH_BYTES);
  if (grid_workspace_bytes(in->ncol, in->nlay, tile) > NOALLOC_SCRATCH_BYTES) return MGCM_RT_ERR_SCRATCH;
  if (!c->d_scratch) return MGCM_RT_ERR_DEVICE;  // the address lookup at init failed
  return run_grid(c, in, out, path, bands, c->d_scratch, tile, (hipStream_t)stream) == hipSuccess
             ? MGCM_RT_OK
             : MGCM_RT_ERR_DEVICE;
}

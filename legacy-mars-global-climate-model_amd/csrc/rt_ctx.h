// rt_ctx.h -- host-side context of the C ABI, shared by the entry-point files.
#pragma once
#include "../../include/mgcm_rt.h"
#include "rt_internal.h"

#define HIP_TRY(x)                                   \
  do {                                               \
    if ((x) != hipSuccess) return MGCM_RT_ERR_DEVICE; \
  } while (0)

static_assert(mgcm::MAX_NLAY == MGCM_RT_MAX_NLAY, "layer limit of the kernels and of the C ABI");

struct mgcm_rt_ctx {
  int device;
  mgcm::RtConst *d_const;
  double *d_kv, *d_ki, *d_planck;
  float *d_kv32, *d_ki32;
  mgcm::DevTables dt;
  int *d_flag;
  // continuum-point records of mgcm_rt_run (grow-only)
  double *d_clr;
  size_t clr_bytes;
  // device staging of the host-pointer path (grow-only)
  size_t host_bytes;
  double *d_host;
  // batched-path state (rt_batched.hip)
  void *batch;
  // mgcm_rt_run_host's streams, pinned slots and copy threads (rt_host.hip)
  void *host;
  // aerosol producers' size grid and optical tables (rt_aerosol.hip)
  double *d_aer;
  // device address of the no-allocation entry point's scratch array on this
  // context's device (rt_noalloc.hip; not owned: it lives until the process ends)
  double *d_scratch;
  // column kernel tile of single-grid calls: band slots per column (1, 2, 4,
  // or TILE_BANDS: one band per wave; 0 = chosen per grid by auto_tile) and
  // the device's CU count
  int tile;
  int ncu;
};

namespace mgcm {
// tile: band slots per column of the column kernel (1, 2, 4 or TILE_BANDS)
hipError_t launch_plain(const DevTables &dt, const RunArgs &a, int tile, hipStream_t st);
hipError_t launch_mixed(const DevTables &dt, const RunArgs &a, int tile, hipStream_t st);
hipError_t launch_plain_tile(const DevTables &dt, const RunArgs &a, int tile, hipStream_t st);  // 2, 4, bands
hipError_t launch_mixed_tile(const DevTables &dt, const RunArgs &a, int tile, hipStream_t st);
// The precision study's column kernels (rt_study.hip): policy = P_* bits;
// hipErrorInvalidValue outside the 25..36-layer class or for a policy not built.
hipError_t launch_study(int policy, const DevTables &dt, const RunArgs &a, hipStream_t st);
// waves, tile: the column kernel instance's waves per SIMD (1 or 2) and band
// slots per column (1, 2 or 4; one wave per SIMD at tile 1 only); kind 0 only.
// nwork counts 4-column items.
hipError_t launch_batch_plain(int cls, int kind, int waves, int tile, const DevTables &dt, const RunArgs *grids,
                              const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st);
hipError_t launch_batch_mixed(int cls, int kind, int waves, int tile, const DevTables &dt, const RunArgs *grids,
                              const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st);
hipError_t launch_batch_plain_tile(int cls, int kind, int waves, int tile, const DevTables &dt, const RunArgs *grids,
                                   const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st);
hipError_t launch_batch_mixed_tile(int cls, int kind, int waves, int tile, const DevTables &dt, const RunArgs *grids,
                                   const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st);
size_t clr_workspace_bytes(int ncol, int nlay);
// The workspace of a single-grid call at a tile: the continuum-point records,
// plus the band records at TILE_BANDS.
size_t grid_workspace_bytes(int ncol, int nlay, int tile);
// The tile of a single-grid call: the context's, or auto_tile's choice among
// the tiles whose workspace fits in max_ws bytes.
int grid_tile(const mgcm_rt_ctx *ctx, int ncol, int nlay, size_t max_ws = (size_t)-1);
// Splits the precision argument into the path (MGCM_RT_FP64/MIXED, or
// MGCM_RT_STUDY with its policy bits in path >> 8) and the band mask;
// MGCM_RT_ERR_PRECISION for anything else.
int decode_mode(int precision, int *path, int *bands);
int check_columns(const mgcm_rt_columns *in, const mgcm_rt_fluxes *out);
void fill_args(RunArgs &a, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out, int *flag);
int run_batched_impl(mgcm_rt_ctx *ctx, int ngrids, const mgcm_rt_columns *grids, const mgcm_rt_fluxes *outs,
                     int precision, hipStream_t st);
void batch_free(void *state);
// Looks up the scratch array of mgcm_rt_run_noalloc on the current device
// (once per device) into ctx->d_scratch.
int noalloc_setup(mgcm_rt_ctx *ctx);
// Waits for stream st only, then reads and clears the context's range flag.
int check_on_stream(mgcm_rt_ctx *ctx, hipStream_t st);
int batched_tuning_impl(const mgcm_rt_ctx *ctx, int maxcls, mgcm_rt_batch_class *info);
// mgcm_rt_run's launches on a grid whose arguments are checked, at a tile,
// with a workspace ws of grid_workspace_bytes(in->ncol, in->nlay, tile) bytes.
hipError_t run_grid(const mgcm_rt_ctx *c, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out, int path,
                    int bands, double *ws, int tile, hipStream_t st);
void host_free(void *state);
// DSOLFLUX + DNVFLUX of a grid (rt_surface.hip); dnvflux may be null.
hipError_t launch_dsolflux(const RtConst *rc, int ncol, const double *sol, const double *acosz,
                           const double *detau, const double *diffvt, double *directsol, double *dnvflux,
                           hipStream_t st);
// DSOLFLUX of many grids: a persistent launch over (grid, first column) items of
// 16 columns; grids whose RunArgs::directsol is null are skipped.
constexpr int DSOLFLUX_ITEM_COLS = 16;
hipError_t launch_dsolflux_batch(const RtConst *rc, const RunArgs *grids, const uint2 *work, int nwork, int nblocks,
                                 hipStream_t st);
// The context's grow-only device block for the host-pointer entry points
// (shared with mgcm_rt_run_host's staging); null when it cannot grow.
double *host_staging(mgcm_rt_ctx *ctx, size_t bytes);
}  // namespace mgcm

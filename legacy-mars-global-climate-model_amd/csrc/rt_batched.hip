// rt_batched.hip -- many independent grids (ragged column and layer counts) in
// one call: mgcm_rt_run_batched.
//
// Plan (rebuilt when the batch's shape changes):
//   * grids are grouped by layer-count class (kernel instance NLM);
//   * each class gets a work list of (grid, first column) 4-column blocks,
//     grids with more layers first, columns in order;
//   * one persistent kernel per class walks its list round-robin; the tuner
//     picks the number of resident blocks per class on the first call of a
//     shape (timed on a prefix of the list);
//   * the class launches are captured once into a HIP graph.
// Per call: the grids' argument blocks (device pointers) go to the device with
// one asynchronous copy on the caller's stream, then the graph is launched on
// that stream.  Nothing is allocated once a shape has been planned.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "rt_column.h"
#include "rt_ctx.h"

namespace mgcm {

namespace {

// Per layer class, three work lists: kind 0 = column kernel (4-column blocks),
// kind 1 = VIS continuum points (9-column blocks), kind 2 = IR continuum
// points (12-column blocks).
constexpr int NKIND = 3;
constexpr int KIND_COLS[NKIND] = {4, 9, 12};

struct ClassPlan {
  int nwork[NKIND] = {0, 0, 0};
  size_t work_off[NKIND] = {0, 0, 0};  // offsets of the lists in d_work
  size_t lds = 0;                      // dynamic LDS per column-kernel block
  int nblocks[NKIND] = {0, 0, 0};      // persistent grid sizes (kind 0 from the tuner)
};

struct BatchState {
  std::vector<long long> sig;  // precision, ngrids, (ncol, nlay) per grid
  ClassPlan cls[NCLASS];
  RunArgs *d_grids = nullptr;
  RunArgs *h_grids = nullptr;  // pinned staging of the argument blocks
  size_t grids_cap = 0;
  uint2 *d_work = nullptr;
  size_t work_cap = 0;
  double *d_clr = nullptr;            // continuum-point records of every grid
  size_t clr_cap = 0;
  std::vector<size_t> clr_off;        // per grid, in doubles
  hipEvent_t uploaded = nullptr;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

void drop_graph(BatchState *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

hipError_t launch_class(int precision, int k, int kind, const DevTables &dt, const BatchState *b, int nwork,
                        int nblocks, hipStream_t st) {
  const ClassPlan &p = b->cls[k];
  const uint2 *work = b->d_work + p.work_off[kind];
  return (precision & 0xf) == MGCM_RT_FP64
             ? launch_batch_plain(k, kind, dt, b->d_grids, work, nwork, nblocks, p.lds, st)
             : launch_batch_mixed(k, kind, dt, b->d_grids, work, nwork, nblocks, p.lds, st);
}

// Resident one-wave blocks per CU for `lds` bytes of LDS at 2 waves per SIMD.
int resident_per_cu(size_t lds) {
  const int by_lds = (int)((160 * 1024) / std::max<size_t>(lds, 1));
  return std::max(1, std::min(8, by_lds));
}

// Picks the persistent grid size of one class: 1x, 2x or 4x the resident
// capacity, timed on a prefix of the work list (the prefix's outputs are
// written here and rewritten by the real run).
int tune_class(int precision, int k, const DevTables &dt, const BatchState *b, int ncu, hipStream_t st) {
  const ClassPlan &p = b->cls[k];
  const int cap = ncu * resident_per_cu(p.lds);
  if (p.nwork[0] <= cap) return p.nwork[0];
  const int prefix = std::min(p.nwork[0], 8 * cap);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return cap;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return cap;
  }
  int best = cap;
  float best_ms = 1e30f;
  const int cand[3] = {cap, 2 * cap, 4 * cap};
  for (int c : cand) {
    const int nb = std::min(c, prefix);
    if (launch_class(precision, k, 0, dt, b, prefix, nb, st) != hipSuccess) break;  // warm-up
    (void)hipEventRecord(e0, st);
    if (launch_class(precision, k, 0, dt, b, prefix, nb, st) != hipSuccess) break;
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) break;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return std::min(best, p.nwork[0]);
}

int plan(BatchState *b, int ngrids, const mgcm_rt_columns *grids) {
  drop_graph(b);
  // work lists, grids with more layers first
  std::vector<int> order(ngrids);
  for (int i = 0; i < ngrids; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return grids[x].nlay > grids[y].nlay; });
  std::vector<uint2> work;
  for (int k = 0; k < NCLASS; k++) {
    ClassPlan &p = b->cls[k];
    p = ClassPlan();
    int maxlay = 0;
    for (int kind = 0; kind < NKIND; kind++) {
      p.work_off[kind] = work.size();
      for (int i : order) {
        if (grids[i].ncol == 0 || nl_class(grids[i].nlay + 1) != k) continue;
        maxlay = std::max(maxlay, grids[i].nlay);
        for (int col = 0; col < grids[i].ncol; col += KIND_COLS[kind])
          work.push_back(make_uint2((unsigned)i, (unsigned)col));
      }
      p.nwork[kind] = (int)(work.size() - p.work_off[kind]);
    }
    p.lds = maxlay ? stage_bytes(maxlay) : 0;
  }
  // continuum-point records, grid after grid
  b->clr_off.assign(ngrids, 0);
  size_t clr_total = 0;
  for (int i = 0; i < ngrids; i++) {
    b->clr_off[i] = clr_total;
    clr_total += clr_workspace_bytes(grids[i].ncol, grids[i].nlay) / sizeof(double);
  }
  if (clr_total * sizeof(double) > b->clr_cap) {
    if (b->d_clr) (void)hipFree(b->d_clr);
    b->d_clr = nullptr;
    b->clr_cap = 0;
    if (hipMalloc(&b->d_clr, clr_total * sizeof(double)) != hipSuccess) return MGCM_RT_ERR_NOMEM;
    b->clr_cap = clr_total * sizeof(double);
  }
  if (work.size() > b->work_cap) {
    if (b->d_work) (void)hipFree(b->d_work);
    b->d_work = nullptr;
    b->work_cap = 0;
    if (hipMalloc(&b->d_work, work.size() * sizeof(uint2)) != hipSuccess) return MGCM_RT_ERR_NOMEM;
    b->work_cap = work.size();
  }
  if (hipMemcpy(b->d_work, work.data(), work.size() * sizeof(uint2), hipMemcpyHostToDevice) != hipSuccess)
    return MGCM_RT_ERR_DEVICE;
  if ((size_t)ngrids > b->grids_cap) {
    if (b->d_grids) (void)hipFree(b->d_grids);
    if (b->h_grids) (void)hipHostFree(b->h_grids);
    b->d_grids = nullptr;
    b->h_grids = nullptr;
    b->grids_cap = 0;
    if (hipMalloc(&b->d_grids, ngrids * sizeof(RunArgs)) != hipSuccess) return MGCM_RT_ERR_NOMEM;
    if (hipHostMalloc(&b->h_grids, ngrids * sizeof(RunArgs), hipHostMallocDefault) != hipSuccess)
      return MGCM_RT_ERR_NOMEM;
    b->grids_cap = ngrids;
  }
  if (!b->uploaded && hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming) != hipSuccess)
    return MGCM_RT_ERR_DEVICE;
  if (!b->cap && hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess) return MGCM_RT_ERR_DEVICE;
  return MGCM_RT_OK;
}

int upload_args(BatchState *b, mgcm_rt_ctx *c, int ngrids, const mgcm_rt_columns *grids,
                const mgcm_rt_fluxes *outs, int bands, hipStream_t st) {
  // the previous call's copy out of the pinned staging must have finished
  if (hipEventSynchronize(b->uploaded) != hipSuccess) return MGCM_RT_ERR_DEVICE;
  for (int i = 0; i < ngrids; i++) {
    fill_args(b->h_grids[i], &grids[i], &outs[i], c->d_flag);
    b->h_grids[i].bands = bands;
    b->h_grids[i].clr = b->d_clr + b->clr_off[i];
  }
  if (hipMemcpyAsync(b->d_grids, b->h_grids, ngrids * sizeof(RunArgs), hipMemcpyHostToDevice, st) != hipSuccess)
    return MGCM_RT_ERR_DEVICE;
  if (hipEventRecord(b->uploaded, st) != hipSuccess) return MGCM_RT_ERR_DEVICE;
  return MGCM_RT_OK;
}

}  // namespace

void batch_free(void *p) {
  BatchState *b = (BatchState *)p;
  if (!b) return;
  drop_graph(b);
  if (b->d_grids) (void)hipFree(b->d_grids);
  if (b->h_grids) (void)hipHostFree(b->h_grids);
  if (b->d_work) (void)hipFree(b->d_work);
  if (b->d_clr) (void)hipFree(b->d_clr);
  if (b->uploaded) (void)hipEventDestroy(b->uploaded);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  delete b;
}

int run_batched_impl(mgcm_rt_ctx *c, int ngrids, const mgcm_rt_columns *grids, const mgcm_rt_fluxes *outs,
                     int precision, hipStream_t st) {
  if (ngrids < 0 || (ngrids > 0 && (!grids || !outs))) return MGCM_RT_ERR_ARG;
  int path, bands;
  const int sm = decode_mode(precision, &path, &bands);
  if (sm != MGCM_RT_OK) return sm;
  long long total = 0;
  for (int i = 0; i < ngrids; i++) {
    const int s = check_columns(&grids[i], &outs[i]);
    if (s != MGCM_RT_OK) return s;
    total += grids[i].ncol;
  }
  if (total == 0) return MGCM_RT_OK;
  if (!c->batch) c->batch = new (std::nothrow) BatchState();
  BatchState *b = (BatchState *)c->batch;
  if (!b) return MGCM_RT_ERR_NOMEM;

  std::vector<long long> sig;
  sig.reserve(2 + 2 * (size_t)ngrids);
  sig.push_back(precision);
  sig.push_back(ngrids);
  for (int i = 0; i < ngrids; i++) {
    sig.push_back(grids[i].ncol);
    sig.push_back(grids[i].nlay);
  }
  const bool fresh = sig != b->sig;
  if (fresh) {
    b->sig.clear();
    const int s = plan(b, ngrids, grids);
    if (s != MGCM_RT_OK) return s;
  }
  int s = upload_args(b, c, ngrids, grids, outs, bands, st);
  if (s != MGCM_RT_OK) return s;
  if (fresh) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || ncu <= 0)
      return MGCM_RT_ERR_DEVICE;
    // continuum points before the column kernels that weight them
    for (int k = 0; k < NCLASS; k++)
      for (int kind = 1; kind < NKIND; kind++)
        if ((bands & kind) && b->cls[k].nwork[kind] > 0) {
          if (launch_class(precision, k, kind, c->dt, b, b->cls[k].nwork[kind],
                           std::min(b->cls[k].nwork[kind], 16 * ncu), st) != hipSuccess)
            return MGCM_RT_ERR_DEVICE;
        }
    for (int k = 0; k < NCLASS; k++) {
      ClassPlan &p = b->cls[k];
      if (p.nwork[0] > 0) p.nblocks[0] = tune_class(precision, k, c->dt, b, ncu, st);
      for (int kind = 1; kind < NKIND; kind++) p.nblocks[kind] = std::min(p.nwork[kind], 16 * ncu);
    }
    // capture the class launches; the argument upload stays outside the graph
    if (hipStreamSynchronize(st) != hipSuccess) return MGCM_RT_ERR_DEVICE;
    if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) return MGCM_RT_ERR_DEVICE;
    bool ok = true;
    for (int kind = 1; kind < NKIND && ok; kind++)
      for (int k = 0; k < NCLASS && ok; k++)
        if ((bands & kind) && b->cls[k].nwork[kind] > 0)
          ok = launch_class(precision, k, kind, c->dt, b, b->cls[k].nwork[kind], b->cls[k].nblocks[kind], b->cap) ==
               hipSuccess;
    for (int k = 0; k < NCLASS && ok; k++)
      if (b->cls[k].nwork[0] > 0)
        ok = launch_class(precision, k, 0, c->dt, b, b->cls[k].nwork[0], b->cls[k].nblocks[0], b->cap) == hipSuccess;
    hipGraph_t g = nullptr;
    const bool ended = hipStreamEndCapture(b->cap, &g) == hipSuccess;
    if (!ok || !ended || !g) {
      if (g) (void)hipGraphDestroy(g);
      return MGCM_RT_ERR_DEVICE;
    }
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
      drop_graph(b);
      return MGCM_RT_ERR_DEVICE;
    }
    b->sig = sig;
  }
  if (hipGraphLaunch(b->exec, st) != hipSuccess) return MGCM_RT_ERR_DEVICE;
  // surface fluxes for TEMPGR of the grids that ask for them (code/comp3.f:336-348)
  for (int i = 0; i < ngrids; i++)
    if (outs[i].directsol && grids[i].ncol > 0 &&
        launch_dsolflux(c->d_const, grids[i].ncol, grids[i].sol, grids[i].acosz, outs[i].detau, outs[i].diffvt,
                        outs[i].directsol, outs[i].dnvflux, st) != hipSuccess)
      return MGCM_RT_ERR_DEVICE;
  return MGCM_RT_OK;
}

}  // namespace mgcm

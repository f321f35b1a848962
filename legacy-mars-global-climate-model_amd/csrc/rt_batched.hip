// rt_batched.hip -- many independent grids (ragged column and layer counts) in
// one call: mgcm_rt_run_batched.
//
// Plan (rebuilt when the batch's shape changes):
//   * grids are grouped by layer-count class (kernel instance NLM);
//   * each class gets a work list of (grid, first column) 4-column blocks,
//     grids with more layers first, columns in order;
//   * one persistent kernel per class walks its list round-robin; on the first
//     call of a shape the tuner times, per class, the column kernel's tiles
//     (1, 2 or 4 band slots per column: 4, 2 or 1 columns per wave) and
//     instances (one or two waves per SIMD) at multiples of their resident
//     capacity on a prefix of the list and keeps the fastest;
//   * DSOLFLUX of the grids that ask for surface fluxes is one more persistent
//     launch over (grid, 16 columns) items;
//   * the launches are captured once into a HIP graph.
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

// Tuner candidates of the column kernel: (band slots per column, waves per
// SIMD, multiple of the instance's resident capacity).
constexpr int NCAND = MGCM_RT_NCAND;
constexpr int CAND_TILE[NCAND] = {1, 1, 1, 1, 1, 1, 2, 2, 4, 4};
constexpr int CAND_WAVES[NCAND] = {2, 2, 2, 1, 1, 1, 2, 2, 2, 2};
constexpr int CAND_MULT[NCAND] = {1, 2, 4, 1, 2, 4, 1, 2, 1, 2};

struct ClassPlan {
  int nwork[NKIND] = {0, 0, 0};
  size_t work_off[NKIND] = {0, 0, 0};  // offsets of the lists in d_work
  size_t lds = 0;                      // dynamic LDS per column-kernel block
  int nblocks[NKIND] = {0, 0, 0};      // persistent grid sizes (kind 0 from the tuner)
  int tile = 1;                        // column kernel tile (from the tuner)
  int waves = MGCM_FP64_WAVES;         // column kernel instance (from the tuner)
  int tuned_prefix = 0;                // work items the candidates were timed on (0: not timed)
  float cand_ms[NCAND] = {};           // per candidate; 0 where not timed
};

struct BatchState {
  std::vector<long long> sig;  // precision, ngrids, (ncol, nlay) per grid
  ClassPlan cls[NCLASS];
  RunArgs *d_grids = nullptr;
  RunArgs *h_grids = nullptr;  // pinned staging of the argument blocks
  size_t grids_cap = 0;
  uint2 *d_work = nullptr;
  size_t work_cap = 0;
  size_t dsol_off = 0;                // DSOLFLUX items (grid, first column) in d_work
  int dsol_nwork = 0;
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

hipError_t launch_class(int precision, int k, int kind, int waves, int tile, const DevTables &dt,
                        const BatchState *b, int nwork, int nblocks, hipStream_t st) {
  const ClassPlan &p = b->cls[k];
  const uint2 *work = b->d_work + p.work_off[kind];
  return (precision & 0xf) == MGCM_RT_FP64
             ? launch_batch_plain(k, kind, waves, tile, dt, b->d_grids, work, nwork, nblocks, p.lds, st)
             : launch_batch_mixed(k, kind, waves, tile, dt, b->d_grids, work, nwork, nblocks, p.lds, st);
}

// Picks the column kernel's tile, instance and persistent grid size of one
// class: each candidate (band slots per column, waves per SIMD, multiple of
// that instance's resident capacity, at most one block per sub-item) runs
// twice on a prefix of the work list, the second run timed; the fastest is
// kept (only the candidates of tile force_tile when it is not 0).  The
// prefix's outputs are written here and rewritten by the real run.
void tune_class(int precision, int k, const DevTables &dt, BatchState *b, int ncu, int force_tile,
                hipStream_t st) {
  ClassPlan &p = b->cls[k];
  const int cap2 = ncu * resident_per_cu(p.lds, 2);
  p.tile = force_tile ? force_tile : 1;
  p.waves = MGCM_FP64_WAVES;
  p.nblocks[0] = std::min(p.nwork[0] * p.tile, cap2);
  p.tuned_prefix = 0;
  for (int c = 0; c < NCAND; c++) p.cand_ms[c] = 0.f;
  const int prefix = std::min(p.nwork[0], 8 * cap2);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return;
  }
  auto blocks_of = [&](int c) {
    return std::min(CAND_MULT[c] * ncu * resident_per_cu(p.lds, CAND_WAVES[c]), prefix * CAND_TILE[c]);
  };
  float best_ms = 1e30f;
  for (int c = 0; c < NCAND; c++) {
    const int s = CAND_TILE[c], w = CAND_WAVES[c], nb = blocks_of(c);
    if (force_tile && s != force_tile) continue;
    bool repeat = false;  // the same launch as an earlier candidate
    for (int e = 0; e < c; e++) repeat = repeat || (CAND_TILE[e] == s && CAND_WAVES[e] == w && blocks_of(e) == nb);
    if (repeat) continue;
    if (launch_class(precision, k, 0, w, s, dt, b, prefix, nb, st) != hipSuccess) break;  // warm-up
    (void)hipEventRecord(e0, st);
    if (launch_class(precision, k, 0, w, s, dt, b, prefix, nb, st) != hipSuccess) break;
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) break;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    p.cand_ms[c] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      p.tile = s;
      p.waves = w;
      p.nblocks[0] = nb;
    }
  }
  p.tuned_prefix = prefix;
  p.nblocks[0] = std::min(p.nblocks[0], p.nwork[0] * p.tile);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
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
  // DSOLFLUX items of every grid (skipped at run time where not asked for)
  b->dsol_off = work.size();
  for (int i = 0; i < ngrids; i++)
    for (int col = 0; col < grids[i].ncol; col += DSOLFLUX_ITEM_COLS) work.push_back(make_uint2((unsigned)i, (unsigned)col));
  b->dsol_nwork = (int)(work.size() - b->dsol_off);
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
  if ((path & 0xf) == MGCM_RT_STUDY) return MGCM_RT_ERR_PRECISION;  // single grids only
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
  sig.reserve(3 + 2 * (size_t)ngrids);
  sig.push_back(precision);
  const int force_tile = c->tile == TILE_BANDS ? 0 : c->tile;  // single-grid calls only
  sig.push_back(force_tile);
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
          if (launch_class(precision, k, kind, 0, 1, c->dt, b, b->cls[k].nwork[kind],
                           std::min(b->cls[k].nwork[kind], 16 * ncu), st) != hipSuccess)
            return MGCM_RT_ERR_DEVICE;
        }
    for (int k = 0; k < NCLASS; k++) {
      ClassPlan &p = b->cls[k];
      if (p.nwork[0] > 0) tune_class(precision, k, c->dt, b, ncu, force_tile, st);
      for (int kind = 1; kind < NKIND; kind++) p.nblocks[kind] = std::min(p.nwork[kind], 16 * ncu);
    }
    // capture the class launches; the argument upload stays outside the graph
    if (hipStreamSynchronize(st) != hipSuccess) return MGCM_RT_ERR_DEVICE;
    if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) return MGCM_RT_ERR_DEVICE;
    bool ok = true;
    for (int kind = 1; kind < NKIND && ok; kind++)
      for (int k = 0; k < NCLASS && ok; k++)
        if ((bands & kind) && b->cls[k].nwork[kind] > 0)
          ok = launch_class(precision, k, kind, 0, 1, c->dt, b, b->cls[k].nwork[kind], b->cls[k].nblocks[kind],
                            b->cap) == hipSuccess;
    for (int k = 0; k < NCLASS && ok; k++)
      if (b->cls[k].nwork[0] > 0)
        ok = launch_class(precision, k, 0, b->cls[k].waves, b->cls[k].tile, c->dt, b, b->cls[k].nwork[0],
                          b->cls[k].nblocks[0], b->cap) == hipSuccess;
    // surface fluxes for TEMPGR of the grids that ask for them (code/comp3.f:336-348)
    if (ok && b->dsol_nwork > 0)
      ok = launch_dsolflux_batch(c->d_const, b->d_grids, b->d_work + b->dsol_off, b->dsol_nwork,
                                 std::min(b->dsol_nwork, 4 * ncu), b->cap) == hipSuccess;
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
  return MGCM_RT_OK;
}

int batched_tuning_impl(const mgcm_rt_ctx *c, int maxcls, mgcm_rt_batch_class *info) {
  const BatchState *b = (const BatchState *)c->batch;
  if (!b || b->sig.empty()) return 0;
  int n = 0;
  for (int k = 0; k < NCLASS; k++) {
    const ClassPlan &p = b->cls[k];
    if (p.nwork[0] == 0) continue;
    if (n < maxcls && info) {
      mgcm_rt_batch_class &o = info[n];
      o.max_layers = KCLASS_NL[k] - 1;
      o.items = p.nwork[0];
      o.tile = p.tile;
      o.waves_per_simd = p.waves;
      o.blocks = p.nblocks[0];
      o.tuned_items = p.tuned_prefix;
      for (int cnd = 0; cnd < NCAND; cnd++) {
        o.cand_tile[cnd] = CAND_TILE[cnd];
        o.cand_waves[cnd] = CAND_WAVES[cnd];
        o.cand_mult[cnd] = CAND_MULT[cnd];
        o.cand_ms[cnd] = p.cand_ms[cnd];
      }
    }
    n++;
  }
  return n;
}

}  // namespace mgcm

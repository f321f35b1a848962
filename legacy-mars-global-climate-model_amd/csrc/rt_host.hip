// rt_host.hip -- mgcm_rt_run_host, the host-array form of mgcm_rt_run that the
// Fortran host's COMP3 replacement calls (code/comp3.f:637-738 over a grid).
//
// The grid is cut into chunks of whole columns.  Chunk i goes
//   host arrays -> pinned slot (host threads) -> device (copy stream, DMA)
//   -> the radiation on the chunk (run stream i % 2)
//   -> pinned slot (copy stream, DMA) -> host arrays (host threads),
// with two pinned slots each way.  The host keeps two chunks queued ahead of
// the one it drains, so its copies overlap the runs.  The two run streams, each
// with its own continuum-point workspace, let chunk i+1's kernels start while
// chunk i's last waves finish.  Uploads and downloads share one stream: a
// process has 4 hardware queues, and a fifth stream would put both run
// streams on one queue.  Every chunk has its own region of the device staging
// buffer, so the device side has no reuse hazards; pinned slots and workspaces
// are handed back by stream order and events.  A chunk is a sub-grid: its
// columns are computed exactly as in a whole-grid run (columns are
// independent), so the outputs are the same bits.
//
// A call costs the chunks' runs plus one chunk's upload and host copy before
// the first run and one chunk's download and host copy after the last.
#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <new>
#include <string.h>
#include <thread>
#include <vector>

#include "rt_ctx.h"

using namespace mgcm;

namespace mgcm {
namespace {

constexpr int HOST_THREADS = 8;          // copy threads, the caller included
constexpr int PIPE_CHUNKS = 8;           // chunks of a large grid
constexpr size_t PIPE_MIN_COLS = 4096;   // below this a grid is one chunk
constexpr size_t PIN_SLOT_BYTES = 64ull << 20;  // cap of one pinned slot (in or out)
// pinned slots: the download slots non-coherent, as the host reads them only
// after the download's event
constexpr unsigned PIN_IN_FLAGS = hipHostMallocDefault, PIN_OUT_FLAGS = hipHostMallocNonCoherent;

// A fixed team of threads that splits a byte range of copies.  run() hands the
// job to every worker, takes part 0 itself and returns when all parts are done.
class CopyTeam {
 public:
  explicit CopyTeam(int n) : n_(n) {
    try {
      for (int p = 1; p < n; p++) th_.emplace_back([this, p] { work(p); });
    } catch (...) {  // a thread that would not start: end the others, then report
      stop();
      throw;
    }
  }
  ~CopyTeam() { stop(); }
  int size() const { return n_; }
  void run(const std::function<void(int)> &job) {
    {
      std::lock_guard<std::mutex> l(m_);
      job_ = &job;
      pending_ = n_ - 1;
      gen_++;
    }
    go_.notify_all();
    job(0);
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [this] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  void work(int p) {
    unsigned seen = 0;
    for (;;) {
      const std::function<void(int)> *job;
      {
        std::unique_lock<std::mutex> l(m_);
        go_.wait(l, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        job = job_;
      }
      (*job)(p);
      std::lock_guard<std::mutex> l(m_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  void stop() {
    {
      std::lock_guard<std::mutex> l(m_);
      stop_ = true;
    }
    go_.notify_all();
    for (auto &t : th_) t.join();
    th_.clear();
  }
  const int n_;
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable go_, done_;
  const std::function<void(int)> *job_ = nullptr;
  unsigned gen_ = 0;
  int pending_ = 0;
  bool stop_ = false;
};

struct Seg {
  char *dst;
  const char *src;
  size_t bytes;
};

constexpr size_t TEAM_MIN_BYTES = 1u << 20;  // smaller copies stay on the calling thread

// Copies the concatenation of segs split into equal byte shares over the team.
void team_copy(CopyTeam &team, const std::vector<Seg> &segs) {
  size_t total = 0;
  for (const Seg &s : segs) total += s.bytes;
  if (total == 0) return;
  if (total < TEAM_MIN_BYTES) {  // waking the team would cost more than the copy
    for (const Seg &s : segs) memcpy(s.dst, s.src, s.bytes);
    return;
  }
  const int n = team.size();
  team.run([&](int p) {
    size_t lo = total * p / n, hi = total * (p + 1) / n, base = 0;
    for (const Seg &s : segs) {
      const size_t a = std::max(lo, base), b = std::min(hi, base + s.bytes);
      if (a < b) memcpy(s.dst + (a - base), s.src + (a - base), b - a);
      base += s.bytes;
      if (base >= hi) break;
    }
  });
}

struct HostPipe {
  CopyTeam *team = nullptr;
  hipStream_t up = nullptr, run[2] = {};  // up: the copy stream, both directions
  hipEvent_t ev_up[2] = {}, ev_run[4] = {}, ev_down[2] = {};  // ev_run: chunk i % 4
  double *clr[2] = {};  // continuum-point workspace of each run stream
  size_t clr_bytes = 0;
  char *pin_in[2] = {}, *pin_out[2] = {};
  size_t pin_in_bytes = 0, pin_out_bytes = 0;
};

HostPipe *pipe_get(mgcm_rt_ctx *c) {
  if (c->host) return (HostPipe *)c->host;
  HostPipe *h = new (std::nothrow) HostPipe();
  if (!h) return nullptr;
  c->host = h;  // freed by host_free, also when set-up stops half way
  bool ok = hipStreamCreateWithFlags(&h->up, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&h->run[0], hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&h->run[1], hipStreamNonBlocking) == hipSuccess;
  for (int s = 0; ok && s < 4; s++) ok = hipEventCreateWithFlags(&h->ev_run[s], hipEventDisableTiming) == hipSuccess;
  for (int s = 0; ok && s < 2; s++)
    ok = hipEventCreateWithFlags(&h->ev_up[s], hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&h->ev_down[s], hipEventDisableTiming) == hipSuccess;
  if (ok) {
    const unsigned hw = std::thread::hardware_concurrency();
    try {
      h->team = new CopyTeam(std::max(1, std::min<int>(HOST_THREADS, hw ? (int)hw : 1)));
    } catch (...) {
      h->team = nullptr;
    }
    ok = h->team != nullptr;
  }
  if (!ok) {
    host_free(h);
    c->host = nullptr;
    return nullptr;
  }
  return h;
}

// Grow-only pinned slots (called with the device idle).
bool pin_grow(char *(&slot)[2], size_t &have, size_t need, unsigned flags) {
  if (need <= have) return true;
  for (char *&p : slot) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
  }
  have = 0;
  for (char *&p : slot)
    if (hipHostMalloc((void **)&p, need, flags) != hipSuccess) return false;
  have = need;
  return true;
}

// One per-column array of a grid: its host pointer at column 0 and its doubles
// per column.  A chunk's region holds the grid's arrays for the chunk's columns
// in this order: inputs, then DETAU/CUMTAUV (in and out: sun-down columns keep
// the caller's values), then the outputs.
struct Col {
  const double *h_in;  // host source (inputs and DETAU/CUMTAUV)
  double *h_out;       // host destination (DETAU/CUMTAUV and outputs)
  size_t pc;
};

}  // namespace

void host_free(void *p) {
  HostPipe *h = (HostPipe *)p;
  if (!h) return;
  delete h->team;
  for (int s = 0; s < 2; s++) {
    if (h->pin_in[s]) (void)hipHostFree(h->pin_in[s]);
    if (h->pin_out[s]) (void)hipHostFree(h->pin_out[s]);
    if (h->clr[s]) (void)hipFree(h->clr[s]);
    if (h->ev_up[s]) (void)hipEventDestroy(h->ev_up[s]);
    if (h->ev_down[s]) (void)hipEventDestroy(h->ev_down[s]);
    if (h->run[s]) (void)hipStreamDestroy(h->run[s]);
  }
  for (hipEvent_t e : h->ev_run)
    if (e) (void)hipEventDestroy(e);
  if (h->up) (void)hipStreamDestroy(h->up);
  delete h;
}

}  // namespace mgcm

namespace {
int run_host(mgcm_rt_ctx *c, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out, int precision) {
  if (!c) return MGCM_RT_ERR_ARG;
  int s = check_columns(in, out);
  if (s != MGCM_RT_OK) return s;
  int path, bands;
  s = decode_mode(precision, &path, &bands);
  if (s != MGCM_RT_OK) return s;
  if (in->ncol == 0) return MGCM_RT_OK;
  const size_t ncol = in->ncol, nlev = 2 * (size_t)in->nlay + 3, nl = in->nlay + 1, nio = (size_t)NG * NV;

  // the grid's arrays in region order
  std::vector<Col> a = {{in->plev, nullptr, nlev},        {in->tlev, nullptr, nlev},
                        {in->pmid, nullptr, nlev},        {in->tmid, nullptr, nlev},
                        {in->qh2o, nullptr, nlev},        {in->tauref, nullptr, nlev + 1},
                        {in->taurefcld, nullptr, nlev + 1}, {in->als, nullptr, 1},
                        {in->albi, nullptr, 1},           {in->acosz, nullptr, 1}};
  // optional per-level aerosol optics: the non-null arrays in mgcm_rt_aerosol order
  const mgcm_rt_aerosol *hae = in->aerosol;
  mgcm_rt_aerosol dae = {};
  const double **aer_dev[14] = {&dae.qxv_dst, &dae.qsv_dst, &dae.gv_dst, &dae.qxi_dst, &dae.qsi_dst, &dae.gi_dst,
                                &dae.qxv_cld, &dae.qsv_cld, &dae.gv_cld, &dae.qxi_cld, &dae.qsi_cld, &dae.gi_cld,
                                &dae.qextref_dst, &dae.qextref_cld};
  int aer_at[14];
  for (int i = 0; i < 14; i++) aer_at[i] = -1;
  if (hae) {
    const double *src[14] = {hae->qxv_dst, hae->qsv_dst, hae->gv_dst, hae->qxi_dst, hae->qsi_dst, hae->gi_dst,
                             hae->qxv_cld, hae->qsv_cld, hae->gv_cld, hae->qxi_cld, hae->qsi_cld, hae->gi_cld,
                             hae->qextref_dst, hae->qextref_cld};
    for (int i = 0; i < 14; i++) {
      if (!src[i]) continue;
      aer_at[i] = (int)a.size();
      a.push_back({src[i], nullptr, (i < 12 ? (((i / 3) % 2 == 0) ? NV : NI) : 1) * (nlev + 1)});
    }
  }
  const int io0 = (int)a.size();
  a.push_back({out->detau, out->detau, nio});
  a.push_back({out->cumtauv, out->cumtauv, nio});
  const int io1 = (int)a.size();
  for (double *o : {out->fmnetv, out->fluxupv, out->fluxdnv, out->fmneti, out->fluxupi, out->fluxdni})
    a.push_back({nullptr, o, nl});
  for (double *o : {out->nfluxtopv, out->nfluxtopi, out->diffvt}) a.push_back({nullptr, o, 1});
  const int hr_at = out->hr ? (int)a.size() : -1;
  if (out->hr) a.push_back({nullptr, out->hr, nl});
  const int ds_at = out->directsol ? (int)a.size() : -1;
  if (out->directsol) a.push_back({nullptr, out->directsol, 1});
  const int dn_at = (out->directsol && out->dnvflux) ? (int)a.size() : -1;
  if (dn_at >= 0) a.push_back({nullptr, out->dnvflux, 1});
  const int na = (int)a.size();
  std::vector<size_t> off(na + 1, 0);  // doubles per column before array j
  for (int j = 0; j < na; j++) off[j + 1] = off[j] + a[j].pc;
  const size_t region_pc = off[na], in_pc = off[io1], out_pc = region_pc - off[io0];

  // chunking: a large grid in PIPE_CHUNKS pieces, each within a pinned slot
  size_t K = ncol;
  if (ncol > PIPE_MIN_COLS) K = std::max(PIPE_MIN_COLS, (ncol + PIPE_CHUNKS - 1) / PIPE_CHUNKS);
  K = std::min(K, std::max<size_t>(64, PIN_SLOT_BYTES / (8 * std::max(in_pc, out_pc))));
  K = std::min(K, ncol);
  const size_t nchunk = (ncol + K - 1) / K;

  // the staging below belongs to this entry point alone, and its previous call
  // returned only after all of its work had ended
  HostPipe *h = pipe_get(c);
  if (!h) return MGCM_RT_ERR_NOMEM;
  if (!pin_grow(h->pin_in, h->pin_in_bytes, K * in_pc * 8, PIN_IN_FLAGS) ||
      !pin_grow(h->pin_out, h->pin_out_bytes, K * out_pc * 8, PIN_OUT_FLAGS))
    return MGCM_RT_ERR_NOMEM;
  const size_t bytes = ncol * region_pc * 8;
  if (bytes > c->host_bytes) {
    if (c->d_host) (void)hipFree(c->d_host);
    c->d_host = nullptr;
    c->host_bytes = 0;
    if (hipMalloc(&c->d_host, bytes) != hipSuccess) return MGCM_RT_ERR_NOMEM;
    c->host_bytes = bytes;
  }
  // each chunk runs at its own tile; the workspace fits the full and the last chunk
  const int tile_k = grid_tile(c, (int)K, in->nlay), last = (int)(ncol - (nchunk - 1) * K);
  const int tile_last = grid_tile(c, last, in->nlay);
  const size_t ws = std::max(grid_workspace_bytes((int)K, in->nlay, tile_k),
                             grid_workspace_bytes(last, in->nlay, tile_last));
  if (ws > h->clr_bytes) {
    for (double *&p : h->clr) {
      if (p) (void)hipFree(p);
      p = nullptr;
    }
    h->clr_bytes = 0;
    for (double *&p : h->clr)
      if (hipMalloc((void **)&p, ws) != hipSuccess) return MGCM_RT_ERR_NOMEM;
    h->clr_bytes = ws;
  }

  auto chunk_cols = [&](size_t i) { return std::min(K, ncol - i * K); };
  auto region = [&](size_t i) { return c->d_host + i * K * region_pc; };  // chunks before i are all K wide
  std::vector<Seg> segs;
  segs.reserve(na);  // no allocation once copies are in flight
  // host -> pinned in-slot i % 2 -> device, then the run; the slot is free once
  // the upload of chunk i-2 has ended
  auto start = [&](size_t i) -> int {
    const int slot = (int)(i & 1);
    const size_t k = chunk_cols(i), col0 = i * K;
    double *r = region(i);
    if (i >= 2 && hipEventSynchronize(h->ev_up[slot]) != hipSuccess) return MGCM_RT_ERR_DEVICE;
    segs.clear();
    for (int j = 0; j < io1; j++)
      segs.push_back({h->pin_in[slot] + k * off[j] * 8, (const char *)(a[j].h_in + col0 * a[j].pc), k * a[j].pc * 8});
    team_copy(*h->team, segs);
    mgcm_rt_columns din = *in;
    mgcm_rt_fluxes dout = {};
    auto dev = [&](int j) { return r + k * off[j]; };
    din.ncol = (int)k;
    din.plev = dev(0); din.tlev = dev(1); din.pmid = dev(2); din.tmid = dev(3); din.qh2o = dev(4);
    din.tauref = dev(5); din.taurefcld = dev(6); din.als = dev(7); din.albi = dev(8); din.acosz = dev(9);
    if (hae) {
      for (int q = 0; q < 14; q++) *aer_dev[q] = aer_at[q] >= 0 ? dev(aer_at[q]) : nullptr;
      din.aerosol = &dae;
    }
    dout.detau = dev(io0); dout.cumtauv = dev(io0 + 1);
    dout.fmnetv = dev(io1); dout.fluxupv = dev(io1 + 1); dout.fluxdnv = dev(io1 + 2);
    dout.fmneti = dev(io1 + 3); dout.fluxupi = dev(io1 + 4); dout.fluxdni = dev(io1 + 5);
    dout.nfluxtopv = dev(io1 + 6); dout.nfluxtopi = dev(io1 + 7); dout.diffvt = dev(io1 + 8);
    dout.hr = hr_at >= 0 ? dev(hr_at) : nullptr;
    dout.directsol = ds_at >= 0 ? dev(ds_at) : nullptr;
    dout.dnvflux = dn_at >= 0 ? dev(dn_at) : nullptr;
    if (hipMemcpyAsync(r, h->pin_in[slot], k * in_pc * 8, hipMemcpyHostToDevice, h->up) != hipSuccess ||
        hipEventRecord(h->ev_up[slot], h->up) != hipSuccess || hipStreamWaitEvent(h->run[slot], h->ev_up[slot], 0) != hipSuccess)
      return MGCM_RT_ERR_DEVICE;
    if (run_grid(c, &din, &dout, path, bands, h->clr[slot], k == K ? tile_k : tile_last, h->run[slot]) != hipSuccess)
      return MGCM_RT_ERR_DEVICE;
    return hipEventRecord(h->ev_run[i & 3], h->run[slot]) == hipSuccess ? MGCM_RT_OK : MGCM_RT_ERR_DEVICE;
  };
  // device -> pinned out-slot i % 2 (free: chunk i-2 has been drained), after
  // chunk i's run and chunk i+2's upload on the copy stream.  A
  // pitched copy, which the runtime hands to the DMA engine; a plain one it
  // does with a copy kernel that slows the next chunk's run.
  auto fetch = [&](size_t i) -> int {
    const int slot = (int)(i & 1);
    const size_t k = chunk_cols(i), row = out_pc * 8;  // k rows of one column's worth of bytes
    return hipStreamWaitEvent(h->up, h->ev_run[i & 3], 0) == hipSuccess &&
                   hipMemcpy2DAsync(h->pin_out[slot], row, region(i) + k * off[io0], row, row, k,
                                    hipMemcpyDeviceToHost, h->up) == hipSuccess &&
                   hipEventRecord(h->ev_down[slot], h->up) == hipSuccess
               ? MGCM_RT_OK
               : MGCM_RT_ERR_DEVICE;
  };
  // pinned out-slot i % 2 -> the caller's arrays
  auto drain = [&](size_t i) -> int {
    const int slot = (int)(i & 1);
    const size_t k = chunk_cols(i), col0 = i * K;
    if (hipEventSynchronize(h->ev_down[slot]) != hipSuccess) return MGCM_RT_ERR_DEVICE;
    segs.clear();
    for (int j = io0; j < na; j++)
      segs.push_back({(char *)(a[j].h_out + col0 * a[j].pc), h->pin_out[slot] + k * (off[j] - off[io0]) * 8,
                      k * a[j].pc * 8});
    team_copy(*h->team, segs);
    return MGCM_RT_OK;
  };
  // two chunks ahead: chunk i+2 is uploaded and queued before chunk i is
  // fetched and drained, so the device never waits for the host's copies
  int st = MGCM_RT_OK;
  for (size_t i = 0; i < std::min<size_t>(2, nchunk) && st == MGCM_RT_OK; i++) st = start(i);
  for (size_t i = 0; i < nchunk && st == MGCM_RT_OK; i++) {
    if (i + 2 < nchunk) st = start(i + 2);
    if (st == MGCM_RT_OK) st = fetch(i);
    if (st == MGCM_RT_OK) st = drain(i);
  }
  if (st != MGCM_RT_OK) {  // nothing in flight may outlive the call
    (void)hipDeviceSynchronize();
    return st;
  }
  // every run of the call has ended (its outputs were downloaded); the flag
  // holds them and any earlier run of the context not yet checked
  return check_on_stream(c, h->up);
}
}  // namespace

extern "C" int mgcm_rt_run_host(mgcm_rt_ctx *c, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out,
                                int precision) {
  try {
    return run_host(c, in, out, precision);
  } catch (...) {  // host memory exhausted: nothing of the call may stay in flight
    (void)hipDeviceSynchronize();
    return MGCM_RT_ERR_NOMEM;
  }
}

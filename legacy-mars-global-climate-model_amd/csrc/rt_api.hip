// rt_api.hip -- the C ABI of include/mgcm_rt.h: context set-up, runs, and the
// init-time table builders (Planck table, LAGINTERP) as device kernels.
#include <new>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "rt_column.h"
#include "rt_ctx.h"
#include "rt_libm.h"

using namespace mgcm;

extern "C" const char *mgcm_rt_strerror(int s) {
  switch (s) {
    case MGCM_RT_OK: return "ok";
    case MGCM_RT_ERR_ARG: return "invalid argument (null pointer or negative size)";
    case MGCM_RT_ERR_NLAY: return "layer count outside 1..99";
    case MGCM_RT_ERR_DEVICE: return "HIP runtime error";
    case MGCM_RT_ERR_NOMEM: return "out of device memory";
    case MGCM_RT_ERR_RANGE: return "temperature outside the Planck table (50.0..900.0 K) or a NaN input";
    case MGCM_RT_ERR_PRECISION: return "unknown precision mode";
    case MGCM_RT_ERR_IO: return "data file missing, unreadable or not in the expected record layout";
    case MGCM_RT_ERR_SCRATCH:
      return "grid needs more than the no-allocation entry point's 128 MiB scratch (run it in column slices)";
    default: return "unknown status";
  }
}

// Re-lay a Fortran k-table CO2(7,51,10,nb,17) into [nb][51][10][7][16] (gas
// Gauss points fastest): the 16 points of one (T, p, H2O) cell share a
// 128-byte line, which the 16 lanes of a column read together.
static void relayout_ktable(const double *src, int nb, double *dst, float *dst32) {
  for (int nw = 0; nw < nb; nw++)
    for (int p = 0; p < NP; p++)
      for (int h = 0; h < NH; h++)
        for (int t = 0; t < NT; t++)
          for (int g = 0; g < NGAS; g++) {
            const size_t s = (size_t)t + NT * ((size_t)p + NP * ((size_t)h + NH * ((size_t)nw + (size_t)nb * g)));
            const size_t d = ((((size_t)nw * NP + p) * NH + h) * NT + t) * NGAS + g;
            dst[d] = src[s];
            dst32[d] = (float)src[s];
          }
}

extern "C" int mgcm_rt_init(const mgcm_rt_tables *t, mgcm_rt_ctx **out) {
  if (!t || !out) return MGCM_RT_ERR_ARG;
  const double *req[] = {t->co2v, t->co2i, t->pfgasref, t->tgasref, t->wrefh2o, t->gwv, t->gwi,
                         t->fzerov, t->fzeroi, t->tauray, t->dwni, t->planckir,
                         t->qxv_dst, t->qsv_dst, t->gv_dst, t->qxi_dst, t->qsi_dst, t->gi_dst,
                         t->qxv_cld, t->qsv_cld, t->gv_cld, t->qxi_cld, t->qsi_cld, t->gi_cld};
  for (const double *p : req)
    if (!p) return MGCM_RT_ERR_ARG;
  *out = nullptr;
  mgcm_rt_ctx *c = new (std::nothrow) mgcm_rt_ctx();
  if (!c) return MGCM_RT_ERR_NOMEM;
  if (hipGetDevice(&c->device) != hipSuccess) { delete c; return MGCM_RT_ERR_DEVICE; }
  if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
      c->ncu <= 0) {
    delete c;
    return MGCM_RT_ERR_DEVICE;
  }
  c->tile = 0;

  RtConst h;
  memset(&h, 0, sizeof(h));
  memcpy(h.pfgasref, t->pfgasref, sizeof(h.pfgasref));
  h.pf_ascending = 1;
  for (int n = 2; n < NP - 1; n++)
    if (!(h.pfgasref[n - 1] <= h.pfgasref[n])) h.pf_ascending = 0;
  memcpy(h.tgasref, t->tgasref, sizeof(h.tgasref));
  memcpy(h.wrefh2o, t->wrefh2o, sizeof(h.wrefh2o));
  for (int b = 0; b < NB; b++)
    for (int g = 0; g < NG; g++) h.gw[b][g] = (b < NV) ? t->gwv[g + NG * b] : t->gwi[g + NG * (b - NV)];
  for (int b = 0; b < NV; b++) {
    h.fzero[b] = t->fzerov[b];
    h.tauray[b] = t->tauray[b];
    h.qx_dst[b] = t->qxv_dst[b]; h.qs_dst[b] = t->qsv_dst[b]; h.g_dst[b] = t->gv_dst[b];
    h.qx_cld[b] = t->qxv_cld[b]; h.qs_cld[b] = t->qsv_cld[b]; h.g_cld[b] = t->gv_cld[b];
  }
  for (int b = 0; b < NI; b++) {
    h.fzero[NV + b] = t->fzeroi[b];
    h.dwni[b] = t->dwni[b];
    h.qx_dst[NV + b] = t->qxi_dst[b]; h.qs_dst[NV + b] = t->qsi_dst[b]; h.g_dst[NV + b] = t->gi_dst[b];
    h.qx_cld[NV + b] = t->qxi_cld[b]; h.qs_cld[NV + b] = t->qsi_cld[b]; h.g_cld[NV + b] = t->gi_cld[b];
  }
  h.qextref_vis = t->qxv_dst[5];  // QXVDST(K, L_NREFV) with L_NREFV = 6 (code/optcv.f90:114)
  h.qextref_dst = t->qextref_dst;
  h.qextref_cld = t->qextref_cld;

  const size_t nkv = (size_t)NV * NP * NH * NT * NGAS, nki = (size_t)NI * NP * NH * NT * NGAS;
  std::vector<double> kv(nkv), ki(nki);
  std::vector<float> kv32(nkv), ki32(nki);
  relayout_ktable(t->co2v, NV, kv.data(), kv32.data());
  relayout_ktable(t->co2i, NI, ki.data(), ki32.data());
  bool ok = hipMalloc(&c->d_const, sizeof(RtConst)) == hipSuccess &&
            hipMalloc(&c->d_kv, nkv * sizeof(double)) == hipSuccess &&
            hipMalloc(&c->d_ki, nki * sizeof(double)) == hipSuccess &&
            hipMalloc(&c->d_kv32, nkv * sizeof(float)) == hipSuccess &&
            hipMalloc(&c->d_ki32, nki * sizeof(float)) == hipSuccess &&
            hipMalloc(&c->d_planck, (size_t)NPLANCK * NI * sizeof(double)) == hipSuccess &&
            hipMalloc(&c->d_flag, sizeof(int)) == hipSuccess;
  if (ok)
    ok = hipMemcpy(c->d_const, &h, sizeof(h), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_kv, kv.data(), nkv * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_ki, ki.data(), nki * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_kv32, kv32.data(), nkv * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_ki32, ki32.data(), nki * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_planck, t->planckir, (size_t)NPLANCK * NI * sizeof(double), hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemset(c->d_flag, 0, sizeof(int)) == hipSuccess;
  if (!ok) {
    mgcm_rt_destroy(c);
    return MGCM_RT_ERR_DEVICE;
  }
  c->dt.c = c->d_const;
  c->dt.kv = c->d_kv;
  c->dt.ki = c->d_ki;
  c->dt.planck = c->d_planck;
  c->dt.kv32 = c->d_kv32;
  c->dt.ki32 = c->d_ki32;
  // the no-allocation entry point's scratch on this device; a failed lookup
  // leaves it unusable (MGCM_RT_ERR_DEVICE there), the other entries unaffected
  (void)noalloc_setup(c);
  *out = c;
  return MGCM_RT_OK;
}

extern "C" int mgcm_rt_destroy(mgcm_rt_ctx *c) {
  if (!c) return MGCM_RT_ERR_ARG;
  void *bufs[] = {c->d_const, c->d_kv, c->d_ki,   c->d_kv32, c->d_ki32,
                  c->d_planck, c->d_flag, c->d_host, c->d_clr, c->d_aer};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  if (c->batch) mgcm::batch_free(c->batch);
  if (c->host) mgcm::host_free(c->host);
  delete c;
  return MGCM_RT_OK;
}

namespace mgcm {
int decode_mode(int precision, int *path, int *bands) {
  const int p = precision & 0xf, pol = (precision >> 8) & 0x1f;
  const int skip = precision & ~0xf & ~(0x1f << 8);
  if ((p != MGCM_RT_FP64 && p != MGCM_RT_MIXED && p != MGCM_RT_STUDY) || (p != MGCM_RT_STUDY && pol) ||
      (p == MGCM_RT_STUDY && !pol) || (skip & ~(MGCM_RT_SKIP_VIS | MGCM_RT_SKIP_IR)) ||
      skip == (MGCM_RT_SKIP_VIS | MGCM_RT_SKIP_IR))
    return MGCM_RT_ERR_PRECISION;
  *path = p | (pol << 8);
  *bands = ((skip & MGCM_RT_SKIP_VIS) ? 0 : 1) | ((skip & MGCM_RT_SKIP_IR) ? 0 : 2);
  return MGCM_RT_OK;
}

int check_columns(const mgcm_rt_columns *in, const mgcm_rt_fluxes *out) {
  if (!in || !out) return MGCM_RT_ERR_ARG;
  if (in->ncol < 0) return MGCM_RT_ERR_ARG;
  if (in->nlay < 1 || in->nlay > MAX_NLAY) return MGCM_RT_ERR_NLAY;
  if (in->ncol == 0) return MGCM_RT_OK;
  const void *req[] = {in->plev, in->tlev, in->pmid, in->tmid, in->qh2o, in->tauref, in->taurefcld,
                       in->als, in->albi, in->acosz, out->fmnetv, out->fluxupv, out->fluxdnv,
                       out->fmneti, out->fluxupi, out->fluxdni, out->nfluxtopv, out->nfluxtopi,
                       out->diffvt, out->detau, out->cumtauv};
  for (const void *p : req)
    if (!p) return MGCM_RT_ERR_ARG;
  return MGCM_RT_OK;
}

void fill_args(RunArgs &a, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out, int *flag) {
  a.ncol = in->ncol;
  a.nlay = in->nlay;
  a.plev = in->plev; a.tlev = in->tlev; a.pmid = in->pmid; a.tmid = in->tmid; a.qh2o = in->qh2o;
  a.tauref = in->tauref; a.taurefcld = in->taurefcld;
  a.als = in->als; a.albi = in->albi; a.acosz = in->acosz;
  for (int b = 0; b < NV; b++) a.sol[b] = in->sol[b];
  a.fmnetv = out->fmnetv; a.fluxupv = out->fluxupv; a.fluxdnv = out->fluxdnv;
  a.fmneti = out->fmneti; a.fluxupi = out->fluxupi; a.fluxdni = out->fluxdni;
  a.nfluxtopv = out->nfluxtopv; a.nfluxtopi = out->nfluxtopi; a.diffvt = out->diffvt;
  a.detau = out->detau; a.cumtauv = out->cumtauv; a.hr = out->hr;
  a.flag = flag;
  a.bands = 3;
  a.clr = nullptr;
  a.part = nullptr;
  a.tp = nullptr;
  const mgcm_rt_aerosol *ae = in->aerosol;
  const double *aq[12] = {nullptr};
  if (ae) {
    const double *src[12] = {ae->qxv_dst, ae->qsv_dst, ae->gv_dst, ae->qxi_dst, ae->qsi_dst, ae->gi_dst,
                             ae->qxv_cld, ae->qsv_cld, ae->gv_cld, ae->qxi_cld, ae->qsi_cld, ae->gi_cld};
    for (int i = 0; i < 12; i++) aq[i] = src[i];
  }
  for (int i = 0; i < 12; i++) a.aq[i] = aq[i];
  a.qref_dst = ae ? ae->qextref_dst : nullptr;
  a.qref_cld = ae ? ae->qextref_cld : nullptr;
  a.directsol = out->directsol;
  a.dnvflux = out->directsol ? out->dnvflux : nullptr;
}
}  // namespace mgcm

namespace mgcm {
hipError_t run_grid(const mgcm_rt_ctx *c, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out, int path,
                    int bands, double *ws, int tile, hipStream_t st) {
  RunArgs a;
  fill_args(a, in, out, c->d_flag);
  a.bands = bands;
  a.clr = ws;
  a.part = tile == TILE_BANDS ? ws + clr_workspace_bytes(in->ncol, in->nlay) / sizeof(double) : nullptr;
  a.tp = a.part ? a.part + (size_t)in->ncol * NB * part_rec(in->nlay + 1) : nullptr;
  const int p = path & 0xf;
  hipError_t e = (p == MGCM_RT_FP64)    ? launch_plain(c->dt, a, tile, st)
                 : (p == MGCM_RT_MIXED) ? launch_mixed(c->dt, a, tile, st)
                                        : launch_study(path >> 8, c->dt, a, st);
  // surface fluxes for TEMPGR from this run's DETAU/DIFFVT (code/comp3.f:336-348)
  if (e == hipSuccess && out->directsol)
    e = launch_dsolflux(c->d_const, in->ncol, in->sol, in->acosz, out->detau, out->diffvt, out->directsol,
                        out->dnvflux, st);
  return e;
}
}  // namespace mgcm

extern "C" int mgcm_rt_run(mgcm_rt_ctx *c, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out,
                           int precision, void *stream) {
  if (!c) return MGCM_RT_ERR_ARG;
  int s = check_columns(in, out);
  if (s != MGCM_RT_OK) return s;
  int path, bands;
  s = decode_mode(precision, &path, &bands);
  if (s != MGCM_RT_OK) return s;
  if (in->ncol == 0) return MGCM_RT_OK;
  const int tile = grid_tile(c, in->ncol, in->nlay);
  const size_t need = grid_workspace_bytes(in->ncol, in->nlay, tile);
  if (need > c->clr_bytes) {  // grow-only workspace
    if (c->d_clr) {  // the context's earlier runs were issued on this stream (or have ended)
      if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return MGCM_RT_ERR_DEVICE;
      (void)hipFree(c->d_clr);
    }
    c->d_clr = nullptr;
    c->clr_bytes = 0;
    if (hipMalloc(&c->d_clr, need) != hipSuccess) return MGCM_RT_ERR_NOMEM;
    c->clr_bytes = need;
  }
  return run_grid(c, in, out, path, bands, c->d_clr, tile, (hipStream_t)stream) == hipSuccess ? MGCM_RT_OK
                                                                                              : MGCM_RT_ERR_DEVICE;
}

extern "C" int mgcm_rt_check(mgcm_rt_ctx *c) {
  if (!c) return MGCM_RT_ERR_ARG;
  int f = 0;
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(&f, c->d_flag, sizeof(int), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemset(c->d_flag, 0, sizeof(int)));
  return f ? MGCM_RT_ERR_RANGE : MGCM_RT_OK;
}

namespace mgcm {
int check_on_stream(mgcm_rt_ctx *c, hipStream_t st) {
  int f = 0;
  HIP_TRY(hipMemcpyAsync(&f, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemsetAsync(c->d_flag, 0, sizeof(int), st));
  HIP_TRY(hipStreamSynchronize(st));
  return f ? MGCM_RT_ERR_RANGE : MGCM_RT_OK;
}
}  // namespace mgcm

extern "C" int mgcm_rt_check_stream(mgcm_rt_ctx *c, void *stream) {
  if (!c) return MGCM_RT_ERR_ARG;
  return check_on_stream(c, (hipStream_t)stream);
}

extern "C" int mgcm_rt_run_batched(mgcm_rt_ctx *c, int ngrids, const mgcm_rt_columns *grids,
                                   const mgcm_rt_fluxes *outs, int precision, void *stream) {
  if (!c || ngrids < 0 || (ngrids > 0 && (!grids || !outs))) return MGCM_RT_ERR_ARG;
  int path, bands;
  const int sm = decode_mode(precision, &path, &bands);
  if (sm != MGCM_RT_OK) return sm;
  if ((path & 0xf) == MGCM_RT_STUDY) return MGCM_RT_ERR_PRECISION;  // single grids only
  for (int i = 0; i < ngrids; i++) {
    const int s = check_columns(&grids[i], &outs[i]);
    if (s != MGCM_RT_OK) return s;
  }
  return run_batched_impl(c, ngrids, grids, outs, precision, (hipStream_t)stream);
}

extern "C" int mgcm_rt_batched_tuning(const mgcm_rt_ctx *c, int maxcls, mgcm_rt_batch_class *info) {
  if (!c || maxcls < 0 || (maxcls > 0 && !info)) return MGCM_RT_ERR_ARG;
  return batched_tuning_impl(c, maxcls, info);
}

extern "C" int mgcm_rt_set_tile(mgcm_rt_ctx *c, int band_slots) {
  if (!c || !(band_slots == 0 || band_slots == 1 || band_slots == 2 || band_slots == 4 || band_slots == TILE_BANDS))
    return MGCM_RT_ERR_ARG;
  c->tile = band_slots;
  return MGCM_RT_OK;
}

extern "C" int mgcm_rt_tile_for(const mgcm_rt_ctx *c, int ncol, int nlay) {
  if (!c || ncol < 0 || nlay < 1 || nlay > MAX_NLAY) return MGCM_RT_ERR_ARG;
  return grid_tile(c, ncol, nlay);
}

// ---------------------------------------------------------------------------
// Init-time tables.

// SETSPI's Planck table (code/setspi.f90:91-105): one thread per (band, T).
__global__ void k_planck(double *planck, const double *bwni, const double *dwni) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPLANCK * NI) return;
  const int nw = i % NI, n = i / NI;  // n = nt - 500
  const double x[12] = {-0.981560634246719, -0.904117256370475, -0.769902674194305,
                        -0.587317954286617, -0.367831498998180, -0.125233408511469,
                        0.125233408511469,  0.367831498998180,  0.587317954286617,
                        0.769902674194305,  0.904117256370475,  0.981560634246719};
  const double w[12] = {0.047175336386512, 0.106939325995318, 0.160078328543346,
                        0.203167426723066, 0.233492536538355, 0.249147045813403,
                        0.249147045813403, 0.233492536538355, 0.203167426723066,
                        0.160078328543346, 0.106939325995318, 0.047175336386512};
  const double c1 = 3.741832e-16, c2 = 1.438786e-2;
  const double a = 1.0e-2 / bwni[nw + 1], b = 1.0e-2 / bwni[nw];
  const double bpa = (b + a) / 2.0, bma = (b - a) / 2.0;
  const double tt = (double)(n + 500) / 1.0e1;
  double ans = 0.0;
  for (int m = 0; m < 12; m++) {
    const double y = bma * x[m] + bpa;
    const double y5 = y * y * y * y * y;  // Y**5 as the reference build evaluates it: products left to right
    ans = ans + w[m] * c1 / (y5 * (mgcm_libm::mgcm_exp(c2 / (y * tt)) - 1.0));
  }
  planck[(size_t)n * NI + nw] = ans * bma / (PI * dwni[nw]);
}

extern "C" int mgcm_rt_planck_table(double *planckir, double *dwni, double *wnoi) {
  if (!planckir || !dwni || !wnoi) return MGCM_RT_ERR_ARG;
  const double bwni[6] = {10.000, 166.667, 416.667, 833.333, 1250.000, 2222.222};  // code/setspi.f90:70-75
  for (int m = 0; m < NI; m++) {
    wnoi[m] = 0.5 * (bwni[m + 1] + bwni[m]);
    dwni[m] = bwni[m + 1] - bwni[m];
  }
  double *d = nullptr;
  const size_t n = (size_t)NPLANCK * NI;
  HIP_TRY(hipMalloc(&d, (n + 16) * sizeof(double)));
  double *db = d + n, *dd = d + n + 8;
  int st = MGCM_RT_OK;
  if (hipMemcpy(db, bwni, 6 * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dd, dwni, NI * 8, hipMemcpyHostToDevice) != hipSuccess) {
    st = MGCM_RT_ERR_DEVICE;
  } else {
    hipLaunchKernelGGL(k_planck, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, d, db, dd);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(planckir, d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      st = MGCM_RT_ERR_DEVICE;
  }
  (void)hipFree(d);
  return st;
}

// LAGINTERP (code/laginterp.f90:74-236) + LAGRANGE (code/lagrange.f90:17-31).
// The 4-point interpolations in log10 p run on the device, one thread per
// (T, H2O, band, g) table row of 11 pressures -> 51.  The two libm steps
// around them -- log10 of the raw k (floor -200) and 10**ans -- use the host C
// library, as the reference does: the refined table then equals the reference's
// bit for bit.  That matters: on the reference's k-distribution a one-ulp
// change of a k-coefficient moves the IR fluxes of thin layers by ~1e-9.
__device__ double lagrange4(double x, const double *xi, const double *yi) {
  const double fm1 = x - xi[0], fm2 = x - xi[1], fm3 = x - xi[2], fm4 = x - xi[3];
  return fm2 * fm3 * fm4 * yi[0] / ((xi[0] - xi[1]) * (xi[0] - xi[2]) * (xi[0] - xi[3])) +
         fm1 * fm3 * fm4 * yi[1] / ((xi[1] - xi[0]) * (xi[1] - xi[2]) * (xi[1] - xi[3])) +
         fm1 * fm2 * fm4 * yi[2] / ((xi[2] - xi[0]) * (xi[2] - xi[1]) * (xi[2] - xi[3])) +
         fm1 * fm2 * fm3 * yi[3] / ((xi[3] - xi[0]) * (xi[3] - xi[1]) * (xi[3] - xi[2]));
}

// lg: log10 of the raw table (7,11,10,nb,17); ans: log10 k on (7,51,10,nb,17).
__global__ void k_laginterp(const double *lg, const double *pref, const double *pint, int nb, double *ans) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (t, h, nw, g) with t fastest
  const int nrow = NT * NH * nb * NG;
  if (i >= nrow) return;
  const int t = i % NT, h = (i / NT) % NH, nw = (i / (NT * NH)) % nb, g = i / (NT * NH * nb);
  const int npref = 11;
  auto RAW = [&](int p) {
    return lg[(size_t)t + NT * ((size_t)p + npref * ((size_t)h + NH * ((size_t)nw + (size_t)nb * g)))];
  };
  auto OUT = [&](int p) -> double & {
    return ans[(size_t)t + NT * ((size_t)p + NP * ((size_t)h + NH * ((size_t)nw + (size_t)nb * g)))];
  };
  double xi[4], yi[4];
  for (int m = 0; m < 5; m++) {  // first interval, one-sided stencil
    for (int j = 0; j < 4; j++) { xi[j] = pref[j]; yi[j] = RAW(j); }
    OUT(m) = lagrange4(pint[m], xi, yi);
  }
  for (int n = 2; n <= npref - 2; n++)  // interior (1-based n)
    for (int m = 1; m <= 5; m++) {
      const int ii = (n - 1) * 5 + m;
      for (int j = 0; j < 4; j++) { xi[j] = pref[n - 2 + j]; yi[j] = RAW(n - 2 + j); }
      OUT(ii - 1) = lagrange4(pint[ii - 1], xi, yi);
    }
  {
    const int n = npref - 1;  // last interval
    for (int m = 1; m <= 5; m++) {
      const int ii = (n - 1) * 5 + m;
      for (int j = 0; j < 4; j++) { xi[j] = pref[n - 3 + j]; yi[j] = RAW(n - 3 + j); }
      OUT(ii - 1) = lagrange4(pint[ii - 1], xi, yi);
    }
  }
  OUT(NP - 1) = RAW(npref - 1);  // the last point is copied (code/laginterp.f90:162,231)
}

extern "C" int mgcm_rt_laginterp(const double *pgref, const double *raw, int nbands, double *pint, double *out) {
  if (!pgref || !raw || !pint || !out || nbands < 1 || nbands > NB) return MGCM_RT_ERR_ARG;
  // code/laginterp.f90:30-41: log10 p = -6.0 .. 4.0 in steps of 0.2
  static const double pin[NP] = {-6.0, -5.8, -5.6, -5.4, -5.2, -5.0, -4.8, -4.6, -4.4, -4.2, -4.0,
                                 -3.8, -3.6, -3.4, -3.2, -3.0, -2.8, -2.6, -2.4, -2.2, -2.0, -1.8,
                                 -1.6, -1.4, -1.2, -1.0, -0.8, -0.6, -0.4, -0.2, 0.0,  0.2,  0.4,
                                 0.6,  0.8,  1.0,  1.2,  1.4,  1.6,  1.8,  2.0,  2.2,  2.4,  2.6,
                                 2.8,  3.0,  3.2,  3.4,  3.6,  3.8,  4.0};
  memcpy(pint, pin, sizeof(pin));
  double pref[11];
  for (int n = 0; n < 11; n++) pref[n] = log10(pgref[n]);
  const size_t nraw = (size_t)NT * 11 * NH * nbands * NG, nout = (size_t)NT * NP * NH * nbands * NG;
  std::vector<double> lg(nraw);
  for (size_t q = 0; q < nraw; q++) lg[q] = (raw[q] > 1.0e-200) ? log10(raw[q]) : -200.0;  // laginterp.f90:74-98
  double *d = nullptr;
  HIP_TRY(hipMalloc(&d, (nraw + nout + 11 + NP) * sizeof(double)));
  double *dlg = d, *dans = d + nraw, *dpref = d + nraw + nout, *dpin = dpref + 11;
  int st = MGCM_RT_OK;
  if (hipMemcpy(dlg, lg.data(), nraw * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dpref, pref, 11 * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dpin, pin, NP * 8, hipMemcpyHostToDevice) != hipSuccess)
    st = MGCM_RT_ERR_DEVICE;
  if (st == MGCM_RT_OK) {
    const int nrow = NT * NH * nbands * NG;
    hipLaunchKernelGGL(k_laginterp, dim3((nrow + 127) / 128), dim3(128), 0, nullptr, dlg, dpref, dpin, nbands, dans);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, dans, nout * 8, hipMemcpyDeviceToHost) != hipSuccess)
      st = MGCM_RT_ERR_DEVICE;
  }
  (void)hipFree(d);
  if (st == MGCM_RT_OK)
    for (size_t q = 0; q < nout; q++) out[q] = pow(10.0, out[q]);  // 10.0**ans
  return st;
}

// Device results of the transcendental functions of the FP64 path, for the
// parity check against the host C library (mgcm_rt_libm_check).
__global__ void k_libm(const double *x, int n, double *e, double *l, double *q, double *ln) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  e[i] = mgcm_libm::mgcm_exp(x[i]);
  l[i] = (x[i] > 0.0) ? mgcm_libm::mgcm_log10(x[i]) : 0.0;
  q[i] = (x[i] >= 0.0) ? sqrt(x[i]) : 0.0;
  ln[i] = (x[i] > 0.0) ? mgcm_libm::mgcm_log(x[i]) : 0.0;
}

extern "C" int mgcm_rt_libm_check(const double *x, int n, double *exp_out, double *log10_out, double *sqrt_out,
                                  double *log_out) {
  if (n < 0 || (n > 0 && (!x || !exp_out || !log10_out || !sqrt_out || !log_out))) return MGCM_RT_ERR_ARG;
  if (n == 0) return MGCM_RT_OK;
  double *d = nullptr;
  const size_t b = (size_t)n * sizeof(double);
  HIP_TRY(hipMalloc(&d, 5 * b));
  int st = MGCM_RT_OK;
  if (hipMemcpy(d, x, b, hipMemcpyHostToDevice) != hipSuccess) st = MGCM_RT_ERR_DEVICE;
  if (st == MGCM_RT_OK) {
    hipLaunchKernelGGL(k_libm, dim3((n + 255) / 256), dim3(256), 0, nullptr, d, n, d + n, d + 2 * (size_t)n,
                       d + 3 * (size_t)n, d + 4 * (size_t)n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(exp_out, d + n, b, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(log10_out, d + 2 * (size_t)n, b, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(sqrt_out, d + 3 * (size_t)n, b, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(log_out, d + 4 * (size_t)n, b, hipMemcpyDeviceToHost) != hipSuccess)
      st = MGCM_RT_ERR_DEVICE;
  }
  (void)hipFree(d);
  return st;
}

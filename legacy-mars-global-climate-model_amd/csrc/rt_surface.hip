// rt_surface.hip -- DSOLFLUX (code/dsolflux.f) and COMP3's surface VIS flux for
// TEMPGR (code/comp3.f:336-348) over a grid, from the DETAU/DIFFVT outputs of the
// radiation: one thread per column, the glibc-exact exp, the reference's
// summation order, so the result equals the reference's bit for bit.
// Also the radiation's history diagnostics (k_history): COMP3's REAL*4 flux
// fields and the heating rate with its non-LTE factor.
#include "rt_ctx.h"
#include "rt_libm.h"

namespace mgcm {

struct SurfArgs {
  int ncol;
  double sol[NV];
  const double *acosz, *detau, *diffvt;
  double *directsol, *dnvflux;
};

// DSOL_COLS columns per 64-lane block.  The 119 terms of a column (its DETAU
// read contiguously across lanes) are formed in parallel into LDS, then one
// lane per column adds them in the reference's order (band, then Gauss point),
// so the sum is the reference's bit for bit.  A term past DETAU > 5 adds +0,
// which leaves the non-negative sum unchanged.
constexpr int DSOL_COLS = 16, DSOL_TERMS = NG * NV;
__device__ __forceinline__ void dsolflux_block(const RtConst *__restrict__ rc, const SurfArgs &a, int col0,
                                               double (*term)[DSOL_TERMS]) {
  const int lane = threadIdx.x;
  for (int t = lane; t < DSOL_COLS * DSOL_TERMS; t += 64) {
    const int cc = t / DSOL_TERMS, e = t - cc * DSOL_TERMS;  // e = ng*NV + nw: DETAU(nw, ng) of the column
    const int col = col0 + cc;
    double v = 0.0;
    if (col < a.ncol) {
      const double ubar0 = a.acosz[col];
      const int nw = e % NV, ng = e / NV;
      const double d = a.detau[(size_t)col * DSOL_TERMS + e];
      if (ubar0 >= F32LIT(1.0e-5) && d <= 5.0) {  // code/dsolflux.f:28-45; sun down: code/comp3.f:336-345
        const double factor = ubar0 * a.sol[nw];
        v = (ng < NG - 1) ? factor * mgcm_libm::mgcm_exp(-d / ubar0) * rc->gw[nw][ng] * (1.0 - rc->fzero[nw])
                          : factor * mgcm_libm::mgcm_exp(-d / ubar0) * rc->fzero[nw];
      }
    }
    term[cc][e] = v;
  }
  __syncthreads();
  const int col = col0 + lane;
  if (lane < DSOL_COLS && col < a.ncol) {
    double directsol = 0.0;
    for (int nw = 0; nw < NV; nw++)
      for (int ng = 0; ng < NG; ng++) directsol = directsol + term[lane][ng * NV + nw];
    a.directsol[col] = directsol;
    if (a.dnvflux) a.dnvflux[col] = a.diffvt[col] + directsol;  // DNVFLUX = DNDIFFV + DIRECTSOL
  }
}

__global__ __launch_bounds__(64) void k_dsolflux(const RtConst *__restrict__ rc, SurfArgs a) {
  __shared__ double term[DSOL_COLS][DSOL_TERMS];
  dsolflux_block(rc, a, blockIdx.x * DSOL_COLS, term);
}

// The batched path: a persistent grid of blocks walks a work list of (grid,
// first column) 16-column items; grids that asked for no surface fluxes
// (directsol null, uniform over an item) are passed over.
__global__ __launch_bounds__(64) void k_dsolflux_batch(const RtConst *__restrict__ rc,
                                                       const RunArgs *__restrict__ grids,
                                                       const uint2 *__restrict__ work, int nwork) {
  __shared__ double term[DSOL_COLS][DSOL_TERMS];
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const uint2 wk = work[w];
    const RunArgs &g = grids[wk.x];
    if (!g.directsol) continue;
    SurfArgs a;
    a.ncol = g.ncol;
    for (int b = 0; b < NV; b++) a.sol[b] = g.sol[b];
    a.acosz = g.acosz;
    a.detau = g.detau;
    a.diffvt = g.diffvt;
    a.directsol = g.directsol;
    a.dnvflux = g.dnvflux;
    __syncthreads();  // the previous item's reads of term are done
    dsolflux_block(rc, a, (int)wk.y, term);
  }
}

hipError_t launch_dsolflux_batch(const RtConst *rc, const RunArgs *grids, const uint2 *work, int nwork, int nblocks,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_dsolflux_batch, dim3(nblocks), dim3(64), 0, st, rc, grids, work, nwork);
  return hipGetLastError();
}

hipError_t launch_dsolflux(const RtConst *rc, int ncol, const double *sol, const double *acosz,
                           const double *detau, const double *diffvt, double *directsol, double *dnvflux,
                           hipStream_t st) {
  SurfArgs a;
  a.ncol = ncol;
  for (int b = 0; b < NV; b++) a.sol[b] = sol[b];
  a.acosz = acosz;
  a.detau = detau;
  a.diffvt = diffvt;
  a.directsol = directsol;
  a.dnvflux = dnvflux;
  hipLaunchKernelGGL(k_dsolflux, dim3((ncol + DSOL_COLS - 1) / DSOL_COLS), dim3(64), 0, st, rc, a);
  return hipGetLastError();
}

}  // namespace mgcm

extern "C" int mgcm_rt_dsolflux(mgcm_rt_ctx *c, int ncol, const double *sol, const double *acosz,
                                const double *detau, const double *diffvt, double *directsol, double *dnvflux,
                                void *stream) {
  if (!c || !sol || ncol < 0) return MGCM_RT_ERR_ARG;
  if (ncol == 0) return MGCM_RT_OK;
  if (!acosz || !detau || !directsol || (dnvflux && !diffvt)) return MGCM_RT_ERR_ARG;
  return mgcm::launch_dsolflux(c->d_const, ncol, sol, acosz, detau, diffvt, directsol, dnvflux,
                               (hipStream_t)stream) == hipSuccess
             ? MGCM_RT_OK
             : MGCM_RT_ERR_DEVICE;
}

extern "C" int mgcm_rt_dsolflux_host(mgcm_rt_ctx *c, int ncol, const double *sol, const double *acosz,
                                     const double *detau, const double *diffvt, double *directsol,
                                     double *dnvflux) {
  if (!c || !sol || ncol < 0) return MGCM_RT_ERR_ARG;
  if (ncol == 0) return MGCM_RT_OK;
  if (!acosz || !detau || !directsol || (dnvflux && !diffvt)) return MGCM_RT_ERR_ARG;
  using namespace mgcm;
  const size_t n = (size_t)ncol, nd = n * NG * NV;
  double *p = host_staging(c, (n + nd + (dnvflux ? 3 : 1) * n) * sizeof(double));
  if (!p) return MGCM_RT_ERR_NOMEM;
  double *d_acosz = p, *d_detau = p + n, *d_dir = d_detau + nd;
  double *d_dif = dnvflux ? d_dir + n : nullptr, *d_dnv = dnvflux ? d_dir + 2 * n : nullptr;
  HIP_TRY(hipMemcpy(d_acosz, acosz, n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_detau, detau, nd * sizeof(double), hipMemcpyHostToDevice));
  if (dnvflux) HIP_TRY(hipMemcpy(d_dif, diffvt, n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(launch_dsolflux(c->d_const, ncol, sol, d_acosz, d_detau, d_dif, d_dir, d_dnv, nullptr));
  HIP_TRY(hipMemcpy(directsol, d_dir, n * sizeof(double), hipMemcpyDeviceToHost));
  if (dnvflux) HIP_TRY(hipMemcpy(dnvflux, d_dnv, n * sizeof(double), hipMemcpyDeviceToHost));
  return MGCM_RT_OK;
}

namespace mgcm {

struct HistArgs {
  int ncol, nlay;
  const double *plev, *acosz, *als;
  const double *fmnetv, *fluxupv, *fluxdnv, *fmneti, *fluxupi, *fluxdni, *nfluxtopv, *nfluxtopi;
  float *fuptopv, *fdntopv, *fupsurfv, *fdnsurfv, *fuptopir, *fupsurfir, *fdnsurfir, *surfalb;
  double *hr_nlte;
  const double *p, *sigma;
  double ptrop;
};

// One thread per (column, heating slot l = 0..nlay): coalesced over the
// [ncol][nlay+1] arrays; the l = 0 thread of a column also writes its REAL*4
// history fields.
__global__ __launch_bounds__(256) void k_history(HistArgs a) {
  const int nl = a.nlay + 1, nlev = 2 * a.nlay + 3;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)a.ncol * nl) return;
  const int col = (int)(i / nl), l = (int)(i - (long long)col * nl);
  const size_t row = (size_t)col * nl;
  const bool sun = a.acosz[col] >= F32LIT(1.0e-5);  // code/comp3.f:637
  if (l == 0) {
    // code/comp3.f:654,703-706 (VIS, zero when the sun is down: :690-695) and :745-747 (IR)
    a.fuptopv[col] = sun ? (float)a.fluxupv[row] : 0.0f;
    a.fdntopv[col] = sun ? (float)a.fluxdnv[row] : 0.0f;
    a.fupsurfv[col] = sun ? (float)a.fluxupv[row + nl - 1] : 0.0f;
    a.fdnsurfv[col] = sun ? (float)a.fluxdnv[row + nl - 1] : 0.0f;
    a.fuptopir[col] = (float)a.fluxupi[row];
    a.fupsurfir[col] = (float)a.fluxupi[row + nl - 1];
    a.fdnsurfir[col] = (float)a.fluxdni[row + nl - 1];
    if (a.surfalb && sun) a.surfalb[col] = (float)a.als[col];
  }
  if (a.hr_nlte) {
    // SUNTOT(K+1), IRTOTAL(K+1) with K = 2l+2 (code/comp3.f:677-681,735-738)
    const double nv = a.fmnetv[row + l], ni = a.fmneti[row + l];
    const double sunt = sun ? (l == 0 ? nv - a.nfluxtopv[col] : nv - a.fmnetv[row + l - 1]) : 0.0;
    const double irt = (l == 0) ? ni - a.nfluxtopi[col] : ni - a.fmneti[row + l - 1];
    // code/comp3.f:765-771: the non-LTE factor on the solar part
    const double pn = a.sigma[2 * l + 1] * (a.p[col] - a.ptrop) + a.ptrop;
    const double sunlte = ((sunt * F32LIT(2.2e4)) * pn) / (F32LIT(1.0) + F32LIT(2.2e4) * pn);
    const double *plev = a.plev + (size_t)col * nlev;
    const double dp = (l == 0) ? (plev[0] + plev[1]) : (plev[2 * l + 2] - plev[2 * l]);
    a.hr_nlte[row + l] = (((sunlte + irt) * GRAV) / ((CPAIR * SCALEP) * dp)) * SECONDS_PER_SOL;
  }
}

}  // namespace mgcm

extern "C" int mgcm_rt_history_fields(const mgcm_rt_columns *in, const mgcm_rt_fluxes *out,
                                      const mgcm_rt_history *h, void *stream) {
  if (!in || !out || !h || in->ncol < 0) return MGCM_RT_ERR_ARG;
  if (in->nlay < 1 || in->nlay > MGCM_RT_MAX_NLAY) return MGCM_RT_ERR_NLAY;
  if (in->ncol == 0) return MGCM_RT_OK;
  if (!in->acosz || !out->fluxupv || !out->fluxdnv || !out->fluxupi || !out->fluxdni || !h->fuptopv ||
      !h->fdntopv || !h->fupsurfv || !h->fdnsurfv || !h->fuptopir || !h->fupsurfir || !h->fdnsurfir)
    return MGCM_RT_ERR_ARG;
  if (h->surfalb && !in->als) return MGCM_RT_ERR_ARG;
  if (h->hr_nlte && (!h->p || !h->sigma || !in->plev || !out->fmnetv || !out->fmneti || !out->nfluxtopv ||
                     !out->nfluxtopi))
    return MGCM_RT_ERR_ARG;
  mgcm::HistArgs a;
  a.ncol = in->ncol;
  a.nlay = in->nlay;
  a.plev = in->plev;
  a.acosz = in->acosz;
  a.als = in->als;
  a.fmnetv = out->fmnetv; a.fluxupv = out->fluxupv; a.fluxdnv = out->fluxdnv;
  a.fmneti = out->fmneti; a.fluxupi = out->fluxupi; a.fluxdni = out->fluxdni;
  a.nfluxtopv = out->nfluxtopv; a.nfluxtopi = out->nfluxtopi;
  a.fuptopv = h->fuptopv; a.fdntopv = h->fdntopv; a.fupsurfv = h->fupsurfv; a.fdnsurfv = h->fdnsurfv;
  a.fuptopir = h->fuptopir; a.fupsurfir = h->fupsurfir; a.fdnsurfir = h->fdnsurfir;
  a.surfalb = h->surfalb;
  a.hr_nlte = h->hr_nlte;
  a.p = h->p;
  a.sigma = h->sigma;
  a.ptrop = h->ptrop;
  const long long n = (long long)in->ncol * (in->nlay + 1);
  hipLaunchKernelGGL(mgcm::k_history, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? MGCM_RT_OK : MGCM_RT_ERR_DEVICE;
}
static_assert(mgcm::DSOLFLUX_ITEM_COLS == mgcm::DSOL_COLS, "batched DSOLFLUX item width");

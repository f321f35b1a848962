// rt_aerosol.hip -- the aerosol producers of the radiation's inputs:
//   * INITCLD's radiative size grid and optical tables (code/initcld.f:131-184):
//     the 20-bin grid rad_rt/radb_rt, and Qext/Qscat/g per bin and band read
//     from data/Dust_{vis,ir}_wolff2010_JD_12bands.dat and
//     data/waterCoated_{vis,ir}_JD_12bands.dat (host code, once);
//   * OPT_DST (code/opt_dst.f) and OPT_CLD (code/opt_cld.f): per column and
//     layer, the log-normal size distribution of the dust (or ice-coated dust)
//     tracers, its cross-section weights over the 20 bins (erf), and the
//     weighted Qext/Qscat/g per band, the reference extinction and TAUREF (on the
//     device, whole columns packed into 256-thread blocks, one thread per
//     layer).
// The outputs are the per-level arrays of mgcm_rt_aerosol plus TAUREF/TAUREFCLD.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rt_ctx.h"
#include "rt_libm.h"

namespace mgcm {
namespace aer {

constexpr int NBIN = 20, NRATIO = 15, NTRACE = 6;
// offsets (doubles) into the device table block
constexpr int OFF_RADB = 0;                                  // radb_rt(21)
constexpr int OFF_DV = 32;                                   // dust VIS (3, 7, 20)
constexpr int OFF_DI = OFF_DV + 3 * NV * NBIN;               // dust IR (3, 5, 20)
constexpr int OFF_CV = OFF_DI + 3 * NI * NBIN;               // cloud VIS (3, 7, 20, 15)
constexpr int OFF_CI = OFF_CV + 3 * NV * NBIN * NRATIO;      // cloud IR (3, 5, 20, 15)
constexpr int TABLE_DOUBLES = OFF_CI + 3 * NI * NBIN * NRATIO;
// The device copy keeps the block's offsets, but each cloud table is held per
// core ratio: ratio j's (3, nb, 20) slice, laid out as the dust tables are, at
// OFF_C? + j * 3 * nb * 20, so that a layer's 20 bins of a band are adjacent.

// code/modules.f90:1112-1114 (REAL*4)
__constant__ float cor_ratio[NRATIO] = {0.1f, 0.2f, 0.25f, 0.3f, 0.35f, 0.4f, 0.45f, 0.5f,
                                        0.55f, 0.6f, 0.65f, 0.7f, 0.8f, 0.9f, 0.99f};

struct AerOut {
  double *qxv, *qsv, *gv, *qxi, *qsi, *gi, *qextref, *tauref, *tau2;
};

// OPT_DST (CLOUD = false) / OPT_CLD (CLOUD = true): a block of AER_BLOCK
// threads takes cpb = AER_BLOCK / nlay whole columns, thread cc*nlay + l-1 on
// layer l of its column cc (24 layers: 10 columns, 240 of 256 threads busy).
// qtr: [ncol][6][nlay]; pl: [ncol][nlev].
constexpr int AER_BLOCK = 256;
template <bool CLOUD>
__global__ __launch_bounds__(AER_BLOCK) void k_opt_aer(const double *__restrict__ tab, int ncol, int nlay,
                                                      const double *__restrict__ qtr,
                                                      const double *__restrict__ pl_all, AerOut o) {
  __shared__ double s_tau[AER_BLOCK][4];
  __shared__ int s_on[AER_BLOCK];
  const int cpb = AER_BLOCK / nlay, tid = threadIdx.x;
  const int cc = tid / nlay, col = blockIdx.x * cpb + cc;
  const bool active = cc < cpb && col < ncol;  // no early return: the block synchronises below
  const int lane = active ? tid - cc * nlay : nlay;  // layer - 1; nlay for idle threads
  const int nlevels = 2 * nlay + 3, ld = nlevels + 1;
  const double athird = (double)(1.0f / 3.0f);        // code/initcld.f:57 (REAL(4) 1./3.)
  const double dpden_dt = 2.5e3, dpden_ice = 917.;    // code/initdt.f:38, code/initcld.f:61
  const double dev = CLOUD ? (double)0.3087f : (double)0.63676f;  // code/initcld.f:93-95
  const double dev2 = 1. / ((double)sqrtf(2.f) * dev);
  const double qs_init = (double)0.99f;
  const size_t cz = active ? (size_t)col : 0;
  const double *q = qtr + cz * NTRACE * nlay;
  const double *pl = pl_all + cz * nlevels;
  double *qxv = o.qxv + cz * NV * ld, *qsv = o.qsv + cz * NV * ld;
  double *gv = o.gv + cz * NV * ld, *qxi = o.qxi + cz * NI * ld;
  double *qsi = o.qsi + cz * NI * ld, *gi = o.gi + cz * NI * ld;
  double *qextref = o.qextref + cz * ld, *tauref = o.tauref + cz * ld;
  auto defaults = [&](int k) {  // code/opt_dst.f:52-69: levels without aerosol
    qextref[k - 1] = 1.;
    tauref[k - 1] = 0.;
    for (int w = 0; w < NV; w++) { qxv[w * ld + k - 1] = 1.; qsv[w * ld + k - 1] = qs_init; gv[w * ld + k - 1] = 0.; }
    for (int w = 0; w < NI; w++) { qxi[w * ld + k - 1] = 1.; qsi[w * ld + k - 1] = qs_init; gi[w * ld + k - 1] = 0.; }
  };
  if (active && lane == 0) {
    for (int k = 1; k <= 3; k++) defaults(k);
    defaults(ld);
  }
  const int l = lane + 1;
  bool on = false;
  if (active) {
    auto QT = [&](int m) { return q[(m - 1) * nlay + l - 1]; };
    double mo, no, cst;
    int irap = 0;
    if (!CLOUD) {  // iMa_dt = 1, iNb_dt = 2 (code/initcld.f:37-38)
      on = QT(1) > (double)1.e-8f && QT(2) > 1.;
      mo = QT(1);
      no = QT(2) + 1.;
      cst = 0.75 / (PI * dpden_dt);
    } else {       // iMa_cld = 3, iNb_cld = 4, iMa_cor = 5
      on = QT(3) + QT(5) > (double)1.e-7f && QT(4) > 1.;
      mo = QT(3) + QT(5);
      no = QT(4);
      cst = 0.;
      if (on) {
        float mtc = (float)(QT(5) / dpden_dt / (QT(3) / dpden_ice + QT(5) / dpden_dt));
        mtc = (float)pow((double)mtc, athird);
        irap = NRATIO;
        for (int i = 1; i <= NRATIO; i++) {  // code/opt_cld.f:104-116
          if (mtc < cor_ratio[i - 1] && i != 1) { irap = i - 1; break; }
          else if (mtc == cor_ratio[i - 1]) { irap = i; break; }
          else if (mtc < cor_ratio[i - 1] && i == 1) { irap = 1; break; }
        }
        const double dens = QT(3) / mo * dpden_ice + QT(5) / mo * dpden_dt;
        cst = 0.75 / (PI * dens);
      }
    }
    if (!on) {
      defaults(2 * l + 2);
      defaults(2 * l + 3);
    } else {
      double rs = pow(mo / no * cst, athird) * mgcm_libm::mgcm_exp(-0.5 * (dev * dev));
      const double ao = no * PI * (rs * rs);
      rs = rs * mgcm_libm::mgcm_exp(1.5 * (dev * dev));
      rs = fmin(fmax(rs, (double)1.e-7f), CLOUD ? (double)100.e-6f : (double)50.e-6f);
      rs = 1. / rs;
      // surf(i) = (erf(bin top) - erf(bin bottom))/2; adjacent bins share an edge value
      double surf[NBIN];
      double e0 = erf(mgcm_libm::mgcm_log(tab[OFF_RADB] * rs) * dev2);
#pragma unroll
      for (int i = 0; i < NBIN; i++) {
        const double e1 = erf(mgcm_libm::mgcm_log(tab[OFF_RADB + i + 1] * rs) * dev2);
        surf[i] = 0.5 * (e1 - e0);
        e0 = e1;
      }
      const double *tv = tab + (CLOUD ? OFF_CV + (irap - 1) * 3 * NV * NBIN : OFF_DV);
      const double *ti = tab + (CLOUD ? OFF_CI + (irap - 1) * 3 * NI * NBIN : OFF_DI);
      constexpr int sv = NV * NBIN, si = NI * NBIN;
      auto idx = [&](int w, int i) { return i + NBIN * w; };
      double qx6 = 0.0, qxi4 = 0.0, qsi4 = 0.0;
      for (int w = 0; w < NV; w++) {
        double x = 0., sc = 0., g = 0.;
        for (int i = 0; i < NBIN; i++) {
          const int oo = idx(w, i);
          x = x + surf[i] * tv[oo];
          sc = sc + surf[i] * tv[oo + sv];
          g = g + surf[i] * tv[oo + 2 * sv];
        }
        sc = fmin(sc, (double)0.99999f * x);  // code/opt_dst.f:136
        for (int k = 2 * l + 2; k <= 2 * l + 3; k++) {
          qxv[w * ld + k - 1] = x;
          qsv[w * ld + k - 1] = sc;
          gv[w * ld + k - 1] = g;
        }
        if (w == L_NREFV - 1) qx6 = x;
      }
      for (int w = 0; w < NI; w++) {
        double x = 0., sc = 0., g = 0.;
        for (int i = 0; i < NBIN; i++) {
          const int oo = idx(w, i);
          x = x + surf[i] * ti[oo];
          sc = sc + surf[i] * ti[oo + si];
          g = g + surf[i] * ti[oo + 2 * si];
        }
        sc = fmin(sc, (double)0.99999f * x);
        for (int k = 2 * l + 2; k <= 2 * l + 3; k++) {
          qxi[w * ld + k - 1] = x;
          qsi[w * ld + k - 1] = sc;
          gi[w * ld + k - 1] = g;
        }
        if (w == 3) { qxi4 = x; qsi4 = sc; }
      }
      // code/opt_dst.f:151-160: reference extinction, TAUREF, diagnostic terms
      for (int nn = 1; nn <= 2; nn++) {
        const int k = 2 * l + 1 + nn;
        qextref[k - 1] = qx6;
        const double mass = 100. * (pl[k - 1] - pl[k - 2]) / GRAV;
        const double t = ao * qx6 * mass;
        tauref[k - 1] = t;
        s_tau[tid][2 * (nn - 1)] = t;
        s_tau[tid][2 * (nn - 1) + 1] = t / qx6 * (qxi4 - qsi4);
      }
    }
  }
  s_on[tid] = on ? 1 : 0;
  __syncthreads();
  if (active && lane == 0 && o.tau2) {  // taudust/taucloud: REAL*4 sums in layer order
    float t1 = 0.f, t2 = 0.f;
    for (int j = tid; j < tid + nlay; j++) {
      if (!s_on[j]) continue;
      for (int nn = 0; nn < 2; nn++) {
        t1 = (float)((double)t1 + s_tau[j][2 * nn]);
        t2 = (float)((double)t2 + s_tau[j][2 * nn + 1]);
      }
    }
    o.tau2[2 * col] = t1;
    o.tau2[2 * col + 1] = t2;
  }
}

// One Fortran formatted record of n fields of width w separated by one blank
// ("(n(Ew.d,x))"); a short line reads as blanks (zero).
static bool read_fields(FILE *f, int n, int w, float *out) {
  char line[512];
  if (!fgets(line, sizeof line, f)) return false;
  const size_t len = strcspn(line, "\r\n");
  for (int j = 0; j < n; j++) {
    char field[64] = {0};
    const size_t a = (size_t)j * (w + 1);
    if (a < len) memcpy(field, line + a, std::min((size_t)w, len - a));
    char *end = nullptr;
    out[j] = strtof(field, &end);  // decimal -> REAL(4), correctly rounded
    for (const char *p = end; *p; p++)
      if (*p != ' ') return false;
  }
  return true;
}

}  // namespace aer
}  // namespace mgcm

using namespace mgcm;
using namespace mgcm::aer;

// INITCLD's radiative size grid (code/initcld.f:131-146): host C library, as
// the reference computes it.
static void aer_grid(double *rad_rt, double *radb_rt) {
  const double athird = (double)(1.0f / 3.0f);
  rad_rt[0] = (double)1.e-7f;
  rad_rt[NBIN - 1] = (double)50.e-6f;
  radb_rt[0] = (double)0.0001e-6f;
  radb_rt[NBIN] = (double)1.e-2f;
  double vrat_rt = log(rad_rt[NBIN - 1] / rad_rt[0]) / (double)(float)(NBIN - 1) * 3.;
  vrat_rt = exp(vrat_rt);
  for (int i = 1; i <= NBIN - 1; i++) {
    rad_rt[i] = rad_rt[i - 1] * pow(vrat_rt, athird);
    radb_rt[i] = pow((2. * vrat_rt) / (vrat_rt + 1.), athird) * rad_rt[i - 1];
  }
}

extern "C" int mgcm_rt_read_aerosol_tables(const char *datadir, double *tables) {
  if (!datadir || !tables) return MGCM_RT_ERR_ARG;
  std::vector<double> t(TABLE_DOUBLES, 0.0);
  double rad[NBIN];
  aer_grid(rad, &t[OFF_RADB]);
  const std::string dir(datadir);
  FILE *f60 = fopen((dir + "/waterCoated_vis_JD_12bands.dat").c_str(), "r");
  FILE *f61 = fopen((dir + "/waterCoated_ir_JD_12bands.dat").c_str(), "r");
  FILE *f62 = fopen((dir + "/Dust_vis_wolff2010_JD_12bands.dat").c_str(), "r");
  FILE *f63 = fopen((dir + "/Dust_ir_wolff2010_JD_12bands.dat").c_str(), "r");
  bool ok = f60 && f61 && f62 && f63;
  float v[7];
  // code/initcld.f:158-167: per core ratio j and bin i, three records per file
  for (int j = 0; ok && j < NRATIO; j++)
    for (int i = 0; ok && i < NBIN; i++)
      for (int kind = 0; ok && kind < 3; kind++) {
        ok = read_fields(f60, NV, 12, v);
        for (int w = 0; ok && w < NV; w++) t[OFF_CV + j + NRATIO * (i + NBIN * (w + NV * kind))] = v[w];
        ok = ok && read_fields(f61, NI, 12, v);
        for (int w = 0; ok && w < NI; w++) t[OFF_CI + j + NRATIO * (i + NBIN * (w + NI * kind))] = v[w];
      }
  // code/initcld.f:169-182 (the IR scaling factor there is 1.0)
  for (int i = 0; ok && i < NBIN; i++)
    for (int kind = 0; ok && kind < 3; kind++) {
      ok = read_fields(f62, NV, 11, v);
      for (int w = 0; ok && w < NV; w++) t[OFF_DV + i + NBIN * (w + NV * kind)] = v[w];
      ok = ok && read_fields(f63, NI, 11, v);
      for (int w = 0; ok && w < NI; w++) t[OFF_DI + i + NBIN * (w + NI * kind)] = v[w];
    }
  for (FILE *f : {f60, f61, f62, f63})
    if (f) fclose(f);
  if (!ok) return MGCM_RT_ERR_IO;
  memcpy(tables, t.data(), sizeof(double) * TABLE_DOUBLES);
  return MGCM_RT_OK;
}

extern "C" int mgcm_rt_set_aerosol_tables(mgcm_rt_ctx *c, const double *tables) {
  if (!c || !tables) return MGCM_RT_ERR_ARG;
  if (!c->d_aer && hipMalloc(&c->d_aer, sizeof(double) * TABLE_DOUBLES) != hipSuccess) {
    c->d_aer = nullptr;
    return MGCM_RT_ERR_NOMEM;
  }
  // the cloud tables per core ratio on the device (see OFF_CV)
  std::vector<double> t(tables, tables + TABLE_DOUBLES);
  for (int j = 0; j < NRATIO; j++)
    for (int kind = 0; kind < 3; kind++)
      for (int i = 0; i < NBIN; i++) {
        for (int w = 0; w < NV; w++)
          t[OFF_CV + j * 3 * NV * NBIN + i + NBIN * (w + NV * kind)] = tables[OFF_CV + j + NRATIO * (i + NBIN * (w + NV * kind))];
        for (int w = 0; w < NI; w++)
          t[OFF_CI + j * 3 * NI * NBIN + i + NBIN * (w + NI * kind)] = tables[OFF_CI + j + NRATIO * (i + NBIN * (w + NI * kind))];
      }
  if (hipMemcpy(c->d_aer, t.data(), sizeof(double) * TABLE_DOUBLES, hipMemcpyHostToDevice) != hipSuccess)
    return MGCM_RT_ERR_DEVICE;
  return MGCM_RT_OK;
}

static int opt_aer(bool cloud, mgcm_rt_ctx *c, const mgcm_rt_tracers *in, const mgcm_rt_aerosol_out *out,
                   void *stream) {
  if (!c || !in || !out) return MGCM_RT_ERR_ARG;
  if (!c->d_aer) return MGCM_RT_ERR_ARG;  // mgcm_rt_set_aerosol_tables first
  if (in->ncol < 0) return MGCM_RT_ERR_ARG;
  if (in->nlay < 1 || in->nlay > MAX_NLAY) return MGCM_RT_ERR_NLAY;
  if (in->ncol == 0) return MGCM_RT_OK;
  const void *req[] = {in->pl, in->qtrace, out->qxv, out->qsv, out->gv, out->qxi, out->qsi, out->gi,
                       out->qextref, out->tauref};
  for (const void *p : req)
    if (!p) return MGCM_RT_ERR_ARG;
  AerOut o{out->qxv, out->qsv, out->gv, out->qxi, out->qsi, out->gi, out->qextref, out->tauref, out->tau2};
  hipStream_t st = (hipStream_t)stream;
  const int cpb = AER_BLOCK / in->nlay;  // whole columns per block (nlay <= 99 < AER_BLOCK)
  const dim3 grid((unsigned)((in->ncol + cpb - 1) / cpb));
  if (cloud)
    hipLaunchKernelGGL(k_opt_aer<true>, grid, dim3(AER_BLOCK), 0, st, c->d_aer, in->ncol, in->nlay,
                       in->qtrace, in->pl, o);
  else
    hipLaunchKernelGGL(k_opt_aer<false>, grid, dim3(AER_BLOCK), 0, st, c->d_aer, in->ncol, in->nlay,
                       in->qtrace, in->pl, o);
  return hipGetLastError() == hipSuccess ? MGCM_RT_OK : MGCM_RT_ERR_DEVICE;
}

extern "C" int mgcm_rt_opt_dst(mgcm_rt_ctx *c, const mgcm_rt_tracers *in, const mgcm_rt_aerosol_out *out,
                               void *stream) {
  return opt_aer(false, c, in, out, stream);
}

// Host-pointer variants: stage the arrays in the context's grow-only device
// block (the one mgcm_rt_run_host uses), run on the null stream, copy back,
// synchronise.
namespace mgcm {
double *host_staging(mgcm_rt_ctx *c, size_t bytes) {
  if (bytes > c->host_bytes) {
    if (c->d_host) (void)hipFree(c->d_host);
    c->d_host = nullptr;
    c->host_bytes = 0;
    if (hipMalloc(&c->d_host, bytes) != hipSuccess) return nullptr;
    c->host_bytes = bytes;
  }
  return c->d_host;
}
}  // namespace mgcm

static int opt_aer_host(bool cloud, mgcm_rt_ctx *c, const mgcm_rt_tracers *in, const mgcm_rt_aerosol_out *out) {
  if (!c || !in || !out) return MGCM_RT_ERR_ARG;
  if (in->ncol < 0) return MGCM_RT_ERR_ARG;
  if (in->nlay < 1 || in->nlay > MAX_NLAY) return MGCM_RT_ERR_NLAY;
  if (in->ncol == 0) return MGCM_RT_OK;
  const size_t ncol = in->ncol, nlev = 2 * (size_t)in->nlay + 3, ld = nlev + 1;
  const size_t nq = ncol * NTRACE * in->nlay, npl = ncol * nlev;
  // outputs in mgcm_rt_aerosol_out order: 3 x VIS, 3 x IR, qextref, tauref, tau2
  double *hout[9] = {out->qxv, out->qsv, out->gv, out->qxi, out->qsi, out->gi, out->qextref, out->tauref,
                     out->tau2};
  const size_t nout[9] = {ncol * NV * ld, ncol * NV * ld, ncol * NV * ld, ncol * NI * ld, ncol * NI * ld,
                          ncol * NI * ld, ncol * ld, ncol * ld, ncol * 2};
  if (!in->pl || !in->qtrace) return MGCM_RT_ERR_ARG;
  for (int i = 0; i < 8; i++)
    if (!hout[i]) return MGCM_RT_ERR_ARG;
  size_t total = nq + npl;
  for (int i = 0; i < 9; i++) total += hout[i] ? nout[i] : 0;
  double *p = mgcm::host_staging(c, total * sizeof(double));
  if (!p) return MGCM_RT_ERR_NOMEM;
  mgcm_rt_tracers din = *in;
  din.qtrace = p;
  if (hipMemcpy(p, in->qtrace, nq * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return MGCM_RT_ERR_DEVICE;
  p += nq;
  din.pl = p;
  if (hipMemcpy(p, in->pl, npl * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return MGCM_RT_ERR_DEVICE;
  p += npl;
  double *dout[9] = {nullptr};
  for (int i = 0; i < 9; i++)
    if (hout[i]) { dout[i] = p; p += nout[i]; }
  const mgcm_rt_aerosol_out o{dout[0], dout[1], dout[2], dout[3], dout[4], dout[5], dout[6], dout[7], dout[8]};
  const int s = opt_aer(cloud, c, &din, &o, nullptr);
  if (s != MGCM_RT_OK) return s;
  if (hipStreamSynchronize(nullptr) != hipSuccess) return MGCM_RT_ERR_DEVICE;
  for (int i = 0; i < 9; i++)
    if (hout[i] && hipMemcpy(hout[i], dout[i], nout[i] * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return MGCM_RT_ERR_DEVICE;
  return MGCM_RT_OK;
}

extern "C" int mgcm_rt_opt_dst_host(mgcm_rt_ctx *c, const mgcm_rt_tracers *in, const mgcm_rt_aerosol_out *out) {
  return opt_aer_host(false, c, in, out);
}
extern "C" int mgcm_rt_opt_cld_host(mgcm_rt_ctx *c, const mgcm_rt_tracers *in, const mgcm_rt_aerosol_out *out) {
  return opt_aer_host(true, c, in, out);
}

extern "C" int mgcm_rt_opt_cld(mgcm_rt_ctx *c, const mgcm_rt_tracers *in, const mgcm_rt_aerosol_out *out,
                               void *stream) {
  return opt_aer(true, c, in, out, stream);
}

// ---------------------------------------------------------------------------
// Prescribed (Conrath) dust: DUSTPROFILE's column profile (code/dustprofile.f:
// 56-113) and FILLTAUCUM (code/filltaucum.f:33-52), one 64-lane wave per
// column.  The exponentials of the Conrath integral and of the mixing ratios,
// and the per-level interpolations, are formed across the lanes; the two
// running sums (the integral and TAUDST) are added by lane 0 in the
// reference's order.  Every operation is IEEE or the glibc-exact exp, so TAUREF
// equals the reference's bit for bit.
namespace mgcm {
namespace aer {
constexpr int NPDST_MAX = 128;

__device__ int jsrchgt(int n, const double *sx, double target) {  // code/jsrchgt.f90:28-66
  if (n < 1) return 0;
  if (target > sx[n - 1]) return n + 1;
  if (target < sx[0]) return 1;
  int jl = 1, jh = n;
  while (jh - jl > 1) {
    const int jm = (jl + jh) / 2;
    if (target > sx[jm - 1]) jl = jm;
    else jh = jm;
  }
  if (target == sx[jh - 1]) jh = jh + 1;
  return jh;
}

__global__ __launch_bounds__(64) void k_dust_tauref(mgcm_rt_dust_columns in, double *tauref_all, double *taudst_out) {
  __shared__ double s_pr[NPDST_MAX], s_t[NPDST_MAX + 8], s_tau[NPDST_MAX + 8], s_tc[2 * MAX_NLAY + 4];
  const int col = blockIdx.x, lane = threadIdx.x;
  const int npdst = in.npdst, nlevels = 2 * in.nlay + 3;
  const double rptau = in.rptau, conr = in.conr[col];
  for (int n = lane; n < npdst; n += 64) s_pr[n] = in.prdst[n];
  __syncthreads();
  // The integral's loop ends at the first n (2..npdst) with PRDST(n) >= RPTAU, or after npdst.
  int nb = npdst + 1;
  for (int n0 = 2; n0 <= npdst; n0 += 64) {
    const int n = n0 + lane;
    const unsigned long long m = __ballot(n <= npdst && s_pr[n - 1] >= rptau);
    if (m) {
      nb = n0 + __builtin_ctzll(m);
      break;
    }
  }
  // E(m) = exp(CONR (1 - RPTAU / PAVE)), PAVE = (PRDST(m) + PRDST(m+1)) / 2 (1-based), for m = 1..npdst-1 with
  // PRDST(m+1) < RPTAU: the integral's term for n = m + 1 and QRDST(m)'s region 1 use the same exponential.
  // The lane with m = nb - 1 forms the integral's last term, whose exponential is also that of QRDST(m)'s
  // region 2 there (PAVE is the same sum, commuted).  The terms n = 2..nb-1 go to s_t[n - 2], padded with
  // +0 to a multiple of 8: a term the reference skips is +0 here too, which leaves the sum (+0 at the
  // start, so never -0) unchanged.
  __shared__ double s_fin;
  const int nt = nb - 2, ntp = (nt + 7) & ~7;
  double ex[2], x2[2];
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int m = 1 + lane + 64 * c;
    const bool below = m <= npdst - 1 && s_pr[m] < rptau;
    ex[c] = 0.;
    x2[c] = 0.;
    if (below) {
      const double pave = 0.5 * (s_pr[m] + s_pr[m - 1]);
      ex[c] = mgcm_libm::mgcm_exp(conr * (1. - (rptau / pave)));
    }
    const bool r2 = m <= npdst - 1 && rptau <= s_pr[m] && rptau >= s_pr[m - 1];
    if (r2 || m == nb - 1) {
      const double pave = 0.5 * (s_pr[m - 1] + rptau), pdif1 = rptau - s_pr[m - 1];
      const double e2 = mgcm_libm::mgcm_exp(conr * (1.0 - (rptau / pave)));
      if (r2) x2[c] = e2 * pdif1 + (s_pr[m] - rptau);
      if (m == nb - 1) s_fin = e2 * pdif1;
    }
    if (m - 1 < ntp) s_t[m - 1] = m - 1 < nt && below ? ex[c] * (s_pr[m] - s_pr[m - 1]) : 0.;
  }
  __syncthreads();
  if (lane == 0) {
    double sum = 0.;
    for (int i = 0; i < ntp; i += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = s_t[i + u];
#pragma unroll
      for (int u = 0; u < 8; u++) sum = sum + v[u];
    }
    sum = sum + s_fin;
    s_tc[0] = in.tautot[col] / sum;
  }
  __syncthreads();
  const double qrdst0 = s_tc[0];
  // QRDST(n) in its three regions, then its TAUDST increment QRDST(n) (PRDST(n+1) - PRDST(n)) at s_t[n - 1],
  // padded with +0 as above.
  const int np1 = npdst - 1, np1p = (np1 + 7) & ~7;
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int n = 1 + lane + 64 * c;
    if (n - 1 < np1p) {
      double dt = 0.;
      if (n <= npdst - 1) {
        double q = 0.;
        if (rptau > s_pr[n]) q = qrdst0 * ex[c];
        if (rptau <= s_pr[n] && rptau >= s_pr[n - 1]) q = qrdst0 * x2[c] / (s_pr[n] - s_pr[n - 1]);
        if (rptau < s_pr[n - 1]) q = qrdst0;
        dt = q * (s_pr[n] - s_pr[n - 1]);
      }
      s_t[n - 1] = dt;
    }
  }
  __syncthreads();
  if (lane == 0) {
    double tau = 0.0;
    s_tau[0] = tau;
    for (int i = 0; i < np1p; i += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = s_t[i + u];
#pragma unroll
      for (int u = 0; u < 8; u++) s_tau[i + u + 1] = tau = tau + v[u];
    }
  }
  __syncthreads();
  if (taudst_out)
    for (int n = lane; n < npdst; n += 64) taudst_out[(size_t)col * npdst + n] = s_tau[n];
  // FILLTAUCUM: TAUCUM at each level k = 4..nlevels by interpolation in TAUDST, then its differences.
  const double pcol = in.p[col];
  for (int k = 4 + lane; k <= nlevels; k += 64) {
    const double pstar = pcol * in.sigma[k - 1] + in.ptrop;
    const double pstar1 = fmax(pstar, s_pr[0]);
    int nstar = jsrchgt(npdst - 1, s_pr, pstar1) - 1;
    if (nstar > npdst - 1) nstar = npdst - 1;
    s_tc[k] = s_tau[nstar - 1] + (pstar1 - s_pr[nstar - 1]) * (s_tau[nstar] - s_tau[nstar - 1]) /
                                     (s_pr[nstar] - s_pr[nstar - 1]);
  }
  __syncthreads();
  double *tauref = tauref_all + (size_t)col * (nlevels + 1);
  for (int k = 1 + lane; k <= nlevels + 1; k += 64)
    tauref[k - 1] = k < 4 || k > nlevels ? 0.0 : s_tc[k] - (k == 4 ? 0.0 : s_tc[k - 1]);
}
}  // namespace aer
}  // namespace mgcm

extern "C" int mgcm_rt_dust_grid(double psf, double ptrop, int npdst, double *prdst) {
  if (!prdst || npdst < 2 || npdst > NPDST_MAX || !(psf > 0.0) || !(ptrop > 0.0)) return MGCM_RT_ERR_ARG;
  const double refpr = pow(5.0 * psf / ptrop, (double)(1.0f / ((float)npdst - 1.0f)));  // code/dustprofile.f:32
  prdst[0] = ptrop;
  for (int n = 2; n <= npdst; n++) prdst[n - 1] = refpr * prdst[n - 2];
  return MGCM_RT_OK;
}

extern "C" int mgcm_rt_dust_tauref(const mgcm_rt_dust_columns *in, double *tauref, double *taudst, void *stream) {
  if (!in || !tauref) return MGCM_RT_ERR_ARG;
  if (in->ncol < 0 || in->npdst < 2 || in->npdst > NPDST_MAX) return MGCM_RT_ERR_ARG;
  if (in->nlay < 1 || in->nlay > MAX_NLAY) return MGCM_RT_ERR_NLAY;
  if (in->ncol == 0) return MGCM_RT_OK;
  if (!in->tautot || !in->conr || !in->p || !in->sigma || !in->prdst) return MGCM_RT_ERR_ARG;
  hipLaunchKernelGGL(k_dust_tauref, dim3(in->ncol), dim3(64), 0, (hipStream_t)stream, *in, tauref,
                     taudst);
  return hipGetLastError() == hipSuccess ? MGCM_RT_OK : MGCM_RT_ERR_DEVICE;
}

// Host-pointer variant (no context: a temporary device block per call).
extern "C" int mgcm_rt_dust_tauref_host(const mgcm_rt_dust_columns *in, double *tauref, double *taudst) {
  if (!in || !tauref) return MGCM_RT_ERR_ARG;
  if (in->ncol < 0 || in->npdst < 2 || in->npdst > NPDST_MAX) return MGCM_RT_ERR_ARG;
  if (in->nlay < 1 || in->nlay > MAX_NLAY) return MGCM_RT_ERR_NLAY;
  if (in->ncol == 0) return MGCM_RT_OK;
  if (!in->tautot || !in->conr || !in->p || !in->sigma || !in->prdst) return MGCM_RT_ERR_ARG;
  const size_t ncol = in->ncol, nlev = 2 * (size_t)in->nlay + 3, np = in->npdst;
  const size_t total = 3 * ncol + nlev + np + ncol * (nlev + 1) + (taudst ? ncol * np : 0);
  double *d = nullptr;
  if (hipMalloc(&d, total * sizeof(double)) != hipSuccess) return MGCM_RT_ERR_NOMEM;
  mgcm_rt_dust_columns din = *in;
  double *p = d;
  bool ok = true;
  auto up = [&](const double *src, size_t n) {
    double *q = p;
    ok = ok && hipMemcpy(q, src, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    p += n;
    return q;
  };
  din.tautot = up(in->tautot, ncol);
  din.conr = up(in->conr, ncol);
  din.p = up(in->p, ncol);
  din.sigma = up(in->sigma, nlev);
  din.prdst = up(in->prdst, np);
  double *dtr = p;
  p += ncol * (nlev + 1);
  double *dtd = taudst ? p : nullptr;
  int s = ok ? mgcm_rt_dust_tauref(&din, dtr, dtd, nullptr) : MGCM_RT_ERR_DEVICE;
  if (s == MGCM_RT_OK && hipStreamSynchronize(nullptr) != hipSuccess) s = MGCM_RT_ERR_DEVICE;
  if (s == MGCM_RT_OK &&
      (hipMemcpy(tauref, dtr, ncol * (nlev + 1) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
       (taudst && hipMemcpy(taudst, dtd, ncol * np * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)))
    s = MGCM_RT_ERR_DEVICE;
  (void)hipFree(d);
  return s;
}

// rt_plain.hip -- FP64 plain path of the column radiative transfer on gfx950.
//
// Work decomposition (one launch each, all on one stream):
//   k_gas<VIS|IR>   one wave = 4 columns x 16 k-distribution Gauss points of one
//                   band; the wave stages the 4 columns' table indices and
//                   aerosol optics in LDS, each lane solves its (column, band, g)
//                   item, and the 16 lanes of a column reduce their weighted
//                   fluxes in a fixed order.
//   k_clear<VIS|IR> one lane = the continuum-only Gauss point 17 of one
//                   (column, band), weighted with FZERO plus the weight of the
//                   TLIMIT-skipped points (read from k_gas's skip masks).
//   k_finalize      one lane = one column: sums the bands in order, forms the
//                   net-flux differences and the layer heating rates.
//
// Every FP64 statement follows the reference's evaluation order; the file is
// compiled with -ffp-contract=off so no product is fused into a sum.
#include "rt_internal.h"
#include "rt_libm.h"

namespace mgcm {

// ---------------------------------------------------------------------------
// Per-column staging in LDS for a gas wave (4 columns).
struct GasLds {
  int *base;     // [4][nsub]  cell offset of (MT, MP, NH2O) within the band table
  int *dnh;      // [4][nsub]  offset of the NH2O+1 neighbour (0 at the top of the grid)
  double *lc;    // [4][nsub][4] bilinear weights LCOEF (code/tpindex.f90:110-113)
  double *wr;    // [4][nsub]  WRATIO
  double *u;     // [4][nsub]  Cmk*dp
  double *c1;    // [4][nsub]  VIS: TRAY+TAEROS+TCLOUD; IR: TAEROS
  double *c2;    // [4][nsub]  IR: TCLOUD
  double *scat;  // [4][nl]    layer scattering optical depth (WBAR numerator)
  double *cosb;  // [4][nl]    layer asymmetry (continuum channel)
  double *bk;    // [4][nlev]  IR: PLANCKIR(nw, int(TLEV(K)*10-499))
};

// TPINDEX (code/tpindex.f90:72-132) for one sub-level.
__device__ void tpindex_dev(const RtConst *__restrict__ rc, double pw, double tw, double qh2o,
                            double lc[4], int &mt, int &mp, int &nh, double &wratio) {
  if (tw <= rc->tgasref[0]) {
    mt = 1;
  } else {
    mt = NT - 1;
    for (int n = 1; n <= NT - 1; n++) {
      if (tw > rc->tgasref[n - 1] && tw <= rc->tgasref[n]) { mt = n; break; }
    }
  }
  const double u = (tw - rc->tgasref[mt - 1]) / (rc->tgasref[mt] - rc->tgasref[mt - 1]);
  const double pwl = mgcm_libm::mgcm_log10(pw);
  mp = NP - 1;
  for (int n = 2; n <= NP - 1; n++) {
    if (pwl <= rc->pfgasref[n - 1]) { mp = n - 1; break; }
  }
  const double t = (pwl - rc->pfgasref[mp - 1]) / (rc->pfgasref[mp] - rc->pfgasref[mp - 1]);
  lc[0] = (1.0 - t) * (1.0 - u);
  lc[1] = t * (1.0 - u);
  lc[2] = t * u;
  lc[3] = (1.0 - t) * u;
  nh = 1;
  wratio = 0.0;
  if (qh2o <= rc->wrefh2o[0]) {
    nh = 1;
  } else if (qh2o >= rc->wrefh2o[NH - 1]) {
    nh = NH;
  } else {
    for (int n = 2; n <= NH; n++) {
      if (qh2o >= rc->wrefh2o[n - 2] && qh2o < rc->wrefh2o[n - 1]) {
        nh = n - 1;
        wratio = (qh2o - rc->wrefh2o[n - 2]) / (rc->wrefh2o[n - 1] - rc->wrefh2o[n - 2]);
        break;
      }
    }
  }
}

// Aerosol and Rayleigh optical depths of one sub-level K (1-based, 2..NLEV),
// OPTCV code/optcv.f90:103-123,184-185 and OPTCI code/optci.f90:113-130,184-185.
struct SubAer {
  double tray, taeros, tcloud, tlk, tck;
};

template <bool VIS>
__device__ __forceinline__ SubAer sub_aerosol(const RtConst *__restrict__ rc, int b,
                                              const double *__restrict__ plev,
                                              const double *__restrict__ tauref,
                                              const double *__restrict__ taurefcld, int K) {
  SubAer s;
  const double dpr = plev[K - 1] - plev[K - 2];
  const double tr = tauref[K - 1] / (VIS ? rc->qextref_vis : rc->qextref_dst);
  const double tc = taurefcld[K - 1] / rc->qextref_cld;
  s.tray = VIS ? rc->tauray[b] * dpr : 0.0;
  s.taeros = tr * rc->qx_dst[b];
  s.tcloud = tc * rc->qx_cld[b];
  s.tlk = tr * rc->qs_dst[b];
  s.tck = tc * rc->qs_cld[b];
  return s;
}

// IR padding level L_LEVELS+1 (code/optci.f90:89-95,182-187): TAUREF there is
// not divided by the reference extinction.
__device__ __forceinline__ void ir_pad(const RtConst *__restrict__ rc, int b,
                                       const double *__restrict__ tauref,
                                       const double *__restrict__ taurefcld, int nlev,
                                       double &tlk, double &tck) {
  tlk = tauref[nlev] * rc->qs_dst[b];
  tck = taurefcld[nlev] * rc->qs_cld[b];
}

// Layer scattering optical depth and asymmetry of the continuum channel:
// VIS code/optcv.f90:194-222, IR code/optci.f90:195-219.
template <bool VIS>
__device__ void layer_aerosol(const RtConst *__restrict__ rc, int b, const double *__restrict__ plev,
                              const double *__restrict__ tauref, const double *__restrict__ taurefcld,
                              int nl, int L, double &scat, double &cosb) {
  const int nlev = 2 * nl + 1;
  const int K = 2 * L + 1;
  const double gd = rc->g_dst[b], gc = rc->g_cld[b];
  if (VIS) {
    const double f9999 = F32LIT(0.9999);
    const SubAer a = sub_aerosol<true>(rc, b, plev, tauref, taurefcld, K);
    if (L < nl) {
      const SubAer a2 = sub_aerosol<true>(rc, b, plev, tauref, taurefcld, K + 1);
      cosb = (gd * a.tlk + gd * a2.tlk + gc * a.tck + gc * a2.tck) /
             (a.tray + a2.tray + a.tlk + a2.tlk + a.tck + a2.tck);
      scat = a.tlk + a2.tlk + a.tck + a2.tck + (a.tray + a2.tray) * f9999;
    } else {
      cosb = (gd * a.tlk + gc * a.tck) / (a.tray + a.tlk + a.tck);
      scat = a.tlk + a.tck + a.tray * f9999;
    }
  } else {
    const SubAer a = sub_aerosol<false>(rc, b, plev, tauref, taurefcld, K);
    double tlk2, tck2;
    if (K + 1 <= nlev) {
      const SubAer a2 = sub_aerosol<false>(rc, b, plev, tauref, taurefcld, K + 1);
      tlk2 = a2.tlk;
      tck2 = a2.tck;
    } else {
      ir_pad(rc, b, tauref, taurefcld, nlev, tlk2, tck2);
    }
    scat = a.tlk + tlk2 + a.tck + tck2;
    cosb = (scat > 0.0) ? (gd * a.tlk + gd * tlk2 + gc * a.tck + gc * tck2) / scat : 0.0;
  }
}

// ---------------------------------------------------------------------------
// One two-stream item.  NLM bounds the radiation layer count (arrays are
// 1-based).  Inputs arrive through the Src functor: dtk(K) gives the sub-level
// optical depth DTAUK(K) and adds the gas part to taugsurf.

template <int NLM>
struct VisState {
  double w0[NLM + 1], cb[NLM + 1], dtau[NLM + 1], tau[NLM + 2], hm[NLM + 1], tm[NLM + 1];
  double lam[NLM + 1], gama[NLM + 1], am[NLM + 1], ap[NLM + 1];
  double as[2 * NLM + 1], ds[2 * NLM + 1];
};

// VIS optics of one item + delta-Eddington scaling (OPTCV layer assembly
// code/optcv.f90:189-274 fused with GETDETAU/GFLUXV code/gfluxv.f90:76-103).
// Returns DETAU; cumtau receives TAUCUMV(L_LEVELS).
template <int NLM, typename Src>
__device__ double vis_optics(int nl, Src &src, const double *__restrict__ scat,
                             const double *__restrict__ cosb, VisState<NLM> &s, double &cumtau) {
  // sub-level 2 (top half of the cap)
  const double d2 = src.dtk(2);
  double tcv = d2;  // TAUCUMV(2)
  double fprev = 0.0, tcv_prev_level = 0.0;
  s.tau[1] = 0.0;
  for (int L = 1; L <= nl - 1; L++) {
    const double a = src.dtk(2 * L + 1), b = src.dtk(2 * L + 2);
    const double dtauv = a + b;
    const double tcA = tcv + a;  // TAUCUMV(2L+1)
    const double tcB = tcA + b;  // TAUCUMV(2L+2)
    const double W = scat[L - 1] / dtauv, C = cosb[L - 1];
    const double F = 1.0 - W * (C * C);
    if (L == 1) {
      s.tau[1] = 0.0 * F;  // TDEL(1)*FACTOR with TDEL(1) = 0
      s.tm[1] = tcv * F;   // TAUCUM(2)
      s.hm[1] = s.tm[1] + (tcA - tcv) * F;
    } else {
      // TAUCUM(2L) = TAU(L); TAUCUM(2L+1) uses the factor of layer L-1
      s.tm[L] = s.tau[L];
      s.hm[L] = s.tm[L] + (tcA - tcv_prev_level) * fprev;
    }
    s.w0[L] = W * (1.0 - C * C) / F;
    s.cb[L] = C / (1.0 + C);
    s.dtau[L] = dtauv * F;
    s.tau[L + 1] = s.tau[L] + s.dtau[L];
    fprev = F;
    tcv_prev_level = tcB;  // TAUCUMV(2L+2), the base of the next layer's mid level
    tcv = tcB;
  }
  {
    const int L = nl;
    const double a = src.dtk(2 * L + 1);
    const double dtauv = a;
    cumtau = tcv + a;  // TAUCUMV(L_LEVELS)
    const double W = scat[L - 1] / dtauv, C = cosb[L - 1];
    const double F = 1.0 - W * (C * C);
    s.w0[L] = W * (1.0 - C * C) / F;
    s.cb[L] = C / (1.0 + C);
    s.dtau[L] = dtauv * F;
    s.tau[L + 1] = s.tau[L] + s.dtau[L];
  }
  return s.tau[nl + 1];
}

// GFLUXV (code/gfluxv.f90:108-280) + DSOLVER/DTRIDGL (code/dsolver.f90,
// code/dtridgl.f90) on a scaled item.  Calls acc(L, fmidp, fmidm) for each
// layer mid-point and returns the top fluxes and DIFFV.
template <int NLM, typename Acc>
__device__ void vis_solve(int nl, VisState<NLM> &s, double u0, double f0pi, double rsf, double btop,
                          double bsurf, Acc &acc, double &fluxup, double &fluxdn, double &diffv) {
  const double s3 = 0x1.bb67aep+0;   // SQRT(3.0) in REAL(4) (code/gfluxv.f90:114-116)
  const double s3h = 0x1.bb67aep-1;  // SQRT(3.0)*0.5 in REAL(4)
  const double f1em10 = F32LIT(1.e-10);
  // bottom-up: layer constants and the reverse-Thomas elimination
  double gama_n1 = 0.0, cpm1_n1 = 0.0, cmm1_n1 = 0.0;
  for (int n = nl; n >= 1; n--) {
    const double w0 = s.w0[n], cb = s.cb[n];
    const double g1 = s3h * (2.0 - w0 * (1.0 + cb));
    const double g2 = (s3 * w0 * 0.5) * (1.0 - cb);
    const double g3 = 0.5 * (1.0 - s3 * cb * u0);
    const double lam = sqrt(g1 * g1 - g2 * g2);
    const double gama = (g1 - lam) / g2;
    const double g4 = 1.0 - g3;
    double denom = lam * lam - 1.0 / (u0 * u0);
    if (denom == 0.0) denom = f1em10;
    const double am = f0pi * w0 * (g4 * (g1 + 1.0 / u0) + g2 * g3) / denom;
    const double ap = f0pi * w0 * (g3 * (g1 - 1.0 / u0) + g2 * g4) / denom;
    const double xtop = mgcm_libm::mgcm_exp(-dmin(s.tau[n] / u0, MAXEXP));
    const double xbot = mgcm_libm::mgcm_exp(-dmin(s.tau[n + 1] / u0, MAXEXP));
    const double cpm1 = ap * xtop, cmm1 = am * xtop;
    const double cp = ap * xbot, cm = am * xbot;
    const double ep = mgcm_libm::mgcm_exp(dmin(MAXEXP, lam * s.dtau[n]));
    const double em = 1.0 / ep;
    const double e1 = ep + gama * em, e2 = ep - gama * em;
    const double e3 = gama * ep + em, e4 = gama * ep - em;
    s.lam[n] = lam;
    s.gama[n] = gama;
    s.am[n] = am;
    s.ap[n] = ap;
    if (n == nl) {  // surface row (dsolver.f90:82-85)
      const int i = 2 * nl;
      const double af = e1 - rsf * e3, bf = e2 - rsf * e4;
      const double df = bsurf - cp + rsf * cm;
      s.as[i] = af / bf;
      s.ds[i] = df / bf;
    } else {
      {  // odd row 2n+1 (dsolver.f90:74-80)
        const int i = 2 * n + 1;
        const double af = 2.0 * (1.0 - gama * gama);
        const double bf = (e1 - e3) * (1.0 + gama_n1);
        const double cf = (e1 + e3) * (gama_n1 - 1.0);
        const double df = e3 * (cpm1_n1 - cp) + e1 * (cm - cmm1_n1);
        const double x = 1.0 / (bf - cf * s.as[i + 1]);
        s.as[i] = af * x;
        s.ds[i] = (df - cf * s.ds[i + 1]) * x;
      }
      {  // even row 2n (dsolver.f90:63-70)
        const int i = 2 * n;
        const double af = (e1 + e3) * (gama_n1 - 1.0);
        const double bf = (e2 + e4) * (gama_n1 - 1.0);
        const double cf = 2.0 * (1.0 - gama_n1 * gama_n1);
        const double df = (gama_n1 - 1.0) * (cpm1_n1 - cp) + (1.0 - gama_n1) * (cm - cmm1_n1);
        const double x = 1.0 / (bf - cf * s.as[i + 1]);
        s.as[i] = af * x;
        s.ds[i] = (df - cf * s.ds[i + 1]) * x;
      }
    }
    if (n == 1) {  // top row (dsolver.f90:54-57)
      const double af = 0.0, bf = gama + 1.0, cf = gama - 1.0, df = btop - cmm1;
      const double x = 1.0 / (bf - cf * s.as[2]);
      s.as[1] = af * x;
      s.ds[1] = (df - cf * s.ds[2]) * x;
    }
    gama_n1 = gama;
    cpm1_n1 = cpm1;
    cmm1_n1 = cmm1;
  }
  // top-down: substitution, unmixing and the layer fluxes
  double xkprev = s.ds[1];
  double xk1_1 = 0.0, xk2_1 = 0.0;
  for (int n = 1; n <= nl; n++) {
    const double xa = (n == 1) ? s.ds[1] : s.ds[2 * n - 1] - s.as[2 * n - 1] * xkprev;
    const double xb = s.ds[2 * n] - s.as[2 * n] * xa;
    xkprev = xb;
    const double xk1 = xa + xb;
    double xk2 = xa - xb;
    if (xk2 != 0.0 && fabs(xk2 / (xa + F32LIT(1.e-20))) < F32LIT(1.e-30)) xk2 = 0.0;
    const double lam = s.lam[n], gama = s.gama[n], am = s.am[n], ap = s.ap[n];
    double fmidp, fmidm;
    if (n < nl) {  // gfluxv.f90:170-206
      const double ep = mgcm_libm::mgcm_exp(dmin(MAXEXP, lam * (s.hm[n] - s.tm[n])));
      const double em = 1.0 / ep;
      const double taumid = s.hm[n];
      const double xm = mgcm_libm::mgcm_exp(-dmin(taumid / u0, MAXEXP));
      const double cpmid = ap * xm, cmmid = am * xm;
      fmidp = xk1 * ep + gama * xk2 * em + cpmid;
      fmidm = xk1 * ep * gama + xk2 * em + cmmid;
      fmidm = fmidm + u0 * f0pi * xm;
    } else {  // bottom layer, full sub-layer (gfluxv.f90:243-280)
      const double ep = mgcm_libm::mgcm_exp(dmin(MAXEXP, lam * (s.tau[nl + 1] - s.tau[nl])));
      const double em = 1.0 / ep;
      const double taumid = dmin(s.tau[nl + 1], MAXEXP);
      const double xm = mgcm_libm::mgcm_exp(-dmin(taumid / u0, MAXEXP));
      const double cpmid = ap * xm, cmmid = am * xm;
      fmidp = xk1 * ep + gama * xk2 * em + cpmid;
      fmidm = xk1 * ep * gama + xk2 * em + cmmid;
      diffv = fmidm;
      fmidm = fmidm + u0 * f0pi * xm;
    }
    acc(n, fmidp, fmidm);
    if (n == 1) { xk1_1 = xk1; xk2_1 = xk2; }
  }
  {  // top of the atmosphere (gfluxv.f90:210-238)
    const double ep = 1.0, em = 1.0;
    fluxup = xk1_1 * ep + s.gama[1] * xk2_1 * em + s.ap[1];
    fluxdn = xk1_1 * ep * s.gama[1] + xk2_1 * em + s.am[1];
    fluxdn = fluxdn + u0 * f0pi * mgcm_libm::mgcm_exp(-dmin(0.0 / u0, MAXEXP));
  }
}

template <int NLM>
struct IrState {
  double w0[NLM + 1], cb[NLM + 1], dtau[NLM + 1], dmid[NLM + 1];
  double lam[NLM + 1], gama[NLM + 1], b0[NLM + 1], b1[NLM + 1], term[NLM + 1];
  double as[2 * NLM + 1], ds[2 * NLM + 1];
};

// IR optics of one item (OPTCI code/optci.f90:189-269).
template <int NLM, typename Src>
__device__ void ir_optics(int nl, Src &src, const double *__restrict__ scat,
                          const double *__restrict__ cosb, IrState<NLM> &s) {
  const double f1em9 = F32LIT(1.0e-9);
  double tcv = src.dtk(2);  // TAUCUMI(2)
  for (int L = 1; L <= nl; L++) {
    const double a = src.dtk(2 * L + 1);
    const double b = (L < nl) ? src.dtk(2 * L + 2) : 0.0;  // DTAUKI(L_LEVELS+1) = 0
    double dt = a + b + 1.e-50;
    double w;
    if (dt > f1em9) {
      w = scat[L - 1] / dt;
    } else {
      w = 0.0;
      dt = f1em9;
    }
    const double tcA = tcv + a;  // TAUCUMI(2L+1)
    s.dmid[L] = tcA - tcv;       // TAUCUMI(2L+1) - TAUCUMI(2L)
    tcv = (L < nl) ? tcA + b : tcA;
    s.w0[L] = w;
    s.cb[L] = cosb[L - 1];
    s.dtau[L] = dt;
  }
}

// GFLUXI (code/gfluxi.f90:69-194) + DSOLVER/DTRIDGL.  bk(K) = Planck value of
// TLEV(K).  Calls acc(L, fmidp, fmidm); returns FTOPUP.
template <int NLM, typename Acc>
__device__ double ir_solve(int nl, IrState<NLM> &s, const double *__restrict__ bk, double rsf,
                           double btop, double bsurf, Acc &acc) {
  double gama_n1 = 0.0, cpm1_n1 = 0.0, cmm1_n1 = 0.0;
  for (int n = nl; n >= 1; n--) {
    const double w0 = s.w0[n], cb = s.cb[n], dtau = s.dtau[n];
    const double alpha = sqrt((1.0 - w0) / (1.0 - w0 * cb));
    const double lam = alpha * (1.0 - w0 * cb) / UBARI;
    double b0, b1;
    if (n < nl) {
      const double bt = bk[2 * n - 1], bb = bk[2 * n + 1];  // TLEV(2L), TLEV(2L+2)
      b1 = (bb - bt) / dtau;
      b0 = bt;
    } else {
      const double bt = bk[2 * n - 1], bb = bk[2 * n];      // TLEV(2L), TLEV(2L+1)
      b1 = (bb - bt) / dtau;
      b0 = bt;
    }
    const double gama = (1.0 - alpha) / (1.0 + alpha);
    const double term = UBARI / (1.0 - w0 * cb);
    const double cp = b0 + b1 * dtau + b1 * term;
    const double cm = b0 + b1 * dtau - b1 * term;
    const double cpm1 = b0 + b1 * term;
    const double cmm1 = b0 - b1 * term;
    const double ep = mgcm_libm::mgcm_exp(dmin(lam * dtau, MAXEXP));
    const double em = 1.0 / ep;
    const double e1 = ep + gama * em, e2 = ep - gama * em;
    const double e3 = gama * ep + em, e4 = gama * ep - em;
    s.lam[n] = lam;
    s.gama[n] = gama;
    s.b0[n] = b0;
    s.b1[n] = b1;
    s.term[n] = term;
    if (n == nl) {
      const int i = 2 * nl;
      const double af = e1 - rsf * e3, bf = e2 - rsf * e4;
      const double df = bsurf - cp + rsf * cm;
      s.as[i] = af / bf;
      s.ds[i] = df / bf;
    } else {
      {
        const int i = 2 * n + 1;
        const double af = 2.0 * (1.0 - gama * gama);
        const double bf = (e1 - e3) * (1.0 + gama_n1);
        const double cf = (e1 + e3) * (gama_n1 - 1.0);
        const double df = e3 * (cpm1_n1 - cp) + e1 * (cm - cmm1_n1);
        const double x = 1.0 / (bf - cf * s.as[i + 1]);
        s.as[i] = af * x;
        s.ds[i] = (df - cf * s.ds[i + 1]) * x;
      }
      {
        const int i = 2 * n;
        const double af = (e1 + e3) * (gama_n1 - 1.0);
        const double bf = (e2 + e4) * (gama_n1 - 1.0);
        const double cf = 2.0 * (1.0 - gama_n1 * gama_n1);
        const double df = (gama_n1 - 1.0) * (cpm1_n1 - cp) + (1.0 - gama_n1) * (cm - cmm1_n1);
        const double x = 1.0 / (bf - cf * s.as[i + 1]);
        s.as[i] = af * x;
        s.ds[i] = (df - cf * s.ds[i + 1]) * x;
      }
    }
    if (n == 1) {
      const double af = 0.0, bf = gama + 1.0, cf = gama - 1.0, df = btop - cmm1;
      const double x = 1.0 / (bf - cf * s.as[2]);
      s.as[1] = af * x;
      s.ds[1] = (df - cf * s.ds[2]) * x;
    }
    gama_n1 = gama;
    cpm1_n1 = cpm1;
    cmm1_n1 = cmm1;
  }
  double xkprev = s.ds[1];
  double xk1_1 = 0.0, xk2_1 = 0.0;
  for (int n = 1; n <= nl; n++) {
    const double xa = (n == 1) ? s.ds[1] : s.ds[2 * n - 1] - s.as[2 * n - 1] * xkprev;
    const double xb = s.ds[2 * n] - s.as[2 * n] * xa;
    xkprev = xb;
    const double xk1 = xa + xb;
    double xk2 = xa - xb;
    if (xk2 != 0.0 && fabs(xk2 / (xa + F32LIT(1.e-20))) < F32LIT(1.e-30)) xk2 = 0.0;
    const double lam = s.lam[n], gama = s.gama[n], b0 = s.b0[n], b1 = s.b1[n], term = s.term[n];
    const double dk = (n < nl) ? s.dmid[n] : s.dtau[n];
    const double ep = mgcm_libm::mgcm_exp(dmin(lam * dk, MAXEXP));
    const double em = 1.0 / ep;
    const double cpmid = b0 + b1 * dk + b1 * term;
    const double cmmid = b0 + b1 * dk - b1 * term;
    double fmidp = xk1 * ep + gama * xk2 * em + cpmid;
    double fmidm = xk1 * ep * gama + xk2 * em + cmmid;
    fmidp = fmidp * PI;
    fmidm = fmidm * PI;
    acc(n, fmidp, fmidm);
    if (n == 1) { xk1_1 = xk1; xk2_1 = xk2; }
  }
  const double ep = 1.0, em = 1.0;
  const double cpmid = s.b0[1] + s.b1[1] * s.term[1];
  const double cmmid = s.b0[1] - s.b1[1] * s.term[1];
  const double fluxup = xk1_1 * ep + s.gama[1] * xk2_1 * em + cpmid;
  const double fluxdn = xk1_1 * ep * s.gama[1] + xk2_1 * em + cmmid;
  return (fluxup - fluxdn) * PI;
}

// ---------------------------------------------------------------------------
// Gas-point source: k-interpolation from the g-fastest table + LDS staging.
template <bool VIS>
struct GasSrc {
  const double *__restrict__ tab;  // band's table base
  const GasLds *lds;
  int c, nsub, g0;  // g0 = g-1
  double taugsurf;
  __device__ __forceinline__ double dtk(int K) {
    const int k = c * nsub + (K - 2);
    const int base = lds->base[k], dnh = lds->dnh[k];
    const double wr = lds->wr[k];
    const double *lc = lds->lc + 4 * k;
    const double *p = tab + base + g0;
    // corners (MT,MP), (MT,MP+1), (MT+1,MP+1), (MT+1,MP)   (code/optcv.f90:142-162)
    const int sT = NGAS, sP = NH * NT * NGAS;
    const double a0 = p[0], a1 = p[dnh];
    const double b0 = p[sP], b1 = p[sP + dnh];
    const double c0 = p[sP + sT], c1 = p[sP + sT + dnh];
    const double d0 = p[sT], d1 = p[sT + dnh];
    const double k1 = a0 + wr * (a1 - a0);
    const double k2 = b0 + wr * (b1 - b0);
    const double k3 = c0 + wr * (c1 - c0);
    const double k4 = d0 + wr * (d1 - d0);
    const double ans = lc[0] * k1 + lc[1] * k2 + lc[2] * k3 + lc[3] * k4;
    const double taugas = lds->u[k] * ans;
    taugsurf = taugsurf + taugas;
    if (VIS) return taugas + lds->c1[k];
    return taugas + lds->c1[k] + lds->c2[k];
  }
};

// 16-lane fixed-order sum; every lane of the group returns lane 0's value.
__device__ __forceinline__ double group16_sum(double v) {
  v = v + __shfl_xor(v, 8, 16);
  v = v + __shfl_xor(v, 4, 16);
  v = v + __shfl_xor(v, 2, 16);
  v = v + __shfl_xor(v, 1, 16);
  return __shfl(v, 0, 16);
}

template <bool VIS>
__device__ void stage_gas_columns(const DevTables &dt, const RunArgs &a, int b, int col0, GasLds &L,
                                  int nl, int nsub, int nlev) {
  const RtConst *rc = dt.c;
  const int lane = threadIdx.x;
  for (int t = lane; t < 4 * nsub; t += 64) {
    const int c = t / nsub, K = t % nsub + 2;
    const int col = col0 + c;
    if (col >= a.ncol) continue;
    const double *plev = a.plev + (size_t)col * nlev;
    int mt, mp, nh;
    double lc[4], wr;
    tpindex_dev(rc, a.pmid[(size_t)col * nlev + K - 1], a.tmid[(size_t)col * nlev + K - 1],
                a.qh2o[(size_t)col * nlev + K - 1], lc, mt, mp, nh, wr);
    L.base[t] = kcell(0, mt, mp, nh);
    L.dnh[t] = (nh < NH) ? NT * NGAS : 0;
    for (int j = 0; j < 4; j++) L.lc[4 * t + j] = lc[j];
    L.wr[t] = wr;
    L.u[t] = CMK * (plev[K - 1] - plev[K - 2]);
    const SubAer s = sub_aerosol<VIS>(rc, b, plev, a.tauref + (size_t)col * (nlev + 1),
                                      a.taurefcld + (size_t)col * (nlev + 1), K);
    if (VIS) {
      L.c1[t] = s.tray + s.taeros + s.tcloud;
    } else {
      L.c1[t] = s.taeros;
      L.c2[t] = s.tcloud;
    }
  }
  for (int t = lane; t < 4 * nl; t += 64) {
    const int c = t / nl, Lr = t % nl + 1;
    const int col = col0 + c;
    if (col >= a.ncol) continue;
    layer_aerosol<VIS>(rc, b, a.plev + (size_t)col * nlev, a.tauref + (size_t)col * (nlev + 1),
                       a.taurefcld + (size_t)col * (nlev + 1), nl, Lr, L.scat[t], L.cosb[t]);
  }
  if (!VIS) {
    const int nw = b - NV;
    for (int t = lane; t < 4 * nlev; t += 64) {
      const int c = t / nlev, K = t % nlev + 1;
      const int col = col0 + c;
      if (col >= a.ncol) continue;
      int nt = (int)(a.tlev[(size_t)col * nlev + K - 1] * 10.0 - 499);
      if (nt < 1 || nt > NPLANCK) {
        atomicOr(a.flag, 1);
        nt = nt < 1 ? 1 : NPLANCK;
      }
      L.bk[t] = dt.planck[(size_t)(nt - 1) * NI + nw];
    }
  }
}

template <int NLM, bool VIS>
__global__ __launch_bounds__(64) void k_gas(DevTables dt, RunArgs a, int band0) {
  extern __shared__ double smem[];
  const int nl = a.nlay + 1, nlev = 2 * a.nlay + 3, nsub = nlev - 1;
  const int b = band0 + blockIdx.y;
  const int col0 = blockIdx.x * 4;
  const int lane = threadIdx.x, c = lane >> 4, g0 = lane & 15;
  const int col = col0 + c;
  const RtConst *rc = dt.c;
  GasLds L;
  {
    double *p = smem;
    L.lc = p; p += 4 * 4 * nsub;
    L.wr = p; p += 4 * nsub;
    L.u = p; p += 4 * nsub;
    L.c1 = p; p += 4 * nsub;
    L.c2 = p; p += 4 * nsub;
    L.scat = p; p += 4 * nl;
    L.cosb = p; p += 4 * nl;
    L.bk = p; p += 4 * nlev;
    int *q = (int *)p;
    L.base = q; q += 4 * nsub;
    L.dnh = q;
  }
  const double fz = rc->fzero[b];
  const bool band_skip = fz >= F32LIT(0.99);
  if (band_skip) {  // FZERO >= 0.99: straight to the continuum point (code/sfluxv.f90:64-65)
    if (g0 == 0 && col < a.ncol) a.skip[(size_t)b * a.ncol + col] = 0u;
    const int nf = 3 * nl + 2;
    for (int t = lane; t < 4 * nf; t += 64) {
      const int cc = col0 + t / nf;
      if (cc < a.ncol) a.gas[((size_t)b * nf + t % nf) * a.ncol + cc] = 0.0;
    }
    return;
  }
  stage_gas_columns<VIS>(dt, a, b, col0, L, nl, nsub, nlev);
  __syncthreads();

  const bool valid = col < a.ncol;
  const int colc = valid ? col : a.ncol - 1;
  const double u0 = VIS ? a.acosz[colc] : 0.0;
  const bool sun = VIS ? (u0 >= F32LIT(1.0e-5)) : true;
  const double gw = rc->gw[b][g0];
  const double onemf = 1.0 - fz;
  const int nw = VIS ? b : b - NV;
  const double dw = VIS ? 0.0 : rc->dwni[nw];
  GasSrc<VIS> src{(VIS ? dt.kv : dt.ki) + (size_t)nw * NP * NH * NT * NGAS, &L, c, nsub, g0, 0.0};
  const double *scat = L.scat + c * nl, *cosb = L.cosb + c * nl;
  const int nf = 3 * nl + 2;
  double *gout = a.gas + (size_t)b * nf * a.ncol;

  bool skip;
  if (VIS) {
    VisState<NLM> s;
    double cumtau = 0.0, detau = 0.0;
    if (valid && sun) detau = vis_optics<NLM>(nl, src, scat, cosb, s, cumtau);
    skip = !(valid && sun) || src.taugsurf < TLIMITS;
    if (valid && sun) {
      a.detau[(size_t)col * NG * NV + g0 * NV + nw] = detau;
      a.cumtauv[(size_t)col * NG * NV + g0 * NV + nw] = cumtau;
    }
    double fluxup = 0.0, fluxdn = 0.0, diffv = 0.0;
    auto acc = [&](int n, double fp, double fm) {
      const double net = skip ? 0.0 : (fp - fm) * gw * onemf;
      const double up = skip ? 0.0 : fp * gw * onemf;
      const double dn = skip ? 0.0 : fm * gw * onemf;
      const double snet = group16_sum(net), sup = group16_sum(up), sdn = group16_sum(dn);
      if (g0 == 0 && valid) {
        gout[(size_t)(n - 1) * a.ncol + col] = snet;
        gout[(size_t)(nl + n - 1) * a.ncol + col] = sup;
        gout[(size_t)(2 * nl + n - 1) * a.ncol + col] = sdn;
      }
    };
    // The 16-lane sums inside acc need every lane of the wave: the solve runs
    // wave-uniformly whenever one lane needs it, and skipped lanes (TLIMIT,
    // sun down, past the last column) contribute exact zeros through acc.
    if (__any(!skip)) {
      const double bsurf = a.als[colc] * u0 * a.sol[nw] * mgcm_libm::mgcm_exp(-dmin(detau / u0, MAXEXP));
      vis_solve<NLM>(nl, s, u0, a.sol[nw], a.als[colc], 0.0, bsurf, acc, fluxup, fluxdn, diffv);
    } else {
      for (int n = 1; n <= nl; n++) acc(n, 0.0, 0.0);
    }
    const double top = skip ? 0.0 : (fluxup - fluxdn) * gw * onemf;
    const double dif = skip ? 0.0 : diffv * gw * onemf;
    const double stop = group16_sum(top), sdif = group16_sum(dif);
    if (g0 == 0 && valid) {
      gout[(size_t)(3 * nl) * a.ncol + col] = stop;
      gout[(size_t)(3 * nl + 1) * a.ncol + col] = sdif;
    }
  } else {
    IrState<NLM> s;
    if (valid) ir_optics<NLM>(nl, src, scat, cosb, s);
    skip = !valid || src.taugsurf < TLIMITI;
    const double *bk = L.bk + c * nlev;
    const double *plev = a.plev + (size_t)colc * nlev;
    auto acc = [&](int n, double fp, double fm) {
      const double net = skip ? 0.0 : (fp - fm) * dw * gw * onemf;
      const double up = skip ? 0.0 : fp * dw * gw * onemf;
      const double dn = skip ? 0.0 : fm * dw * gw * onemf;
      const double snet = group16_sum(net), sup = group16_sum(up), sdn = group16_sum(dn);
      if (g0 == 0 && valid) {
        gout[(size_t)(n - 1) * a.ncol + col] = snet;
        gout[(size_t)(nl + n - 1) * a.ncol + col] = sup;
        gout[(size_t)(2 * nl + n - 1) * a.ncol + col] = sdn;
      }
    };
    double ftopup = 0.0;
    if (__any(!skip)) {  // wave-uniform, as for VIS
      // boundary conditions (code/sfluxi.f90:61-62,81-82)
      const double rsf = a.albi[colc];
      const double bsurf = (1. - rsf) * bk[nlev - 1];
      const double pltop = bk[1];
      const double tautop = s.dtau[1] * plev[1] / (plev[3] - plev[1]);
      const double btop = (1.0 - mgcm_libm::mgcm_exp(-tautop / UBARI)) * pltop;
      ftopup = ir_solve<NLM>(nl, s, bk, rsf, btop, bsurf, acc);
    } else {
      for (int n = 1; n <= nl; n++) acc(n, 0.0, 0.0);
    }
    const double top = skip ? 0.0 : ftopup * dw * gw * onemf;
    const double stop = group16_sum(top);
    if (g0 == 0 && valid) {
      gout[(size_t)(3 * nl) * a.ncol + col] = stop;
      gout[(size_t)(3 * nl + 1) * a.ncol + col] = 0.0;
    }
  }
  // skip mask of the column's 16 points (drives FZERO for the continuum point)
  const unsigned long long m = __ballot(skip);
  if (g0 == 0 && valid) a.skip[(size_t)b * a.ncol + col] = (uint32_t)((m >> (16 * c)) & 0xffffu);
}

// ---------------------------------------------------------------------------
// Continuum-only point g = 17, one lane per (column, band).
template <bool VIS>
struct ClearSrc {
  const RtConst *rc;
  int b;
  const double *plev, *tauref, *taurefcld;
  double taugsurf;
  __device__ __forceinline__ double dtk(int K) {
    const SubAer s = sub_aerosol<VIS>(rc, b, plev, tauref, taurefcld, K);
    if (VIS) return s.taeros + s.tray + s.tcloud;
    return s.taeros + s.tcloud;
  }
};

template <int NLM, bool VIS>
__global__ __launch_bounds__(64) void k_clear(DevTables dt, RunArgs a, int band0) {
  const int b = band0 + blockIdx.y;
  const int col = blockIdx.x * 64 + threadIdx.x;
  if (col >= a.ncol) return;
  const RtConst *rc = dt.c;
  const int nl = a.nlay + 1, nlev = 2 * a.nlay + 3;
  const int nf = 3 * nl + 2;
  const int nw = VIS ? b : b - NV;
  double *cout = a.clear + (size_t)b * nf * a.ncol;
  const double *plev = a.plev + (size_t)col * nlev;
  const double *tauref = a.tauref + (size_t)col * (nlev + 1);
  const double *taurefcld = a.taurefcld + (size_t)col * (nlev + 1);
  // FZERO plus the weight of the TLIMIT-skipped points, in g order
  // (code/sfluxv.f90:72-75, code/sfluxi.f90:72-75)
  double fzero = rc->fzero[b];
  if (!(fzero >= F32LIT(0.99))) {
    const uint32_t m = a.skip[(size_t)b * a.ncol + col];
    for (int g = 0; g < NGAS; g++)
      if (m & (1u << g)) fzero = fzero + (1.0 - rc->fzero[b]) * rc->gw[b][g];
  }
  // layer aerosol optics of this column-band (the continuum channel)
  double scat[NLM], cosb[NLM];
  for (int L = 1; L <= nl; L++) layer_aerosol<VIS>(rc, b, plev, tauref, taurefcld, nl, L, scat[L - 1], cosb[L - 1]);
  ClearSrc<VIS> src{rc, b, plev, tauref, taurefcld, 0.0};
  if (VIS) {
    const double u0 = a.acosz[col];
    if (!(u0 >= F32LIT(1.0e-5))) {
      for (int f = 0; f < nf; f++) cout[(size_t)f * a.ncol + col] = 0.0;
      return;
    }
    VisState<NLM> s;
    double cumtau;
    const double detau = vis_optics<NLM>(nl, src, scat, cosb, s, cumtau);
    a.detau[(size_t)col * NG * NV + 16 * NV + nw] = detau;
    a.cumtauv[(size_t)col * NG * NV + 16 * NV + nw] = cumtau;
    const double bsurf = a.als[col] * u0 * a.sol[nw] * mgcm_libm::mgcm_exp(-dmin(detau / u0, MAXEXP));
    auto acc = [&](int n, double fp, double fm) {
      cout[(size_t)(n - 1) * a.ncol + col] = (fp - fm) * fzero;
      cout[(size_t)(nl + n - 1) * a.ncol + col] = fp * fzero;
      cout[(size_t)(2 * nl + n - 1) * a.ncol + col] = fm * fzero;
    };
    double fluxup, fluxdn, diffv;
    vis_solve<NLM>(nl, s, u0, a.sol[nw], a.als[col], 0.0, bsurf, acc, fluxup, fluxdn, diffv);
    cout[(size_t)(3 * nl) * a.ncol + col] = (fluxup - fluxdn) * fzero;
    cout[(size_t)(3 * nl + 1) * a.ncol + col] = diffv * fzero;
  } else {
    IrState<NLM> s;
    ir_optics<NLM>(nl, src, scat, cosb, s);
    const double dw = rc->dwni[nw];
    double bk[2 * NLM + 1];
    for (int K = 1; K <= nlev; K++) {
      int nt = (int)(a.tlev[(size_t)col * nlev + K - 1] * 10.0 - 499);
      if (nt < 1 || nt > NPLANCK) {
        atomicOr(a.flag, 1);
        nt = nt < 1 ? 1 : NPLANCK;
      }
      bk[K - 1] = dt.planck[(size_t)(nt - 1) * NI + nw];
    }
    const double rsf = a.albi[col];
    const double bsurf = (1. - rsf) * bk[nlev - 1];
    const double tautop = s.dtau[1] * plev[1] / (plev[3] - plev[1]);
    const double btop = (1.0 - mgcm_libm::mgcm_exp(-tautop / UBARI)) * bk[1];
    auto acc = [&](int n, double fp, double fm) {
      cout[(size_t)(n - 1) * a.ncol + col] = (fp - fm) * dw * fzero;
      cout[(size_t)(nl + n - 1) * a.ncol + col] = fp * dw * fzero;
      cout[(size_t)(2 * nl + n - 1) * a.ncol + col] = fm * dw * fzero;
    };
    const double ftopup = ir_solve<NLM>(nl, s, bk, rsf, btop, bsurf, acc);
    cout[(size_t)(3 * nl) * a.ncol + col] = ftopup * dw * fzero;
    cout[(size_t)(3 * nl + 1) * a.ncol + col] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Band sums, net-flux differences and heating (code/comp3.f:677-681,735-738).
__global__ __launch_bounds__(64) void k_finalize(RunArgs a) {
  const int col = blockIdx.x * 64 + threadIdx.x;
  if (col >= a.ncol) return;
  const int nl = a.nlay + 1, nlev = 2 * a.nlay + 3, nf = 3 * nl + 2;
  const bool sun = a.acosz[col] >= F32LIT(1.0e-5);
  const double *plev = a.plev + (size_t)col * nlev;
  auto G = [&](int b, int f) { return a.gas[((size_t)b * nf + f) * a.ncol + col]; };
  auto Cl = [&](int b, int f) { return a.clear[((size_t)b * nf + f) * a.ncol + col]; };
  double topv = 0.0, topi = 0.0, dif = 0.0, prevv = 0.0, previ = 0.0;
  for (int b = 0; b < NV; b++) {
    topv = topv + (G(b, 3 * nl) + Cl(b, 3 * nl));
    dif = dif + (G(b, 3 * nl + 1) + Cl(b, 3 * nl + 1));
  }
  for (int b = NV; b < NB; b++) topi = topi + (G(b, 3 * nl) + Cl(b, 3 * nl));
  if (!sun) { topv = 0.0; dif = 0.0; }
  a.nfluxtopv[col] = topv;
  a.nfluxtopi[col] = topi;
  a.diffvt[col] = dif;
  for (int L = 1; L <= nl; L++) {
    double nv = 0.0, uv = 0.0, dv = 0.0, ni = 0.0, ui = 0.0, di = 0.0;
    if (sun) {
      for (int b = 0; b < NV; b++) {
        nv = nv + (G(b, L - 1) + Cl(b, L - 1));
        uv = uv + (G(b, nl + L - 1) + Cl(b, nl + L - 1));
        dv = dv + (G(b, 2 * nl + L - 1) + Cl(b, 2 * nl + L - 1));
      }
    }
    for (int b = NV; b < NB; b++) {
      ni = ni + (G(b, L - 1) + Cl(b, L - 1));
      ui = ui + (G(b, nl + L - 1) + Cl(b, nl + L - 1));
      di = di + (G(b, 2 * nl + L - 1) + Cl(b, 2 * nl + L - 1));
    }
    const size_t o = (size_t)col * nl + L - 1;
    a.fmnetv[o] = nv; a.fluxupv[o] = uv; a.fluxdnv[o] = dv;
    a.fmneti[o] = ni; a.fluxupi[o] = ui; a.fluxdni[o] = di;
    if (a.hr) {
      // SUNTOT(2L+1), IRTOTAL(2L+1): layer L-1 of the GCM (0 = stratosphere)
      const double sunt = sun ? (L == 1 ? nv - topv : nv - prevv) : 0.0;
      const double irt = (L == 1) ? ni - topi : ni - previ;
      const int l = L - 1;
      const double dp = (l == 0) ? (plev[0] + plev[1]) : (plev[2 * l + 2] - plev[2 * l]);
      a.hr[(size_t)col * nl + l] = (((sunt + irt) * GRAV) / ((CPAIR * SCALEP) * dp)) * SECONDS_PER_SOL;
    }
    prevv = nv;
    previ = ni;
  }
}

// ---------------------------------------------------------------------------
size_t gas_lds_bytes(int nlay) {
  const int nl = nlay + 1, nlev = 2 * nlay + 3, nsub = nlev - 1;
  return sizeof(double) * (size_t)(4 * 4 * nsub + 4 * 4 * nsub + 4 * 2 * nl + 4 * nlev) +
         sizeof(int) * (size_t)(2 * 4 * nsub);
}

template <int NLM>
static hipError_t launch_plain_nlm(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  const size_t lds = gas_lds_bytes(a.nlay);
  const unsigned gx = (unsigned)((a.ncol + 3) / 4);
  const unsigned cx = (unsigned)((a.ncol + 63) / 64);
  hipLaunchKernelGGL((k_gas<NLM, true>), dim3(gx, NV), dim3(64), lds, st, dt, a, 0);
  hipLaunchKernelGGL((k_gas<NLM, false>), dim3(gx, NI), dim3(64), lds, st, dt, a, NV);
  hipLaunchKernelGGL((k_clear<NLM, true>), dim3(cx, NV), dim3(64), 0, st, dt, a, 0);
  hipLaunchKernelGGL((k_clear<NLM, false>), dim3(cx, NI), dim3(64), 0, st, dt, a, NV);
  hipLaunchKernelGGL(k_finalize, dim3(cx), dim3(64), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_plain(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  const int nl = a.nlay + 1;
  if (nl <= 25) return launch_plain_nlm<25>(dt, a, st);
  if (nl <= 37) return launch_plain_nlm<37>(dt, a, st);
  return launch_plain_nlm<MAX_NLAY + 1>(dt, a, st);
}

}  // namespace mgcm

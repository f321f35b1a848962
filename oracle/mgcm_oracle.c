/*
 * mgcm_oracle.c -- CPU FP64 restatement of the Legacy Mars GCM (v24) column
 * radiative transfer.  TEST INFRASTRUCTURE ONLY: the checker for the HIP
 * product path (see mgcm_oracle.h).  Nothing here is shipped or called by the
 * product library.
 *
 * Each routine follows the reference routine named in its comment, operation
 * by operation and in the same evaluation order, so that (built without FMA
 * contraction, -ffp-contract=off) it reproduces the reference Fortran to the
 * last bit on the same libm.  Single-precision literals of the reference are
 * written as (double)<x>f.  Arrays are 1-based (index 0 unused) to mirror the
 * Fortran indexing.
 */
#include "mgcm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NT OR_NT
#define NP OR_NP
#define NH OR_NH
#define NG OR_NG
#define NV OR_NV
#define NI OR_NI
#define MAXLEV 256 /* >= 2*nlay+4 for nlay <= 99 */

/* code/modules.f90:152, :230-232, :333, :364-365, :127-143 */
static const double CMK = 3.51e22;
static const double MAXEXP = 35.0;
static const double UBARI = 0.5;
static const double TLIMITS = 1.0e-3;
static const double TLIMITI = 5.0e-3;
static const double PI = 3.1415926535897932;
static const double GRAV = 3.72;
static const double CPAIR = 7.3594e2;
static const double SCALEP = 1.0e2;
static const double SECONDS_PER_SOL = 88775.0; /* code/newstep.f:132 */

/* Fortran index helpers (1-based). */
#define CO2(tab, nb, t, p, h, w, g) \
  (tab)[((t)-1) + NT * (((p)-1) + NP * (((h)-1) + NH * (((w)-1) + (nb) * ((g)-1))))]
#define AER(a, nlevp1, k, w) (a)[((k)-1) + (nlevp1) * ((w)-1)]
#define PLANCK(tab, w, n) (tab)[((w)-1) + NI * ((n)-1)]
#define GW(gw, g, w) (gw)[((g)-1) + NG * ((w)-1)]

static double fmin2(double a, double b) { return (a < b) ? a : b; } /* Fortran MIN for non-NaN */

/* ------------------------------------------------------------------------ */
/* TPINDEX: code/tpindex.f90:1-137 */
void or_tpindex(double pw, double tw, double qh2o, const double *pref, const double *tref,
                const double *wrefh2o, double lcoef[4], int *mt_out, int *mp_out,
                int *nh2o_out, double *wratio_out)
{
  int mt, mp, nh2o = 1, n;
  double u, t, pwl, wratio = 0.0;
  /* T cell, extrapolating at the ends (tpindex.f90:72-87) */
  if (tw <= tref[0]) {
    mt = 1;
  } else {
    mt = NT - 1;
    for (n = 1; n <= NT - 1; n++) {
      if (tw > tref[n - 1] && tw <= tref[n]) { mt = n; break; }
    }
  }
  u = (tw - tref[mt - 1]) / (tref[mt] - tref[mt - 1]);
  /* p cell in log10 (tpindex.f90:93-106) */
  pwl = log10(pw);
  mp = NP - 1;
  for (n = 2; n <= NP - 1; n++) {
    if (pwl <= pref[n - 1]) { mp = n - 1; break; }
  }
  t = (pwl - pref[mp - 1]) / (pref[mp] - pref[mp - 1]);
  /* tpindex.f90:110-113 */
  lcoef[0] = (1.0 - t) * (1.0 - u);
  lcoef[1] = t * (1.0 - u);
  lcoef[2] = t * u;
  lcoef[3] = (1.0 - t) * u;
  /* water cell, linear in vmr, clamped (tpindex.f90:118-132) */
  if (qh2o <= wrefh2o[0]) {
    nh2o = 1; wratio = 0.0;
  } else if (qh2o >= wrefh2o[NH - 1]) {
    nh2o = NH; wratio = 0.0;
  } else {
    for (n = 2; n <= NH; n++) {
      if (qh2o >= wrefh2o[n - 2] && qh2o < wrefh2o[n - 1]) {
        nh2o = n - 1;
        wratio = (qh2o - wrefh2o[n - 2]) / (wrefh2o[n - 1] - wrefh2o[n - 2]);
        break;
      }
    }
  }
  *mt_out = mt; *mp_out = mp; *nh2o_out = nh2o; *wratio_out = wratio;
}

/* k-coefficient at one (MT,MP,NH2O) cell with the H2O lerp and bilinear T-p
 * combine: code/optcv.f90:142-162 (same form in code/optci.f90:143-163).
 * At NH2O = L_REFH2O the reference reads one element past the H2O axis with
 * WRATIO = 0; the product k + 0*(x-k) equals k for finite x, so the upper
 * neighbour is clamped to the cell itself. */
static double kinterp(const double *tab, int nb, int w, int g, int mt, int mp, int nh,
                      double wratio, const double *lc)
{
  int nh1 = (nh < NH) ? nh + 1 : nh;
  double k1 = CO2(tab, nb, mt, mp, nh, w, g) +
              wratio * (CO2(tab, nb, mt, mp, nh1, w, g) - CO2(tab, nb, mt, mp, nh, w, g));
  double k2 = CO2(tab, nb, mt, mp + 1, nh, w, g) +
              wratio * (CO2(tab, nb, mt, mp + 1, nh1, w, g) - CO2(tab, nb, mt, mp + 1, nh, w, g));
  double k3 = CO2(tab, nb, mt + 1, mp + 1, nh, w, g) +
              wratio * (CO2(tab, nb, mt + 1, mp + 1, nh1, w, g) - CO2(tab, nb, mt + 1, mp + 1, nh, w, g));
  double k4 = CO2(tab, nb, mt + 1, mp, nh, w, g) +
              wratio * (CO2(tab, nb, mt + 1, mp, nh1, w, g) - CO2(tab, nb, mt + 1, mp, nh, w, g));
  return lc[0] * k1 + lc[1] * k2 + lc[2] * k3 + lc[3] * k4;
}

/* ------------------------------------------------------------------------ */
/* DTRIDGL: code/dtridgl.f90:1-41 (bottom-up elimination, top-down substitution) */
void or_dtridgl(int l, const double *af, const double *bf, const double *cf, const double *df,
                double *xk)
{
  double as[MAXLEV * 2], ds[MAXLEV * 2], x;
  int i;
  as[l] = af[l] / bf[l];
  ds[l] = df[l] / bf[l];
  for (i = 2; i <= l; i++) {
    x = 1.0 / (bf[l + 1 - i] - cf[l + 1 - i] * as[l + 2 - i]);
    as[l + 1 - i] = af[l + 1 - i] * x;
    ds[l + 1 - i] = (df[l + 1 - i] - cf[l + 1 - i] * ds[l + 2 - i]) * x;
  }
  xk[1] = ds[1];
  for (i = 2; i <= l; i++) xk[i] = ds[i] - as[i] * xk[i - 1];
}

/* DSOLVER: code/dsolver.f90:1-106 (mixed coefficients Y1=XK1+XK2, Y2=XK1-XK2) */
void or_dsolver(int nl, const double *gama, const double *cp, const double *cm,
                const double *cpm1, const double *cmm1, const double *e1, const double *e2,
                const double *e3, const double *e4, double btop, double bsurf, double rsf,
                double *xk1, double *xk2)
{
  double af[MAXLEV * 2], bf[MAXLEV * 2], cf[MAXLEV * 2], df[MAXLEV * 2], xk[MAXLEV * 2];
  int l = 2 * nl, i, n;
  af[1] = 0.0;
  bf[1] = gama[1] + 1.0;
  cf[1] = gama[1] - 1.0;
  df[1] = btop - cmm1[1];
  n = 0;
  for (i = 2; i <= l - 2; i += 2) { /* dsolver.f90:63-70 */
    n++;
    af[i] = (e1[n] + e3[n]) * (gama[n + 1] - 1.0);
    bf[i] = (e2[n] + e4[n]) * (gama[n + 1] - 1.0);
    cf[i] = 2.0 * (1.0 - gama[n + 1] * gama[n + 1]);
    df[i] = (gama[n + 1] - 1.0) * (cpm1[n + 1] - cp[n]) + (1.0 - gama[n + 1]) * (cm[n] - cmm1[n + 1]);
  }
  n = 0;
  for (i = 3; i <= l - 1; i += 2) { /* dsolver.f90:74-80 */
    n++;
    af[i] = 2.0 * (1.0 - gama[n] * gama[n]);
    bf[i] = (e1[n] - e3[n]) * (1.0 + gama[n + 1]);
    cf[i] = (e1[n] + e3[n]) * (gama[n + 1] - 1.0);
    df[i] = e3[n] * (cpm1[n + 1] - cp[n]) + e1[n] * (cm[n] - cmm1[n + 1]);
  }
  af[l] = e1[nl] - rsf * e3[nl];
  bf[l] = e2[nl] - rsf * e4[nl];
  cf[l] = 0.0;
  df[l] = bsurf - cp[nl] + rsf * cm[nl];
  or_dtridgl(l, af, bf, cf, df, xk);
  for (n = 1; n <= nl; n++) { /* dsolver.f90:91-103 */
    xk1[n] = xk[2 * n - 1] + xk[2 * n];
    xk2[n] = xk[2 * n - 1] - xk[2 * n];
    if (xk2[n] == 0.0) continue;
    if (fabs(xk2[n] / (xk[2 * n - 1] + (double)1.e-20f)) < (double)1.e-30f) xk2[n] = 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* Delta-Eddington scaling shared by GETDETAU (code/getdetau.f90:27-56) and
 * GFLUXV (code/gfluxv.f90:76-103).  Fills w0, cosbar, dtau (1..nl), tau
 * (1..nl+1), taucum (1..nlev) and returns the scaled surface optical depth. */
static double delta_eddington(int nl, const double *dtdel, const double *tdel,
                              const double *taucumin, const double *wdel, const double *cdel,
                              double *w0, double *cosbar, double *dtau, double *tau,
                              double *taucum)
{
  double factor = 1.0 - wdel[1] * (cdel[1] * cdel[1]);
  int l, k;
  tau[1] = tdel[1] * factor;
  taucum[1] = 0.0;
  taucum[2] = taucumin[2] * factor;
  taucum[3] = taucum[2] + (taucumin[3] - taucumin[2]) * factor;
  for (l = 1; l <= nl - 1; l++) {
    factor = 1.0 - wdel[l] * (cdel[l] * cdel[l]);
    w0[l] = wdel[l] * (1.0 - cdel[l] * cdel[l]) / factor;
    cosbar[l] = cdel[l] / (1.0 + cdel[l]);
    dtau[l] = dtdel[l] * factor;
    tau[l + 1] = tau[l] + dtau[l];
    k = 2 * (l + 1);
    taucum[k] = tau[l + 1];
    taucum[k + 1] = taucum[k] + (taucumin[k + 1] - taucumin[k]) * factor;
  }
  l = nl;
  factor = 1.0 - wdel[l] * (cdel[l] * cdel[l]);
  w0[l] = wdel[l] * (1.0 - cdel[l] * cdel[l]) / factor;
  cosbar[l] = cdel[l] / (1.0 + cdel[l]);
  dtau[l] = dtdel[l] * factor;
  tau[l + 1] = tau[l] + dtau[l];
  taucum[2 * l + 1] = tau[l + 1];
  return taucum[2 * l + 1];
}

/* GFLUXV: code/gfluxv.f90:1-283 for one (band, Gauss point) item. */
static void gfluxv(int nl, const double *dtdel, const double *tdel, const double *taucumin,
                   const double *wdel, const double *cdel, double ubar0, double f0pi, double rsf,
                   double btop, double bsurf, double *fmidp, double *fmidm, double *diffv,
                   double *fluxup, double *fluxdn, double *detau)
{
  double w0[MAXLEV], cosbar[MAXLEV], dtau[MAXLEV], tau[MAXLEV], taucum[MAXLEV];
  double lamda[MAXLEV], xk1[MAXLEV], xk2[MAXLEV], g1[MAXLEV], g2[MAXLEV], g3[MAXLEV];
  double gama[MAXLEV], cp[MAXLEV], cm[MAXLEV], cpm1[MAXLEV], cmm1[MAXLEV];
  double e1[MAXLEV], e2[MAXLEV], e3[MAXLEV], e4[MAXLEV];
  const double s3 = (double)sqrtf(3.0f);       /* SQRT(3.0) in REAL(4) */
  const double s3h = (double)(sqrtf(3.0f) * 0.5f); /* SQRT(3.0)*0.5 in REAL(4) */
  const int nlev = 2 * nl + 1;
  double g4, denom, am, ap, ep, em, exptrm, taumid, cpmid, cmmid;
  int l;

  *detau = delta_eddington(nl, dtdel, tdel, taucumin, wdel, cdel, w0, cosbar, dtau, tau, taucum);

  for (l = 1; l <= nl; l++) { /* gfluxv.f90:108-119 (ALPHA is unused) */
    g1[l] = s3h * (2.0 - w0[l] * (1.0 + cosbar[l]));
    g2[l] = (s3 * w0[l] * 0.5) * (1.0 - cosbar[l]);
    g3[l] = 0.5 * (1.0 - s3 * cosbar[l] * ubar0);
    lamda[l] = sqrt(g1[l] * g1[l] - g2[l] * g2[l]);
    gama[l] = (g1[l] - lamda[l]) / g2[l];
  }
  for (l = 1; l <= nl; l++) { /* gfluxv.f90:121-149 */
    g4 = 1.0 - g3[l];
    denom = lamda[l] * lamda[l] - 1.0 / (ubar0 * ubar0);
    if (denom == 0.0) denom = (double)1.e-10f;
    am = f0pi * w0[l] * (g4 * (g1[l] + 1.0 / ubar0) + g2[l] * g3[l]) / denom;
    ap = f0pi * w0[l] * (g3[l] * (g1[l] - 1.0 / ubar0) + g2[l] * g4) / denom;
    cpm1[l] = ap * exp(-fmin2(tau[l] / ubar0, MAXEXP));
    cmm1[l] = am * exp(-fmin2(tau[l] / ubar0, MAXEXP));
    cp[l] = ap * exp(-fmin2(tau[l + 1] / ubar0, MAXEXP));
    cm[l] = am * exp(-fmin2(tau[l + 1] / ubar0, MAXEXP));
  }
  for (l = 1; l <= nl; l++) { /* gfluxv.f90:154-163 */
    exptrm = fmin2(MAXEXP, lamda[l] * dtau[l]);
    ep = exp(exptrm);
    em = 1.0 / ep;
    e1[l] = ep + gama[l] * em;
    e2[l] = ep - gama[l] * em;
    e3[l] = gama[l] * ep + em;
    e4[l] = gama[l] * ep - em;
  }
  or_dsolver(nl, gama, cp, cm, cpm1, cmm1, e1, e2, e3, e4, btop, bsurf, rsf, xk1, xk2);

  for (l = 1; l <= nl - 1; l++) { /* gfluxv.f90:170-206 */
    exptrm = fmin2(MAXEXP, lamda[l] * (taucum[2 * l + 1] - taucum[2 * l]));
    ep = exp(exptrm);
    em = 1.0 / ep;
    g4 = 1.0 - g3[l];
    denom = lamda[l] * lamda[l] - 1.0 / (ubar0 * ubar0);
    if (denom == 0.0) denom = (double)1.e-10f;
    am = f0pi * w0[l] * (g4 * (g1[l] + 1.0 / ubar0) + g2[l] * g3[l]) / denom;
    ap = f0pi * w0[l] * (g3[l] * (g1[l] - 1.0 / ubar0) + g2[l] * g4) / denom;
    taumid = taucum[2 * l + 1];
    cpmid = ap * exp(-fmin2(taumid / ubar0, MAXEXP));
    cmmid = am * exp(-fmin2(taumid / ubar0, MAXEXP));
    fmidp[l] = xk1[l] * ep + gama[l] * xk2[l] * em + cpmid;
    fmidm[l] = xk1[l] * ep * gama[l] + xk2[l] * em + cmmid;
    fmidm[l] = fmidm[l] + ubar0 * f0pi * exp(-fmin2(taumid / ubar0, MAXEXP));
  }
  /* top of the atmosphere (gfluxv.f90:210-238) */
  ep = 1.0;
  em = 1.0;
  g4 = 1.0 - g3[1];
  denom = lamda[1] * lamda[1] - 1.0 / (ubar0 * ubar0);
  if (denom == 0.0) denom = (double)1.e-10f;
  am = f0pi * w0[1] * (g4 * (g1[1] + 1.0 / ubar0) + g2[1] * g3[1]) / denom;
  ap = f0pi * w0[1] * (g3[1] * (g1[1] - 1.0 / ubar0) + g2[1] * g4) / denom;
  cpmid = ap;
  cmmid = am;
  *fluxup = xk1[1] * ep + gama[1] * xk2[1] * em + cpmid;
  *fluxdn = xk1[1] * ep * gama[1] + xk2[1] * em + cmmid;
  *fluxdn = *fluxdn + ubar0 * f0pi * exp(-fmin2(taucum[1] / ubar0, MAXEXP));
  /* bottom layer uses the full sub-layer (gfluxv.f90:243-280) */
  l = nl;
  exptrm = fmin2(MAXEXP, lamda[l] * (taucum[nlev] - taucum[nlev - 1]));
  ep = exp(exptrm);
  em = 1.0 / ep;
  g4 = 1.0 - g3[l];
  denom = lamda[l] * lamda[l] - 1.0 / (ubar0 * ubar0);
  if (denom == 0.0) denom = (double)1.e-10f;
  am = f0pi * w0[l] * (g4 * (g1[l] + 1.0 / ubar0) + g2[l] * g3[l]) / denom;
  ap = f0pi * w0[l] * (g3[l] * (g1[l] - 1.0 / ubar0) + g2[l] * g4) / denom;
  taumid = fmin2(taucum[nlev], MAXEXP);
  cpmid = ap * exp(-fmin2(taumid / ubar0, MAXEXP));
  cmmid = am * exp(-fmin2(taumid / ubar0, MAXEXP));
  fmidp[l] = xk1[l] * ep + gama[l] * xk2[l] * em + cpmid;
  fmidm[l] = xk1[l] * ep * gama[l] + xk2[l] * em + cmmid;
  *diffv = fmidm[l];
  fmidm[l] = fmidm[l] + ubar0 * f0pi * exp(-fmin2(taumid / ubar0, MAXEXP));
}

/* GFLUXI: code/gfluxi.f90:1-197 for one (band, Gauss point) item. */
static void gfluxi(int nl, const double *tlev, const double *planckir, int nw,
                   const double *dtau, const double *taucum, const double *w0,
                   const double *cosbar, double ubari, double rsf, double btop, double bsurf,
                   double *ftopup, double *fmidp, double *fmidm)
{
  double b0[MAXLEV], b1[MAXLEV], alpha[MAXLEV], lamda[MAXLEV], xk1[MAXLEV], xk2[MAXLEV];
  double gama[MAXLEV], cp[MAXLEV], cm[MAXLEV], cpm1[MAXLEV], cmm1[MAXLEV];
  double e1[MAXLEV], e2[MAXLEV], e3[MAXLEV], e4[MAXLEV];
  double term, ep, em, dtauk, cpmid, cmmid, fluxup, fluxdn;
  int l, nt, nt2;

  for (l = 1; l <= nl - 1; l++) { /* gfluxi.f90:69-78 */
    alpha[l] = sqrt((1.0 - w0[l]) / (1.0 - w0[l] * cosbar[l]));
    lamda[l] = alpha[l] * (1.0 - w0[l] * cosbar[l]) / ubari;
    nt2 = (int)(tlev[2 * l + 2] * 10.0 - 499);
    nt = (int)(tlev[2 * l] * 10.0 - 499);
    b1[l] = (PLANCK(planckir, nw, nt2) - PLANCK(planckir, nw, nt)) / dtau[l];
    b0[l] = PLANCK(planckir, nw, nt);
  }
  l = nl; /* gfluxi.f90:82-89 */
  alpha[l] = sqrt((1.0 - w0[l]) / (1.0 - w0[l] * cosbar[l]));
  lamda[l] = alpha[l] * (1.0 - w0[l] * cosbar[l]) / ubari;
  nt = (int)(tlev[2 * l + 1] * 10.0 - 499);
  nt2 = (int)(tlev[2 * l] * 10.0 - 499);
  b1[l] = (PLANCK(planckir, nw, nt) - PLANCK(planckir, nw, nt2)) / dtau[l];
  b0[l] = PLANCK(planckir, nw, nt2);

  for (l = 1; l <= nl; l++) { /* gfluxi.f90:91-106 */
    gama[l] = (1.0 - alpha[l]) / (1.0 + alpha[l]);
    term = ubari / (1.0 - w0[l] * cosbar[l]);
    cp[l] = b0[l] + b1[l] * dtau[l] + b1[l] * term;
    cm[l] = b0[l] + b1[l] * dtau[l] - b1[l] * term;
    cpm1[l] = b0[l] + b1[l] * term;
    cmm1[l] = b0[l] - b1[l] * term;
  }
  for (l = 1; l <= nl; l++) { /* gfluxi.f90:113-123 */
    ep = exp(fmin2(lamda[l] * dtau[l], MAXEXP));
    em = 1.0 / ep;
    e1[l] = ep + gama[l] * em;
    e2[l] = ep - gama[l] * em;
    e3[l] = gama[l] * ep + em;
    e4[l] = gama[l] * ep - em;
  }
  or_dsolver(nl, gama, cp, cm, cpm1, cmm1, e1, e2, e3, e4, btop, bsurf, rsf, xk1, xk2);

  for (l = 1; l <= nl - 1; l++) { /* gfluxi.f90:136-154 */
    dtauk = taucum[2 * l + 1] - taucum[2 * l];
    ep = exp(fmin2(lamda[l] * dtauk, MAXEXP));
    em = 1.0 / ep;
    term = ubari / (1. - w0[l] * cosbar[l]);
    cpmid = b0[l] + b1[l] * dtauk + b1[l] * term;
    cmmid = b0[l] + b1[l] * dtauk - b1[l] * term;
    fmidp[l] = xk1[l] * ep + gama[l] * xk2[l] * em + cpmid;
    fmidm[l] = xk1[l] * ep * gama[l] + xk2[l] * em + cmmid;
    fmidp[l] = fmidp[l] * PI;
    fmidm[l] = fmidm[l] * PI;
  }
  l = nl; /* gfluxi.f90:158-175 */
  ep = exp(fmin2(lamda[l] * dtau[l], MAXEXP));
  em = 1.0 / ep;
  term = ubari / (1. - w0[l] * cosbar[l]);
  cpmid = b0[l] + b1[l] * dtau[l] + b1[l] * term;
  cmmid = b0[l] + b1[l] * dtau[l] - b1[l] * term;
  fmidp[l] = xk1[l] * ep + gama[l] * xk2[l] * em + cpmid;
  fmidm[l] = xk1[l] * ep * gama[l] + xk2[l] * em + cmmid;
  fmidp[l] = fmidp[l] * PI;
  fmidm[l] = fmidm[l] * PI;
  /* top level (gfluxi.f90:179-194) */
  ep = 1.0;
  em = 1.0;
  term = ubari / (1.0 - w0[1] * cosbar[1]);
  cpmid = b0[1] + b1[1] * term;
  cmmid = b0[1] - b1[1] * term;
  fluxup = xk1[1] * ep + gama[1] * xk2[1] * em + cpmid;
  fluxdn = xk1[1] * ep * gama[1] + xk2[1] * em + cmmid;
  *ftopup = (fluxup - fluxdn) * PI;
}

/* ------------------------------------------------------------------------ */
typedef struct {
  /* 1-based scratch for one column */
  double dtaukv[NV + 1][NG + 1][MAXLEV];
  double dtauv[NV + 1][NG + 1][MAXLEV], tauv[NV + 1][NG + 1][MAXLEV];
  double taucumv[NV + 1][NG + 1][MAXLEV], wbarv[NV + 1][NG + 1][MAXLEV];
  double cosbv[NV + 1][NG + 1][MAXLEV];
  double taugsurf[NV + 1][NG + 1];
  int ran[2][NV + 1][NG + 1]; /* executed items, for the op counter */
} col_work;

/* OPTCV (code/optcv.f90:1-284) and OPTCI (code/optci.f90:1-279).  vis=1: VIS
 * with Rayleigh and the 0.9999 Rayleigh albedo; vis=0: IR with the 1e-9 floor. */
static void optics(const or_tables *t, const or_column *c, int vis, col_work *w)
{
  const int nlay = t->nlay, nlev = 2 * nlay + 3, nlayrad = nlay + 1, nlp1 = nlev + 1;
  const int nb = vis ? NV : NI;
  const double *tab = vis ? t->co2v : t->co2i;
  const double *qx = vis ? t->qxvdst : t->qxidst, *qs = vis ? t->qsvdst : t->qsidst;
  const double *gd = vis ? t->gvdst : t->gidst;
  const double *qxc = vis ? t->qxvcld : t->qxicld, *qsc = vis ? t->qsvcld : t->qsicld;
  const double *gc = vis ? t->gvcld : t->gicld;
  static const double f9999 = (double)0.9999f;
  double dpr[MAXLEV], u[MAXLEV], lk[MAXLEV][4], wr[MAXLEV], tauref[MAXLEV], taurefcld[MAXLEV];
  double tray[MAXLEV][NV + 1], taeros[MAXLEV][NV + 1], tcloud[MAXLEV][NV + 1];
  double taureflk[MAXLEV][NV + 1], taucldk[MAXLEV][NV + 1];
  int mt[MAXLEV], mp[MAXLEV], nh[MAXLEV];
  int k, nw, ng, l;

  for (k = 1; k <= nlp1; k++) { tauref[k] = c->tauref[k - 1]; taurefcld[k] = c->taurefcld[k - 1]; }
  for (ng = 1; ng <= NG - 1; ng++)
    for (nw = 1; nw <= nb; nw++) w->taugsurf[nw][ng] = 0.0;

  for (k = 2; k <= nlev; k++) { /* optcv.f90:103-124, optci.f90:113-131 */
    double qref;
    dpr[k] = c->plev[k - 1] - c->plev[k - 2];
    u[k] = CMK * dpr[k];
    or_tpindex(c->pmid[k - 1], c->tmid[k - 1], c->qh2o[k - 1], t->pfgasref, t->tgasref,
               t->wrefh2o, lk[k], &mt[k], &mp[k], &nh[k], &wr[k]);
    qref = vis ? AER(qx, nlp1, k, 6) : t->qextrefdst[k - 1]; /* L_NREFV = 6 */
    tauref[k] = tauref[k] / qref;
    taurefcld[k] = taurefcld[k] / t->qextrefcld[k - 1];
    for (nw = 1; nw <= nb; nw++) {
      tray[k][nw] = vis ? t->tauray[nw - 1] * dpr[k] : 0.0;
      taeros[k][nw] = tauref[k] * AER(qx, nlp1, k, nw);
      tcloud[k][nw] = taurefcld[k] * AER(qxc, nlp1, k, nw);
    }
  }
  for (k = 2; k <= nlev; k++) { /* optcv.f90:128-176, optci.f90:133-176 */
    for (nw = 1; nw <= nb; nw++) {
      double trayaer = tray[k][nw] + taeros[k][nw] + tcloud[k][nw];
      for (ng = 1; ng <= NG - 1; ng++) {
        double ans = kinterp(tab, nb, nw, ng, mt[k], mp[k], nh[k], wr[k], lk[k]);
        double taugas = u[k] * ans;
        w->taugsurf[nw][ng] = w->taugsurf[nw][ng] + taugas;
        w->dtaukv[nw][ng][k] = vis ? taugas + trayaer : taugas + taeros[k][nw] + tcloud[k][nw];
      }
      w->dtaukv[nw][NG][k] = vis ? taeros[k][nw] + tray[k][nw] + tcloud[k][nw]
                                 : taeros[k][nw] + tcloud[k][nw];
    }
  }
  if (!vis) {
    for (nw = 1; nw <= nb; nw++) for (ng = 1; ng <= NG; ng++) w->dtaukv[nw][ng][nlp1] = 0.0;
  }
  for (nw = 1; nw <= nb; nw++) { /* optcv.f90:182-187, optci.f90:182-187 */
    for (k = 2; k <= nlev; k++) {
      taureflk[k][nw] = tauref[k] * AER(qs, nlp1, k, nw);
      taucldk[k][nw] = taurefcld[k] * AER(qsc, nlp1, k, nw);
    }
    if (!vis) { /* IR also fills the padding level L_LEVELS+1 */
      taureflk[nlp1][nw] = tauref[nlp1] * AER(qs, nlp1, nlp1, nw);
      taucldk[nlp1][nw] = taurefcld[nlp1] * AER(qsc, nlp1, nlp1, nw);
    }
  }
  for (nw = 1; nw <= nb; nw++) {
    if (vis) { /* optcv.f90:189-247 */
      for (l = 1; l <= nlay; l++) {
        k = 2 * l + 1;
        w->dtauv[nw][NG][l] = w->dtaukv[nw][NG][k] + w->dtaukv[nw][NG][k + 1];
        w->cosbv[nw][NG][l] =
            (AER(gd, nlp1, k, nw) * taureflk[k][nw] + AER(gd, nlp1, k + 1, nw) * taureflk[k + 1][nw] +
             AER(gc, nlp1, k, nw) * taucldk[k][nw] + AER(gc, nlp1, k + 1, nw) * taucldk[k + 1][nw]) /
            (tray[k][nw] + tray[k + 1][nw] + taureflk[k][nw] + taureflk[k + 1][nw] +
             taucldk[k][nw] + taucldk[k + 1][nw]);
        w->wbarv[nw][NG][l] = (taureflk[k][nw] + taureflk[k + 1][nw] + taucldk[k][nw] +
                               taucldk[k + 1][nw] + (tray[k][nw] + tray[k + 1][nw]) * f9999) /
                              w->dtauv[nw][NG][l];
      }
      l = nlayrad;
      k = 2 * l + 1;
      w->dtauv[nw][NG][l] = w->dtaukv[nw][NG][k];
      w->cosbv[nw][NG][l] = (AER(gd, nlp1, k, nw) * taureflk[k][nw] + AER(gc, nlp1, k, nw) * taucldk[k][nw]) /
                            (tray[k][nw] + taureflk[k][nw] + taucldk[k][nw]);
      w->wbarv[nw][NG][l] = (taureflk[k][nw] + taucldk[k][nw] + tray[k][nw] * f9999) / w->dtauv[nw][NG][l];
      for (ng = 1; ng <= NG - 1; ng++) {
        for (l = 1; l <= nlay; l++) {
          k = 2 * l + 1;
          w->dtauv[nw][ng][l] = w->dtaukv[nw][ng][k] + w->dtaukv[nw][ng][k + 1];
          w->cosbv[nw][ng][l] = w->cosbv[nw][NG][l];
          w->wbarv[nw][ng][l] = (taureflk[k][nw] + taureflk[k + 1][nw] + taucldk[k][nw] +
                                 taucldk[k + 1][nw] + (tray[k][nw] + tray[k + 1][nw]) * f9999) /
                                w->dtauv[nw][ng][l];
        }
        l = nlayrad;
        k = 2 * l + 1;
        w->dtauv[nw][ng][l] = w->dtaukv[nw][ng][k];
        w->cosbv[nw][ng][l] = w->cosbv[nw][NG][l];
        w->wbarv[nw][ng][l] = (taureflk[k][nw] + taucldk[k][nw] + tray[k][nw] * f9999) / w->dtauv[nw][ng][l];
      }
    } else { /* optci.f90:189-242 */
      const double f1em9 = (double)1.0e-9f;
      for (l = 1; l <= nlayrad; l++) {
        double tauac;
        k = 2 * l + 1;
        w->dtauv[nw][NG][l] = w->dtaukv[nw][NG][k] + w->dtaukv[nw][NG][k + 1] + 1.e-50;
        if (w->dtauv[nw][NG][l] > f1em9) {
          w->wbarv[nw][NG][l] = (taureflk[k][nw] + taureflk[k + 1][nw] + taucldk[k][nw] + taucldk[k + 1][nw]) /
                                w->dtauv[nw][NG][l];
        } else {
          w->wbarv[nw][NG][l] = 0.0;
          w->dtauv[nw][NG][l] = f1em9;
        }
        tauac = taureflk[k][nw] + taureflk[k + 1][nw] + taucldk[k][nw] + taucldk[k + 1][nw];
        if (tauac > 0.0) {
          w->cosbv[nw][NG][l] =
              (AER(gd, nlp1, k, nw) * taureflk[k][nw] + AER(gd, nlp1, k + 1, nw) * taureflk[k + 1][nw] +
               AER(gc, nlp1, k, nw) * taucldk[k][nw] + AER(gc, nlp1, k + 1, nw) * taucldk[k + 1][nw]) /
              (taureflk[k][nw] + taureflk[k + 1][nw] + taucldk[k][nw] + taucldk[k + 1][nw]);
        } else {
          w->cosbv[nw][NG][l] = 0.0;
        }
      }
      for (ng = 1; ng <= NG - 1; ng++) {
        for (l = 1; l <= nlayrad; l++) {
          k = 2 * l + 1;
          w->dtauv[nw][ng][l] = w->dtaukv[nw][ng][k] + w->dtaukv[nw][ng][k + 1] + 1.e-50;
          if (w->dtauv[nw][ng][l] > f1em9) {
            w->wbarv[nw][ng][l] = (taureflk[k][nw] + taureflk[k + 1][nw] + taucldk[k][nw] + taucldk[k + 1][nw]) /
                                  w->dtauv[nw][ng][l];
          } else {
            w->wbarv[nw][ng][l] = 0.0;
            w->dtauv[nw][ng][l] = f1em9;
          }
          w->cosbv[nw][ng][l] = w->cosbv[nw][NG][l];
        }
      }
    }
  }
  /* prefix sums (optcv.f90:251-274, optci.f90:246-269) */
  for (nw = 1; nw <= nb; nw++) {
    for (ng = 1; ng <= NG; ng++) {
      w->tauv[nw][ng][1] = 0.0;
      for (l = 1; l <= nlayrad; l++) w->tauv[nw][ng][l + 1] = w->tauv[nw][ng][l] + w->dtauv[nw][ng][l];
      w->taucumv[nw][ng][1] = 0.0;
      for (k = 2; k <= nlev; k++) w->taucumv[nw][ng][k] = w->taucumv[nw][ng][k - 1] + w->dtaukv[nw][ng][k];
    }
  }
}

/* SFLUXV (code/sfluxv.f90:1-169) with per-band Gauss weights GW(g,band).  With
 * the same weights in every band this is the reference's single GWEIGHT. */
static void sfluxv(const or_tables *t, const double *sol, const or_column *c, col_work *w,
                   double *fmnetv, double *fluxupv, double *fluxdnv, double *nfluxtopv,
                   double *diffvt, double *detau)
{
  const int nl = t->nlay + 1;
  const double ubar0 = c->acosz, rsfv = c->als;
  const double f099 = (double)0.99f;
  double fmupv[MAXLEV], fmdv[MAXLEV], fluxup, fluxdn, diffv, eterm, bsurf, btop, fzero, f0pi;
  double scratch_w0[MAXLEV], scratch_c[MAXLEV], scratch_d[MAXLEV], scratch_t[MAXLEV], scratch_tc[MAXLEV];
  int l, nw, ng;

  *nfluxtopv = 0.0;
  for (l = 1; l <= nl; l++) { fmnetv[l] = 0.0; fluxupv[l] = 0.0; fluxdnv[l] = 0.0; }
  *diffvt = 0.0;
  for (nw = 1; nw <= NV; nw++) {
    f0pi = sol[nw - 1];
    fzero = t->fzerov[nw - 1];
    if (!(fzero >= f099)) {
      for (ng = 1; ng <= NG - 1; ng++) {
        const double gw = GW(t->gwv, ng, nw);
        /* GETDETAU (code/getdetau.f90) */
        detau[(nw - 1) + NV * (ng - 1)] =
            delta_eddington(nl, w->dtauv[nw][ng], w->tauv[nw][ng], w->taucumv[nw][ng], w->wbarv[nw][ng],
                            w->cosbv[nw][ng], scratch_w0, scratch_c, scratch_d, scratch_t, scratch_tc);
        if (w->taugsurf[nw][ng] < TLIMITS) {
          fzero = fzero + (1.0 - t->fzerov[nw - 1]) * gw;
          continue;
        }
        btop = 0.0;
        eterm = fmin2(detau[(nw - 1) + NV * (ng - 1)] / ubar0, MAXEXP);
        bsurf = rsfv * ubar0 * sol[nw - 1] * exp(-eterm);
        gfluxv(nl, w->dtauv[nw][ng], w->tauv[nw][ng], w->taucumv[nw][ng], w->wbarv[nw][ng],
               w->cosbv[nw][ng], ubar0, f0pi, rsfv, btop, bsurf, fmupv, fmdv, &diffv, &fluxup, &fluxdn,
               &detau[(nw - 1) + NV * (ng - 1)]);
        w->ran[1][nw][ng] = 1;
        *nfluxtopv = *nfluxtopv + (fluxup - fluxdn) * gw * (1.0 - t->fzerov[nw - 1]);
        for (l = 1; l <= nl; l++) {
          fmnetv[l] = fmnetv[l] + (fmupv[l] - fmdv[l]) * gw * (1.0 - t->fzerov[nw - 1]);
          fluxupv[l] = fluxupv[l] + fmupv[l] * gw * (1.0 - t->fzerov[nw - 1]);
          fluxdnv[l] = fluxdnv[l] + fmdv[l] * gw * (1.0 - t->fzerov[nw - 1]);
        }
        *diffvt = *diffvt + diffv * gw * (1.0 - t->fzerov[nw - 1]);
      }
    }
    /* 17th Gauss point: continuum only (sfluxv.f90:121-162) */
    ng = NG;
    btop = 0.0;
    detau[(nw - 1) + NV * (ng - 1)] =
        delta_eddington(nl, w->dtauv[nw][ng], w->tauv[nw][ng], w->taucumv[nw][ng], w->wbarv[nw][ng],
                        w->cosbv[nw][ng], scratch_w0, scratch_c, scratch_d, scratch_t, scratch_tc);
    eterm = fmin2(detau[(nw - 1) + NV * (ng - 1)] / ubar0, MAXEXP);
    bsurf = rsfv * ubar0 * sol[nw - 1] * exp(-eterm);
    gfluxv(nl, w->dtauv[nw][ng], w->tauv[nw][ng], w->taucumv[nw][ng], w->wbarv[nw][ng],
           w->cosbv[nw][ng], ubar0, f0pi, rsfv, btop, bsurf, fmupv, fmdv, &diffv, &fluxup, &fluxdn,
           &detau[(nw - 1) + NV * (ng - 1)]);
    w->ran[1][nw][ng] = 1;
    *nfluxtopv = *nfluxtopv + (fluxup - fluxdn) * fzero;
    for (l = 1; l <= nl; l++) {
      fmnetv[l] = fmnetv[l] + (fmupv[l] - fmdv[l]) * fzero;
      fluxupv[l] = fluxupv[l] + fmupv[l] * fzero;
      fluxdnv[l] = fluxdnv[l] + fmdv[l] * fzero;
    }
    *diffvt = *diffvt + diffv * fzero;
  }
}

/* SFLUXI (code/sfluxi.f90:1-150) with per-band Gauss weights. */
static void sfluxi(const or_tables *t, const or_column *c, col_work *w, double *fmneti,
                   double *fluxupi, double *fluxdni, double *nfluxtopi)
{
  const int nl = t->nlay + 1, nlev = 2 * t->nlay + 3;
  const double rsfi = c->albi, f099 = (double)0.99f;
  const double *tl1 = c->tlev - 1; /* 1-based TLEV */
  const double *pl1 = c->plev - 1;
  double fmupi[MAXLEV], fmdi[MAXLEV], ftopup, tautop, btop, bsurf, pltop, fzero;
  int l, nw, ng, nts, ntt;

  *nfluxtopi = 0.0;
  for (l = 1; l <= nl; l++) { fmneti[l] = 0.0; fluxupi[l] = 0.0; fluxdni[l] = 0.0; }
  nts = (int)(tl1[nlev] * 10.0 - 499);
  ntt = (int)(tl1[2] * 10.0 - 499);
  for (nw = 1; nw <= NI; nw++) {
    const double dw = t->dwni[nw - 1];
    bsurf = (1. - rsfi) * PLANCK(t->planckir, nw, nts);
    pltop = PLANCK(t->planckir, nw, ntt);
    fzero = t->fzeroi[nw - 1];
    if (!(fzero >= f099)) {
      for (ng = 1; ng <= NG - 1; ng++) {
        const double gw = GW(t->gwi, ng, nw);
        if (w->taugsurf[nw][ng] < TLIMITI) {
          fzero = fzero + (1.0 - t->fzeroi[nw - 1]) * gw;
          continue;
        }
        tautop = w->dtauv[nw][ng][1] * pl1[2] / (pl1[4] - pl1[2]);
        btop = (1.0 - exp(-tautop / UBARI)) * pltop;
        gfluxi(nl, tl1, t->planckir, nw, w->dtauv[nw][ng], w->taucumv[nw][ng], w->wbarv[nw][ng],
               w->cosbv[nw][ng], UBARI, rsfi, btop, bsurf, &ftopup, fmupi, fmdi);
        w->ran[0][nw][ng] = 1;
        *nfluxtopi = *nfluxtopi + ftopup * dw * gw * (1.0 - t->fzeroi[nw - 1]);
        for (l = 1; l <= nl; l++) {
          fmneti[l] = fmneti[l] + (fmupi[l] - fmdi[l]) * dw * gw * (1.0 - t->fzeroi[nw - 1]);
          fluxupi[l] = fluxupi[l] + fmupi[l] * dw * gw * (1.0 - t->fzeroi[nw - 1]);
          fluxdni[l] = fluxdni[l] + fmdi[l] * dw * gw * (1.0 - t->fzeroi[nw - 1]);
        }
      }
    }
    ng = NG; /* sfluxi.f90:114-143 */
    tautop = w->dtauv[nw][ng][1] * pl1[2] / (pl1[4] - pl1[2]);
    btop = (1.0 - exp(-tautop / UBARI)) * pltop;
    gfluxi(nl, tl1, t->planckir, nw, w->dtauv[nw][ng], w->taucumv[nw][ng], w->wbarv[nw][ng],
           w->cosbv[nw][ng], UBARI, rsfi, btop, bsurf, &ftopup, fmupi, fmdi);
    w->ran[0][nw][ng] = 1;
    *nfluxtopi = *nfluxtopi + ftopup * dw * fzero;
    for (l = 1; l <= nl; l++) {
      fmneti[l] = fmneti[l] + (fmupi[l] - fmdi[l]) * dw * fzero;
      fluxupi[l] = fluxupi[l] + fmupi[l] * dw * fzero;
      fluxdni[l] = fluxdni[l] + fmdi[l] * dw * fzero;
    }
  }
}

/* One column: COMP3's radiation block (code/comp3.f:654-738) plus the layer
 * heating rate in K/sol from the net-flux differences SUNTOT/IRTOTAL
 * (code/comp3.f:677-681,735-738) over the layer mass (code/comp3.f:478-484),
 * K/s x 88775 (code/newstep.f:132).  The non-LTE factor and the theta
 * conversion of code/comp3.f:765-771 stay with the caller. */
static void column_run_work(const or_tables *t, const double *sol, const or_column *c, or_out *o,
                            col_work *w)
{
  const int nlay = t->nlay, nl = nlay + 1;
  double fmnetv[MAXLEV], fluxupv[MAXLEV], fluxdnv[MAXLEV], fmneti[MAXLEV], fluxupi[MAXLEV],
      fluxdni[MAXLEV], suntot[MAXLEV], irtot[MAXLEV];
  double nfluxtopv = 0.0, nfluxtopi = 0.0, diffvt = 0.0;
  int l, nw, ng;
  const double sun_up = (double)1.0e-5f; /* code/comp3.f:637 */

  memset(w->ran, 0, sizeof(w->ran));
  if (c->acosz >= sun_up) {
    optics(t, c, 1, w);
    for (nw = 1; nw <= NV; nw++)
      for (ng = 1; ng <= NG; ng++)
        o->cumtauv[(nw - 1) + NV * (ng - 1)] = w->taucumv[nw][ng][2 * nlay + 3];
    sfluxv(t, sol, c, w, fmnetv, fluxupv, fluxdnv, &nfluxtopv, &diffvt, o->detau);
    suntot[1] = fmnetv[1] - nfluxtopv;
    for (l = 2; l <= nl; l++) suntot[l] = fmnetv[l] - fmnetv[l - 1];
  } else {
    for (l = 1; l <= nl; l++) { fmnetv[l] = 0.0; fluxupv[l] = 0.0; fluxdnv[l] = 0.0; suntot[l] = 0.0; }
    nfluxtopv = 0.0;
    diffvt = 0.0;
  }
  optics(t, c, 0, w);
  sfluxi(t, c, w, fmneti, fluxupi, fluxdni, &nfluxtopi);
  irtot[1] = fmneti[1] - nfluxtopi;
  for (l = 2; l <= nl; l++) irtot[l] = fmneti[l] - fmneti[l - 1];

  for (l = 1; l <= nl; l++) {
    o->fmnetv[l - 1] = fmnetv[l]; o->fluxupv[l - 1] = fluxupv[l]; o->fluxdnv[l - 1] = fluxdnv[l];
    o->fmneti[l - 1] = fmneti[l]; o->fluxupi[l - 1] = fluxupi[l]; o->fluxdni[l - 1] = fluxdni[l];
  }
  *o->nfluxtopv = nfluxtopv;
  *o->nfluxtopi = nfluxtopi;
  *o->diffvt = diffvt;
  /* heating: hr[0] = stratosphere (mass PTROP*100/g, PTROP = 2*PLEV(1));
   * hr[j] = GCM layer j, mass (PLEV(2j+3)-PLEV(2j+1))*100/g */
  for (l = 0; l <= nlay; l++) {
    const double dp = (l == 0) ? (c->plev[0] + c->plev[1]) : (c->plev[2 * l + 2] - c->plev[2 * l]);
    const double df = suntot[l + 1] + irtot[l + 1];
    o->hr[l] = ((df * GRAV) / ((CPAIR * SCALEP) * dp)) * SECONDS_PER_SOL;
  }
}

void or_column_run(const or_tables *t, const double *sol, const or_column *c, or_out *o)
{
  col_work *w = (col_work *)malloc(sizeof(col_work));
  column_run_work(t, sol, c, o, w);
  free(w);
}

static void batch_column(const or_batch_in *in, const or_batch_out *out, int nlay, int i,
                         or_column *c, or_out *o)
{
  const int nlev = 2 * nlay + 3, nl = nlay + 1;
  c->plev = in->plev + (size_t)i * nlev;
  c->tlev = in->tlev + (size_t)i * nlev;
  c->pmid = in->pmid + (size_t)i * nlev;
  c->tmid = in->tmid + (size_t)i * nlev;
  c->qh2o = in->qh2o + (size_t)i * nlev;
  c->tauref = in->tauref + (size_t)i * (nlev + 1);
  c->taurefcld = in->taurefcld + (size_t)i * (nlev + 1);
  c->als = in->als[i];
  c->albi = in->albi[i];
  c->acosz = in->acosz[i];
  o->fmnetv = out->fmnetv + (size_t)i * nl;
  o->fluxupv = out->fluxupv + (size_t)i * nl;
  o->fluxdnv = out->fluxdnv + (size_t)i * nl;
  o->fmneti = out->fmneti + (size_t)i * nl;
  o->fluxupi = out->fluxupi + (size_t)i * nl;
  o->fluxdni = out->fluxdni + (size_t)i * nl;
  o->nfluxtopv = out->nfluxtopv + i;
  o->nfluxtopi = out->nfluxtopi + i;
  o->diffvt = out->diffvt + i;
  o->detau = out->detau + (size_t)i * NV * NG;
  o->cumtauv = out->cumtauv + (size_t)i * NV * NG;
  o->hr = out->hr + (size_t)i * (nlay + 1);
}

/* The tables of column i: its own aerosol optics where the batch carries them
 * (OPT_DST/OPT_CLD, code/comp3.f:595-631), else the shared ones. */
static void column_tables(const or_tables *t, const or_batch_in *in, int i, or_tables *tc)
{
  const size_t ld = (size_t)(2 * t->nlay + 4);
  *tc = *t;
#define AERO(f, nb) if (in->f) tc->f = in->f + (size_t)i * (nb) * ld
  AERO(qxvdst, NV); AERO(qsvdst, NV); AERO(gvdst, NV); AERO(qxvcld, NV); AERO(qsvcld, NV); AERO(gvcld, NV);
  AERO(qxidst, NI); AERO(qsidst, NI); AERO(gidst, NI); AERO(qxicld, NI); AERO(qsicld, NI); AERO(gicld, NI);
  AERO(qextrefdst, 1); AERO(qextrefcld, 1);
#undef AERO
}

void or_batch_run(const or_tables *t, const double *sol, const or_batch_in *in,
                  or_batch_out *out, int nthreads)
{
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
  {
    col_work *w = (col_work *)malloc(sizeof(col_work));
    int i;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 8)
#endif
    for (i = 0; i < in->ncol; i++) {
      or_column c;
      or_out o;
      or_tables tc;
      batch_column(in, out, t->nlay, i, &c, &o);
      column_tables(t, in, i, &tc);
      column_run_work(&tc, sol, &c, &o, w);
    }
    free(w);
  }
  (void)nthreads;
}

/* ------------------------------------------------------------------------ */
/* Executed-work counter.  Counts are read off the routines above: FP64
 * add/sub/mul/div per executed statement (comparisons, MIN and integer work
 * are not counted), with exp/sqrt/log10 tallied separately.  nl = NLAYRAD. */
static void count_column(const or_tables *t, const or_column *c, col_work *w, or_opcount *n)
{
  const int nlay = t->nlay, nl = nlay + 1, nlev = 2 * nlay + 3, nsub = nlev - 1;
  int nw, ng;
  /* per sub-level, per optics call: dpr 1, u 1, tpindex (u 2 + t 3 + lcoef 8 + wratio 3 worst) 16,
   * tauref/cld 2 divs, per band tray 1 + taeros 1 + tcloud 1, per (band, g<17) kinterp 4*3+7=19
   * + taugas 1 + taugsurf 1 + dtaukv 1 (+1 IR), continuum 1-2 */
  for (int vis = 1; vis >= 0; vis--) {
    const int nb = vis ? NV : NI;
    if (vis && !(c->acosz >= (double)1.0e-5f)) continue;
    n->logs += nsub;
    n->flops += nsub * (2 + 16 + 2 + 3.0 * nb + nb * (NG - 1) * (19 + 3 + (vis ? 0 : 1)) + nb * 2.0);
    /* scattering tau per band per sub-level 2; layer dtau/wbar/cosb per (band,g,layer) ~
     * dtau 1(+1 IR) + wbar 5 numer + 1 div (+ cosb 13 at g17); prefix sums nl + nlev */
    n->flops += nb * nsub * 2.0 + nb * NG * nl * (7.0 + (vis ? 1 : 0)) + nb * nl * 13.0 +
                nb * NG * (nl + nlev);
  }
  for (nw = 1; nw <= NV; nw++) {
    for (ng = 1; ng <= NG; ng++) {
      if (!(c->acosz >= (double)1.0e-5f)) break;
      /* GETDETAU: 4 + 8 per layer */
      n->flops += 4 + 8.0 * nl;
      if (!w->ran[1][nw][ng]) continue;
      /* GFLUXV: scaling 4+8/layer; G1..GAMA 14/layer; AM/AP/C 2 loops: 2x(17)+4 per layer;
       * E-terms 1+1+8; dsolver rows ~ 9+9 per layer + 2-row unmix 4; dtridgl 7 per row;
       * mid fluxes ~32 per layer; top 20; bottom 34; bsurf/eterm 5; accumulation 3+9/layer */
      n->flops += 4 + 8.0 * nl + 14.0 * nl + 38.0 * nl + 10.0 * nl + 22.0 * nl + 14.0 * nl +
                  32.0 * nl + 20 + 34 + 5 + 3 + 9.0 * nl;
      n->sqrts += nl;
      n->exps += 4.0 * nl + 1 + 3.0 * nl + 2;
    }
  }
  for (nw = 1; nw <= NI; nw++) {
    for (ng = 1; ng <= NG; ng++) {
      if (!w->ran[0][nw][ng]) continue;
      /* GFLUXI: alpha/lamda 7, B0/B1 6, gama/term/C 14, E 10, dsolver+dtridgl 22+14,
       * mid 26, top 12, btop 5, accumulation 4+11/layer */
      n->flops += (7.0 + 6 + 14 + 10 + 22 + 14 + 26 + 11) * nl + 12 + 5 + 4;
      n->sqrts += nl;
      n->exps += 2.0 * nl + 1;
    }
  }
  n->flops += 6.0 * nl + 6.0 * (nlay + 1); /* suntot/irtotal and heating */
}

void or_batch_count(const or_tables *t, const double *sol, const or_batch_in *in,
                    or_opcount *count)
{
  col_work *w = (col_work *)malloc(sizeof(col_work));
  or_batch_out dummy;
  const int nlay = t->nlay, nl = nlay + 1;
  double *buf = (double *)calloc((size_t)(6 * nl + 3 + 2 * NV * NG + nlay + 1), sizeof(double));
  memset(count, 0, sizeof(*count));
  dummy.fmnetv = buf; dummy.fluxupv = buf + nl; dummy.fluxdnv = buf + 2 * nl;
  dummy.fmneti = buf + 3 * nl; dummy.fluxupi = buf + 4 * nl; dummy.fluxdni = buf + 5 * nl;
  dummy.nfluxtopv = buf + 6 * nl; dummy.nfluxtopi = buf + 6 * nl + 1; dummy.diffvt = buf + 6 * nl + 2;
  dummy.detau = buf + 6 * nl + 3; dummy.cumtauv = buf + 6 * nl + 3 + NV * NG;
  dummy.hr = buf + 6 * nl + 3 + 2 * NV * NG;
  for (int i = 0; i < in->ncol; i++) {
    or_column c;
    or_out o;
    or_batch_in one = *in;
    batch_column(&one, &dummy, nlay, i, &c, &o);
    /* outputs all alias the single dummy column */
    o.fmnetv = dummy.fmnetv; o.fluxupv = dummy.fluxupv; o.fluxdnv = dummy.fluxdnv;
    o.fmneti = dummy.fmneti; o.fluxupi = dummy.fluxupi; o.fluxdni = dummy.fluxdni;
    o.nfluxtopv = dummy.nfluxtopv; o.nfluxtopi = dummy.nfluxtopi; o.diffvt = dummy.diffvt;
    o.detau = dummy.detau; o.cumtauv = dummy.cumtauv; o.hr = dummy.hr;
    or_tables tc;
    column_tables(t, in, i, &tc);
    column_run_work(&tc, sol, &c, &o, w);
    count_column(&tc, &c, w, count);
  }
  free(buf);
  free(w);
}

/* ------------------------------------------------------------------------ */
/* Init-time tables. */

/* LAGRANGE: code/lagrange.f90:17-31 */
void or_lagrange(double x, const double xi[4], const double yi[4], double *ans)
{
  double fm1 = x - xi[0], fm2 = x - xi[1], fm3 = x - xi[2], fm4 = x - xi[3];
  *ans = fm2 * fm3 * fm4 * yi[0] / ((xi[0] - xi[1]) * (xi[0] - xi[2]) * (xi[0] - xi[3])) +
         fm1 * fm3 * fm4 * yi[1] / ((xi[1] - xi[0]) * (xi[1] - xi[2]) * (xi[1] - xi[3])) +
         fm1 * fm2 * fm4 * yi[2] / ((xi[2] - xi[0]) * (xi[2] - xi[1]) * (xi[2] - xi[3])) +
         fm1 * fm2 * fm3 * yi[3] / ((xi[3] - xi[0]) * (xi[3] - xi[1]) * (xi[3] - xi[2]));
}

/* LAGINTERP: code/laginterp.f90:1-239.  pgref11: the 11 reference pressures
 * (mbar); co2v8/co2i8: raw tables (7,11,10,nb,17) (overwritten by their log10);
 * outputs pint51 (log10 p grid -6..4), co2v/co2i on (7,51,10,nb,17). */
void or_laginterp(const double *pgref11, const double *co2v8_in, const double *co2i8_in,
                  double *pint51, double *co2v, double *co2i)
{
  const int npref = 11;
  double pref[11];
  int n, m, i, nt, nh, nw, ng, pass;
  for (n = 0; n < OR_NP; n++) pint51[n] = -6.0 + 0.2 * n;
  /* the table is written with D0 literals -6.0D0..4.0D0 in steps of 0.2 */
  {
    static const double pin[OR_NP] = {
        -6.0, -5.8, -5.6, -5.4, -5.2, -5.0, -4.8, -4.6, -4.4, -4.2, -4.0, -3.8, -3.6,
        -3.4, -3.2, -3.0, -2.8, -2.6, -2.4, -2.2, -2.0, -1.8, -1.6, -1.4, -1.2, -1.0,
        -0.8, -0.6, -0.4, -0.2, 0.0,  0.2,  0.4,  0.6,  0.8,  1.0,  1.2,  1.4,  1.6,
        1.8,  2.0,  2.2,  2.4,  2.6,  2.8,  3.0,  3.2,  3.4,  3.6,  3.8,  4.0};
    for (n = 0; n < OR_NP; n++) pint51[n] = pin[n];
  }
  for (n = 0; n < npref; n++) pref[n] = log10(pgref11[n]);
  for (pass = 0; pass < 2; pass++) {
    const int nb = pass ? NV : NI;
    const double *raw = pass ? co2v8_in : co2i8_in;
    double *out = pass ? co2v : co2i;
    const size_t nraw = (size_t)NT * npref * NH * nb * NG;
    double *lg = (double *)malloc(nraw * sizeof(double));
    for (size_t q = 0; q < nraw; q++) lg[q] = (raw[q] > 1.0e-200) ? log10(raw[q]) : -200.0;
#define RAW(t_, p_, h_, w_, g_) lg[((t_)-1) + NT * (((p_)-1) + npref * (((h_)-1) + NH * (((w_)-1) + nb * ((g_)-1))))]
    for (nt = 1; nt <= NT; nt++)
      for (nh = 1; nh <= NH; nh++)
        for (nw = 1; nw <= nb; nw++)
          for (ng = 1; ng <= NG; ng++) {
            double xi[4], yi[4], ans;
            int j;
            n = 1;
            for (m = 1; m <= 5; m++) {
              for (j = 0; j < 4; j++) { xi[j] = pref[n + j - 1]; yi[j] = RAW(nt, n + j, nh, nw, ng); }
              or_lagrange(pint51[m - 1], xi, yi, &ans);
              CO2(out, nb, nt, m, nh, nw, ng) = pow(10.0, ans);
            }
            for (n = 2; n <= npref - 2; n++)
              for (m = 1; m <= 5; m++) {
                i = (n - 1) * 5 + m;
                for (j = 0; j < 4; j++) { xi[j] = pref[n + j - 2]; yi[j] = RAW(nt, n - 1 + j, nh, nw, ng); }
                or_lagrange(pint51[i - 1], xi, yi, &ans);
                CO2(out, nb, nt, i, nh, nw, ng) = pow(10.0, ans);
              }
            n = npref - 1;
            for (m = 1; m <= 5; m++) {
              i = (n - 1) * 5 + m;
              for (j = 0; j < 4; j++) { xi[j] = pref[n + j - 3]; yi[j] = RAW(nt, n - 2 + j, nh, nw, ng); }
              or_lagrange(pint51[i - 1], xi, yi, &ans);
              CO2(out, nb, nt, i, nh, nw, ng) = pow(10.0, ans);
            }
            CO2(out, nb, nt, OR_NP, nh, nw, ng) = pow(10.0, RAW(nt, npref, nh, nw, ng));
          }
#undef RAW
    free(lg);
  }
}

/* SETSPI Planck table: code/setspi.f90:48-105.  planckir (5,8501). */
void or_planck_table(double *planckir, double *dwni, double *wnoi)
{
  static const double bwni[6] = {10.000, 166.667, 416.667, 833.333, 1250.000, 2222.222};
  static const double x[12] = {-0.981560634246719, -0.904117256370475, -0.769902674194305,
                               -0.587317954286617, -0.367831498998180, -0.125233408511469,
                               0.125233408511469,  0.367831498998180,  0.587317954286617,
                               0.769902674194305,  0.904117256370475,  0.981560634246719};
  static const double wq[12] = {0.047175336386512, 0.106939325995318, 0.160078328543346,
                                0.203167426723066, 0.233492536538355, 0.249147045813403,
                                0.249147045813403, 0.233492536538355, 0.203167426723066,
                                0.160078328543346, 0.106939325995318, 0.047175336386512};
  const double c1 = 3.741832e-16, c2 = 1.438786e-2;
  int m, nw, nt;
  for (m = 0; m < NI; m++) {
    wnoi[m] = 0.5 * (bwni[m + 1] + bwni[m]);
    dwni[m] = bwni[m + 1] - bwni[m];
  }
  for (nw = 1; nw <= NI; nw++) {
    const double a = 1.0e-2 / bwni[nw], b = 1.0e-2 / bwni[nw - 1];
    const double bpa = (b + a) / 2.0, bma = (b - a) / 2.0;
    for (nt = 500; nt <= 9000; nt++) {
      const double tt = (double)nt / 1.0e1;
      double ans = 0.0;
      for (m = 0; m < 12; m++) {
        const double y = bma * x[m] + bpa;
        const double y5 = y * y * y * y * y; /* Y**5 as the amdflang build evaluates it: left to right */
        ans = ans + wq[m] * c1 / (y5 * (exp(c2 / (y * tt)) - 1.0));
      }
      PLANCK(planckir, nw, nt - 499) = ans * bma / (PI * dwni[nw - 1]);
    }
  }
}

/* SETSPV: code/setspv.f90:67-105 (REAL(4) literals kept as such). */
void or_setspv(double *wnov, double *dwnv, double *solarf, double *tauray)
{
  static const double bwnv[8] = {2222.22, 3087.37, 4030.63, 5370.57, 7651.11, 12500.00, 25000.00, 41666.67};
  static const float solar[7] = {12.7f, 24.2f, 54.6f, 145.9f, 354.9f, 657.5f, 106.3f};
  const double p0 = 9.423e+6;
  int m;
  for (m = 0; m < NV; m++) {
    double wl;
    wnov[m] = 0.5 * (bwnv[m + 1] + bwnv[m]);
    dwnv[m] = bwnv[m + 1] - bwnv[m];
    solarf[m] = (double)solar[m];
    wl = (double)1.0e+4f / wnov[m];
    tauray[m] = ((double)8.7f / GRAV) *
                ((double)1.527f * (1.0 + (double)0.013f / (wl * wl)) / (((wl * wl) * wl) * wl)) * SCALEP / p0;
  }
}

/* ------------------------------------------------------------------------ */
/* Aerosol producers. */

/* INITCLD: code/initcld.f:131-146 (REAL(4) literals promoted; ATHIRD = 1./3.
 * evaluated in REAL(4), code/initcld.f:57). */
void or_aer_grid(double *rad_rt, double *radb_rt)
{
  const int nbin_rt = 20;
  const double athird = (double)(1.0f / 3.0f);
  double vrat_rt;
  int i;
  rad_rt[0] = (double)1.e-7f;
  rad_rt[nbin_rt - 1] = (double)50.e-6f;
  radb_rt[0] = (double)0.0001e-6f;
  radb_rt[nbin_rt] = (double)1.e-2f;
  vrat_rt = log(rad_rt[nbin_rt - 1] / rad_rt[0]) / (double)(float)(nbin_rt - 1) * 3.;
  vrat_rt = exp(vrat_rt);
  for (i = 1; i <= nbin_rt - 1; i++) {
    rad_rt[i] = rad_rt[i - 1] * pow(vrat_rt, athird);
    radb_rt[i] = pow((2. * vrat_rt) / (vrat_rt + 1.), athird) * rad_rt[i - 1];
  }
}

/* OPT_DST code/opt_dst.f:45-163 and OPT_CLD code/opt_cld.f:45-189. */
void or_opt_aer(int cloud, const or_aer_tables *tb, int ncol, int nlay, const double *qtr,
                const double *pl_all, double *qxv_all, double *qsv_all, double *gv_all, double *qxi_all,
                double *qsi_all, double *gi_all, double *qextref_all, double *tauref_all, double *tau2)
{
  enum { NBIN = 20, NRATIO = 15 };
  static const float cor_ratio[NRATIO] = {0.1f, 0.2f, 0.25f, 0.3f, 0.35f, 0.4f, 0.45f, 0.5f,
                                          0.55f, 0.6f, 0.65f, 0.7f, 0.8f, 0.9f, 0.99f};
  const int nlevels = 2 * nlay + 3, ld = nlevels + 1;
  const double athird = (double)(1.0f / 3.0f);
  const double dpden_dt = 2.5e3, dpden_ice = 917.;
  const double dev = cloud ? (double)0.3087f : (double)0.63676f;
  const double dev2 = 1. / ((double)sqrtf(2.f) * dev);
  const double pi = 3.1415926535897932, grav = 3.72;
  int col;
  for (col = 0; col < ncol; col++) {
    const double *q = qtr + (size_t)col * 6 * nlay; /* QTRACE(l, m) = q[(m-1)*nlay + l-1] */
    const double *pl = pl_all + (size_t)col * nlevels;
    double *qxv = qxv_all + (size_t)col * NV * ld, *qsv = qsv_all + (size_t)col * NV * ld;
    double *gv = gv_all + (size_t)col * NV * ld, *qxi = qxi_all + (size_t)col * NI * ld;
    double *qsi = qsi_all + (size_t)col * NI * ld, *gi = gi_all + (size_t)col * NI * ld;
    double *qextref = qextref_all + (size_t)col * ld, *tauref = tauref_all + (size_t)col * ld;
    float t1 = 0.f, t2 = 0.f;
    int k, l, i, w;
#define QT(l_, m_) q[((m_)-1) * nlay + (l_)-1]
    for (k = 1; k <= ld; k++) {
      qextref[k - 1] = 1.;
      tauref[k - 1] = 0.;
    }
    for (w = 0; w < NV; w++)
      for (k = 1; k <= ld; k++) {
        qxv[w * ld + k - 1] = qextref[k - 1];
        qsv[w * ld + k - 1] = qextref[k - 1] * (double)0.99f;
        gv[w * ld + k - 1] = 0.;
      }
    for (w = 0; w < NI; w++)
      for (k = 1; k <= ld; k++) {
        qxi[w * ld + k - 1] = qextref[k - 1];
        qsi[w * ld + k - 1] = qextref[k - 1] * (double)0.99f;
        gi[w * ld + k - 1] = 0.;
      }
    for (l = 1; l <= nlay; l++) {
      double rs, ao, mo, no, surf[NBIN], e[NBIN + 1], cst;
      int irap = 0, nn;
      if (!cloud) {
        /* iMa_dt = 1, iNb_dt = 2 (code/initcld.f:37-38) */
        if (!(QT(l, 1) > (double)1.e-8f && QT(l, 2) > 1.)) continue;
        mo = QT(l, 1);
        no = QT(l, 2) + 1.;
        cst = 0.75 / (pi * dpden_dt);
      } else {
        /* iMa_cld = 3, iNb_cld = 4, iMa_cor = 5 */
        float mtc;
        double dens;
        if (!(QT(l, 3) + QT(l, 5) > (double)1.e-7f && QT(l, 4) > 1.)) continue;
        mtc = (float)(QT(l, 5) / dpden_dt / (QT(l, 3) / dpden_ice + QT(l, 5) / dpden_dt));
        mtc = (float)pow((double)mtc, athird);
        irap = NRATIO;
        for (i = 1; i <= NRATIO; i++) {
          if (mtc < cor_ratio[i - 1] && i != 1) { irap = i - 1; break; }
          else if (mtc == cor_ratio[i - 1]) { irap = i; break; }
          else if (mtc < cor_ratio[i - 1] && i == 1) { irap = 1; break; }
        }
        mo = QT(l, 3) + QT(l, 5);
        no = QT(l, 4);
        dens = QT(l, 3) / mo * dpden_ice + QT(l, 5) / mo * dpden_dt;
        cst = 0.75 / (pi * dens);
      }
      rs = pow(mo / no * cst, athird) * exp(-0.5 * (dev * dev));
      ao = no * pi * (rs * rs);
      rs = rs * exp(1.5 * (dev * dev));
      rs = fmin(fmax(rs, (double)1.e-7f), cloud ? (double)100.e-6f : (double)50.e-6f);
      rs = 1. / rs;
      for (i = 0; i <= NBIN; i++) e[i] = erf(log(tb->radb_rt[i] * rs) * dev2);
      for (i = 0; i < NBIN; i++) surf[i] = 0.5 * (e[i + 1] - e[i]);
      for (nn = 1; nn <= 2; nn++) {
        const int kk = 2 * l + 1 + nn;
        double mass;
        for (w = 0; w < NV; w++) {
          double x = 0., sc = 0., g = 0.;
          for (i = 0; i < NBIN; i++) {
            const size_t o = cloud ? (size_t)(irap - 1) + NRATIO * (i + NBIN * (size_t)w) : (size_t)i + NBIN * (size_t)w;
            const size_t st = cloud ? (size_t)NRATIO * NBIN * NV : (size_t)NBIN * NV;
            x = x + surf[i] * tb->q_v[o];
            sc = sc + surf[i] * tb->q_v[o + st];
            g = g + surf[i] * tb->q_v[o + 2 * st];
          }
          qxv[w * ld + kk - 1] = x;
          qsv[w * ld + kk - 1] = fmin(sc, (double)0.99999f * x);
          gv[w * ld + kk - 1] = g;
        }
        for (w = 0; w < NI; w++) {
          double x = 0., sc = 0., g = 0.;
          for (i = 0; i < NBIN; i++) {
            const size_t o = cloud ? (size_t)(irap - 1) + NRATIO * (i + NBIN * (size_t)w) : (size_t)i + NBIN * (size_t)w;
            const size_t st = cloud ? (size_t)NRATIO * NBIN * NI : (size_t)NBIN * NI;
            x = x + surf[i] * tb->q_i[o];
            sc = sc + surf[i] * tb->q_i[o + st];
            g = g + surf[i] * tb->q_i[o + 2 * st];
          }
          qxi[w * ld + kk - 1] = x;
          qsi[w * ld + kk - 1] = fmin(sc, (double)0.99999f * x);
          gi[w * ld + kk - 1] = g;
        }
        qextref[kk - 1] = qxv[5 * ld + kk - 1]; /* L_NREFV = 6 */
        mass = 100. * (pl[kk - 1] - pl[kk - 2]) / grav;
        tauref[kk - 1] = ao * qextref[kk - 1] * mass;
        t1 = (float)((double)t1 + tauref[kk - 1]);
        t2 = (float)((double)t2 + tauref[kk - 1] / qextref[kk - 1] * (qxi[3 * ld + kk - 1] - qsi[3 * ld + kk - 1]));
      }
    }
#undef QT
    tau2[2 * col] = t1;
    tau2[2 * col + 1] = t2;
  }
}

/* DUSTPROFILE: code/dustprofile.f:30-37 (1.0/(FLOAT(NPDST)-1.0) evaluated in REAL(4)). */
void or_dust_grid(double psf, double ptrop, int npdst, double *prdst)
{
  const double refpr = pow(5.0 * psf / ptrop, (double)(1.0f / ((float)npdst - 1.0f)));
  int n;
  prdst[0] = ptrop;
  for (n = 2; n <= npdst; n++) prdst[n - 1] = refpr * prdst[n - 2];
}

/* JSRCHGT: code/jsrchgt.f90:28-66, 1-based result. */
static int jsrchgt(int n, const double *sx, double target)
{
  int jl, jh, jm;
  if (n < 1) return 0;
  if (target > sx[n - 1]) return n + 1;
  if (target < sx[0]) return 1;
  jl = 1;
  jh = n;
  while (jh - jl > 1) {
    jm = (jl + jh) / 2;
    if (target > sx[jm - 1]) jl = jm;
    else jh = jm;
  }
  if (target == sx[jh - 1]) jh = jh + 1;
  return jh;
}

void or_dust_columns(int ncol, int nlay, int npdst, const double *prdst, double rptau, double ptrop,
                     const double *tautot, const double *conr_all, const double *p, const double *sigma,
                     double *taudst_all, double *tauref_all, double *taucum_all)
{
  const int nlevels = 2 * nlay + 3;
  int col;
  for (col = 0; col < ncol; col++) {
    const double conr = conr_all[col];
    double *taudst = taudst_all + (size_t)col * npdst;
    double *tauref = tauref_all + (size_t)col * (nlevels + 1);
    double *taucum = taucum_all + (size_t)col * nlevels;
    double qrdst[256];
    double sum = 0., pave, qrdst0;
    int n, k;
    /* code/dustprofile.f:56-71 */
    for (n = 2; n <= npdst; n++) {
      if (prdst[n - 1] < rptau) {
        pave = 0.5 * (prdst[n - 1] + prdst[n - 2]);
        sum = sum + exp(conr * (1. - (rptau / pave))) * (prdst[n - 1] - prdst[n - 2]);
      }
      if (prdst[n - 1] >= rptau) break;
    }
    pave = 0.5 * (rptau + prdst[n - 2]);
    sum = sum + exp(conr * (1. - (rptau / pave))) * (rptau - prdst[n - 2]);
    qrdst0 = tautot[col] / sum;
    /* code/dustprofile.f:75-104 */
    for (n = 1; n <= npdst - 1; n++) {
      if (rptau > prdst[n]) {
        pave = 0.5 * (prdst[n] + prdst[n - 1]);
        qrdst[n - 1] = qrdst0 * exp(conr * (1.0 - (rptau / pave)));
      }
      if (rptau <= prdst[n] && rptau >= prdst[n - 1]) {
        const double pdif1 = rptau - prdst[n - 1], pdif2 = prdst[n] - rptau;
        pave = 0.5 * (prdst[n - 1] + rptau);
        qrdst[n - 1] = qrdst0 * (exp(conr * (1.0 - (rptau / pave))) * pdif1 + pdif2) / (prdst[n] - prdst[n - 1]);
      }
      if (rptau < prdst[n - 1]) qrdst[n - 1] = qrdst0;
    }
    taudst[0] = 0.0;
    for (n = 2; n <= npdst; n++) taudst[n - 1] = taudst[n - 2] + qrdst[n - 2] * (prdst[n - 1] - prdst[n - 2]);
    /* FILLTAUCUM: code/filltaucum.f:33-52 */
    for (n = 1; n <= 3; n++) {
      taucum[n - 1] = 0.0;
      tauref[n - 1] = 0.0;
    }
    for (k = 4; k <= nlevels; k++) {
      const double pstar = p[col] * sigma[k - 1] + ptrop;
      const double pstar1 = fmax(pstar, prdst[0]);
      int nstar = jsrchgt(npdst - 1, prdst, pstar1) - 1;
      if (nstar > npdst - 1) nstar = npdst - 1;
      taucum[k - 1] = taudst[nstar - 1] + (pstar1 - prdst[nstar - 1]) * (taudst[nstar] - taudst[nstar - 1]) /
                                              (prdst[nstar] - prdst[nstar - 1]);
      tauref[k - 1] = taucum[k - 1] - taucum[k - 2];
    }
    tauref[nlevels] = 0.0;
  }
}

/* DSOLFLUX: code/dsolflux.f:18-45. */
double or_dsolflux(const double *sol, double ubar0, const double *gw, const double *fzerov,
                   const double *detau)
{
  double directsol = 0.0, factor;
  int nw, ng;
  for (nw = 1; nw <= NV; nw++) {
    factor = ubar0 * sol[nw - 1];
    for (ng = 1; ng <= NG - 1; ng++)
      if (detau[(nw - 1) + NV * (ng - 1)] <= 5.0)
        directsol = directsol + factor * exp(-detau[(nw - 1) + NV * (ng - 1)] / ubar0) * GW(gw, ng, nw) *
                                    (1.0 - fzerov[nw - 1]);
    ng = NG;
    if (detau[(nw - 1) + NV * (ng - 1)] <= 5.0)
      directsol = directsol + factor * exp(-detau[(nw - 1) + NV * (ng - 1)] / ubar0) * fzerov[nw - 1];
  }
  return directsol;
}

void or_libm(int n, const double *x, double *e, double *l10, double *ln)
{
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < n; i++) {
    e[i] = exp(x[i]);
    l10[i] = x[i] > 0.0 ? log10(x[i]) : 0.0;
    ln[i] = x[i] > 0.0 ? log(x[i]) : 0.0;
  }
}

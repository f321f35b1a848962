/*
 * mgcm_oracle.h -- CPU FP64 restatement of the Legacy Mars GCM column
 * radiative transfer (TEST INFRASTRUCTURE ONLY).
 *
 * This header and oracle/mgcm_oracle.c are the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load them.  The product library never links, loads or calls them.
 *
 * Every array uses the reference's Fortran layout (first index fastest,
 * 1-based in the Fortran; 0-based here).  Dimensions:
 *   nlev    = 2*nlay+3   (L_LEVELS,  code/modules.f90:41)
 *   nlayrad = nlay+1     (L_NLAYRAD, code/modules.f90:221)
 *   k-tables CO2V(7,51,10,7,17), CO2I(7,51,10,5,17) (code/modules.f90:289-290)
 */
#ifndef MGCM_ORACLE_H
#define MGCM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NT 7
#define OR_NP 51
#define OR_NH 10
#define OR_NG 17
#define OR_NV 7
#define OR_NI 5
#define OR_NPLANCK 8501

typedef struct {
  int nlay;
  const double *co2v;     /* (7,51,10,7,17) */
  const double *co2i;     /* (7,51,10,5,17) */
  const double *pfgasref; /* (51) log10 p [mbar] */
  const double *tgasref;  /* (7) K */
  const double *wrefh2o;  /* (10) vmr */
  const double *gwv;      /* (17,7): Gauss weights per VIS band, g fastest */
  const double *gwi;      /* (17,5): Gauss weights per IR band */
  const double *fzerov;   /* (7) */
  const double *fzeroi;   /* (5) */
  const double *tauray;   /* (7) */
  const double *dwni;     /* (5) */
  const double *planckir; /* (5,8501) */
  /* aerosol optics, per sub-level, shared by all columns: (nlev+1, nb) */
  const double *qxvdst, *qsvdst, *gvdst;
  const double *qxvcld, *qsvcld, *gvcld;
  const double *qxidst, *qsidst, *gidst;
  const double *qxicld, *qsicld, *gicld;
  const double *qextrefdst; /* (nlev+1) */
  const double *qextrefcld; /* (nlev+1) */
} or_tables;

typedef struct {
  /* per column, Fortran per-column layout */
  const double *plev, *tlev, *pmid, *tmid, *qh2o; /* (nlev) */
  const double *tauref, *taurefcld;               /* (nlev+1) */
  double als, albi, acosz;                        /* VIS albedo, IR albedo, mu0 */
} or_column;

typedef struct {
  double *fmnetv, *fluxupv, *fluxdnv; /* (nlayrad) */
  double *fmneti, *fluxupi, *fluxdni; /* (nlayrad) */
  double *nfluxtopv, *nfluxtopi, *diffvt;
  double *detau;   /* (7,17)  delta-Eddington surface tau; untouched if sun down */
  double *cumtauv; /* (7,17)  TAUCUMV(L_LEVELS); untouched if sun down */
  double *hr;      /* (nlay+1) heating K/sol, [0]=stratosphere */
} or_out;

/* One column of COMP3's radiation block (code/comp3.f:654-738). sol[7]=SOLARF/RSDIST. */
void or_column_run(const or_tables *t, const double *sol, const or_column *c, or_out *o);

/* Batch over ncol columns stored per-column-contiguous (stride nlev / nlev+1 /
 * nlayrad...).  nthreads<=0 uses all OpenMP threads. */
typedef struct {
  int ncol;
  const double *plev, *tlev, *pmid, *tmid, *qh2o; /* [ncol][nlev] */
  const double *tauref, *taurefcld;               /* [ncol][nlev+1] */
  const double *als, *albi, *acosz;               /* [ncol] */
  /* optional per-column aerosol optics (NULL = the tables' arrays); per column
   * the tables' layout (nlev+1, nb): [ncol][nb][nlev+1], references [ncol][nlev+1] */
  const double *qxvdst, *qsvdst, *gvdst, *qxvcld, *qsvcld, *gvcld;
  const double *qxidst, *qsidst, *gidst, *qxicld, *qsicld, *gicld;
  const double *qextrefdst, *qextrefcld;
} or_batch_in;

typedef struct {
  double *fmnetv, *fluxupv, *fluxdnv, *fmneti, *fluxupi, *fluxdni; /* [ncol][nlayrad] */
  double *nfluxtopv, *nfluxtopi, *diffvt;                          /* [ncol] */
  double *detau, *cumtauv;                                         /* [ncol][17*7] */
  double *hr;                                                      /* [ncol][nlay+1] */
} or_batch_out;

void or_batch_run(const or_tables *t, const double *sol, const or_batch_in *in,
                  or_batch_out *out, int nthreads);

/* Executed FP64 +,-,*,/ per column (transcendentals counted separately). */
typedef struct { double flops, exps, sqrts, logs; } or_opcount;
void or_batch_count(const or_tables *t, const double *sol, const or_batch_in *in,
                    or_opcount *count);

/* Individual reference routines, exposed for per-function tests. */
void or_tpindex(double pw, double tw, double qh2o, const double *pref, const double *tref,
                const double *wrefh2o, double lcoef[4], int *mt, int *mp, int *nh2o,
                double *wratio);
void or_dtridgl(int l, const double *af, const double *bf, const double *cf, const double *df,
                double *xk);
void or_dsolver(int nl, const double *gama, const double *cp, const double *cm,
                const double *cpm1, const double *cmm1, const double *e1, const double *e2,
                const double *e3, const double *e4, double btop, double bsurf, double rsf,
                double *xk1, double *xk2);
void or_lagrange(double x, const double xi[4], const double yi[4], double *ans);
void or_laginterp(const double *pgref11, const double *co2v8, const double *co2i8,
                  double *pint51, double *co2v, double *co2i);
void or_planck_table(double *planckir, double *dwni, double *wnoi);
void or_setspv(double *wnov, double *dwnv, double *solarf, double *tauray);

/* Aerosol producers.  INITCLD's radiative size grid (code/initcld.f:131-146):
 * rad_rt(20), radb_rt(21). */
void or_aer_grid(double *rad_rt, double *radb_rt);

/* OPT_DST (cloud = 0, code/opt_dst.f) or OPT_CLD (cloud = 1, code/opt_cld.f) on
 * ncol columns of nlay layers.  Tables as INITCLD reads them (REAL*4 values):
 * dust q_v (20,7,3), q_i (20,5,3); cloud q_v (15,20,7,3), q_i (15,20,5,3); the
 * last index is Qext, Qscat, g.  qtr: [ncol][6][nlay] QTRACE(L_LAYERS, NTRACE);
 * pl: [ncol][nlev].  Outputs [ncol][nb][nlev+1], [ncol][nlev+1], tau2 [ncol][2]
 * (taudust / taucloud, REAL*4 sums). */
typedef struct {
  const double *radb_rt;
  const double *q_v, *q_i;
} or_aer_tables;
void or_opt_aer(int cloud, const or_aer_tables *tb, int ncol, int nlay, const double *qtr,
                const double *pl, double *qxv, double *qsv, double *gv, double *qxi, double *qsi,
                double *gi, double *qextref, double *tauref, double *tau2);

/* DUSTPROFILE's reference pressure grid PRDST(npdst) (code/dustprofile.f:30-37). */
void or_dust_grid(double psf, double ptrop, int npdst, double *prdst);

/* DUSTPROFILE's column profile (code/dustprofile.f:56-113) and FILLTAUCUM
 * (code/filltaucum.f:33-52) for ncol columns: tautot[ncol] the opacity at
 * RPTAU, conr[ncol] the Conrath parameter, p[ncol] = P(J,I), sigma[nlev].
 * Outputs taudst[ncol][npdst], tauref[ncol][nlev+1], taucum[ncol][nlev]. */
void or_dust_columns(int ncol, int nlay, int npdst, const double *prdst, double rptau, double ptrop,
                     const double *tautot, const double *conr, const double *p, const double *sigma,
                     double *taudst, double *tauref, double *taucum);

/* DSOLFLUX (code/dsolflux.f:18-45) of one column: gw (17,7) per-band Gauss
 * weights (the reference's single GWEIGHT repeated), detau (7,17). */
double or_dsolflux(const double *sol, double ubar0, const double *gw, const double *fzerov,
                   const double *detau);

/* The host C library's exp, log10 and log of n arguments (the functions the
 * reference's EXP, LOG10 and LOG call); entries with x <= 0 give 0 for the logs.
 * The reference the device's restatements are compared with. */
void or_libm(int n, const double *x, double *e, double *l10, double *ln);

#ifdef __cplusplus
}
#endif
#endif

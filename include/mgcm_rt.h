/*
 * mgcm_rt.h -- C ABI of the MI355X column radiative transfer for the Legacy
 * Mars GCM (v24).  A Fortran host binds it with ISO_C_BINDING
 * (legacy-mars-global-climate-model_amd/fortran/mgcm_rt_mod.f90); Python binds
 * it with ctypes (legacy-mars-global-climate-model_amd/mgcm_rt/).
 *
 * What it replaces (reference files, relative to the reference tree):
 *   - the per-column radiation block of COMP3: OPTCV (code/comp3.f:658-662),
 *     SFLUXV (code/comp3.f:672-675), OPTCI (code/comp3.f:724-727) and SFLUXI
 *     (code/comp3.f:731-733), with the net-flux differencing of
 *     code/comp3.f:677-681,735-738 turned into layer heating rates;
 *   - the table set-up of RADSETUP (code/main.f:211): SETSPI's Planck table
 *     (code/setspi.f90:91-105), LAGINTERP's reads of the k-coefficient files
 *     (code/laginterp.f90:59-69) and its 11 -> 51 pressure refinement
 *     (code/laginterp.f90:30-236).
 *
 * Arrays use the reference's Fortran layouts (first index fastest), one column
 * after another: nlev = 2*nlay+3 (L_LEVELS), nlayrad = nlay+1 (L_NLAYRAD).
 * Every entry point returns MGCM_RT_OK (0) or a negative status; none aborts.
 *
 * Contexts share no mutable state: several may run at once on different
 * streams or host threads.  One context's workspaces (the continuum-point
 * records, the host-pointer staging) are reused by each of its calls, so calls
 * on the same context must not overlap: issue them on one stream, or
 * synchronise between them.  A context's tables live on the HIP device that was
 * current at mgcm_rt_init; call it from that device.
 */
#ifndef MGCM_RT_H
#define MGCM_RT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGCM_RT_NT 7      /* L_NTREF   (code/modules.f90:228) */
#define MGCM_RT_NP 51     /* L_PINT    (code/modules.f90:234) */
#define MGCM_RT_NH 10     /* L_REFH2O  (code/modules.f90:236) */
#define MGCM_RT_NG 17     /* L_NGAUSS  (code/modules.f90:226) */
#define MGCM_RT_NV 7      /* L_NSPECTV (code/modules.f90:225) */
#define MGCM_RT_NI 5      /* L_NSPECTI (code/modules.f90:224) */
#define MGCM_RT_NPLANCK 8501
/* 2*NLAYRAD rows within DSOLVER's NMAX = 201 (code/dsolver.f90:12) and NLL <=
 * 101 of GFLUXI (code/gfluxi.f90:65): at most 99 GCM layers, 100 radiation layers */
#define MGCM_RT_MAX_NLAY 99

enum {
  MGCM_RT_OK = 0,
  MGCM_RT_ERR_ARG = -1,      /* null pointer or bad size */
  MGCM_RT_ERR_NLAY = -2,     /* nlay outside 1..MGCM_RT_MAX_NLAY */
  MGCM_RT_ERR_DEVICE = -3,   /* HIP runtime failure */
  MGCM_RT_ERR_NOMEM = -4,
  MGCM_RT_ERR_RANGE = -5,    /* a temperature outside the Planck table (50.0..900.0 K), or a NaN in
                                PLEV, TLEV, PMID, TMID, QH2O, TAUREF, TAUREFCLD, ALS, ALBI or ACOSZ
                                of a column (the per-level aerosol arrays are not checked) */
  MGCM_RT_ERR_PRECISION = -6, /* unknown precision mode */
  MGCM_RT_ERR_IO = -7,        /* data file missing, unreadable or not in the expected record layout */
  MGCM_RT_ERR_SCRATCH = -8    /* grid larger than mgcm_rt_run_noalloc's scratch holds */
};

enum {
  MGCM_RT_FP64 = 0,       /* plain path: every operation in FP64, reference evaluation order */
  MGCM_RT_MIXED = 1,      /* k-tables read in FP32, all arithmetic FP64 */
  /* Precision study (diagnostic; mgcm_rt_run/_run_host/_run_noalloc, grids of
   * 25..36 layers): MGCM_RT_STUDY | component bits -- the column kernel with
   * those components in cheaper arithmetic, the continuum point g = 17 in FP64.
   * Built combinations: INTERP32 (with TAB32), TAU32, TRANS32, all four, FASTDIV,
   * FASTDIV | TAB32.  Other combinations return MGCM_RT_ERR_DEVICE. */
  MGCM_RT_STUDY = 2,
  MGCM_RT_P_TAB32 = 0x100,    /* k-tables stored in FP32 */
  MGCM_RT_P_INTERP32 = 0x200, /* k-interpolation arithmetic in FP32 */
  MGCM_RT_P_TAU32 = 0x400,    /* sub-level optical depths rounded to FP32 */
  MGCM_RT_P_TRANS32 = 0x800,  /* layer transmissions exp(-lambda dtau), exp(-tau/mu0) in FP32 */
  MGCM_RT_P_FASTDIV = 0x1000, /* FP64 division by reciprocal + 2 Newton steps (not IEEE-rounded) */
  /* OR-ed into the precision argument: compute one part only.  The other
   * part's fluxes are zero, its DETAU/CUMTAUV untouched, and the heating holds
   * the computed part alone (the IR surrogate's exact VIS and fallback IR). */
  MGCM_RT_SKIP_VIS = 0x10,
  MGCM_RT_SKIP_IR = 0x20
};

typedef struct mgcm_rt_ctx mgcm_rt_ctx;

/* Radiation tables: the module state of radcommon_h (code/modules.f90:247-367)
 * plus the aerosol optical constants of INI_OPTDST/INI_OPTCLD, which are the
 * same at every level and column.  Host pointers; copied by mgcm_rt_init. */
typedef struct {
  const double *co2v;     /* CO2V(7,51,10,7,17)  k-coefficients, VIS */
  const double *co2i;     /* CO2I(7,51,10,5,17)  k-coefficients, IR  */
  const double *pfgasref; /* PFGASREF(51)  log10 p [mbar] */
  const double *tgasref;  /* TGASREF(7)    K */
  const double *wrefh2o;  /* WREFH2O(10)   vmr */
  const double *gwv;      /* GW(17,7)  Gauss weights per VIS band (GWEIGHT per band) */
  const double *gwi;      /* GW(17,5)  Gauss weights per IR band */
  const double *fzerov;   /* FZEROV(7) */
  const double *fzeroi;   /* FZEROI(5) */
  const double *tauray;   /* TAURAY(7) */
  const double *dwni;     /* DWNI(5)   cm^-1 */
  const double *planckir; /* PLANCKIR(5,8501) */
  const double *qxv_dst, *qsv_dst, *gv_dst; /* (7) dust Qext, Qscat, g per VIS band */
  const double *qxi_dst, *qsi_dst, *gi_dst; /* (5) per IR band */
  const double *qxv_cld, *qsv_cld, *gv_cld; /* (7) water-ice cloud */
  const double *qxi_cld, *qsi_cld, *gi_cld; /* (5) */
  double qextref_dst;     /* QEXTREFDST (IR dust reference extinction) */
  double qextref_cld;     /* QEXTREFCLD */
} mgcm_rt_tables;

/* Per-level aerosol optical properties of each column, as OPT_DST and OPT_CLD
 * produce them for active dust and water-ice clouds (code/opt_dst.f:114-158,
 * code/opt_cld.f; COMP3 hands them to OPTCV/OPTCI, code/comp3.f:595-631).  Any
 * pointer may be NULL: that quantity then takes the tables' level-constant band
 * value (INI_OPTDST/INI_OPTCLD).  Per column the reference's layout
 * Q(L_LEVELS+1, nspect): [ncol][nspect][nlev+1].  The VIS dust reference
 * extinction is QXVDST(K, L_NREFV = 6) (code/optcv.f90:114).  Device pointers for
 * mgcm_rt_run/mgcm_rt_run_batched, host pointers for mgcm_rt_run_host. */
typedef struct {
  const double *qxv_dst, *qsv_dst, *gv_dst; /* [ncol][7][nlev+1] dust Qext, Qscat, g */
  const double *qxi_dst, *qsi_dst, *gi_dst; /* [ncol][5][nlev+1] */
  const double *qxv_cld, *qsv_cld, *gv_cld; /* [ncol][7][nlev+1] water ice */
  const double *qxi_cld, *qsi_cld, *gi_cld; /* [ncol][5][nlev+1] */
  const double *qextref_dst;                /* [ncol][nlev+1] QEXTREFDST (OPTCI's QREFV) */
  const double *qextref_cld;                /* [ncol][nlev+1] QEXTREFCLD */
} mgcm_rt_aerosol;

/* One grid of columns.  Device pointers for mgcm_rt_run, host pointers for
 * mgcm_rt_run_host.  sol[] is SOL(NW) = SOLARF(NW)/RSDIST (code/comp3.f:300-304).
 * aerosol: optional per-level aerosol optics (NULL = the tables' constants). */
typedef struct {
  int ncol;
  int nlay;
  const double *plev, *tlev, *pmid, *tmid, *qh2o; /* [ncol][nlev]    */
  const double *tauref, *taurefcld;               /* [ncol][nlev+1]  */
  const double *als;                              /* [ncol] VIS surface albedo (RSFV) */
  const double *albi;                             /* [ncol] IR albedo 1-emissivity (RSFI) */
  const double *acosz;                            /* [ncol] cos(zenith) (UBAR0); < 1e-5 = sun down */
  double sol[MGCM_RT_NV];
  const mgcm_rt_aerosol *aerosol;
} mgcm_rt_columns;

/* Outputs: SFLUXV/SFLUXI results per column and the layer heating rates.
 * detau/cumtauv (DETAU, CUMTAUV of code/comp3.f:664-668) are left untouched
 * for sun-down columns, as the reference leaves its module arrays.  hr may be
 * NULL.  hr[0] is the stratosphere (mass PTROP*100/g), hr[l] GCM layer l, in
 * K/sol: (SUNTOT+IRTOTAL)*g/(Cp*100*dp)*88775 (code/comp3.f:765-771 without
 * the non-LTE factor and the theta conversion, which stay with the caller).
 * directsol/dnvflux: the surface fluxes COMP3 derives for TEMPGR
 * (code/comp3.f:336-348, code/dsolflux.f), computed after the radiation when
 * given (0 for sun-down columns). */
typedef struct {
  double *fmnetv, *fluxupv, *fluxdnv; /* [ncol][nlayrad] */
  double *fmneti, *fluxupi, *fluxdni; /* [ncol][nlayrad] */
  double *nfluxtopv, *nfluxtopi;      /* [ncol] */
  double *diffvt;                     /* [ncol] diffuse VIS down at the surface */
  double *detau;                      /* [ncol][17][7] DETAU(nw,ng) */
  double *cumtauv;                    /* [ncol][17][7] TAUCUMV(L_LEVELS,nw,ng) */
  double *hr;                         /* [ncol][nlay+1] K/sol, optional */
  double *directsol;                  /* [ncol] DSOLFLUX's direct solar flux at the surface, optional */
  double *dnvflux;                    /* [ncol] DIFFVT + directsol (COMP3's DNVFLUX), optional */
} mgcm_rt_fluxes;

/* Replaces RADSETUP's table state (code/radsetup.f90:26-46, code/modules.f90:
 * 247-367): uploads and re-lays out the tables on the current HIP device. */
int mgcm_rt_init(const mgcm_rt_tables *tables, mgcm_rt_ctx **ctx);
int mgcm_rt_destroy(mgcm_rt_ctx *ctx);

/* Replaces the per-column OPTCV+SFLUXV+OPTCI+SFLUXI sequence of COMP3
 * (code/comp3.f:637-738) for a whole grid in one call.  Device pointers,
 * asynchronous on `stream` (a hipStream_t; NULL = default stream), no host
 * synchronisation and no allocation after the first call at a given size, so
 * it can be captured in a HIP graph. */
int mgcm_rt_run(mgcm_rt_ctx *ctx, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out,
                int precision, void *stream);

/* mgcm_rt_run for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same kernels and results bit for bit, with the continuum-point records (and
 * at tile 12 the band records and TPINDEX results) in a static device array
 * of MGCM_RT_NOALLOC_SCRATCH_BYTES (128 MiB, held by every device the library
 * runs on until the process ends) instead of the context's workspace.  Nothing
 * is allocated, ever.  A grid needing more than the array -- ncol * 12 *
 * (2*nlay+4) * 8 bytes, plus ncol * (12 (3*nlay+7) + 5 (2*nlay+2)) * 8 at
 * tile 12 (the automatic choice takes tile 12 only where it fits): 26,886
 * columns at 24 layers, 3,461 at 99 -- returns MGCM_RT_ERR_SCRATCH; run it in column
 * slices (every array is per column, so a slice is a pointer offset).  All
 * calls on one device share the array: calls that overlap on a device (any
 * context, any stream) race, so issue them on one stream or synchronise
 * between them.  No counterpart in the reference. */
#define MGCM_RT_NOALLOC_SCRATCH_BYTES (128ull << 20)
int mgcm_rt_run_noalloc(mgcm_rt_ctx *ctx, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out,
                        int precision, void *stream);

/* Same with host pointers: copies in, runs, copies out, synchronises, and
 * reports MGCM_RT_ERR_RANGE if any column was out of the tables' range.  It
 * ends with a range check on its own stream: like mgcm_rt_check, the result
 * covers this call and any earlier run of the context not yet checked.
 * A grid above 4,096 columns goes in chunks of whole columns: library threads
 * (up to 8) copy each chunk to pinned staging while the device runs the one
 * before, so the arrays must not be touched by other threads during the call.
 * The context keeps its staging (device, up to 256 MB pinned) until destroyed. */
int mgcm_rt_run_host(mgcm_rt_ctx *ctx, const mgcm_rt_columns *in, const mgcm_rt_fluxes *out,
                     int precision);

/* Range flag of the context's runs since the last check (synchronises the
 * whole device, then clears the flag); MGCM_RT_OK or MGCM_RT_ERR_RANGE.  The
 * reference has no such check: it reads outside PLANCKIR silently
 * (code/sfluxi.f90:54-55). */
int mgcm_rt_check(mgcm_rt_ctx *ctx);

/* The same waiting for `stream` only (a hipStream_t; NULL = default stream):
 * it covers the runs issued on that stream, and other runs that have ended. */
int mgcm_rt_check_stream(mgcm_rt_ctx *ctx, void *stream);

/* Many independent grids (ragged layer counts) in one persistent launch.
 * grids[i] / outs[i] hold device pointers.  Columns of all grids are packed
 * into one work list per layer class; the launch shape comes from the
 * context's tuner (mgcm_rt_batched_tuning).  Grids whose outs[i].directsol is
 * set also get DSOLFLUX/DNVFLUX, inside the same graph.  The reference calls
 * the radiation per column (code/comp3.f:411-415,637-738). */
int mgcm_rt_run_batched(mgcm_rt_ctx *ctx, int ngrids, const mgcm_rt_columns *grids,
                        const mgcm_rt_fluxes *outs, int precision, void *stream);

/* The batched tuner's choice per layer class of the last planned batch: the
 * column kernel's tile (band slots per column), instance (waves per SIMD) and
 * persistent grid size, and the time of each candidate (tile x waves per SIMD
 * x multiple of that instance's resident capacity) on the first tuned_items
 * work items (0 ms: not timed -- the candidate repeats an earlier launch).
 * Returns the number of classes (fills at most maxcls).  No counterpart in the
 * reference (it has no batched path). */
#define MGCM_RT_NCAND 10
typedef struct {
  int max_layers;     /* grids with nlay <= max_layers (and above the class below) */
  int items;          /* 4-column work items of the class */
  int tile;           /* chosen band slots per column: 1, 2 or 4 */
  int waves_per_simd; /* chosen instance: 1 or 2 */
  int blocks;         /* chosen persistent grid size */
  int tuned_items;    /* items the candidates were timed on */
  int cand_tile[MGCM_RT_NCAND], cand_waves[MGCM_RT_NCAND], cand_mult[MGCM_RT_NCAND];
  float cand_ms[MGCM_RT_NCAND];
} mgcm_rt_batch_class;
int mgcm_rt_batched_tuning(const mgcm_rt_ctx *ctx, int maxcls, mgcm_rt_batch_class *info);

/* The column kernel's tile for single-grid calls (mgcm_rt_run, _noalloc,
 * _host): band_slots = 1 (4 columns per wave, each wave solving one band of a
 * column per 16-lane row at a time), 2 (2 columns, 2 bands at a time), 4
 * (1 column, 4 bands at a time) or 12 (all 12 bands of a column at once, each
 * in its own wave of 4 columns; a finishing kernel adds the bands' sums, and
 * the workspace grows by (12 (3 nlay + 7) + 5 (2 nlay + 2)) * 8 bytes per
 * column); 0
 * (the default) lets the library choose per grid from its column count,
 * layer count and the device's CU count.  Every tile gives the same bits;
 * more slots give a small grid more waves at some extra work per column
 * (empty band slots, staging repeated per wave).  MGCM_RT_ERR_ARG for
 * another value.  No counterpart in the reference (its radiation is serial
 * per column, code/comp3.f:411-415).  A tile of 1, 2 or 4 also restricts the
 * batched tuner to that tile's candidates (the next mgcm_rt_run_batched
 * re-plans); 12 leaves the batched tuner its own choice. */
int mgcm_rt_set_tile(mgcm_rt_ctx *ctx, int band_slots);
/* The tile a single-grid call of ncol columns x nlay layers runs with (1, 2,
 * 4 or 12), or MGCM_RT_ERR_ARG. */
int mgcm_rt_tile_for(const mgcm_rt_ctx *ctx, int ncol, int nlay);

/* Init-time table builders (host in/out pointers), both equal to the
 * reference's tables bit for bit:
 * SETSPI's Planck table PLANCKIR(5,8501) (code/setspi.f90:91-105), on the device;
 * LAGINTERP (code/laginterp.f90): raw tables (7,11,10,nb,17) on the 11
 * reference pressures pgref[11] (mbar) -> (7,51,10,nb,17) on log10 p =
 * -6.0..4.0 step 0.2, returned in pint[51].  The interpolations run on the
 * device; log10 of the raw k and 10**ans use the host C library, as the
 * reference's do. */
int mgcm_rt_planck_table(double *planckir, double *dwni, double *wnoi);
int mgcm_rt_laginterp(const double *pgref, const double *raw, int nbands, double *pint,
                      double *out);

/* Replaces LAGINTERP's reads of the k-coefficient files (code/laginterp.f90:
 * 59-69, data/CO2H2O_V_12_95_INTEL with nbands = 7, data/CO2H2O_IR_12_95_INTEL
 * with nbands = 5): Fortran sequential unformatted records with 4-byte
 * little-endian length markers.  raw receives CO2(7,11,10,nbands,17), the
 * input of mgcm_rt_laginterp; fzero receives FZERO(nbands).  Host code, no
 * device needed.  MGCM_RT_ERR_IO if the file is missing or its records are
 * short, unframed or truncated. */
int mgcm_rt_read_ktable(const char *path, int nbands, double *raw, double *fzero);

/* ---- Aerosol producers of the radiation's inputs (active dust and clouds) ----
 *
 * Replace INITCLD's set-up of the radiative size grid and optical tables
 * (code/initcld.f:131-184).  mgcm_rt_read_aerosol_tables reads
 * datadir/Dust_vis_wolff2010_JD_12bands.dat, Dust_ir_wolff2010_JD_12bands.dat,
 * waterCoated_vis_JD_12bands.dat and waterCoated_ir_JD_12bands.dat (host code,
 * no device needed) into tables[MGCM_RT_AER_TABLE_DOUBLES]: radb_rt(21) at 0;
 * dust Qext/Qscat/g(20,7,3) at 32 and (20,5,3) at 452; ice-coated dust
 * (15,20,7,3) at 752 and (15,20,5,3) at 7052 (REAL*4 values, first index
 * fastest); MGCM_RT_ERR_IO for a missing or malformed file.
 * mgcm_rt_set_aerosol_tables uploads such a block to the context's device. */
#define MGCM_RT_AER_TABLE_DOUBLES 11552
int mgcm_rt_read_aerosol_tables(const char *datadir, double *tables);
int mgcm_rt_set_aerosol_tables(mgcm_rt_ctx *ctx, const double *tables);

/* Tracers of a grid of columns (device pointers): pl[ncol][nlev] is PL(L_LEVELS)
 * in mbar; qtrace[ncol][6][nlay] is QTRACE(L_LAYERS, NTRACE) of each column:
 * 1 dust mass, 2 dust number, 3 ice mass, 4 ice number, 5 dust-core mass,
 * 6 vapour mixing ratio (code/initcld.f:37-48). */
typedef struct {
  int ncol, nlay;
  const double *pl;
  const double *qtrace;
} mgcm_rt_tracers;

/* Outputs of OPT_DST / OPT_CLD (device pointers), the layouts of mgcm_rt_aerosol:
 * Qext/Qscat/g [ncol][7|5][nlev+1]; qextref and tauref [ncol][nlev+1] (the
 * reference extinction, = Qext of VIS band 6, and TAUREF / TAUREFCLD); tau2
 * [ncol][2] the diagnostics taudust/taucloud (REAL*4 sums), optional. */
typedef struct {
  double *qxv, *qsv, *gv;
  double *qxi, *qsi, *gi;
  double *qextref, *tauref;
  double *tau2;
} mgcm_rt_aerosol_out;

/* OPT_DST (code/opt_dst.f) and OPT_CLD (code/opt_cld.f) for every column and
 * layer, asynchronous on stream.  Needs mgcm_rt_set_aerosol_tables first. */
int mgcm_rt_opt_dst(mgcm_rt_ctx *ctx, const mgcm_rt_tracers *in, const mgcm_rt_aerosol_out *out, void *stream);
int mgcm_rt_opt_cld(mgcm_rt_ctx *ctx, const mgcm_rt_tracers *in, const mgcm_rt_aerosol_out *out, void *stream);
/* The same with host pointers (copies in, runs, copies out, synchronises), for a
 * host that keeps QTRACE and PL in host memory, as COMP3 does. */
int mgcm_rt_opt_dst_host(mgcm_rt_ctx *ctx, const mgcm_rt_tracers *in, const mgcm_rt_aerosol_out *out);
int mgcm_rt_opt_cld_host(mgcm_rt_ctx *ctx, const mgcm_rt_tracers *in, const mgcm_rt_aerosol_out *out);

/* Prescribed (Conrath) dust, the TAUREF of runs without active dust.
 * mgcm_rt_dust_grid: DUSTPROFILE's pressure grid PRDST(npdst) from PSF and PTROP
 * (code/dustprofile.f:30-37; host code, npdst <= 128).
 * mgcm_rt_dust_tauref: DUSTPROFILE's Conrath profile of each column
 * (code/dustprofile.f:56-113) and FILLTAUCUM's TAUREF on the sub-levels
 * (code/filltaucum.f:33-52), device pointers, asynchronous on stream: tauref
 * [ncol][nlev+1]; taudst [ncol][npdst] (TAUDST, optional).  The Conrath
 * parameter is the caller's per column (DUSTPROFILE: CONRNU, or its latitude and
 * season formula when VARY_CONR is set). */
typedef struct {
  int ncol, nlay;
  const double *tautot; /* [ncol] TAUTOTJI: dust opacity at RPTAU */
  const double *conr;   /* [ncol] Conrath parameter */
  const double *p;      /* [ncol] P(J,I): surface pressure - PTROP, mbar */
  const double *sigma;  /* [nlev] SIGMA(L_LEVELS) */
  const double *prdst;  /* [npdst] PRDST */
  int npdst;
  double rptau, ptrop;  /* RPTAU, PTROP (mbar) */
} mgcm_rt_dust_columns;
int mgcm_rt_dust_grid(double psf, double ptrop, int npdst, double *prdst);
int mgcm_rt_dust_tauref(const mgcm_rt_dust_columns *in, double *tauref, double *taudst, void *stream);
/* The same with host pointers (copies in, runs, copies out, synchronises). */
int mgcm_rt_dust_tauref_host(const mgcm_rt_dust_columns *in, double *tauref, double *taudst);

/* DSOLFLUX (code/dsolflux.f:18-45) and COMP3's surface VIS flux for TEMPGR
 * (code/comp3.f:336-348) from the DETAU and DIFFVT outputs of mgcm_rt_run:
 * directsol[ncol], the direct solar flux at the surface (0 where the sun is
 * down), and dnvflux[ncol] = DIFFVT + directsol (optional; needs diffvt).  sol:
 * host SOL(7); acosz, detau [ncol][17][7], diffvt: device.  The Gauss weights
 * and FZEROV are the context's.  Asynchronous on stream. */
int mgcm_rt_dsolflux(mgcm_rt_ctx *ctx, int ncol, const double *sol, const double *acosz, const double *detau,
                     const double *diffvt, double *directsol, double *dnvflux, void *stream);
/* The same with host pointers (copies in, runs on the null stream, copies out,
 * synchronises): COMP3's per-step TEMPGR loop (code/comp3.f:311-351) over a
 * host-resident DETAU in one call. */
int mgcm_rt_dsolflux_host(mgcm_rt_ctx *ctx, int ncol, const double *sol, const double *acosz, const double *detau,
                          const double *diffvt, double *directsol, double *dnvflux);

/* Radiation diagnostics of the history file, from the outputs of mgcm_rt_run.
 * The REAL*4 fields COMP3 stores per column (code/comp3.f:654,703-706,745-747;
 * declared REAL*4 in code/modules.f90:510-513,984-986) and MHISTV writes as two
 * records, "fuptopv, fdntopv, fupsurfv, fdnsurfv" and "fuptopir, fupsurfir,
 * fdnsurfir", then "surfalb" (code/mhistv.f:200-202; read back by
 * tutorial/historymod.f90:93-95).  Each field is [ncol], rounded from FP64 to
 * nearest as the REAL*4 assignment does; with columns ordered J fastest
 * (col = (I-1)*L_J + J-1) a field is the history's (L_J, L_I) array.  Sun-down
 * columns get zero VIS fields and keep their surfalb entry, as in COMP3.
 * hr_nlte [ncol][nlay+1] (optional; needs p, sigma): the heating rate of hr with
 * COMP3's non-LTE factor on the solar part, SUNTOT(K+1)*2.2E4*PNLTE/(1+2.2E4*PNLTE),
 * PNLTE = SIGMA(K)*(P(J,I)-PTROP)+PTROP, K = 2l+2 (code/comp3.f:765-780): the
 * radiative temperature tendency COMP3 applies (QRAD times the theta factor), in
 * K/sol, the unit of the history's DHEAT (code/newstep.f:132).  All pointers
 * device (sigma: [nlev] device), asynchronous on stream. */
typedef struct {
  float *fuptopv, *fdntopv, *fupsurfv, *fdnsurfv; /* [ncol] */
  float *fuptopir, *fupsurfir, *fdnsurfir;        /* [ncol] */
  float *surfalb;                                 /* [ncol] ALS, optional */
  double *hr_nlte;                                /* [ncol][nlay+1] K/sol, optional */
  const double *p;                                /* [ncol] P(J,I) = surface pressure - PTROP, mbar */
  const double *sigma;                            /* [nlev] SIGMA(L_LEVELS) */
  double ptrop;                                   /* PTROP, mbar */
} mgcm_rt_history;
int mgcm_rt_history_fields(const mgcm_rt_columns *in, const mgcm_rt_fluxes *out, const mgcm_rt_history *hist,
                           void *stream);

/* Diagnostic: the device's exp, log10, sqrt and log on n host arguments (x > 0
 * for log10 and log, x >= 0 for sqrt; other entries return 0).  The FP64 path
 * reproduces the reference's libm (exp, log and log10 of glibc 2.35, correctly
 * rounded sqrt); this entry lets a host compare them bit for bit.  No
 * counterpart in the reference. */
int mgcm_rt_libm_check(const double *x, int n, double *exp_out, double *log10_out, double *sqrt_out,
                       double *log_out);

const char *mgcm_rt_strerror(int status);

#ifdef __cplusplus
}
#endif
#endif

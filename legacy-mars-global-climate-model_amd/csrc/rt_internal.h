// rt_internal.h -- shared definitions of the MI355X column radiation kernels.
//
// Bands are numbered 0..6 (VIS, L_NSPECTV) then 7..11 (IR, L_NSPECTI).  Gauss
// points g = 1..16 are the k-distribution points, g = 17 the continuum-only
// ("clear") channel (code/modules.f90:221-239).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgcm {

constexpr int NT = 7, NP = 51, NH = 10, NG = 17, NGAS = 16, NV = 7, NI = 5, NB = NV + NI;
constexpr int L_NREFV = 6;  // reference VIS band, 1-based (code/modules.f90:238)
constexpr int NPLANCK = 8501;
constexpr int MAX_NLAY = 99;  // DSOLVER: 2*NLAYRAD <= NMAX = 201 (code/dsolver.f90:12)

// code/modules.f90:152, :230-232, :333, :364-365, :127-143
constexpr double CMK = 3.51e22;
constexpr double MAXEXP = 35.0;
constexpr double UBARI = 0.5;
constexpr double TLIMITS = 1.0e-3;
constexpr double TLIMITI = 5.0e-3;
constexpr double PI = 3.1415926535897932;
constexpr double GRAV = 3.72;
constexpr double CPAIR = 7.3594e2;
constexpr double SCALEP = 1.0e2;
constexpr double SECONDS_PER_SOL = 88775.0;  // code/newstep.f:132

// REAL(4) literals of the reference, promoted to REAL(8).
#define F32LIT(x) ((double)(x##f))

// Small read-only tables; one copy per context in device memory, read with
// wave-uniform (scalar) loads.
struct RtConst {
  double pfgasref[NP];
  double tgasref[NT];
  double wrefh2o[NH];
  double gw[NB][NG];  // Gauss weights per band
  double fzero[NB];
  double tauray[NV];
  double dwni[NI];
  double qx_dst[NB], qs_dst[NB], g_dst[NB];
  double qx_cld[NB], qs_cld[NB], g_cld[NB];
  double qextref_vis;  // QXVDST(K, L_NREFV): VIS dust reference extinction
  double qextref_dst;  // QEXTREFDST: IR dust reference extinction
  double qextref_cld;  // QEXTREFCLD
  int pf_ascending;    // PFGASREF(2..L_PINT-1) does not decrease (TPINDEX's MP by counting)
};

// Per-run argument block (device pointers).
struct RunArgs {
  int ncol, nlay;
  const double *plev, *tlev, *pmid, *tmid, *qh2o, *tauref, *taurefcld;
  const double *als, *albi, *acosz;
  double sol[NV];
  double *fmnetv, *fluxupv, *fluxdnv, *fmneti, *fluxupi, *fluxdni;
  double *nfluxtopv, *nfluxtopi, *diffvt, *detau, *cumtauv, *hr;
  int *flag;   // range flag
  int bands;   // bit 0: VIS bands, bit 1: IR bands
  double *clr; // continuum-point records, [ncol][NB][2*nl+2] (workspace)
  // per-level aerosol optics of each column, or null for the tables' band
  // constants: QXVDST, QSVDST, GVDST, QXIDST, QSIDST, GIDST, then the same six
  // of the water-ice cloud, [ncol][nb][nlev+1]; QEXTREFDST, QEXTREFCLD [ncol][nlev+1]
  const double *aq[12];
  const double *qref_dst, *qref_cld;
  // DSOLFLUX's direct solar flux at the surface and DNVFLUX, [ncol], or null
  // (the batched path's DSOLFLUX kernel reads them here)
  double *directsol, *dnvflux;
  double *part; // one-band-per-wave tile: each band's row sums, [ncol][NB][3*nl+4] (workspace)
  double *tp;   // ... and TPINDEX's results, [ncol][5][2*nl] (workspace)
};

struct DevTables {
  const RtConst *c;
  const double *kv;      // [NV][NP][NH][NT][16]  k-coefficients, gas g fastest
  const double *ki;      // [NI][NP][NH][NT][16]
  const double *planck;  // [NPLANCK][NI]  PLANCKIR(5,8501)
  const float *kv32;     // FP32 copies for the mixed path
  const float *ki32;
};

__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }

// Offset of the gas-g row of cell (mt, mp, nh) (all 1-based) in a band's table.
__device__ __forceinline__ int kcell(int nw0, int mt, int mp, int nh) {
  return (((nw0 * NP + (mp - 1)) * NH + (nh - 1)) * NT + (mt - 1)) * NGAS;
}

}  // namespace mgcm

// rt_column.h -- the fused column kernel of the radiative transfer on gfx950,
// one launch per grid, instantiated for the plain FP64 path (rt_plain.hip) and
// for the mixed-precision path (rt_mixed.hip).
//
// One wave (64 lanes, four 16-lane rows) owns C = 4/S columns for the whole
// radiation call, S in {1, 2, 4} band slots per column (the tile: C columns x
// S bands x 16 Gauss points; S = 1 for large grids, more slots give small
// grids more waves; the smallest grids run each band of 4 columns in a wave of
// its own, TILE_BANDS, and add the bands in a finishing kernel):
//   1. stages the C columns' TPINDEX results (band independent) in LDS;
//   2. in rounds of S bands (the 7 VIS bands first, then the 5 IR bands; a
//      round is all VIS or all IR): row (c, s) stages band round*S+s of
//      column c -- continuum and layer optics -- then lane (c, s, g) solves
//      the (column, band, Gauss point g = 1..16) item -- OPTCV/OPTCI layer
//      assembly, delta-Eddington scaling, GFLUXV/GFLUXI with DSOLVER/DTRIDGL
//      -- and the 16 lanes of a row sum their weighted layer fluxes in a fixed
//      order; the column's slot-0 row adds its S rows' sums in band order to
//      per-column accumulators held in registers (lane g keeps the levels
//      L == g+1 mod 16), so every tile gives the same bits;
//   3. solves the continuum-only Gauss point 17 of every (column, band) -- the
//      VIS ones in one pass, the IR ones in a second -- weighted with FZERO
//      plus the weight of the band's TLIMIT-skipped points (code/sfluxv.f90:
//      72-75,121-162, code/sfluxi.f90:72-75,114-143);
//   4. writes the fluxes and forms SUNTOT/IRTOTAL and the layer heating
//      (code/comp3.f:677-681,735-738).
// Between the stages, only the continuum-point records (written by the
// continuum-point kernel, read in stage 3) go through global memory -- and, at
// one band per wave, the TPINDEX results and the band records; the per-item
// tridiagonal state lives in private (scratch) memory.
#pragma once
#include <type_traits>

#include "rt_item.h"

namespace mgcm {

// LDS staging of a wave: [4] is per column (at most 4) for the TPINDEX
// results and skip masks, per 16-lane row (the row's band) for the band optics.
struct Stage {
  double *t, *u, *wr, *uk;     // [4][nsub]  column: TPINDEX weights T (log p) and U (T), WRATIO, Cmk*dp
  double *c1, *c2;             // [4][nsub]  row: VIS: TRAY+TAEROS+TCLOUD; IR: TAEROS, TCLOUD
  double *scat, *cosb, *cbar;  // [4][nl]    row: layer scattering tau, asymmetry, COSBAR = g/(1+g)
  double *bk;                  // [4][nlev]  row: IR: PLANCKIR(nw, int(TLEV(K)*10-499))
  int *base;                   // [4][nsub]  column: table cell offset; bit 0 set when NH2O = L_REFH2O
  uint32_t *skip;              // [4][NB]    column: TLIMIT skip masks
};

// Dynamic LDS of one wave: the exp table, then the staging.
inline size_t stage_bytes(int nlay) {
  const size_t nl = nlay + 1, nlev = 2 * (size_t)nlay + 3, nsub = nlev - 1;
  return mgcm_libm::LDS_EXPTAB_BYTES + sizeof(double) * (4 * nsub * 6 + 4 * nl * 3 + 4 * nlev) +
         sizeof(int) * (4 * nsub + 4 * NB);
}

// Continuum-point records in RunArgs::clr: per column clr_rec(nl) slots of NB
// doubles, slot k holding value k of every band (band index fastest):
// FMIDP(1..nl), FMIDM(1..nl), then two top terms.  The lanes of a column write
// and read adjacent doubles.
__host__ __device__ inline int clr_rec(int nl) { return 2 * nl + 2; }
inline size_t clr_bytes(int ncol, int nlay) { return (size_t)ncol * NB * clr_rec(nlay + 1) * sizeof(double); }

// Band records in RunArgs::part (the one-band-per-wave tile, TILE_BANDS): per
// column and band, the band's 16-point sums of the net, upward and downward
// flux at levels 1..nl, then its top term (NFLUXTOPV or NFLUXTOPI part), its
// DIFFVT part and its TLIMIT skip mask; written by the band's wave, added in
// band order by the finishing kernel.
__host__ __device__ inline int part_rec(int nl) { return 3 * nl + 4; }
// ... and TPINDEX's results in RunArgs::tp, computed once per column for the 12
// band waves: per column T, U, WRATIO, Cmk*dp and the cell offset (as a double),
// each for the sub-levels 2..L_LEVELS.
constexpr int TP_FIELDS = 5;
inline size_t part_bytes(int ncol, int nlay) {
  return (size_t)ncol * (NB * part_rec(nlay + 1) + TP_FIELDS * (2 * nlay + 2)) * sizeof(double);
}

__device__ inline Stage carve(double *p, int nsub, int nl, int nlev) {
  p += mgcm_libm::LDS_EXPTAB_BYTES / sizeof(double);
  Stage S;
  S.t = p; p += 4 * nsub;
  S.u = p; p += 4 * nsub;
  S.wr = p; p += 4 * nsub;
  S.uk = p; p += 4 * nsub;
  S.c1 = p; p += 4 * nsub;
  S.c2 = p; p += 4 * nsub;
  S.scat = p; p += 4 * nl;
  S.cosb = p; p += 4 * nl;
  S.cbar = p; p += 4 * nl;
  S.bk = p; p += 4 * nlev;
  S.base = (int *)p;
  S.skip = (uint32_t *)(S.base + 4 * nsub);
  return S;
}

// A staging loop of a wave over t = lane, lane + 64, ... < n, four values of t
// at a time: ld(t) reads the inputs of all four before use(t, inputs) takes
// the first, so their memory latencies overlap.
template <typename Load, typename Use>
__device__ __forceinline__ void staged_loop(int lane, int n, Load &&ld, Use &&use) {
  using In = decltype(ld(0));
  for (int t0 = 0; t0 < n; t0 += 256) {
    In v[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int t = t0 + lane + 64 * i;
      if (t < n) v[i] = ld(t);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int t = t0 + lane + 64 * i;
      if (t < n) use(t, v[i]);
    }
  }
}

// TPINDEX (code/tpindex.f90:72-132) of a sub-level without the four bilinear
// products; they are formed where used, with the same expressions as
// code/tpindex.f90:110-113.
static __device__ void tp_stage(const RtConst *__restrict__ rc, double pw, double tw, double qh2o, double &t,
                         double &u, int &mt, int &mp, int &nh, double &wratio) {
  if (tw <= rc->tgasref[0]) {
    mt = 1;
  } else {
    mt = NT - 1;
    for (int n = 1; n <= NT - 1; n++) {
      if (tw > rc->tgasref[n - 1] && tw <= rc->tgasref[n]) { mt = n; break; }
    }
  }
  u = (tw - rc->tgasref[mt - 1]) / (rc->tgasref[mt] - rc->tgasref[mt - 1]);
  const double pwl = mgcm_libm::mgcm_log10(pw);
  if (rc->pf_ascending) {
    // MP = N-1 for the first N = 2..L_PINT-1 with PWL <= PFGASREF(N), else
    // L_PINT-1: on an ascending grid, one plus the number of N that fail the
    // test (a NaN fails them all, as in the scan)
    int fail = 0;
#pragma unroll
    for (int n = 2; n <= NP - 1; n++) fail += !(pwl <= rc->pfgasref[n - 1]);
    mp = 1 + fail;
  } else {
    mp = NP - 1;
    for (int n = 2; n <= NP - 1; n++) {
      if (pwl <= rc->pfgasref[n - 1]) { mp = n - 1; break; }
    }
  }
  t = (pwl - rc->pfgasref[mp - 1]) / (rc->pfgasref[mp] - rc->pfgasref[mp - 1]);
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

// Gas point source: k-interpolation (code/optcv.f90:142-167, code/optci.f90:
// 143-167) from the g-fastest table; the 16 lanes of a column read one 128-byte
// line per corner.
template <bool VIS, int P = 0>
struct GasSrc {
  using kt = typename std::conditional<(P & P_TAB32) != 0, float, double>::type;  // table storage type
  const kt *__restrict__ tab;      // band's table
  const double *t_, *u_, *wr_, *uk_, *c1_, *c2_;  // this column's (row's) staged sub-levels (LDS)
  const int *base_;
  double taugsurf;
  __device__ GasSrc(const kt *tab_, const Stage &S, int c, int row, int nsub, int g0)
      : tab(tab_ + g0), t_(S.t + c * nsub - 2), u_(S.u + c * nsub - 2), wr_(S.wr + c * nsub - 2),
        uk_(S.uk + c * nsub - 2), c1_(S.c1 + row * nsub - 2), c2_(S.c2 + row * nsub - 2),
        base_(S.base + c * nsub - 2), taugsurf(0.0) {}
  // the 8 corners (p0,T0),(p1,T0),(p1,T1),(p0,T1) x (NH2O, NH2O+1) of sub-level K;
  // issued a layer ahead of their use (fetch), combined in K order (finish)
  struct Raw {
    kt a0, a1, b0, b1, c0, c1, d0, d1;
  };
  __device__ __forceinline__ Raw fetch(int K) const {
    const int braw = base_[K];
    const int dnh = (braw & 1) ? 0 : NT * NGAS;  // NH2O = L_REFH2O: WRATIO = 0, upper = itself
    const kt *p = tab + (braw & ~1);
    const int sT = NGAS, sP = NH * NT * NGAS;
    return Raw{p[0], p[dnh], p[sP], p[sP + dnh], p[sP + sT], p[sP + sT + dnh], p[sT], p[sT + dnh]};
  }
  // TPINDEX-weighted combination of the corners (code/optcv.f90:146-160)
  template <class R>
  __device__ __forceinline__ R taugas_of(int K, R a0, R a1, R b0, R b1, R c0, R c1, R d0, R d1) const {
    const R t = (R)t_[K], u = (R)u_[K], wr = (R)wr_[K];
    const R one = 1;
    const R lc0 = (one - t) * (one - u), lc1 = t * (one - u), lc2 = t * u, lc3 = (one - t) * u;
    const R k1 = a0 + wr * (a1 - a0);
    const R k2 = b0 + wr * (b1 - b0);
    const R k3 = c0 + wr * (c1 - c0);
    const R k4 = d0 + wr * (d1 - d0);
    const R ans = lc0 * k1 + lc1 * k2 + lc2 * k3 + lc3 * k4;
    return (R)uk_[K] * ans;
  }
  // DTAUK(K) = gas + continuum (code/optcv.f90:160-167, code/optci.f90:160-167):
  // FP64 unless the policy moves the interpolation (P_INTERP32) or the
  // resulting optical depth (P_TAU32) to FP32
  __device__ __forceinline__ double finish(int K, const Raw &r) {
    double taugas;
    if (P & P_INTERP32)
      taugas = (double)taugas_of<float>(K, (float)r.a0, (float)r.a1, (float)r.b0, (float)r.b1, (float)r.c0,
                                        (float)r.c1, (float)r.d0, (float)r.d1);
    else
      taugas = taugas_of<double>(K, r.a0, r.a1, r.b0, r.b1, r.c0, r.c1, r.d0, r.d1);
    taugsurf = taugsurf + taugas;
    const double d = VIS ? taugas + c1_[K] : taugas + c1_[K] + c2_[K];
    return (P & P_TAU32) ? (double)(float)d : d;
  }
  // TAUGSURF of the FP64 table tab64 (this lane's g), summed in the order of
  // the plain path: the mixed path settles TLIMIT decisions near the threshold
  // with it, so that both paths skip the same Gauss points.
  __device__ double taugsurf_fp64(const double *__restrict__ tab64, int nlev) const {
    double sum = 0.0;
    for (int K = 2; K <= nlev; K++) {
      const int braw = base_[K];
      const int dnh = (braw & 1) ? 0 : NT * NGAS;
      const double *p = tab64 + (braw & ~1);
      const int sT = NGAS, sP = NH * NT * NGAS;
      sum = sum + taugas_of<double>(K, p[0], p[dnh], p[sP], p[sP + dnh], p[sP + sT], p[sP + sT + dnh], p[sT],
                                    p[sT + dnh]);
    }
    return sum;
  }
};

// TLIMIT test of a gas point (code/sfluxv.f90:72, code/sfluxi.f90:72).  The
// mixed path's sum over FP32 table values is within a relative 1e-6 of the
// FP64-table one; closer than 1e-4 to the limit it is recomputed from the FP64
// table by the lanes that are that close (only active lanes: the staged cell
// offsets of an inactive one, past the grid's last column, are never written),
// so both paths skip the same Gauss points.
template <int P, class Src>
__device__ __forceinline__ bool below_tlimit(const Src &src, const double *tab64, int nlev, double tlim,
                                             bool active) {
  double tg = src.taugsurf;
  if (P & (P_TAB32 | P_INTERP32)) {
    const bool near = active && fabs(tg - tlim) <= 1.0e-4 * tlim;
    if (near) tg = src.taugsurf_fp64(tab64, nlev);
  }
  return tg < tlim;
}

// Continuum-only point: sub-level optical depths straight from the column's
// profiles (code/optcv.f90:172-173, code/optci.f90:173-174).  A sub-level's
// profile values are fetched a layer ahead, as the gas points' table corners
// are; the aerosol depths of the last odd and even sub-level are kept for the
// layer optics (ClearLayerAer), which would otherwise compute them again.
template <bool VIS>
struct ClearSrc {
  const Aer &ae;
  int b;
  const double *plev, *tauref, *taurefcld;
  double taugsurf;
  SubAer last[2];  // [K & 1]
  struct Raw {
    double pk, pk1, tr, tc;  // PLEV(K), PLEV(K-1), TAUREF(K), TAUREFCLD(K)
  };
  __device__ __forceinline__ Raw fetch(int K) const {
    return Raw{plev[K - 1], plev[K - 2], tauref[K - 1], taurefcld[K - 1]};
  }
  __device__ __forceinline__ double finish(int K, const Raw &r) {
    const SubAer s = sub_aerosol_of<VIS>(ae, b, K, r.pk, r.pk1, r.tr, r.tc);
    last[K & 1] = s;
    if (VIS) return s.taeros + s.tray + s.tcloud;
    return s.taeros + s.tcloud;
  }
};

// Layer aerosol optics of a continuum point from the sub-levels 2L+1 and 2L+2
// its source has just finished (layer_aerosol's values), keeping what the
// solve needs later: COSBAR (VIS) or COSB (IR) of each layer in keep[L-1].
template <bool VIS>
struct ClearLayerAer {
  const ClearSrc<VIS> &src;
  int nl;
  double *keep;
  __device__ __forceinline__ void get(int L, double &scat, double &cosb) const {
    double tlk2 = 0.0, tck2 = 0.0, tray2 = 0.0;
    if (L < nl) {
      const SubAer &a2 = src.last[0];  // K = 2L+2
      tlk2 = a2.tlk;
      tck2 = a2.tck;
      tray2 = a2.tray;
    } else if (!VIS) {
      ir_pad(src.ae, src.b, src.tauref, src.taurefcld, 2 * nl + 1, tlk2, tck2);
    }
    layer_aerosol_of<VIS>(src.ae, src.b, nl, L, src.last[1], tlk2, tck2, tray2, scat, cosb);
    keep[L - 1] = VIS ? cosb / (1.0 + cosb) : cosb;  // COSBAR of code/gfluxv.f90:86
  }
  __device__ __forceinline__ double scat(int L) const {
    double sc, cb;
    get(L, sc, cb);
    return sc;
  }
};

// Cross-lane move of a double inside 16-lane rows (DPP, no LDS round trip).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// 16-lane fixed-order sum over a row: pairs (lane, lane^1), then quads
// (lane^2), then the two quads of each half (half-mirror), then the two halves
// (mirror).  Each step adds two equal-order partial sums, so every lane of the
// row ends with the same bits.  Needs all 64 lanes active.
__device__ __forceinline__ double group16_sum(double v) {
  v = v + dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v = v + dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v = v + dpp_f64<0x141>(v);  // row_half_mirror
  v = v + dpp_f64<0x140>(v);  // row_mirror
  return v;
}

__device__ __forceinline__ int planck_index(double T, int *flag) {
  int nt = (int)(T * 10.0 - 499);  // code/sfluxi.f90:54-55, code/gfluxi.f90:73-74
  if (nt < 1 || nt > NPLANCK) {
    atomicOr(flag, 1);
    nt = nt < 1 ? 1 : NPLANCK;
  }
  return nt;
}

// Per-column accumulators, distributed over the column's 16 lanes: lane g owns
// radiation levels n with (n-1) % 16 == g.
template <int SL>
struct Acc {
  double vnet[SL], vup[SL], vdn[SL], inet[SL], iup[SL], idn[SL];
};

// Branch-free: the owner lane's slot takes the sum, every other value is kept.
template <int SL>
__device__ __forceinline__ void acc_add(double *arr, int n, int g0, double v) {
  const int slot = (n - 1) >> 4;
  const bool own = ((n - 1) & 15) == g0;
#pragma unroll
  for (int j = 0; j < SL; j++) arr[j] = (own && j == slot) ? arr[j] + v : arr[j];
}

// The row sums of a column's S band slots (rows 16 apart, bands in increasing
// order) added one after the other, as band after band: on the column's slot-0
// row the result is ((acc + v_0) + v_1) + ...; the other rows' results are
// not used.  The reads are wave-wide (every lane takes part).
template <int S>
__device__ __forceinline__ double tile_sum(double acc, double v) {
  acc = acc + v;
#pragma unroll
  for (int j = 1; j < S; j++) acc = acc + __shfl(v, (int)threadIdx.x + 16 * j);
  return acc;
}

template <int S, int SL>
__device__ __forceinline__ void tile_add(double *arr, int n, int g0, double v) {
  acc_add<SL>(arr, n, g0, v);
#pragma unroll
  for (int j = 1; j < S; j++) acc_add<SL>(arr, n, g0, __shfl(v, (int)threadIdx.x + 16 * j));
}

// Band rounds of the tile with S slots per column: VIS rounds, then IR rounds.
template <int S>
struct Tile {
  static_assert(S == 1 || S == 2 || S == 4, "band slots per column");
  static constexpr int C = 4 / S;                        // columns per wave
  static constexpr int RV = (NV + S - 1) / S, RI = (NI + S - 1) / S;
  // band of slot s in round k, or -1 past the last band of the part
  __host__ __device__ static constexpr int band(int k, int s) {
    return k < RV ? (k * S + s < NV ? k * S + s : -1) : ((k - RV) * S + s < NI ? NV + (k - RV) * S + s : -1);
  }
};

// Tile code of one band per wave: the 4 columns of a wave at one band slot
// each, each of the NB band rounds in its own wave (column_block with ONE),
// the bands' sums added by a finishing kernel.
constexpr int TILE_BANDS = NB;

#ifndef MGCM_FP64_WAVES
#define MGCM_FP64_WAVES 2
#endif
// auto_tile's latency of a one-band wave and cost of the finishing launch, in
// hundredths of an S = 1 wave's latency
#ifndef MGCM_BANDS_LATENCY
#define MGCM_BANDS_LATENCY 12
#endif
#ifndef MGCM_BANDS_FINISH
#define MGCM_BANDS_FINISH 4
#endif

// Stages 3 and 4 of a wave's columns (the column's lead row; lane g0 holds the
// levels n with (n-1) % 16 == g0).
template <int SL>
__device__ __forceinline__ void column_tail(const DevTables &dt, const RunArgs &a, const Stage &St, Acc<SL> &acc,
                                            double topv, double difv, double topi, int c, int col, int colc,
                                            bool valid, bool lead, bool sun, int g0, int nl,
                                            const double *plev) {
  const RtConst *rc = dt.c;
  // 3. continuum point g = 17 of every (column, band): solved by k_rt_clear
  //    into a.clr, weighted here with FZERO plus the weights of the band's
  //    TLIMIT-skipped points, in g order, and added band by band
  //    (code/sfluxv.f90:72-75,121-162, code/sfluxi.f90:72-75,114-143); by the
  //    column's slot-0 row
  if (lead) {
    double wz[NB];
    for (int b = 0; b < NB; b++) {
      double fzero = rc->fzero[b];
      if (!(fzero >= F32LIT(0.99))) {
        const uint32_t mk = St.skip[c * NB + b];
        for (int g = 0; g < NGAS; g++)
          if (mk & (1u << g)) fzero = fzero + (1.0 - rc->fzero[b]) * rc->gw[b][g];
      }
      wz[b] = fzero;
    }
    const int rec = clr_rec(nl);
    const double *cr = a.clr + (size_t)colc * NB * rec;
    if (a.bands & 1) {
      for (int n = g0 + 1; n <= nl; n += 16) {  // this lane's levels
        double sn = 0.0, su = 0.0, sd = 0.0;
        for (int b = 0; b < NV; b++) {
          const double fp = cr[(n - 1) * NB + b], fm = cr[(nl + n - 1) * NB + b];
          sn = sn + (sun ? (fp - fm) * wz[b] : 0.0);
          su = su + (sun ? fp * wz[b] : 0.0);
          sd = sd + (sun ? fm * wz[b] : 0.0);
        }
        acc_add<SL>(acc.vnet, n, g0, sn);
        acc_add<SL>(acc.vup, n, g0, su);
        acc_add<SL>(acc.vdn, n, g0, sd);
      }
      double st = 0.0, sdf = 0.0;
      for (int b = 0; b < NV; b++) {
        st = st + (sun ? cr[2 * nl * NB + b] * wz[b] : 0.0);
        sdf = sdf + (sun ? cr[(2 * nl + 1) * NB + b] * wz[b] : 0.0);
      }
      topv = topv + st;
      difv = difv + sdf;
    }
    if (a.bands & 2) {
      for (int n = g0 + 1; n <= nl; n += 16) {  // this lane's levels
        double sn = 0.0, su = 0.0, sd = 0.0;
        for (int b = NV; b < NB; b++) {
          const double dw = rc->dwni[b - NV];
          const double fp = cr[(n - 1) * NB + b], fm = cr[(nl + n - 1) * NB + b];
          sn = sn + (valid ? (fp - fm) * dw * wz[b] : 0.0);
          su = su + (valid ? fp * dw * wz[b] : 0.0);
          sd = sd + (valid ? fm * dw * wz[b] : 0.0);
        }
        acc_add<SL>(acc.inet, n, g0, sn);
        acc_add<SL>(acc.iup, n, g0, su);
        acc_add<SL>(acc.idn, n, g0, sd);
      }
      double st = 0.0;
      for (int b = NV; b < NB; b++) st = st + (valid ? cr[2 * nl * NB + b] * rc->dwni[b - NV] * wz[b] : 0.0);
      topi = topi + st;
    }
  }

  // 4. outputs, net-flux differences and heating (slot-0 rows)
  __syncthreads();
  double *fv = St.t;  // reuse the staging: [C][nl] net VIS, [C][nl] net IR
  double *fi = St.t + 4 * nl;
  if (!sun) { topv = 0.0; difv = 0.0; }
  if (lead) {
#pragma unroll
    for (int j = 0; j < SL; j++) {  // this lane's levels n = 16 j + g0 + 1
      const int n = 16 * j + g0 + 1;
      if (n <= nl) {
        const double nv = sun ? acc.vnet[j] : 0.0, uv = sun ? acc.vup[j] : 0.0, dv = sun ? acc.vdn[j] : 0.0;
        const double ni = acc.inet[j], ui = acc.iup[j], di = acc.idn[j];
        fv[c * nl + n - 1] = nv;
        fi[c * nl + n - 1] = ni;
        if (valid) {
          const size_t o = (size_t)col * nl + n - 1;
          a.fmnetv[o] = nv; a.fluxupv[o] = uv; a.fluxdnv[o] = dv;
          a.fmneti[o] = ni; a.fluxupi[o] = ui; a.fluxdni[o] = di;
        }
      }
    }
    if (g0 == 0 && valid) {
      a.nfluxtopv[col] = topv;
      a.nfluxtopi[col] = topi;
      a.diffvt[col] = difv;
    }
  }
  __syncthreads();
  if (a.hr && valid && lead) {
    for (int l = g0; l <= a.nlay; l += 16) {
      const int L = l + 1;  // SUNTOT(2L+1), IRTOTAL(2L+1)
      const double nv = fv[c * nl + L - 1], ni = fi[c * nl + L - 1];
      const double sunt = sun ? (L == 1 ? nv - topv : nv - fv[c * nl + L - 2]) : 0.0;
      const double irt = (L == 1) ? ni - topi : ni - fi[c * nl + L - 2];
      const double dp = (l == 0) ? (plev[0] + plev[1]) : (plev[2 * l + 2] - plev[2 * l]);
      a.hr[(size_t)col * nl + l] = (((sunt + irt) * GRAV) / ((CPAIR * SCALEP) * dp)) * SECONDS_PER_SOL;
    }
  }
}

// TPINDEX of one sub-level from the column's profiles: the inputs ...
struct TpIn {
  double pm, tm, qw, p0, p1, tr, tc;  // PMID, TMID, QH2O, PLEV(K-1), PLEV(K), TAUREF, TAUREFCLD at K
};
__device__ __forceinline__ TpIn tp_inputs(const RunArgs &a, int cl, int K, int nlev) {
  const size_t o = (size_t)cl * nlev + K - 1, oa = (size_t)cl * (nlev + 1) + K - 1;
  return TpIn{a.pmid[o], a.tmid[o], a.qh2o[o], a.plev[o - 1], a.plev[o], a.tauref[oa], a.taurefcld[oa]};
}
// ... and the results: T and U weights, WRATIO, Cmk*dp and the table cell
// offset (bit 0 set when NH2O = L_REFH2O).  A NaN among the column's inputs
// sets the range flag: the values this sub-level reads, plus (once per column)
// TAUREF/TAUREFCLD at levels 1 and L_LEVELS+1 and the column scalars (TLEV
// reaches the flag through the Planck index).
struct TpOut {
  double t, u, wr, uk;
  int base;
};
__device__ __forceinline__ TpOut tp_sublevel(const RtConst *rc, const RunArgs &a, int cl, int K, int nlev,
                                             const TpIn &v) {
  TpOut r;
  int mt, mp, nh;
  tp_stage(rc, v.pm, v.tm, v.qw, r.t, r.u, mt, mp, nh, r.wr);
  r.uk = CMK * (v.p1 - v.p0);
  r.base = kcell(0, mt, mp, nh) | (nh == NH ? 1 : 0);
  bool nan = isnan(v.pm) || isnan(v.tm) || isnan(v.qw) || isnan(v.p0) || isnan(v.p1) || isnan(v.tr) || isnan(v.tc);
  if (K == 2) {
    const size_t oa = (size_t)cl * (nlev + 1) + K - 1, nsub = nlev - 1;
    nan = nan || isnan(a.tauref[oa - 1]) || isnan(a.taurefcld[oa - 1]) || isnan(a.tauref[oa + nsub]) ||
          isnan(a.taurefcld[oa + nsub]) || isnan(a.acosz[cl]) || isnan(a.als[cl]) || isnan(a.albi[cl]);
  }
  if (nan) atomicOr(a.flag, 1);
  return r;
}

// P: the precision policy (P_* bits; 0 = plain FP64, P_TAB32 = the mixed path,
// whose k-tables are read in FP32 -- half the gather bytes and corner
// registers -- with all arithmetic, the interpolation included, in FP64).
// S: band slots per column (the tile); the block owns Tile<S>::C columns.
// The whole radiation call of columns col0..col0+C-1 of grid a, by one wave.
// The exp table must be in LDS; the staging after it is rewritten.
// D: how many layers ahead the solves load the item state (vis_solve).
// ONE (S = 1 only): round kr alone, its band's sums written to a.part for
// column_finish instead of stages 3 and 4.
template <int NLM, int P, int S, int D = 1, bool ONE = false>
__device__ __forceinline__ void column_block(const DevTables &dt, const RunArgs &a, int col0, double *smem,
                                             int kr = 0) {
  static_assert(!ONE || S == 1, "one band per wave: one column per row");
  constexpr bool T32 = (P & P_TAB32) != 0;
  using kt = typename std::conditional<T32, float, double>::type;
  using TL = Tile<S>;
  constexpr int C = TL::C;
  constexpr int SL = (NLM + 15) / 16;
  if (col0 >= a.ncol) return;  // a batched sub-item past the grid's last column (block-uniform)
  const int nl = a.nlay + 1, nlev = 2 * a.nlay + 3, nsub = nlev - 1;
  const int lane = threadIdx.x, row = lane >> 4, g0 = lane & 15;
  const int c = row / S, slot = row % S;  // this row's column in the block and band slot
  const bool lead = slot == 0;            // the row that holds the column's sums
  const int col = col0 + c;
  const bool valid = col < a.ncol;
  const int colc = valid ? col : a.ncol - 1;
  const RtConst *rc = dt.c;
  Stage St = carve(smem, nsub, nl, nlev);

  // 1. TPINDEX of every sub-level K = 2..L_LEVELS of the C columns (one band
  //    per wave: read from the results k_tp wrote once per column)
  if constexpr (ONE) {
    struct TpRes {
      double v[TP_FIELDS];
    };
    staged_loop(
        lane, C * nsub,
        [&](int t) {
          const int cl = col0 + t / nsub, i = t % nsub;
          TpRes r{};
          if (cl < a.ncol) {
            const double *q = a.tp + (size_t)cl * TP_FIELDS * nsub + i;
#pragma unroll
            for (int f = 0; f < TP_FIELDS; f++) r.v[f] = q[f * nsub];
          }
          return r;
        },
        [&](int t, const TpRes &r) {
          if (col0 + t / nsub >= a.ncol) return;
          St.t[t] = r.v[0];
          St.u[t] = r.v[1];
          St.wr[t] = r.v[2];
          St.uk[t] = r.v[3];
          St.base[t] = (int)r.v[4];
        });
  } else {
    staged_loop(
        lane, C * nsub,
        [&](int t) {
          const int cl = col0 + t / nsub, K = t % nsub + 2;
          return cl < a.ncol ? tp_inputs(a, cl, K, nlev) : TpIn{};
        },
        [&](int t, const TpIn &v) {
          const int cl = col0 + t / nsub, K = t % nsub + 2;
          if (cl >= a.ncol) return;
          const TpOut r = tp_sublevel(rc, a, cl, K, nlev, v);
          St.t[t] = r.t;
          St.u[t] = r.u;
          St.wr[t] = r.wr;
          St.uk[t] = r.uk;
          St.base[t] = r.base;
        });
  }

  Acc<SL> acc;
#pragma unroll
  for (int j = 0; j < SL; j++) acc.vnet[j] = acc.vup[j] = acc.vdn[j] = acc.inet[j] = acc.iup[j] = acc.idn[j] = 0.0;
  double topv = 0.0, difv = 0.0, topi = 0.0;

  const double *plev = a.plev + (size_t)colc * nlev;
  const double u0 = a.acosz[colc];
  const bool sun = u0 >= F32LIT(1.0e-5);  // code/comp3.f:637

  // band of row rr in round k, or -1 where the row has none or the band is
  // not solved: FZERO >= 0.99 (code/sfluxv.f90:64-65, code/sfluxi.f90:67-68)
  // or a part the caller skipped (it contributes nothing)
  auto row_band = [&](int k, int rr) {
    const int b = TL::band(k, rr % S);
    if (b < 0) return -1;
    if (rc->fzero[b] >= F32LIT(0.99) || !(a.bands & (b < NV ? 1 : 2))) return -1;
    return b;
  };

  // 2. k-distribution points, S bands per column per round
  const int k_end = ONE ? kr + 1 : TL::RV + TL::RI;
  for (int k = ONE ? kr : 0; k < k_end; k++) {
    const bool VISR = k < TL::RV;
    const int bnom = TL::band(k, slot);  // this row's band slot (-1: none)
    const int bact = row_band(k, row);   // ... if solved
    const int b = bact >= 0 ? bact : (VISR ? 0 : NV);  // a band index to read constants with
    __syncthreads();
    if (!__any(bact >= 0)) {  // no row solves a band this round
      if (g0 == 0 && bnom >= 0) St.skip[c * NB + bnom] = 0u;
      continue;
    }
    for (int t = lane; t < 4 * nsub; t += 64) {  // continuum optics of each row's band
      const int rr = t / nsub, K = t % nsub + 2;
      const int cl = col0 + rr / S, bb = row_band(k, rr);
      if (cl >= a.ncol || bb < 0) continue;
      const SubIn v = sub_inputs(a.plev + (size_t)cl * nlev, a.tauref + (size_t)cl * (nlev + 1),
                                 a.taurefcld + (size_t)cl * (nlev + 1), K);
      const Aer ae = column_aerosol(a, rc, cl);
      if (VISR) {
        const SubAer s = sub_aerosol_of<true>(ae, bb, K, v.pk, v.pk1, v.trk, v.tck);
        St.c1[t] = s.tray + s.taeros + s.tcloud;
      } else {
        const SubAer s = sub_aerosol_of<false>(ae, bb, K, v.pk, v.pk1, v.trk, v.tck);
        St.c1[t] = s.taeros;
        St.c2[t] = s.tcloud;
      }
    }
    for (int t = lane; t < 4 * nl; t += 64) {
      const int rr = t / nl, L = t % nl + 1;
      const int cl = col0 + rr / S, bb = row_band(k, rr);
      if (cl >= a.ncol || bb < 0) continue;
      const LayerIn v = layer_inputs(a.plev + (size_t)cl * nlev, a.tauref + (size_t)cl * (nlev + 1),
                                     a.taurefcld + (size_t)cl * (nlev + 1), nlev, L);
      const Aer ae = column_aerosol(a, rc, cl);
      double sc, cb;
      if (VISR) layer_aerosol<true>(ae, bb, nl, L, v, sc, cb);
      else layer_aerosol<false>(ae, bb, nl, L, v, sc, cb);
      St.scat[t] = sc;
      St.cosb[t] = cb;
      St.cbar[t] = cb / (1.0 + cb);  // COSBAR of the delta-Eddington scaling (code/gfluxv.f90:86)
    }
    if (!VISR) {
      for (int t = lane; t < 4 * nlev; t += 64) {
        const int rr = t / nlev, K = t % nlev + 1;
        const int cl = col0 + rr / S, bb = row_band(k, rr);
        if (cl >= a.ncol || bb < 0) continue;
        const int nt = planck_index(a.tlev[(size_t)cl * nlev + K - 1], a.flag);
        St.bk[t] = dt.planck[(size_t)(nt - 1) * NI + (bb - NV)];
      }
    }
    __syncthreads();

    const double fz = rc->fzero[b];
    const double gw = rc->gw[b][g0];
    const double onemf = 1.0 - fz;
    const double *scat = St.scat + row * nl, *cosb = St.cosb + row * nl, *cbar = St.cbar + row * nl;
    bool skip;
    if (VISR) {
      const int nw = b;
      const size_t toff = (size_t)nw * NP * NH * NT * NGAS;
      GasSrc<true, P> src(T32 ? (const kt *)(dt.kv32 + toff) : (const kt *)(dt.kv + toff), St, c, row, nsub, g0);
      VisState<NLM> s;
      double cumtau = 0.0, detau = 0.0;
      const bool go = valid && sun && bact >= 0;
      if (go) detau = vis_optics<NLM>(nl, src, StagedLayerAer{scat, cosb}, s, cumtau);
      skip = !go || below_tlimit<P>(src, dt.kv + toff + g0, nlev, TLIMITS, go);
      if (go) {  // GETDETAU output and TAUCUMV(L_LEVELS) (code/comp3.f:664-668)
        a.detau[(size_t)col * NG * NV + g0 * NV + nw] = detau;
        a.cumtauv[(size_t)col * NG * NV + g0 * NV + nw] = cumtau;
      }
      // the 16-lane sums need the whole wave: the solve runs wave-uniformly and
      // skipped lanes (TLIMIT, sun down, past the last column, no band) contribute zeros
      if (__any(!skip)) {
        auto accf = [&](int n, double fp, double fm) {
          const double net = skip ? 0.0 : (fp - fm) * gw * onemf;
          const double up = skip ? 0.0 : fp * gw * onemf;
          const double dn = skip ? 0.0 : fm * gw * onemf;
          tile_add<S, SL>(acc.vnet, n, g0, group16_sum(net));
          tile_add<S, SL>(acc.vup, n, g0, group16_sum(up));
          tile_add<S, SL>(acc.vdn, n, g0, group16_sum(dn));
        };
        double fluxup = 0.0, fluxdn = 0.0, diffv = 0.0;
        const double bsurf = a.als[colc] * u0 * a.sol[nw] * rt_exp35(-dmin(detau / u0, MAXEXP));
        vis_solve<NLM, P, D>(nl, s, cbar, u0, a.sol[nw], a.als[colc], 0.0, bsurf, accf, fluxup, fluxdn, diffv);
        topv = tile_sum<S>(topv, group16_sum(skip ? 0.0 : (fluxup - fluxdn) * gw * onemf));
        difv = tile_sum<S>(difv, group16_sum(skip ? 0.0 : diffv * gw * onemf));
      }
    } else {
      const int nw = b - NV;
      const double dw = rc->dwni[nw];
      const size_t toff = (size_t)nw * NP * NH * NT * NGAS;
      GasSrc<false, P> src(T32 ? (const kt *)(dt.ki32 + toff) : (const kt *)(dt.ki + toff), St, c, row, nsub, g0);
      IrState<NLM> s;
      const bool go = valid && bact >= 0;
      if (go) ir_optics<NLM>(nl, src, StagedLayerAer{scat, cosb}, s);
      skip = !go || below_tlimit<P>(src, dt.ki + toff + g0, nlev, TLIMITI, go);
      if (__any(!skip)) {
        const double *bk = St.bk + row * nlev;
        auto accf = [&](int n, double fp, double fm) {
          const double net = skip ? 0.0 : (fp - fm) * dw * gw * onemf;
          const double up = skip ? 0.0 : fp * dw * gw * onemf;
          const double dn = skip ? 0.0 : fm * dw * gw * onemf;
          tile_add<S, SL>(acc.inet, n, g0, group16_sum(net));
          tile_add<S, SL>(acc.iup, n, g0, group16_sum(up));
          tile_add<S, SL>(acc.idn, n, g0, group16_sum(dn));
        };
        // boundary conditions (code/sfluxi.f90:61-62,81-82)
        const double rsf = a.albi[colc];
        const double bsurf = (1. - rsf) * bk[nlev - 1];
        const double tautop = s.wd[1].y * plev[1] / (plev[3] - plev[1]);
        const double btop = (1.0 - rt_exp(-tautop * 2.0)) * bk[1];  // /UBARI, UBARI = 0.5
        const double ftopup = ir_solve<NLM, P, D>(nl, s, cosb, bk, rsf, btop, bsurf, accf);
        topi = tile_sum<S>(topi, group16_sum(skip ? 0.0 : ftopup * dw * gw * onemf));
      }
    }
    const unsigned long long m = __ballot(skip);
    if (g0 == 0 && bnom >= 0) St.skip[c * NB + bnom] = bact >= 0 ? (uint32_t)((m >> (16 * row)) & 0xffffu) : 0u;
  }
  __syncthreads();
  if constexpr (ONE) {
    // the band's record: its sums added to zero (0 + v is v, or +0 for v = -0),
    // so the finishing kernel's sums in band order, which start from +0 and so
    // are never -0, equal the one-wave kernel's bit for bit
    const int b = TL::band(kr, 0);
    if (valid) {
      double *r = a.part + ((size_t)col * NB + b) * part_rec(nl);
      const bool vis = b < NV;
#pragma unroll
      for (int j = 0; j < SL; j++) {
        const int n = 16 * j + g0 + 1;
        if (n <= nl) {
          r[n - 1] = vis ? acc.vnet[j] : acc.inet[j];
          r[nl + n - 1] = vis ? acc.vup[j] : acc.iup[j];
          r[2 * nl + n - 1] = vis ? acc.vdn[j] : acc.idn[j];
        }
      }
      if (g0 == 0) {
        r[3 * nl] = vis ? topv : topi;
        r[3 * nl + 1] = vis ? difv : 0.0;
        r[3 * nl + 2] = (double)St.skip[c * NB + b];
      }
    }
  } else {
    column_tail<SL>(dt, a, St, acc, topv, difv, topi, c, col, colc, valid, lead, sun, g0, nl, plev);
  }
}

// Stages 3 and 4 of the one-band-per-wave tile, for the columns
// col0..col0+3 (one per 16-lane row, as at S = 1): the band records of
// column_block<..., ONE = true> added in band order, then column_tail.
template <int NLM>
__device__ __forceinline__ void column_finish(const DevTables &dt, const RunArgs &a, int col0, double *smem) {
  constexpr int SL = (NLM + 15) / 16;
  const int nl = a.nlay + 1, nlev = 2 * a.nlay + 3, nsub = nlev - 1;
  const int lane = threadIdx.x, c = lane >> 4, g0 = lane & 15;
  const int col = col0 + c;
  const bool valid = col < a.ncol;
  const int colc = valid ? col : a.ncol - 1;
  Stage St = carve(smem, nsub, nl, nlev);
  const double *plev = a.plev + (size_t)colc * nlev;
  const bool sun = a.acosz[colc] >= F32LIT(1.0e-5);  // code/comp3.f:637
  Acc<SL> acc;
#pragma unroll
  for (int j = 0; j < SL; j++) acc.vnet[j] = acc.vup[j] = acc.vdn[j] = acc.inet[j] = acc.iup[j] = acc.idn[j] = 0.0;
  double topv = 0.0, difv = 0.0, topi = 0.0;
  const int pr = part_rec(nl);
  const double *rc0 = a.part + (size_t)colc * NB * pr;
#pragma unroll
  for (int b = 0; b < NV; b++) {
    const double *r = rc0 + b * pr;
#pragma unroll
    for (int j = 0; j < SL; j++) {
      const int n = 16 * j + g0 + 1;
      if (n <= nl) {
        acc.vnet[j] = acc.vnet[j] + r[n - 1];
        acc.vup[j] = acc.vup[j] + r[nl + n - 1];
        acc.vdn[j] = acc.vdn[j] + r[2 * nl + n - 1];
      }
    }
    topv = topv + r[3 * nl];
    difv = difv + r[3 * nl + 1];
    if (g0 == 0) St.skip[c * NB + b] = (uint32_t)r[3 * nl + 2];
  }
#pragma unroll
  for (int b = NV; b < NB; b++) {
    const double *r = rc0 + b * pr;
#pragma unroll
    for (int j = 0; j < SL; j++) {
      const int n = 16 * j + g0 + 1;
      if (n <= nl) {
        acc.inet[j] = acc.inet[j] + r[n - 1];
        acc.iup[j] = acc.iup[j] + r[nl + n - 1];
        acc.idn[j] = acc.idn[j] + r[2 * nl + n - 1];
      }
    }
    topi = topi + r[3 * nl];
    if (g0 == 0) St.skip[c * NB + b] = (uint32_t)r[3 * nl + 2];
  }
  __syncthreads();
  column_tail<SL>(dt, a, St, acc, topv, difv, topi, c, col, colc, valid, true, sun, g0, nl, plev);
}

// The continuum-only point g = 17 of (column, band) items, one per lane: a
// VIS wave takes 9 columns x 7 bands, an IR wave 12 columns x 5 bands.  Each
// item writes its unweighted mid-layer fluxes FMIDP/FMIDM and its top terms
// (VIS: FLUXUP-FLUXDN and DIFFV, IR: FTOPUP) to its record in a.clr; the
// column kernel weights and adds them.  The exp table must be in LDS.
template <bool VIS>
struct ClearShape {
  static constexpr int nb = VIS ? NV : NI, cols = VIS ? 9 : 12;
};

template <int NLM, bool VIS>
__device__ __forceinline__ void clear_block(const DevTables &dt, const RunArgs &a, int col0) {
  constexpr int nbp = ClearShape<VIS>::nb, per = ClearShape<VIS>::cols;
  const int lane = threadIdx.x, cc = lane / nbp, bb = lane % nbp;
  const int col = col0 + cc;
  if (cc >= per || col >= a.ncol) return;  // no cross-lane work below
  const int nl = a.nlay + 1, nlev = 2 * a.nlay + 3;
  const RtConst *rc = dt.c;
  const int b = (VIS ? 0 : NV) + bb;
  const double *pl = a.plev + (size_t)col * nlev;
  const double *tr = a.tauref + (size_t)col * (nlev + 1);
  const double *tc = a.taurefcld + (size_t)col * (nlev + 1);
  double *rec = a.clr + (size_t)col * NB * clr_rec(nl) + b;  // slot k of this item: rec[k * NB]
  const Aer ae = column_aerosol(a, rc, col);
  // layer aerosol optics computed as the optics pass reaches each layer; the
  // solve's COSBAR (VIS) or COSB (IR) kept per layer
  double keep[NLM];
  auto accf = [&](int n, double fp, double fm) {
    rec[(n - 1) * NB] = fp;
    rec[(nl + n - 1) * NB] = fm;
  };
  if (VIS) {
    const double mu = a.acosz[col];
    if (!(mu >= F32LIT(1.0e-5))) return;  // code/comp3.f:637
    ClearSrc<true> src{ae, b, pl, tr, tc, 0.0};
    const ClearLayerAer<true> la{src, nl, keep};
    VisState<NLM> s;
    double cumtau = 0.0;
    const double detau = vis_optics<NLM>(nl, src, la, s, cumtau);
    a.detau[(size_t)col * NG * NV + 16 * NV + b] = detau;
    a.cumtauv[(size_t)col * NG * NV + 16 * NV + b] = cumtau;
    double fluxup = 0.0, fluxdn = 0.0, diffv = 0.0;
    const double bsurf = a.als[col] * mu * a.sol[b] * rt_exp35(-dmin(detau / mu, MAXEXP));
    vis_solve<NLM>(nl, s, keep, mu, a.sol[b], a.als[col], 0.0, bsurf, accf, fluxup, fluxdn, diffv);
    rec[2 * nl * NB] = fluxup - fluxdn;
    rec[(2 * nl + 1) * NB] = diffv;
  } else {
    const int nw = b - NV;
    ClearSrc<false> src{ae, b, pl, tr, tc, 0.0};
    const ClearLayerAer<false> la{src, nl, keep};
    IrState<NLM> s;
    ir_optics<NLM>(nl, src, la, s);
    double bk[2 * NLM + 1];
    for (int K = 1; K <= nlev; K++)
      bk[K - 1] = dt.planck[(size_t)(planck_index(a.tlev[(size_t)col * nlev + K - 1], a.flag) - 1) * NI + nw];
    const double rsf = a.albi[col];
    const double bsurf = (1. - rsf) * bk[nlev - 1];
    const double tautop = s.wd[1].y * pl[1] / (pl[3] - pl[1]);
    const double btop = (1.0 - rt_exp(-tautop * 2.0)) * bk[1];
    rec[2 * nl * NB] = ir_solve<NLM>(nl, s, keep, bk, rsf, btop, bsurf, accf);
  }
}

template <int NLM, bool VIS>
__global__ __launch_bounds__(64) void k_rt_clear(DevTables dt, RunArgs a) {
  mgcm_libm::load_lds_exptab(threadIdx.x, 64);
  __syncthreads();
  clear_block<NLM, VIS>(dt, a, blockIdx.x * ClearShape<VIS>::cols);
}

// Both parts in one launch: blocks below nvis take VIS items, the rest IR
// items, so the two halves share one dispatch and fill each other's tail.
// Two waves per SIMD: the register budget of 256 schedules the item's loads
// further ahead than the compiler's default occupancy target (4) does.
template <int NLM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_rt_clear2(DevTables dt, RunArgs a, int nvis) {
  mgcm_libm::load_lds_exptab(threadIdx.x, 64);
  __syncthreads();
  if ((int)blockIdx.x < nvis) clear_block<NLM, true>(dt, a, blockIdx.x * ClearShape<true>::cols);
  else clear_block<NLM, false>(dt, a, (blockIdx.x - nvis) * ClearShape<false>::cols);
}

template <int NLM, bool VIS>
__global__ __launch_bounds__(64) void k_rt_clear_batch(DevTables dt, const RunArgs *__restrict__ grids,
                                                       const uint2 *__restrict__ work, int nwork) {
  mgcm_libm::load_lds_exptab(threadIdx.x, 64);
  __syncthreads();
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const uint2 wk = work[w];
    const RunArgs a = grids[wk.x];
    clear_block<NLM, VIS>(dt, a, (int)wk.y);
  }
}

// Waves per SIMD of the single-grid column kernel: two up to 36 layers, one
// above (512 registers per lane for the larger item state; the batched tuner
// measures the one-wave instance 7-8% faster at 48 and 64 layers and the
// two-wave one 2-16% faster at 16..36 layers).
__host__ __device__ constexpr int column_waves(int nlm) { return nlm <= 37 ? MGCM_FP64_WAVES : 1; }
// Layers ahead that the solves load the item state: two at one wave per SIMD
// (its 512 registers hold the third buffer; 4-6% faster from 48 layers up),
// one at two waves (a third buffer spills there).
__host__ __device__ constexpr int state_ahead(int waves) { return waves == 1 ? 2 : 1; }

// One grid: block b owns columns Cb..Cb+C-1 (C = 4/S).
template <int NLM, int P, int S>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(column_waves(NLM), column_waves(NLM))))
void k_rt(DevTables dt, RunArgs a) {
  extern __shared__ double smem[];
  mgcm_libm::load_lds_exptab(threadIdx.x, 64);
  __syncthreads();
  column_block<NLM, P, S, state_ahead(column_waves(NLM))>(dt, a, blockIdx.x * Tile<S>::C, smem);
}

// TPINDEX of every sub-level of a grid for the one-band-per-wave tile: one
// thread per (column, sub-level), results to a.tp.
static __global__ __launch_bounds__(256) void k_tp(DevTables dt, RunArgs a) {
  const int nlev = 2 * a.nlay + 3, nsub = nlev - 1;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)a.ncol * nsub) return;
  const int cl = (int)(idx / nsub), i = (int)(idx % nsub), K = i + 2;
  const TpOut r = tp_sublevel(dt.c, a, cl, K, nlev, tp_inputs(a, cl, K, nlev));
  double *q = a.tp + (size_t)cl * TP_FIELDS * nsub + i;  // field f of the sub-level at q[f * nsub]
  q[0] = r.t;
  q[nsub] = r.u;
  q[2 * nsub] = r.wr;
  q[3 * nsub] = r.uk;
  q[4 * nsub] = (double)r.base;
}

// The one-band-per-wave tile, in one launch with the continuum points (the
// two are independent until k_rt_finish reads both): blocks below nclear are
// continuum-point blocks (VIS below nvis, as k_rt_clear2); block nclear + w
// solves band round w / ngroups (at one band slot per column, round k is band
// k) of columns 4 (w % ngroups) .. +3 and writes the band's records;
// k_rt_finish then adds them per column.  Blocks of one band are adjacent, so
// the waves in flight share few k-table slices.
template <int NLM, int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(column_waves(NLM), column_waves(NLM))))
void k_rt_band(DevTables dt, RunArgs a, int ngroups, int nvis, int nclear) {
  extern __shared__ double smem[];
  mgcm_libm::load_lds_exptab(threadIdx.x, 64);
  __syncthreads();
  const int blk = (int)blockIdx.x;
  if (blk < nclear) {
    if (blk < nvis) clear_block<NLM, true>(dt, a, blk * ClearShape<true>::cols);
    else clear_block<NLM, false>(dt, a, (blk - nvis) * ClearShape<false>::cols);
    return;
  }
  const int w = blk - nclear, k = w / ngroups, grp = w % ngroups;
  column_block<NLM, P, 1, state_ahead(column_waves(NLM)), true>(dt, a, grp * 4, smem, k);
}
template <int NLM>
__global__ __launch_bounds__(64) void k_rt_finish(DevTables dt, RunArgs a) {
  extern __shared__ double smem[];
  column_finish<NLM>(dt, a, (int)blockIdx.x * 4, smem);
}

// Many grids of one layer-count class: a persistent grid of blocks walks the
// work list (grid index, first column of 4) round-robin, largest layer counts
// first; with S band slots per column each item is S sub-items of 4/S columns.
// No counters: every block's share is fixed by its index.
template <int NLM, int P, int S, int D>
__device__ __forceinline__ void batch_items(const DevTables &dt, const RunArgs *__restrict__ grids,
                                            const uint2 *__restrict__ work, int nwork, double *smem) {
  mgcm_libm::load_lds_exptab(threadIdx.x, 64);
  for (int v = blockIdx.x; v < nwork * S; v += gridDim.x) {
    __syncthreads();  // the previous item's reads of the staging are done
    const uint2 wk = work[v / S];
    const RunArgs a = grids[wk.x];
    column_block<NLM, P, S, D>(dt, a, (int)wk.y + (v % S) * Tile<S>::C, smem);
  }
}
// The instances the batched tuner chooses between per layer class: two waves
// per SIMD (at most 256 registers per lane) at every tile ...
template <int NLM, int P, int S>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MGCM_FP64_WAVES)))
void k_rt_batch(DevTables dt, const RunArgs *__restrict__ grids, const uint2 *__restrict__ work, int nwork) {
  extern __shared__ double smem[];
  batch_items<NLM, P, S, state_ahead(MGCM_FP64_WAVES)>(dt, grids, work, nwork, smem);
}
// ... and one wave per SIMD (512 registers per lane: fewer spills of the
// item state, half the waves in flight), one band slot per column.
template <int NLM, int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_rt_batch_w1(DevTables dt, const RunArgs *__restrict__ grids, const uint2 *__restrict__ work, int nwork) {
  extern __shared__ double smem[];
  batch_items<NLM, P, 1, state_ahead(1)>(dt, grids, work, nwork, smem);
}

// The continuum-point kernels of a grid (their records feed the column
// kernel's sums).
template <int NLM>
static void launch_clear(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  const int nvis = (a.ncol + 8) / 9, nir = (a.ncol + 11) / 12;
  if (a.bands == 3)
    hipLaunchKernelGGL((k_rt_clear2<NLM>), dim3(nvis + nir), dim3(64), mgcm_libm::LDS_EXPTAB_BYTES, st, dt, a, nvis);
  else if (a.bands & 1)
    hipLaunchKernelGGL((k_rt_clear<NLM, true>), dim3(nvis), dim3(64), mgcm_libm::LDS_EXPTAB_BYTES, st, dt, a);
  else if (a.bands & 2)
    hipLaunchKernelGGL((k_rt_clear<NLM, false>), dim3(nir), dim3(64), mgcm_libm::LDS_EXPTAB_BYTES, st, dt, a);
}

template <int NLM, int P, int S>
static hipError_t launch_nlm(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  launch_clear<NLM>(dt, a, st);
  const size_t lds = stage_bytes(a.nlay);
  constexpr int C = Tile<S>::C;
  const unsigned gx = (unsigned)((a.ncol + C - 1) / C);
  hipLaunchKernelGGL((k_rt<NLM, P, S>), dim3(gx), dim3(64), lds, st, dt, a);
  return hipGetLastError();
}

// One grid: the continuum-point kernels, then the column kernel (4/S columns
// per wave), each instance sized for the grid's layer count.  a.clr must hold
// clr_bytes(a.ncol, a.nlay).
template <int P, int S>
hipError_t launch_column(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  const int nl = a.nlay + 1;
  if (nl <= 17) return launch_nlm<17, P, S>(dt, a, st);
  if (nl <= 25) return launch_nlm<25, P, S>(dt, a, st);
  if (nl <= 37) return launch_nlm<37, P, S>(dt, a, st);
  if (nl <= 49) return launch_nlm<49, P, S>(dt, a, st);
  if (nl <= 65) return launch_nlm<65, P, S>(dt, a, st);
  return launch_nlm<MAX_NLAY + 1, P, S>(dt, a, st);
}

// One grid at one band per wave (TILE_BANDS): TPINDEX of every sub-level
// (k_tp), then the continuum points and NB waves per 4 columns in one launch,
// then the finishing kernel.  a.clr must hold clr_bytes(a.ncol, a.nlay), a.part (with
// a.tp after the band records) part_bytes(a.ncol, a.nlay).
template <int NLM, int P>
static hipError_t launch_bands_nlm(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  const size_t lds = stage_bytes(a.nlay);
  const int ngroups = (a.ncol + 3) / 4;
  const long long ntp = (long long)a.ncol * (2 * a.nlay + 2);
  hipLaunchKernelGGL(k_tp, dim3((unsigned)((ntp + 255) / 256)), dim3(256), 0, st, dt, a);
  const int nvis = (a.bands & 1) ? (a.ncol + ClearShape<true>::cols - 1) / ClearShape<true>::cols : 0;
  const int nir = (a.bands & 2) ? (a.ncol + ClearShape<false>::cols - 1) / ClearShape<false>::cols : 0;
  hipLaunchKernelGGL((k_rt_band<NLM, P>), dim3((unsigned)(nvis + nir + ngroups * (Tile<1>::RV + Tile<1>::RI))),
                     dim3(64), lds, st, dt, a, ngroups, nvis, nvis + nir);
  hipLaunchKernelGGL((k_rt_finish<NLM>), dim3((unsigned)ngroups), dim3(64), lds, st, dt, a);
  return hipGetLastError();
}
template <int P>
hipError_t launch_bands(const DevTables &dt, const RunArgs &a, hipStream_t st) {
  const int nl = a.nlay + 1;
  if (nl <= 17) return launch_bands_nlm<17, P>(dt, a, st);
  if (nl <= 25) return launch_bands_nlm<25, P>(dt, a, st);
  if (nl <= 37) return launch_bands_nlm<37, P>(dt, a, st);
  if (nl <= 49) return launch_bands_nlm<49, P>(dt, a, st);
  if (nl <= 65) return launch_bands_nlm<65, P>(dt, a, st);
  return launch_bands_nlm<MAX_NLAY + 1, P>(dt, a, st);
}

// Resident one-wave blocks per CU for `lds` bytes of LDS at `waves` waves per
// SIMD (4 SIMDs per CU, 160 KiB of LDS).
inline int resident_per_cu(size_t lds, int waves) {
  const int by_lds = (int)((160 * 1024) / (lds > 0 ? lds : 1));
  const int by_waves = 4 * waves;
  return by_lds < 1 ? 1 : (by_lds < by_waves ? by_lds : by_waves);
}

// Band slots per column of a single-grid call when the caller leaves the
// choice to the library: the tile with the least (rounds of resident waves) x
// (latency of one wave), S = 1 on ties.  A wave's latency at S = 1, 2, 4 is
// 1 : 0.75 : 0.53 (measured on the MI355X with one resident round: 540
// columns of 24 layers take 0.715, 0.535 and 0.376 ms); the band rounds alone
// (12, 7, 4) would say 1 : 0.58 : 0.33 -- the staging, the continuum-point
// sums and the empty band slots do not shrink with S.  Large grids fill the
// device at S = 1, the least work per column.  One band per wave (bands_ok:
// its band records fit the caller's workspace) costs a twelfth of S = 1's
// wave latency plus its TPINDEX staging, and a finishing launch.
inline int auto_tile(int ncol, int nlay, int ncu, bool bands_ok = true) {
  const int nl = nlay + 1;
  const int nlm = nl <= 17 ? 17 : nl <= 25 ? 25 : nl <= 37 ? 37 : nl <= 49 ? 49 : nl <= 65 ? 65 : MAX_NLAY + 1;
  const long long slots = (long long)ncu * resident_per_cu(stage_bytes(nlay), column_waves(nlm));
  constexpr int tiles[4] = {1, 2, 4, TILE_BANDS}, latency[4] = {100, 75, 53, MGCM_BANDS_LATENCY};
  constexpr int per_wave[4] = {4, 2, 1, 4};  // columns per wave
  constexpr int waves_per[4] = {1, 1, 1, TILE_BANDS};  // waves per wave's columns
  int best = 1;
  long long best_cost = -1;
  for (int i = 0; i < (bands_ok ? 4 : 3); i++) {
    const long long waves = ((long long)ncol + per_wave[i] - 1) / per_wave[i] * waves_per[i];
    const long long cost = ((waves + slots - 1) / slots) * latency[i] + (tiles[i] == TILE_BANDS ? MGCM_BANDS_FINISH : 0);
    if (best_cost < 0 || cost < best_cost) {
      best = tiles[i];
      best_cost = cost;
    }
  }
  return best;
}

// Layer-count classes of the batched path: kernel instance NLM = KCLASS_NL[k]
// serves grids with nlay + 1 <= NLM.
constexpr int NCLASS = 6;
constexpr int KCLASS_NL[NCLASS] = {17, 25, 37, 49, 65, MAX_NLAY + 1};
inline int nl_class(int nl) {
  for (int k = 0; k < NCLASS; k++)
    if (nl <= KCLASS_NL[k]) return k;
  return NCLASS - 1;
}

// kind 0: column kernel (waves per SIMD 1 or 2; one wave per SIMD at S = 1
// only); 1: VIS continuum points; 2: IR continuum points (S = 1 instances)
template <int NLM, int P, int S>
static hipError_t launch_batch_nlm(int kind, int waves, const DevTables &dt, const RunArgs *grids, const uint2 *work,
                                   int nwork, int nblocks, size_t lds, hipStream_t st) {
  if constexpr (S == 1) {
    if (kind == 1)
      hipLaunchKernelGGL((k_rt_clear_batch<NLM, true>), dim3(nblocks), dim3(64), mgcm_libm::LDS_EXPTAB_BYTES, st, dt,
                         grids, work, nwork);
    else if (kind == 2)
      hipLaunchKernelGGL((k_rt_clear_batch<NLM, false>), dim3(nblocks), dim3(64), mgcm_libm::LDS_EXPTAB_BYTES, st,
                         dt, grids, work, nwork);
    else if (waves == 1)
      hipLaunchKernelGGL((k_rt_batch_w1<NLM, P>), dim3(nblocks), dim3(64), lds, st, dt, grids, work, nwork);
    else
      hipLaunchKernelGGL((k_rt_batch<NLM, P, 1>), dim3(nblocks), dim3(64), lds, st, dt, grids, work, nwork);
  } else {
    if (kind != 0 || waves != MGCM_FP64_WAVES) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_rt_batch<NLM, P, S>), dim3(nblocks), dim3(64), lds, st, dt, grids, work, nwork);
  }
  return hipGetLastError();
}

template <int P, int S>
hipError_t launch_batch_class(int cls, int kind, int waves, const DevTables &dt, const RunArgs *grids,
                              const uint2 *work, int nwork, int nblocks, size_t lds, hipStream_t st) {
  switch (cls) {
    case 0: return launch_batch_nlm<17, P, S>(kind, waves, dt, grids, work, nwork, nblocks, lds, st);
    case 1: return launch_batch_nlm<25, P, S>(kind, waves, dt, grids, work, nwork, nblocks, lds, st);
    case 2: return launch_batch_nlm<37, P, S>(kind, waves, dt, grids, work, nwork, nblocks, lds, st);
    case 3: return launch_batch_nlm<49, P, S>(kind, waves, dt, grids, work, nwork, nblocks, lds, st);
    case 4: return launch_batch_nlm<65, P, S>(kind, waves, dt, grids, work, nwork, nblocks, lds, st);
    default: return launch_batch_nlm<MAX_NLAY + 1, P, S>(kind, waves, dt, grids, work, nwork, nblocks, lds, st);
  }
}

}  // namespace mgcm

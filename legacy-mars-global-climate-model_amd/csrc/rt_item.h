// rt_item.h -- one two-stream item of the column radiative transfer on the
// device: the aerosol/Rayleigh optics of a sub-level and of a
// layer, and the VIS (GFLUXV) and IR (GFLUXI) solves with DSOLVER/DTRIDGL.
// Every FP64 statement follows the reference's evaluation order; files that
// include this header are compiled with -ffp-contract=off.
#pragma once
#include <type_traits>

#include "rt_internal.h"
#include "rt_libm.h"

namespace mgcm {
// exp of the column kernels: glibc's bits, table in LDS (see rt_libm.h)
__device__ __forceinline__ double rt_exp(double x) { return mgcm_libm::mgcm_exp_lds(x); }
// exp of an argument clamped to [-35, 35] (MAXEXP): branch-free, same bits
__device__ __forceinline__ double rt_exp35(double x) { return mgcm_libm::mgcm_exp_lds_mid(x); }

// Precision policy of the column kernel, a template parameter P of its parts.
// P = 0 is the plain path: every operation FP64 in the reference's order.  The
// other bits move one component of the work to cheaper arithmetic; the
// reduced-precision forms and the precision study are combinations of them.
enum : int {
  P_TAB32 = 1,     // k-tables stored in FP32 (read as FP32, interpolated in FP64)
  P_INTERP32 = 2,  // k-interpolation arithmetic in FP32 (OPTCV/OPTCI's corner combination)
  P_TAU32 = 4,     // sub-level optical depths DTAUK rounded to FP32
  P_TRANS32 = 8,   // layer transmissions exp(-lambda dtau), exp(-tau/mu0) in FP32
  P_FASTDIV = 16,  // FP64 division as a * (1/b), 1/b from v_rcp_f64 and two Newton steps
};

// 1/b to within 1 ulp: the hardware reciprocal estimate refined by two
// Newton-Raphson steps (each doubles its correct bits).
__device__ __forceinline__ double rcp_nr(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  return __builtin_fma(r, e, r);
}
// Division of the two-stream solves under policy P (IEEE a/b on the plain path).
template <int P>
__device__ __forceinline__ double pdiv(double a, double b) {
  if (P & P_FASTDIV) return a * rcp_nr(b);
  return a / b;
}
// exp of a clamped argument under policy P.
template <int P>
__device__ __forceinline__ double pexp35(double x) {
  if (P & P_TRANS32) return (double)expf((float)x);
  return rt_exp35(x);
}
}  // namespace mgcm

namespace mgcm {

// Aerosol optical properties of one column (code/optcv.f90:114-123,
// code/optci.f90:124-130,182-187): the caller's per-level arrays where given
// (OPT_DST/OPT_CLD for active dust and water-ice clouds, code/comp3.f:595-631),
// else the tables' level-constant band values (INI_OPTDST/INI_OPTCLD).  b: band
// 0..NB-1 (VIS first); K: 1-based level 1..L_LEVELS+1.
struct Aer {
  const RtConst *rc;
  const double *q[12];            // RunArgs::aq ([ncol][nb][nlev+1]), or null: wave-uniform
  const double *qref_d, *qref_c;  // RunArgs::qref_dst/_cld ([ncol][nlev+1]), or null
  size_t offv, offi, offr;        // this column's offsets: col*NV*ld, col*NI*ld, col*ld
  int ld;                         // nlev + 1
  // kind: 0 Qext, 1 Qscat, 2 g; cld: 0 dust, 1 water ice
  __device__ __forceinline__ double get(int kind, int cld, int b, int K, double cst) const {
    const double *p = q[cld * 6 + (b < NV ? 0 : 3) + kind];
    return p ? p[(b < NV ? offv + (size_t)b * ld : offi + (size_t)(b - NV) * ld) + K - 1] : cst;
  }
  __device__ __forceinline__ double qx_dst(int b, int K) const { return get(0, 0, b, K, rc->qx_dst[b]); }
  __device__ __forceinline__ double qs_dst(int b, int K) const { return get(1, 0, b, K, rc->qs_dst[b]); }
  __device__ __forceinline__ double g_dst(int b, int K) const { return get(2, 0, b, K, rc->g_dst[b]); }
  __device__ __forceinline__ double qx_cld(int b, int K) const { return get(0, 1, b, K, rc->qx_cld[b]); }
  __device__ __forceinline__ double qs_cld(int b, int K) const { return get(1, 1, b, K, rc->qs_cld[b]); }
  __device__ __forceinline__ double g_cld(int b, int K) const { return get(2, 1, b, K, rc->g_cld[b]); }
  // Qextref(K) = QXVDST(K, L_NREFV) (code/optcv.f90:114), QREFV(K) = QEXTREFDST(K), QEXTREFCLD(K)
  __device__ __forceinline__ double qref_vis(int K) const {
    return q[0] ? q[0][offv + (size_t)(L_NREFV - 1) * ld + K - 1] : rc->qextref_vis;
  }
  __device__ __forceinline__ double qref_dst(int K) const { return qref_d ? qref_d[offr + K - 1] : rc->qextref_dst; }
  __device__ __forceinline__ double qref_cld(int K) const { return qref_c ? qref_c[offr + K - 1] : rc->qextref_cld; }
};

// The base pointers stay the run's (uniform across the wave); only the
// column's offsets are per lane.
__device__ __forceinline__ Aer column_aerosol(const RunArgs &a, const RtConst *rc, int col) {
  Aer e;
  e.rc = rc;
  e.ld = 2 * a.nlay + 4;
  for (int i = 0; i < 12; i++) e.q[i] = a.aq[i];
  e.qref_d = a.qref_dst;
  e.qref_c = a.qref_cld;
  e.offr = (size_t)col * e.ld;
  e.offv = e.offr * NV;
  e.offi = e.offr * NI;
  return e;
}

// Aerosol and Rayleigh optical depths of one sub-level K (1-based, 2..NLEV),
// OPTCV code/optcv.f90:103-123,184-185 and OPTCI code/optci.f90:113-130,184-185.
struct SubAer {
  double tray, taeros, tcloud, tlk, tck;
};

// ... from the sub-level's profile values: PLEV(K), PLEV(K-1), TAUREF(K), TAUREFCLD(K)
template <bool VIS>
__device__ __forceinline__ SubAer sub_aerosol_of(const Aer &ae, int b, int K, double pk, double pk1, double trk,
                                                 double tck) {
  SubAer s;
  const double dpr = pk - pk1;
  const double tr = trk / (VIS ? ae.qref_vis(K) : ae.qref_dst(K));
  const double tc = tck / ae.qref_cld(K);
  s.tray = VIS ? ae.rc->tauray[b] * dpr : 0.0;
  s.taeros = tr * ae.qx_dst(b, K);
  s.tcloud = tc * ae.qx_cld(b, K);
  s.tlk = tr * ae.qs_dst(b, K);
  s.tck = tc * ae.qs_cld(b, K);
  return s;
}
// The profile values one sub-level's aerosol depths read.
struct SubIn {
  double pk, pk1, trk, tck;  // PLEV(K), PLEV(K-1), TAUREF(K), TAUREFCLD(K)
};
__device__ __forceinline__ SubIn sub_inputs(const double *__restrict__ plev, const double *__restrict__ tauref,
                                            const double *__restrict__ taurefcld, int K) {
  return SubIn{plev[K - 1], plev[K - 2], tauref[K - 1], taurefcld[K - 1]};
}

// IR padding level L_LEVELS+1 (code/optci.f90:89-95,182-187): TAUREF there is
// not divided by the reference extinction.
__device__ __forceinline__ void ir_pad(const Aer &ae, int b, const double *__restrict__ tauref,
                                       const double *__restrict__ taurefcld, int nlev,
                                       double &tlk, double &tck) {
  tlk = tauref[nlev] * ae.qs_dst(b, nlev + 1);
  tck = taurefcld[nlev] * ae.qs_cld(b, nlev + 1);
}

// Layer scattering optical depth and asymmetry of the continuum channel:
// VIS code/optcv.f90:194-222, IR code/optci.f90:195-219, from the aerosol
// depths a of sub-level K = 2L+1 and a2 of K+1 (L < NL; for the bottom IR
// layer a2 is the padding level's, ir_pad).
template <bool VIS>
__device__ __forceinline__ void layer_aerosol_of(const Aer &ae, int b, int nl, int L, const SubAer &a,
                                                 double tlk2, double tck2, double tray2, double &scat,
                                                 double &cosb) {
  const int K = 2 * L + 1;
  const double gd = ae.g_dst(b, K), gc = ae.g_cld(b, K);
  if (VIS) {
    const double f9999 = F32LIT(0.9999);
    if (L < nl) {
      const SubAer a2{tray2, 0.0, 0.0, tlk2, tck2};
      const double gd2 = ae.g_dst(b, K + 1), gc2 = ae.g_cld(b, K + 1);
      cosb = (gd * a.tlk + gd2 * a2.tlk + gc * a.tck + gc2 * a2.tck) /
             (a.tray + a2.tray + a.tlk + a2.tlk + a.tck + a2.tck);
      scat = a.tlk + a2.tlk + a.tck + a2.tck + (a.tray + a2.tray) * f9999;
    } else {
      cosb = (gd * a.tlk + gc * a.tck) / (a.tray + a.tlk + a.tck);
      scat = a.tlk + a.tck + a.tray * f9999;
    }
  } else {
    const double gd2 = ae.g_dst(b, K + 1), gc2 = ae.g_cld(b, K + 1);
    scat = a.tlk + tlk2 + a.tck + tck2;
    cosb = (scat > 0.0) ? (gd * a.tlk + gd2 * tlk2 + gc * a.tck + gc2 * tck2) / scat : 0.0;
  }
}
// ... from the profile values of sub-levels K = 2L+1 and K+1 (LayerIn: PLEV(K-1),
// PLEV(K), PLEV(K+1) where K < NLEV, TAUREF and TAUREFCLD at K and K+1; the IR
// padding level's are those at K+1 = NLEV+1).
struct LayerIn {
  double p0, p1, p2, tr1, tc1, tr2, tc2;
};
__device__ __forceinline__ LayerIn layer_inputs(const double *__restrict__ plev, const double *__restrict__ tauref,
                                                const double *__restrict__ taurefcld, int nlev, int L) {
  const int K = 2 * L + 1;
  return LayerIn{plev[K - 2], plev[K - 1], K < nlev ? plev[K] : 0.0, tauref[K - 1], taurefcld[K - 1], tauref[K],
                 taurefcld[K]};
}
template <bool VIS>
__device__ __forceinline__ void layer_aerosol(const Aer &ae, int b, int nl, int L, const LayerIn &v, double &scat,
                                              double &cosb) {
  const int nlev = 2 * nl + 1;
  const int K = 2 * L + 1;
  const SubAer a = sub_aerosol_of<VIS>(ae, b, K, v.p1, v.p0, v.tr1, v.tc1);
  double tlk2 = 0.0, tck2 = 0.0, tray2 = 0.0;
  if (K + 1 <= nlev) {  // L < NL
    const SubAer a2 = sub_aerosol_of<VIS>(ae, b, K + 1, v.p2, v.p1, v.tr2, v.tc2);
    tlk2 = a2.tlk;
    tck2 = a2.tck;
    tray2 = a2.tray;
  } else if (!VIS) {  // ir_pad with TAUREF(NLEV+1), TAUREFCLD(NLEV+1)
    tlk2 = v.tr2 * ae.qs_dst(b, nlev + 1);
    tck2 = v.tc2 * ae.qs_cld(b, nlev + 1);
  }
  layer_aerosol_of<VIS>(ae, b, nl, L, a, tlk2, tck2, tray2, scat, cosb);
}

// Layer scattering optical depth and asymmetry (SCAT, COSB of layer L) as the
// optics pass reads them, from arrays staged per column (the continuum points
// compute them as the pass reaches each layer: ClearLayerAer).
struct StagedLayerAer {
  const double *sc, *cb;
  __device__ __forceinline__ void get(int L, double &scat, double &cosb) const {
    scat = sc[L - 1];
    cosb = cb[L - 1];
  }
  __device__ __forceinline__ double scat(int L) const { return sc[L - 1]; }
};


// ---------------------------------------------------------------------------
// One two-stream item.  NLM bounds the radiation layer count (arrays are
// 1-based).  Inputs arrive through the Src functor: finish(K, fetch(K)) gives the sub-level
// optical depth DTAUK(K) and adds the gas part to taugsurf.

// DSOLVER's clip of XK2 (code/dsolver.f90:101): 0 where XK2 /= 0 and
// |XK2/(XA+1E-20)| < 1E-30.  Where |XK2| > |XA+1E-20| * 2^-99 (a power of two:
// exact while the product is normal) the quotient's magnitude exceeds
// 2^-99 > 1E-30 for certain and XK2 stays; only when some lane of the wave is not
// in that case is the reference's division evaluated, for the whole wave.
__device__ __forceinline__ double xk2_clip(double xa, double xk2) {
  const double d = xa + F32LIT(1.e-20);
  const double t = fabs(d) * 0x1p-99;
  const bool keep = fabs(xk2) > t && t >= 0x1p-1022;
  if (__all(keep || xk2 == 0.0)) return xk2;
  return (xk2 != 0.0 && fabs(xk2 / d) < F32LIT(1.e-30)) ? 0.0 : xk2;
}

// Two FP64 values that both passes of a solve read together: one 16-byte
// private-memory access instead of two 8-byte ones.
struct alignas(16) Pair {
  double x, y;
};

// Per-item state of a VIS item (1-based layer arrays): wt[L] = {scaled
// single-scattering albedo W0, scaled optical depth TAU at the layer top},
// dh[L] = {scaled layer optical depth DTAU, scaled optical depth HM at the
// layer mid-point (TAUCUM(2L+1) of code/gfluxv.f90:76-103)}, and row[i] =
// {AS, DS}, the reverse-Thomas coefficients of the 2*NL rows.  TAUCUM(2L) =
// TAU(L) for L >= 2; TM1 holds TAUCUM(2), TBOT the optical depth TAU(NL+1) of
// the surface.
template <int NLM>
struct VisState {
  Pair wt[NLM + 1], dh[NLM + 1];
  double tm1, tbot;
  Pair row[2 * NLM + 1];
};

// Two-stream constants of one VIS layer (code/gfluxv.f90:108-135): the
// quadrature gammas, LAMDA, GAMA and the particular-solution amplitudes AM/AP.
// Evaluated in the upward elimination and again in the downward flux pass;
// identical inputs give identical bits.
struct VisLayer {
  double g1, g2, g3, lam, gama, am, ap;
};
template <int P = 0>
__device__ __forceinline__ VisLayer vis_layer(double w0, double cb, double u0, double f0pi) {
  const double s3 = 0x1.bb67aep+0;   // SQRT(3.0) in REAL(4) (code/gfluxv.f90:114-116)
  const double s3h = 0x1.bb67aep-1;  // SQRT(3.0)*0.5 in REAL(4)
  VisLayer v;
  v.g1 = s3h * (2.0 - w0 * (1.0 + cb));
  v.g2 = (s3 * w0 * 0.5) * (1.0 - cb);
  v.g3 = 0.5 * (1.0 - s3 * cb * u0);
  v.lam = sqrt(v.g1 * v.g1 - v.g2 * v.g2);
  v.gama = pdiv<P>(v.g1 - v.lam, v.g2);
  const double g4 = 1.0 - v.g3;
  double denom = v.lam * v.lam - pdiv<P>(1.0, u0 * u0);
  if (denom == 0.0) denom = F32LIT(1.e-10);
  v.am = pdiv<P>(f0pi * w0 * (g4 * (v.g1 + pdiv<P>(1.0, u0)) + v.g2 * v.g3), denom);
  v.ap = pdiv<P>(f0pi * w0 * (v.g3 * (v.g1 - pdiv<P>(1.0, u0)) + v.g2 * g4), denom);
  return v;
}

// Compile-time buffer index of the layer loops below.  Each loop reads a
// later layer's inputs ahead of their use into its own buffer and runs as many
// layers per trip as it has buffers, so no value is copied between buffers at
// the loop's end (such a copy would wait there for the loads it moves).
using Buf0 = std::integral_constant<int, 0>;
using Buf1 = std::integral_constant<int, 1>;

// f(buffer k, n - k) for k = 0..NB-1: NB layers of a downward loop from n ...
template <typename F, int... K>
__device__ __forceinline__ void layers_down(F &f, int n, std::integer_sequence<int, K...>) {
  (f(std::integral_constant<int, K>{}, n - K), ...);
}
// ... and f(buffer k, n + k) for an upward one.
template <typename F, int... K>
__device__ __forceinline__ void layers_up(F &f, int n, std::integer_sequence<int, K...>) {
  (f(std::integral_constant<int, K>{}, n + K), ...);
}
// A downward loop over layers nl..1 in trips of NB layers, then the last
// nl % NB layers.
template <int NB, typename F>
__device__ __forceinline__ void sweep_down(F &f, int nl) {
  int n = nl;
  for (; n >= NB; n -= NB) layers_down(f, n, std::make_integer_sequence<int, NB>{});
  if constexpr (NB > 2) {
    if (n == 2) {
      layers_down(f, 2, std::make_integer_sequence<int, 2>{});
      return;
    }
  }
  if (n == 1) layers_down(f, 1, std::make_integer_sequence<int, 1>{});
}
// An upward loop over layers 1..nl in trips of NB layers, then the rest.
template <int NB, typename F>
__device__ __forceinline__ void sweep_up(F &f, int nl) {
  int n = 1;
  for (; n + NB - 1 <= nl; n += NB) layers_up(f, n, std::make_integer_sequence<int, NB>{});
  if constexpr (NB > 2) {
    if (n + 1 == nl) {
      layers_up(f, n, std::make_integer_sequence<int, 2>{});
      return;
    }
  }
  if (n == nl) layers_up(f, n, std::make_integer_sequence<int, 1>{});
}

// VIS optics of one item + delta-Eddington scaling (OPTCV layer assembly
// code/optcv.f90:189-274 fused with GETDETAU/GFLUXV code/gfluxv.f90:76-103).
// Returns DETAU; cumtau receives TAUCUMV(L_LEVELS).
template <int NLM, typename Src, typename LA>
__device__ __forceinline__ double vis_optics(int nl, Src &src, const LA &la, VisState<NLM> &s, double &cumtau) {
  // sub-level inputs are fetched one layer ahead of their use
  typename Src::Raw r2 = src.fetch(2), ra[2], rb[2];
  ra[0] = src.fetch(3);
  rb[0] = src.fetch(4);
  // sub-level 2 (top half of the cap)
  const double d2 = src.finish(2, r2);
  double tcv = d2;  // TAUCUMV(2)
  double fprev = 0.0, tcv_prev_level = 0.0, tauL = 0.0;  // tauL = TAU(L), carried in a register
  // layers 1..nl-1 (two sub-levels each)
  auto mid = [&](auto buf, int L) {
    constexpr int X = decltype(buf)::value;
    const double a = src.finish(2 * L + 1, ra[X]), b = src.finish(2 * L + 2, rb[X]);
    ra[1 - X] = src.fetch(2 * L + 3);                   // L = nl-1: the bottom sub-level L_LEVELS
    if (L < nl - 1) rb[1 - X] = src.fetch(2 * L + 4);
    const double dtauv = a + b;
    const double tcA = tcv + a;  // TAUCUMV(2L+1)
    const double tcB = tcA + b;  // TAUCUMV(2L+2)
    double Sc, C;
    la.get(L, Sc, C);
    const double W = Sc / dtauv;
    const double F = 1.0 - W * (C * C);
    double hm;
    if (L == 1) {
      tauL = 0.0 * F;      // TAU(1) = TDEL(1)*FACTOR with TDEL(1) = 0
      s.tm1 = tcv * F;     // TAUCUM(2)
      hm = s.tm1 + (tcA - tcv) * F;
    } else {
      // TAUCUM(2L) = TAU(L); TAUCUM(2L+1) uses the factor of layer L-1
      hm = tauL + (tcA - tcv_prev_level) * fprev;
    }
    const double dtau = dtauv * F;
    s.wt[L] = Pair{W * (1.0 - C * C) / F, tauL};
    s.dh[L] = Pair{dtau, hm};
    tauL = tauL + dtau;
    fprev = F;
    tcv_prev_level = tcB;  // TAUCUMV(2L+2), the base of the next layer's mid level
    tcv = tcB;
  };
  // the bottom layer nl (one sub-level)
  auto bottom = [&](auto buf) {
    constexpr int X = decltype(buf)::value;
    const int L = nl;
    const double a = src.finish(2 * L + 1, ra[X]);
    const double dtauv = a;
    cumtau = tcv + a;  // TAUCUMV(L_LEVELS)
    double Sc, C;
    la.get(L, Sc, C);
    const double W = Sc / dtauv;
    const double F = 1.0 - W * (C * C);
    const double dtau = dtauv * F;
    s.wt[L] = Pair{W * (1.0 - C * C) / F, tauL};
    s.dh[L] = Pair{dtau, 0.0};  // the bottom layer has no HM
    tauL = tauL + dtau;
    s.tbot = tauL;
  };
  int L = 1;
  for (; L + 1 <= nl - 1; L += 2) {
    mid(Buf0{}, L);
    mid(Buf1{}, L + 1);
  }
  if (L <= nl - 1) {
    mid(Buf0{}, L);
    bottom(Buf1{});
  } else {
    bottom(Buf0{});
  }
  return tauL;
}

// GFLUXV (code/gfluxv.f90:108-280) + DSOLVER/DTRIDGL (code/dsolver.f90,
// code/dtridgl.f90) on a scaled item.  Calls acc(L, fmidp, fmidm) for each
// layer mid-point and returns the top fluxes and DIFFV.
// D: how many layers ahead the two passes load the item state (1 or 2; two
// suits one wave per SIMD, whose 512 registers hold the third buffer).
template <int NLM, int P = 0, int D = 1, typename Acc>
__device__ __forceinline__ void vis_solve(int nl, VisState<NLM> &s, const double *__restrict__ cbar, double u0,
                                          double f0pi, double rsf, double btop,
                          double bsurf, Acc &acc, double &fluxup, double &fluxdn, double &diffv) {
  // bottom-up: layer constants and the reverse-Thomas elimination.  AS/DS of
  // the row below and the transmission at the layer bottom (the top of the
  // layer below, the same argument) are carried in registers.
  double gama_n1 = 0.0, cpm1_n1 = 0.0, cmm1_n1 = 0.0, as1 = 0.0, ds1 = 0.0;
  double xbot = pexp35<P>(-dmin(pdiv<P>(s.tbot, u0), MAXEXP));
  // layer n - D's state is loaded while layer n is eliminated, into its own buffer
  constexpr int NBUF = D + 1;
  struct In {
    Pair wt;
    double cb, dtau;
  } in[NBUF];
  auto ld = [&](int m) { return In{s.wt[m], cbar[m - 1], s.dh[m].x}; };
#pragma unroll
  for (int j = 0; j < D; j++)
    if (nl - j >= 1) in[j] = ld(nl - j);
  auto elim = [&](auto buf, int n) {
    constexpr int X = decltype(buf)::value;
    const double w0c = in[X].wt.x, cbc = in[X].cb, tauc = in[X].wt.y, dtauc = in[X].dtau;
    if (n - D >= 1) in[(X + D) % NBUF] = ld(n - D);
    const VisLayer v = vis_layer<P>(w0c, cbc, u0, f0pi);
    const double lam = v.lam, gama = v.gama, am = v.am, ap = v.ap;
    const double xtop = pexp35<P>(-dmin(pdiv<P>(tauc, u0), MAXEXP));
    const double cpm1 = ap * xtop, cmm1 = am * xtop;
    const double cp = ap * xbot, cm = am * xbot;
    const double ep = pexp35<P>(dmin(MAXEXP, lam * dtauc));
    const double em = pdiv<P>(1.0, ep);
    const double e1 = ep + gama * em, e2 = ep - gama * em;
    const double e3 = gama * ep + em, e4 = gama * ep - em;
    if (n == nl) {  // surface row (dsolver.f90:82-85)
      const int i = 2 * nl;
      const double af = e1 - rsf * e3, bf = e2 - rsf * e4;
      const double df = bsurf - cp + rsf * cm;
      as1 = pdiv<P>(af, bf);
      ds1 = pdiv<P>(df, bf);
      s.row[i] = Pair{as1, ds1};
    } else {
      {  // odd row 2n+1 (dsolver.f90:74-80)
        const int i = 2 * n + 1;
        const double af = 2.0 * (1.0 - gama * gama);
        const double bf = (e1 - e3) * (1.0 + gama_n1);
        const double cf = (e1 + e3) * (gama_n1 - 1.0);
        const double df = e3 * (cpm1_n1 - cp) + e1 * (cm - cmm1_n1);
        const double x = pdiv<P>(1.0, bf - cf * as1);
        as1 = af * x;
        ds1 = (df - cf * ds1) * x;
        s.row[i] = Pair{as1, ds1};
      }
      {  // even row 2n (dsolver.f90:63-70)
        const int i = 2 * n;
        const double af = (e1 + e3) * (gama_n1 - 1.0);
        const double bf = (e2 + e4) * (gama_n1 - 1.0);
        const double cf = 2.0 * (1.0 - gama_n1 * gama_n1);
        const double df = (gama_n1 - 1.0) * (cpm1_n1 - cp) + (1.0 - gama_n1) * (cm - cmm1_n1);
        const double x = pdiv<P>(1.0, bf - cf * as1);
        as1 = af * x;
        ds1 = (df - cf * ds1) * x;
        s.row[i] = Pair{as1, ds1};
      }
    }
    if (n == 1) {  // top row (dsolver.f90:54-57)
      const double af = 0.0, bf = gama + 1.0, cf = gama - 1.0, df = btop - cmm1;
      const double x = pdiv<P>(1.0, bf - cf * as1);
      s.row[1] = Pair{af * x, (df - cf * ds1) * x};
    }
    gama_n1 = gama;
    cpm1_n1 = cpm1;
    cmm1_n1 = cmm1;
    xbot = xtop;
  };
  sweep_down<NBUF>(elim, nl);
  // top-down: substitution, unmixing and the layer fluxes.  The row pair,
  // layer state and sub-level depth of layer n + D are loaded while layer n is
  // solved, into its own buffer.
  struct Sub {
    Pair r1, r2;
    double w0, cb, hm, tt;
  } sub[NBUF];
  auto lds = [&](int m) {  // m >= 2
    const Pair r1 = s.row[2 * m - 1], r2 = s.row[2 * m];
    const Pair wt = s.wt[m];
    return Sub{r1, r2, wt.x, cbar[m - 1], s.dh[m].y, wt.y};
  };
  sub[0] = Sub{s.row[1], s.row[2], s.wt[1].x, cbar[0], s.dh[1].y, s.tm1};
#pragma unroll
  for (int j = 1; j < D; j++)
    if (1 + j <= nl) sub[j] = lds(1 + j);
  double xkprev = sub[0].r1.y;
  double xk1_1 = 0.0, xk2_1 = 0.0, gama_1 = 0.0, am_1 = 0.0, ap_1 = 0.0;
  const double tau_bot = s.tbot, tau_lastl = s.wt[nl].y;
  auto subst = [&](auto buf, int n) {
    constexpr int X = decltype(buf)::value;
    const double as1c = sub[X].r1.x, ds1c = sub[X].r1.y, as2c = sub[X].r2.x, ds2c = sub[X].r2.y;
    const double w0c = sub[X].w0, cbc = sub[X].cb, hmc = sub[X].hm, ttc = sub[X].tt;
    if (n + D <= nl) sub[(X + D) % NBUF] = lds(n + D);
    const double xa = (n == 1) ? ds1c : ds1c - as1c * xkprev;
    const double xb = ds2c - as2c * xa;
    xkprev = xb;
    const double xk1 = xa + xb;
    double xk2 = xa - xb;
    xk2 = xk2_clip(xa, xk2);  // dsolver.f90:101
    const VisLayer v = vis_layer<P>(w0c, cbc, u0, f0pi);
    const double lam = v.lam, gama = v.gama, am = v.am, ap = v.ap;
    double fmidp, fmidm;
    if (n < nl) {  // gfluxv.f90:170-206
      const double ep = pexp35<P>(dmin(MAXEXP, lam * (hmc - ttc)));
      const double em = pdiv<P>(1.0, ep);
      const double taumid = hmc;
      const double xm = pexp35<P>(-dmin(pdiv<P>(taumid, u0), MAXEXP));
      const double cpmid = ap * xm, cmmid = am * xm;
      fmidp = xk1 * ep + gama * xk2 * em + cpmid;
      fmidm = xk1 * ep * gama + xk2 * em + cmmid;
      fmidm = fmidm + u0 * f0pi * xm;
    } else {  // bottom layer, full sub-layer (gfluxv.f90:243-280)
      const double ep = pexp35<P>(dmin(MAXEXP, lam * (tau_bot - tau_lastl)));
      const double em = pdiv<P>(1.0, ep);
      const double taumid = dmin(tau_bot, MAXEXP);
      const double xm = pexp35<P>(-dmin(pdiv<P>(taumid, u0), MAXEXP));
      const double cpmid = ap * xm, cmmid = am * xm;
      fmidp = xk1 * ep + gama * xk2 * em + cpmid;
      fmidm = xk1 * ep * gama + xk2 * em + cmmid;
      diffv = fmidm;
      fmidm = fmidm + u0 * f0pi * xm;
    }
    acc(n, fmidp, fmidm);
    if (n == 1) { xk1_1 = xk1; xk2_1 = xk2; gama_1 = gama; am_1 = am; ap_1 = ap; }
  };
  sweep_up<NBUF>(subst, nl);
  {  // top of the atmosphere (gfluxv.f90:210-238)
    const double ep = 1.0, em = 1.0;
    fluxup = xk1_1 * ep + gama_1 * xk2_1 * em + ap_1;
    fluxdn = xk1_1 * ep * gama_1 + xk2_1 * em + am_1;
    fluxdn = fluxdn + u0 * f0pi * pexp35<P>(-dmin(pdiv<P>(0.0, u0), MAXEXP));
  }
}

// Per-item state of an IR item: wd[L] = {W0 (= WBARI), DTAU (= DTAUI)}, the
// half-layer optical depth DMID(L) = TAUCUMI(2L+1)-TAUCUMI(2L) and row[i] =
// {AS, DS}, the reverse-Thomas coefficients.
template <int NLM>
struct IrState {
  Pair wd[NLM + 1];
  double dmid[NLM + 1];
  Pair row[2 * NLM + 1];
};

// Hemispheric-mean constants of one IR layer (code/gfluxi.f90:69-93): LAMDA,
// GAMA, TERM and the linear-in-tau Planck source B0 + B1*tau.  bt/bb: Planck
// values at the top and bottom temperature levels of the layer.
struct IrLayer {
  double lam, gama, term, b0, b1;
};
template <int P = 0>
__device__ __forceinline__ IrLayer ir_layer(double w0, double cb, double dtau, double bt, double bb) {
  IrLayer v;
  const double alpha = sqrt(pdiv<P>(1.0 - w0, 1.0 - w0 * cb));
  v.lam = alpha * (1.0 - w0 * cb) * 2.0;  // /UBARI with UBARI = 0.5: exact
  v.b1 = pdiv<P>(bb - bt, dtau);
  v.b0 = bt;
  v.gama = pdiv<P>(1.0 - alpha, 1.0 + alpha);
  v.term = pdiv<P>(UBARI, 1.0 - w0 * cb);
  return v;
}

// IR optics of one item (OPTCI code/optci.f90:189-269).
template <int NLM, typename Src, typename LA>
__device__ __forceinline__ void ir_optics(int nl, Src &src, const LA &la, IrState<NLM> &s) {
  const double f1em9 = F32LIT(1.0e-9);
  // sub-level inputs are fetched one layer ahead of their use
  typename Src::Raw r2 = src.fetch(2), ra[2], rb[2];
  ra[0] = src.fetch(3);
  rb[0] = src.fetch(4);
  double tcv = src.finish(2, r2);  // TAUCUMI(2)
  auto layer = [&](auto buf, int L) {
    constexpr int X = decltype(buf)::value;
    const double a = src.finish(2 * L + 1, ra[X]);
    const double b = (L < nl) ? src.finish(2 * L + 2, rb[X]) : 0.0;  // DTAUKI(L_LEVELS+1) = 0
    if (L < nl) ra[1 - X] = src.fetch(2 * L + 3);
    if (L < nl - 1) rb[1 - X] = src.fetch(2 * L + 4);
    double dt = a + b + 1.e-50;
    const double sc = la.scat(L);  // every layer: a live provider keeps its COSB here
    double w;
    if (dt > f1em9) {
      w = sc / dt;
    } else {
      w = 0.0;
      dt = f1em9;
    }
    const double tcA = tcv + a;  // TAUCUMI(2L+1)
    s.dmid[L] = tcA - tcv;       // TAUCUMI(2L+1) - TAUCUMI(2L)
    tcv = (L < nl) ? tcA + b : tcA;
    s.wd[L] = Pair{w, dt};
  };
  int L = 1;
  for (; L + 1 <= nl; L += 2) {
    layer(Buf0{}, L);
    layer(Buf1{}, L + 1);
  }
  if (L == nl) layer(Buf0{}, L);
}

// GFLUXI (code/gfluxi.f90:69-194) + DSOLVER/DTRIDGL.  bk(K) = Planck value of
// TLEV(K).  Calls acc(L, fmidp, fmidm); returns FTOPUP.
// D: as in vis_solve.
template <int NLM, int P = 0, int D = 1, typename Acc>
__device__ __forceinline__ double ir_solve(int nl, IrState<NLM> &s, const double *__restrict__ cosb,
                                           const double *__restrict__ bk, double rsf,
                           double btop, double bsurf, Acc &acc) {
  double gama_n1 = 0.0, cpm1_n1 = 0.0, cmm1_n1 = 0.0, as1 = 0.0, ds1 = 0.0;
  // layer n - D's state is loaded while layer n is eliminated, into its own buffer
  constexpr int NBUF = D + 1;
  struct In {
    Pair wd;
    double cb, bt, bb;
  } in[NBUF];
  // TLEV(2L), TLEV(2L+2); the bottom layer uses TLEV(2L), TLEV(2L+1)
  auto ld = [&](int m) { return In{s.wd[m], cosb[m - 1], bk[2 * m - 1], m < nl ? bk[2 * m + 1] : bk[2 * m]}; };
#pragma unroll
  for (int j = 0; j < D; j++)
    if (nl - j >= 1) in[j] = ld(nl - j);
  auto elim = [&](auto buf, int n) {
    constexpr int X = decltype(buf)::value;
    const double dtau = in[X].wd.y, w0c = in[X].wd.x, cbc = in[X].cb, btc = in[X].bt, bbc = in[X].bb;
    if (n - D >= 1) in[(X + D) % NBUF] = ld(n - D);
    const IrLayer v = ir_layer<P>(w0c, cbc, dtau, btc, bbc);
    const double lam = v.lam, gama = v.gama, term = v.term, b0 = v.b0, b1 = v.b1;
    const double cp = b0 + b1 * dtau + b1 * term;
    const double cm = b0 + b1 * dtau - b1 * term;
    const double cpm1 = b0 + b1 * term;
    const double cmm1 = b0 - b1 * term;
    const double ep = pexp35<P>(dmin(lam * dtau, MAXEXP));
    const double em = pdiv<P>(1.0, ep);
    const double e1 = ep + gama * em, e2 = ep - gama * em;
    const double e3 = gama * ep + em, e4 = gama * ep - em;
    if (n == nl) {
      const int i = 2 * nl;
      const double af = e1 - rsf * e3, bf = e2 - rsf * e4;
      const double df = bsurf - cp + rsf * cm;
      as1 = pdiv<P>(af, bf);
      ds1 = pdiv<P>(df, bf);
      s.row[i] = Pair{as1, ds1};
    } else {
      {
        const int i = 2 * n + 1;
        const double af = 2.0 * (1.0 - gama * gama);
        const double bf = (e1 - e3) * (1.0 + gama_n1);
        const double cf = (e1 + e3) * (gama_n1 - 1.0);
        const double df = e3 * (cpm1_n1 - cp) + e1 * (cm - cmm1_n1);
        const double x = pdiv<P>(1.0, bf - cf * as1);
        as1 = af * x;
        ds1 = (df - cf * ds1) * x;
        s.row[i] = Pair{as1, ds1};
      }
      {
        const int i = 2 * n;
        const double af = (e1 + e3) * (gama_n1 - 1.0);
        const double bf = (e2 + e4) * (gama_n1 - 1.0);
        const double cf = 2.0 * (1.0 - gama_n1 * gama_n1);
        const double df = (gama_n1 - 1.0) * (cpm1_n1 - cp) + (1.0 - gama_n1) * (cm - cmm1_n1);
        const double x = pdiv<P>(1.0, bf - cf * as1);
        as1 = af * x;
        ds1 = (df - cf * ds1) * x;
        s.row[i] = Pair{as1, ds1};
      }
    }
    if (n == 1) {
      const double af = 0.0, bf = gama + 1.0, cf = gama - 1.0, df = btop - cmm1;
      const double x = pdiv<P>(1.0, bf - cf * as1);
      s.row[1] = Pair{af * x, (df - cf * ds1) * x};
    }
    gama_n1 = gama;
    cpm1_n1 = cpm1;
    cmm1_n1 = cmm1;
  };
  sweep_down<NBUF>(elim, nl);
  // top-down; layer n + D's rows and state are loaded while layer n is solved, into its own buffer
  struct Sub {
    Pair r1, r2, wd;
    double cb, dm, bt, bb;
  } sub[NBUF];
  auto lds = [&](int m) {
    return Sub{s.row[2 * m - 1], s.row[2 * m], s.wd[m], cosb[m - 1], s.dmid[m], bk[2 * m - 1],
               m < nl ? bk[2 * m + 1] : bk[2 * m]};
  };
#pragma unroll
  for (int j = 0; j < D; j++)
    if (1 + j <= nl) sub[j] = lds(1 + j);
  double xkprev = sub[0].r1.y;
  double xk1_1 = 0.0, xk2_1 = 0.0, gama_1 = 0.0, b0_1 = 0.0, b1_1 = 0.0, term_1 = 0.0;
  auto subst = [&](auto buf, int n) {
    constexpr int X = decltype(buf)::value;
    const double as1c = sub[X].r1.x, ds1c = sub[X].r1.y, as2c = sub[X].r2.x, ds2c = sub[X].r2.y;
    const double w0c = sub[X].wd.x, cbc = sub[X].cb, dtc = sub[X].wd.y, dmc = sub[X].dm, btc = sub[X].bt,
                 bbc = sub[X].bb;
    if (n + D <= nl) sub[(X + D) % NBUF] = lds(n + D);
    const double xa = (n == 1) ? ds1c : ds1c - as1c * xkprev;
    const double xb = ds2c - as2c * xa;
    xkprev = xb;
    const double xk1 = xa + xb;
    double xk2 = xa - xb;
    xk2 = xk2_clip(xa, xk2);  // dsolver.f90:101
    const IrLayer v = ir_layer<P>(w0c, cbc, dtc, btc, bbc);
    const double lam = v.lam, gama = v.gama, b0 = v.b0, b1 = v.b1, term = v.term;
    const double dk = (n < nl) ? dmc : dtc;
    const double ep = pexp35<P>(dmin(lam * dk, MAXEXP));
    const double em = pdiv<P>(1.0, ep);
    const double cpmid = b0 + b1 * dk + b1 * term;
    const double cmmid = b0 + b1 * dk - b1 * term;
    double fmidp = xk1 * ep + gama * xk2 * em + cpmid;
    double fmidm = xk1 * ep * gama + xk2 * em + cmmid;
    fmidp = fmidp * PI;
    fmidm = fmidm * PI;
    acc(n, fmidp, fmidm);
    if (n == 1) { xk1_1 = xk1; xk2_1 = xk2; gama_1 = gama; b0_1 = b0; b1_1 = b1; term_1 = term; }
  };
  sweep_up<NBUF>(subst, nl);
  const double ep = 1.0, em = 1.0;
  const double cpmid = b0_1 + b1_1 * term_1;
  const double cmmid = b0_1 - b1_1 * term_1;
  const double fluxup = xk1_1 * ep + gama_1 * xk2_1 * em + cpmid;
  const double fluxdn = xk1_1 * ep * gama_1 + xk2_1 * em + cmmid;
  return (fluxup - fluxdn) * PI;
}

// ---------------------------------------------------------------------------
}  // namespace mgcm

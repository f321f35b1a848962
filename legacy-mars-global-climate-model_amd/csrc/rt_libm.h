// rt_libm.h -- exp and log10 that reproduce the host C library the reference
// links against (glibc 2.35, x86-64), usable on the device and on the host.
//
// The reference's FP64 answer is ill-conditioned in places: the IR floor
// DTAUI = 1e-9 makes B1 = dB/dtau huge in near-empty top layers, and in VIS
// LAMDA^2 - 1/UBAR0^2 can come close to zero.  There a 1-ulp change of one exp
// or log10 moves fluxes by 1e-6 relative, so rel 1e-10 parity needs the same
// transcendental results as the reference's libm, not merely accurate ones.
//
//  * mgcm_exp: glibc's table-driven exp (sysdeps/ieee754/dbl-64/e_exp.c, the
//    ARM optimized-routines method): x = k ln2/128 + r, exp(x) = 2^(k/128) *
//    (1 + tail + r + C2 r^2 + ... + C5 r^5) with a 128-entry table of 2^(k/128)
//    (rt_exp_table.inc, generated from the definition by gen_exp_table.py).
//    On FMA-capable x86-64 hosts glibc runs its FMA build; the fused
//    operations below are the ones that build makes, so results are identical
//    bit for bit (checked against libm on the host: 2.85M arguments, 0 differences).
//  * mgcm_log: glibc's table-driven log (sysdeps/ieee754/dbl-64/e_log.c, the
//    same ARM optimized-routines family): a polynomial near 1, elsewhere a
//    128-entry {1/c, log c} table (rt_glibc_log_table.inc, gen_log_table.c)
//    and a degree-6 polynomial for log1p.  Like glibc it is not correctly
//    rounded (about 1.4 in 10^4 results differ from the nearest double); the
//    restatement returns glibc's bits, misroundings included.
//  * mgcm_log10: glibc's fdlibm wrapper (sysdeps/ieee754/dbl-64/e_log10.c):
//    x = 2^e m, log10(x) = e*log10_2lo + ivln10*log(m) + e*log10_2hi, over
//    mgcm_log, so it returns glibc's log10 bit for bit.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MGCM_HD __host__ __device__ __forceinline__
#define MGCM_FMA(a, b, c) __builtin_fma((a), (b), (c))
#else
#include <math.h>
#define MGCM_HD inline
#define MGCM_FMA(a, b, c) fma((a), (b), (c))
#endif

namespace mgcm_libm {

MGCM_HD uint64_t asu(double x) { return __builtin_bit_cast(uint64_t, x); }
MGCM_HD double asd(uint64_t u) { return __builtin_bit_cast(double, u); }

#if defined(__HIPCC__)
__device__ __constant__ static const uint64_t kExpTab[256] = {
#include "rt_exp_table.inc"
};
static const uint64_t kExpTabHost[256] = {
#include "rt_exp_table.inc"
};
MGCM_HD uint64_t exp_tab(int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return kExpTab[i];
#else
  return kExpTabHost[i];
#endif
}
#else
static const uint64_t kExpTabHost[256] = {
#include "rt_exp_table.inc"
};
MGCM_HD uint64_t exp_tab(int i) { return kExpTabHost[i]; }
#endif

#if defined(__HIPCC__)
__device__ __constant__ static const uint64_t kGlogTab[256] = {
#include "rt_glibc_log_table.inc"
};
static const uint64_t kGlogTabHost[256] = {
#include "rt_glibc_log_table.inc"
};
MGCM_HD uint64_t glog_tab(int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return kGlogTab[i];
#else
  return kGlogTabHost[i];
#endif
}
#else
static const uint64_t kGlogTabHost[256] = {
#include "rt_glibc_log_table.inc"
};
MGCM_HD uint64_t glog_tab(int i) { return kGlogTabHost[i]; }
#endif

MGCM_HD double exp_special(double tmp, uint64_t sbits, uint64_t ki) {
  double scale, y;
  if ((ki & 0x80000000u) == 0) {  // k > 0: the exponent of scale may have overflowed
    sbits -= 1009ull << 52;
    scale = asd(sbits);
    y = 0x1p1009 * MGCM_FMA(scale, tmp, scale);
    return y;
  }
  sbits += 1022ull << 52;  // k < 0: round once before scaling into the subnormal range
  scale = asd(sbits);
  y = scale + scale * tmp;
  if (y < 1.0) {
    double hi, lo;
    lo = scale - y + scale * tmp;
    hi = 1.0 + y;
    lo = 1.0 - hi + y + lo;
    y = (hi + lo) - 1.0;
    if (y == 0.0) y = 0.0;
  }
  return 0x1p-1022 * y;
}

template <class Tab>
MGCM_HD double exp_core(double x, Tab tab) {
  const double InvLn2N = 0x1.71547652b82fep7, NegLn2hiN = -0x1.62e42fefa0000p-8,
               NegLn2loN = -0x1.cf79abc9e3b3ap-47, Shift = 0x1.8p52;
  const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
               C5 = 0x1.1111167a4d017p-7;
  uint32_t abstop = (uint32_t)(asu(x) >> 52) & 0x7ffu;
  if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {
    if (abstop - 0x3c9u >= 0x80000000u) return 1.0 + x;  // |x| < 2^-54
    if (abstop >= 0x409u) {                                // |x| >= 1024
      if (asu(x) == 0xfff0000000000000ull) return 0.0;
      if (abstop >= 0x7ffu) return 1.0 + x;
      return (asu(x) >> 63) ? 0.0 : __builtin_inf();
    }
    abstop = 0;  // 512 <= |x| < 1024
  }
  double kd = MGCM_FMA(InvLn2N, x, Shift);
  const uint64_t ki = asu(kd);
  kd -= Shift;
  const double r = MGCM_FMA(kd, NegLn2loN, MGCM_FMA(kd, NegLn2hiN, x));
  const int idx = (int)(2 * (ki % 128));
  const uint64_t top = ki << 45;
  const double tail = asd(tab(idx));
  const uint64_t sbits = tab(idx + 1) + top;
  const double r2 = r * r;
  const double p1 = MGCM_FMA(r, C3, C2), p2 = MGCM_FMA(r, C5, C4);
  const double tmp = MGCM_FMA(r2 * r2, p2, MGCM_FMA(r2, p1, tail + r));
  if (abstop == 0) return exp_special(tmp, sbits, ki);
  const double scale = asd(sbits);
  return MGCM_FMA(scale, tmp, scale);
}

struct TabConst {
  MGCM_HD uint64_t operator()(int i) const { return exp_tab(i); }
};
MGCM_HD double mgcm_exp(double x) { return exp_core(x, TabConst{}); }

#if defined(__HIPCC__)
// The same exp with its 2 KiB table read from the first 256 entries of the
// launching kernel's dynamic LDS, which the kernel fills with exp_tab() before
// the first call (load_lds_exptab).  LDS reads keep the table lookups out of
// the vector-memory queue, where they would wait behind pending stores.
extern __shared__ uint64_t mgcm_lds_exptab[];
struct TabLds {
  __device__ __forceinline__ uint64_t operator()(int i) const { return mgcm_lds_exptab[i]; }
};
__device__ __forceinline__ double mgcm_exp_lds(double x) { return exp_core(x, TabLds{}); }
// The same bits for |x| < 512 without the special-case branch: there the
// general path is exact, including |x| < 2^-54, where glibc returns 1 + x and
// the general path rounds fma(1, x, 1) = 1 + x.  Callers guarantee the range.
__device__ __forceinline__ double mgcm_exp_lds_mid(double x) {
  const double InvLn2N = 0x1.71547652b82fep7, NegLn2hiN = -0x1.62e42fefa0000p-8,
               NegLn2loN = -0x1.cf79abc9e3b3ap-47, Shift = 0x1.8p52;
  const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
               C5 = 0x1.1111167a4d017p-7;
  double kd = MGCM_FMA(InvLn2N, x, Shift);
  const uint64_t ki = asu(kd);
  kd -= Shift;
  const double r = MGCM_FMA(kd, NegLn2loN, MGCM_FMA(kd, NegLn2hiN, x));
  const int idx = (int)(2 * (ki % 128));
  const uint64_t top = ki << 45;
  const double tail = asd(mgcm_lds_exptab[idx]);
  const uint64_t sbits = mgcm_lds_exptab[idx + 1] + top;
  const double r2 = r * r;
  const double p1 = MGCM_FMA(r, C3, C2), p2 = MGCM_FMA(r, C5, C4);
  const double tmp = MGCM_FMA(r2 * r2, p2, MGCM_FMA(r2, p1, tail + r));
  const double scale = asd(sbits);
  return MGCM_FMA(scale, tmp, scale);
}
// Fills the table; every thread of the block calls it, then the block synchronises.
__device__ __forceinline__ void load_lds_exptab(int tid, int nthreads) {
  for (int i = tid; i < 256; i += nthreads) mgcm_lds_exptab[i] = exp_tab(i);
}
constexpr int LDS_EXPTAB_BYTES = 256 * 8;
#endif

// glibc's double log (sysdeps/ieee754/dbl-64/e_log.c, the Arm
// optimized-routines method, FMA build).  Near 1 (x in [1 - 2^-4, 1 + 0x1.09p-4))
// a degree-12 polynomial in r = x - 1 with the r^2/2 term split exactly;
// elsewhere x = 2^k z, z in [0x1.6p-1, 0x1.6p0), log(x) = k ln2 + logc +
// log1p(r), r = z invc - 1, with {invc, logc} of z's subinterval from
// rt_glibc_log_table.inc (gen_log_table.c) and a degree-6 polynomial for
// log1p(r) - r.  The fused operations are those the FMA build makes.
MGCM_HD double glibc_log(double x) {
  const double Ln2hi = 0x1.62e42fefa3800p-1, Ln2lo = 0x1.ef35793c76730p-45;
  const double A0 = -0x1.0000000000001p-1, A1 = 0x1.555555551305bp-2, A2 = -0x1.fffffffeb459p-3,
               A3 = 0x1.999b324f10111p-3, A4 = -0x1.55575e506c89fp-3;
  uint64_t ix = asu(x);
  const uint64_t lo1 = 0x3fee000000000000ull, hi1 = 0x3ff1090000000000ull;  // 1 - 2^-4, 1 + 0x1.09p-4
  if (ix - lo1 < hi1 - lo1) {
    if (ix == 0x3ff0000000000000ull) return 0.0;
    const double B0 = -0x1p-1, B1 = 0x1.5555555555577p-2, B2 = -0x1.ffffffffffdcbp-3, B3 = 0x1.999999995dd0cp-3,
                 B4 = -0x1.55555556745a7p-3, B5 = 0x1.24924a344de3p-3, B6 = -0x1.fffffa4423d65p-4,
                 B7 = 0x1.c7184282ad6cap-4, B8 = -0x1.999eb43b068ffp-4, B9 = 0x1.78182f7afd085p-4,
                 B10 = -0x1.5521375d145cdp-4;
    const double r = x - 1.0, r2 = r * r, r3 = r * r2;
    const double p3 = MGCM_FMA(r3, B10, MGCM_FMA(r2, B9, MGCM_FMA(r, B8, B7)));
    const double p2 = MGCM_FMA(r3, p3, MGCM_FMA(r2, B6, MGCM_FMA(r, B5, B4)));
    const double p1 = MGCM_FMA(r3, p2, MGCM_FMA(r2, B3, MGCM_FMA(r, B2, B1)));
    double w = r * 0x1p27;
    const double rhi = r + w - w, rlo = r - rhi;
    w = rhi * rhi * B0;  // exact
    const double hi = r + w;
    double lo = r - hi + w;
    lo = MGCM_FMA(B0 * rlo, rhi + r, lo);
    const double y = MGCM_FMA(r3, p1, lo);
    return y + hi;
  }
  const uint32_t top = (uint32_t)(ix >> 48);
  if (top - 0x0010u >= 0x7ff0u - 0x0010u) {
    if (ix * 2 == 0) return -1.0 / 0.0;                                   // log(+-0) = -inf
    if (ix == 0x7ff0000000000000ull) return x;                            // log(inf) = inf
    if ((top & 0x8000u) || (top & 0x7ff0u) == 0x7ff0u) return (x - x) / (x - x);  // x < 0 or NaN
    ix = asu(x * 0x1p52);  // subnormal: normalise
    ix -= 52ull << 52;
  }
  const uint64_t off = 0x3fe6000000000000ull;
  const uint64_t tmp = ix - off;
  const int i = (int)((tmp >> 45) % 128);
  const int k = (int)((int64_t)tmp >> 52);
  const double z = asd(ix - (tmp & (0xfffull << 52)));
  const double invc = asd(glog_tab(2 * i)), logc = asd(glog_tab(2 * i + 1));
  const double r = MGCM_FMA(z, invc, -1.0), kd = (double)k;
  const double w = MGCM_FMA(kd, Ln2hi, logc);
  const double hi = w + r;
  const double lo = MGCM_FMA(kd, Ln2lo, (w - hi) + r);
  const double r2 = r * r;
  const double p = MGCM_FMA(r2, MGCM_FMA(r, A4, A3), MGCM_FMA(r, A2, A1));
  return MGCM_FMA(r * r2, p, MGCM_FMA(r2, A0, lo)) + hi;
}

// The reference's LOG (the C library's log).
MGCM_HD double mgcm_log(double x) { return glibc_log(x); }

// glibc's log10 (sysdeps/ieee754/dbl-64/e_log10.c, fdlibm; built without FMA):
// x = 2^e m, log10(x) = e*log10_2lo + ivln10*log(m) + e*log10_2hi.
MGCM_HD double mgcm_log10(double x) {
  const double two54 = 1.80143985094819840000e+16, ivln10 = 4.34294481903251816668e-01,
               log10_2hi = 3.01029995663611771306e-01, log10_2lo = 3.69423907715893078616e-13;
  uint64_t u = asu(x);
  int32_t hx = (int32_t)(u >> 32);
  int32_t k = 0;
  if (hx < 0x00100000) {
    if (((hx & 0x7fffffff) | (uint32_t)u) == 0) return -two54 / __builtin_fabs(x);
    if (hx < 0) return (x - x) / (x - x);
    k -= 54;
    x *= two54;
    u = asu(x);
    hx = (int32_t)(u >> 32);
  }
  if (hx >= 0x7ff00000) return x + x;
  k += (hx >> 20) - 1023;
  const int32_t i = (int32_t)(((uint32_t)k & 0x80000000u) >> 31);
  hx = (hx & 0x000fffff) | ((0x3ff - i) << 20);
  const double y = (double)(k + i);
  const double m = asd(((uint64_t)(uint32_t)hx << 32) | (u & 0xffffffffull));
  const double z = y * log10_2lo + ivln10 * glibc_log(m);
  return z + y * log10_2hi;
}

}  // namespace mgcm_libm

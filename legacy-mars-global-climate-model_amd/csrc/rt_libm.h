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
//  * mgcm_log10: glibc's fdlibm wrapper (sysdeps/ieee754/dbl-64/e_log10.c):
//    x = 2^e m, log10(x) = e*log10_2lo + ivln10*log(m) + e*log10_2hi, over a
//    correctly rounded log (a 129-point table and a double-double atanh series,
//    rt_log_table.inc generated by gen_log_table.py).  It equals glibc's
//    log10 except where glibc's own log misrounds (about 3 in 10^4 arguments).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MGCM_HD __host__ __device__ __forceinline__
#define MGCM_FMA(a, b, c) __builtin_fma((a), (b), (c))
#else
#include <math.h>
#define MGCM_HD static inline
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
__device__ __constant__ static const uint64_t kLogTab[258] = {
#include "rt_log_table.inc"
};
static const uint64_t kLogTabHost[258] = {
#include "rt_log_table.inc"
};
MGCM_HD uint64_t log_tab(int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return kLogTab[i];
#else
  return kLogTabHost[i];
#endif
}
#else
static const uint64_t kLogTabHost[258] = {
#include "rt_log_table.inc"
};
MGCM_HD uint64_t log_tab(int i) { return kLogTabHost[i]; }
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

// ---- double-double helpers ----
struct DD {
  double h, l;
};
MGCM_HD DD dd_two_sum(double a, double b) {
  const double s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
MGCM_HD DD dd_fast(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}
MGCM_HD DD dd_prod(double a, double b) {
  const double p = a * b;
  return {p, MGCM_FMA(a, b, -p)};
}
MGCM_HD DD dd_add(DD a, DD b) {
  const DD s = dd_two_sum(a.h, b.h);
  return dd_fast(s.h, s.l + a.l + b.l);
}
MGCM_HD DD dd_mul(DD a, DD b) {
  const DD p = dd_prod(a.h, b.h);
  return dd_fast(p.h, p.l + (a.h * b.l + a.l * b.h));
}
MGCM_HD DD dd_div(DD a, DD b) {
  const double q1 = a.h / b.h;
  const DD p = dd_mul({q1, 0.0}, b);
  const DD r = dd_add(a, {-p.h, -p.l});
  return dd_fast(q1, r.h / b.h);
}

// k*ln2 + log(m) for m in [0.70703125, 1.4140625), rounded once from
// double-double (2^-100 relative): correctly rounded but for ties closer than
// that.  m = c * (1+s)/(1-s) with c the table point below m (above it under 1,
// so c = 1 on both sides of 1): log(m) = log(c) + 2 atanh(s), |s| < 1/256, the
// atanh series to s^13; log(c) is a double-double table (rt_log_table.inc).
MGCM_HD double log_dd_round(double m, int k) {
  int idx;
  double c;
  if (m >= 1.0) {
    const int j = (int)((m - 1.0) * 128.0);
    idx = j;
    c = 1.0 + j * 0x1p-7;
  } else {
    const int j = (int)((1.0 - m) * 256.0);
    idx = 53 + j;
    c = 1.0 - j * 0x1p-8;
  }
  const DD s = dd_div({m - c, 0.0}, dd_two_sum(m, c));  // m - c is exact
  const DD t = dd_mul(s, s);
  DD p = {0.0, 0.0};
  // unrolled: the coefficients 1/(2j+1) fold to double-double constants at
  // compile time (the same IEEE operations, evaluated once)
#pragma unroll
  for (int j = 6; j >= 1; j--) p = dd_add(dd_mul(p, t), dd_div({1.0, 0.0}, {2.0 * j + 1.0, 0.0}));
  p = dd_add(dd_mul(p, t), {1.0, 0.0});
  DD L = dd_mul(s, p);
  L.h *= 2.0;
  L.l *= 2.0;
  const DD logc = {asd(log_tab(2 * idx)), asd(log_tab(2 * idx + 1))};
  const DD ln2 = {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
  const DD r = dd_add(dd_add(dd_mul({(double)k, 0.0}, ln2), logc), L);
  return r.h + r.l;
}

// Correctly rounded natural log of a positive normal double in [0.5, 2).
MGCM_HD double cr_log_unit(double m) {
  int k = 0;
  if (m < 0.70703125) { m *= 2.0; k = -1; }
  else if (m >= 1.4140625) { m *= 0.5; k = 1; }
  return log_dd_round(m, k);
}

// Correctly rounded natural log of a positive normal double.
MGCM_HD double mgcm_log(double x) {
  const uint64_t u = asu(x);
  int k = (int)((u >> 52) & 0x7ff) - 1023;
  double m = asd((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
  if (m >= 1.4140625) { m *= 0.5; k += 1; }
  return log_dd_round(m, k);
}

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
  const double z = y * log10_2lo + ivln10 * cr_log_unit(m);
  return z + y * log10_2hi;
}

}  // namespace mgcm_libm

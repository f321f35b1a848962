// gen_log_table.c -- writes rt_glibc_log_table.inc, the 128-entry {invc, logc}
// table of the table-driven natural log in rt_libm.h (the Arm
// optimized-routines method that glibc 2.28+ ships as its double log,
// sysdeps/ieee754/dbl-64/e_log.c, with N = 128 subintervals).
//
//   gcc -O2 -fopenmp -ffp-contract=off -mfma gen_log_table.c -lquadmath -lm -o gen_log_table
//   ./gen_log_table > rt_glibc_log_table.inc
//
// The method: x = 2^k z with z in [0x1.6p-1, 0x1.6p0), split by the top 7
// mantissa bits of (bits(x) - bits(0x1.6p-1)) into 128 subintervals;
// log(x) = k ln2 + logc + log1p(z invc - 1).  Entry i holds a double invc
// within 2^29 ulps of 1/centre(i) such that
//   (1) logc = -log(invc) rounded to double is a multiple of 2^-43 (so that
//       k*Ln2hi + logc is exact), and
//   (2) |-log(invc) - logc| < 2^-66.
// About 250 doubles per subinterval satisfy (1) and (2).  The published
// method picks one of them by a further criterion on the no-FMA reduction
// (z - chi - clo) * invc; rather than restate that criterion, this program
// keeps the candidate whose complete log (the function below, the same
// operations as rt_libm.h) returns the host C library's bits on 400,000
// pseudo-random arguments over 61 binades of the subinterval.  Exactly one
// candidate passes in every subinterval on glibc 2.35 (x86-64, FMA build).
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t asu(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double asd(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

// log1p(r) - r polynomial on |r| < 0x1.fp-9 and ln2 split: published constants
static const double A[5] = {-0x1.0000000000001p-1, 0x1.555555551305bp-2, -0x1.fffffffeb459p-3,
                            0x1.999b324f10111p-3, -0x1.55575e506c89fp-3};
static const double Ln2hi = 0x1.62e42fefa3800p-1, Ln2lo = 0x1.ef35793c76730p-45;
static const uint64_t OFF = 0x3fe6000000000000ull;

// The table path of the log for one candidate entry (rt_libm.h glibc_log_tab).
static double table_log(double x, double invc, double logc) {
  const uint64_t ix = asu(x), tmp = ix - OFF;
  const int k = (int)((int64_t)tmp >> 52);
  const double z = asd(ix - (tmp & (0xfffull << 52)));
  const double r = fma(z, invc, -1.0), kd = (double)k;
  const double w = fma(kd, Ln2hi, logc), hi = w + r, lo = fma(kd, Ln2lo, (w - hi) + r);
  const double r2 = r * r;
  return fma(r * r2, fma(r2, fma(r, A[4], A[3]), fma(r, A[2], A[1])), fma(r2, A[0], lo)) + hi;
}

#define MAXC 512
#define NARG 400000

int main(void) {
  static double out_inv[128], out_logc[128];
  static int nfound[128], ncand[128];
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < 128; i++) {
    const uint64_t zlo_b = (i < 80) ? OFF + ((uint64_t)i << 45) : 0x3ff0000000000000ull + ((uint64_t)(i - 80) << 45);
    const double zlo = asd(zlo_b), zhi = asd(zlo_b + (1ull << 45));
    const __float128 centre = ((__float128)zlo + (__float128)zhi) / 2;
    const double v0 = (double)(1 / centre);
    const double u = nextafter(v0, 2.0) - v0;
    const __float128 q43 = ldexpq(1.0Q, -43), eps = ldexpq(1.0Q, -66);
    const __float128 a = (__float128)v0 - ldexpq(1.0Q, 29) * u, b = (__float128)v0 + ldexpq(1.0Q, 29) * u;
    const long long m0 = (long long)ceilq(-logq(b) / q43), m1 = (long long)floorq(-logq(a) / q43);
    const __float128 step = expq(-q43);
    double cinv[MAXC], clog[MAXC];
    int nc = 0;
    __float128 iv = 0;
    // walk the multiples m 2^-43 of the window; the double nearest exp(-m 2^-43)
    // is the only candidate for m
    for (long long m = m0; m <= m1; m++) {
      iv = ((m - m0) & 4095) == 0 ? expq(-(__float128)m * q43) : iv * step;
      const __float128 js = (iv - (__float128)v0) / u, jr = roundq(js);
      if (fabsq((jr - js) * u / iv) >= 2 * eps) continue;  // cheap bound first
      const double cand = (double)((__float128)v0 + jr * u);
      const __float128 err = -logq((__float128)cand) - (__float128)m * q43;
      if (fabsq(err) < eps && nc < MAXC) {
        cinv[nc] = cand;
        clog[nc] = (double)((__float128)m * q43);
        nc++;
      }
    }
    ncand[i] = nc;
    // arguments of this subinterval in binades 2^-30..2^30, off the near-1 path
    double *xs = malloc(sizeof(double) * NARG), *ys = malloc(sizeof(double) * NARG);
    uint64_t s = 0x9e3779b97f4a7c15ull ^ (uint64_t)i;
    for (int n = 0; n < NARG;) {
      s ^= s << 13; s ^= s >> 7; s ^= s << 17;
      const double x = ldexp(asd(zlo_b + (s & ((1ull << 45) - 1))), (int)((s >> 45) % 61) - 30);
      if (asu(x) - asu(1.0 - 0x1p-4) < asu(1.0 + 0x1.09p-4) - asu(1.0 - 0x1p-4)) continue;
      xs[n] = x;
      ys[n] = log(x);
      n++;
    }
    for (int c = 0; c < nc; c++) {
      int ok = 1;
      for (int n = 0; n < NARG && ok; n++) ok = asu(table_log(xs[n], cinv[c], clog[c])) == asu(ys[n]);
      if (ok) {
        if (nfound[i] == 0) { out_inv[i] = cinv[c]; out_logc[i] = clog[c]; }
        nfound[i]++;
      }
    }
    free(xs);
    free(ys);
  }
  int bad = 0;
  for (int i = 0; i < 128; i++) {
    if (nfound[i] != 1) {
      fprintf(stderr, "subinterval %d: %d of %d candidates match the host log\n", i, nfound[i], ncand[i]);
      bad = 1;
    }
  }
  if (bad) return 1;
  printf("// generated by gen_log_table.c: {bits(invc), bits(logc)} of subintervals i = 0..127\n");
  for (int i = 0; i < 128; i++) printf("0x%016llxULL, 0x%016llxULL,\n", (unsigned long long)asu(out_inv[i]),
                                       (unsigned long long)asu(out_logc[i]));
  return 0;
}

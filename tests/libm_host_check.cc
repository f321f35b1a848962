// Host build of the device's libm restatements (rt_libm.h) compared bit for bit
// with the host C library on n pseudo-random arguments: exp over its finite
// range, log and log10 over positive doubles of every binade, pressures
// 1e-7..1e5 and values near 1.  Prints the difference counts; exit 1 on any.
#include "rt_libm.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

static bool same(double a, double b) { return std::memcmp(&a, &b, 8) == 0 || (a != a && b != b); }

int main(int argc, char **argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 10000000;
  uint64_t s = 0x243f6a8885a308d3ull;
  long de = 0, dl = 0, d10 = 0;
  for (long i = 0; i < n; i++) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    const double u = (double)(s >> 11) * 0x1p-53;
    double x;
    switch (i % 4) {
      case 0: x = mgcm_libm::asd(s & 0x7fefffffffffffffull); break;  // any non-negative finite
      case 1: x = 1e-7 * std::pow(10.0, 12.0 * u); break;
      case 2: x = 0.9 + 0.2 * u; break;
      default: x = -745.0 + 1454.0 * u; break;  // exp's finite range
    }
    if (!same(std::exp(x), mgcm_libm::mgcm_exp(x))) de++;
    if (x > 0) {
      if (!same(std::log(x), mgcm_libm::mgcm_log(x))) dl++;
      if (!same(std::log10(x), mgcm_libm::mgcm_log10(x))) d10++;
    }
  }
  const double sp[] = {0.0, -0.0, 1.0, -1.0, __builtin_inf(), -__builtin_inf(), 4.9e-324, 1e-310,
                       2.2250738585072014e-308, 0.5, 2.0, 1.7976931348623157e308, __builtin_nan("")};
  for (double x : sp) {
    if (!same(std::exp(x), mgcm_libm::mgcm_exp(x))) de++;
    if (!same(std::log(x), mgcm_libm::mgcm_log(x))) dl++;
    if (!same(std::log10(x), mgcm_libm::mgcm_log10(x))) d10++;
  }
  std::printf("%ld arguments: exp %ld, log %ld, log10 %ld differ\n", n, de, dl, d10);
  return (de || dl || d10) ? 1 : 0;
}

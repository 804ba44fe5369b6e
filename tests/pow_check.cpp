// Host check of pow_pos (pyminiweather_amd/csrc/pmw_pow.h) against a 113-bit
// reference (quadmath powq, rounded once to double) and glibc's pow.
// Usage: pow_check N lo hi y  -> prints "N mismatch_cr mismatch_glibc
// glibc_vs_cr special_mismatches max_rel_err_of_pow_pos_fast"
#include <quadmath.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include "pmw_pow.h"
static inline uint64_t next(uint64_t* s) {
  uint64_t x = *s;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return *s = x;
}
int main(int argc, char** argv) {
  if (argc < 5) return 2;
  const long long n = atoll(argv[1]);
  const double lo = atof(argv[2]), hi = atof(argv[3]), y = atof(argv[4]);
  uint64_t s = 0x9E3779B97F4A7C15ULL;
  long long bad_cr = 0, bad_glibc = 0, glibc_cr = 0;
  double fast_err = 0.0;
  for (long long i = 0; i < n; ++i) {
    const double u = (double)(next(&s) >> 11) * 0x1p-53;
    const double x = exp(log(lo) + u * (log(hi) - log(lo)));
    const double mine = pmw::pow_pos(x, y);
    const __float128 exact = powq((__float128)x, (__float128)y);
    const double cr = (double)exact;
    const double fe = fabs((double)(((__float128)pmw::pow_pos_fast(x, y) - exact) / exact));
    if (fe > fast_err) fast_err = fe;
    const double g = pow(x, y);
    bad_cr += mine != cr;
    bad_glibc += mine != g;
    glibc_cr += g != cr;
  }
  const double specials[] = {0.0, -0.0, -1.0, INFINITY, -INFINITY, NAN};
  int special_bad = 0;
  for (double x : specials) {
    const double a = pmw::pow_pos(x, y), b = pow(x, y);
    special_bad += !((std::isnan(a) && std::isnan(b)) || a == b);
    const double f = pmw::pow_pos_fast(x, y);  // NaN for these, never a wild read
    special_bad += !(std::isnan(f) || (x == INFINITY && f == INFINITY));
  }
  printf("%lld %lld %lld %lld %d %.17g\n", n, bad_cr, bad_glibc, glibc_cr, special_bad, fast_err);
  return 0;
}

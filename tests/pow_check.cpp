// Host check of pow_pos (pyminiweather_amd/csrc/pmw_pow.h) against a 113-bit
// reference (quadmath powq, rounded once to double) and glibc's pow.
// Usage: pow_check N lo hi y  -> prints "N mismatch_cr mismatch_glibc
// glibc_vs_cr special_mismatches max_rel_err_of_pow_pos_fast"
//        pow_check N lo hi y q EMAX [RXERR] -> pow_pos_near == pow_pos count
//        (reference reciprocal scaled by 1 +- RXERR); prints "N diff inrange"
//        pow_check N lo hi y near EMAX -> pow_pos_near continued from a
// reference at x0 to x0 (1 + e), |e| log-uniform in [1e-15, EMAX] with a
// random sign (one argument in 16 equal to x0); prints "N near_vs_cr
// near_vs_pow_pos pow_pos_vs_cr fallbacks"
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
  if (argc >= 7 && argv[5][0] == 'e') {  // max relative error before rounding
    const double emax = atof(argv[6]);
    double pp_err = 0.0, near_err = 0.0;
    for (long long i = 0; i < n; ++i) {
      const double u = (double)(next(&s) >> 11) * 0x1p-53;
      const double x0 = exp(log(lo) + u * (log(hi) - log(lo)));
      const double v = (double)(next(&s) >> 11) * 0x1p-53;
      double e = exp(log(1e-15) + v * (log(emax) - log(1e-15)));
      if (next(&s) & 1) e = -e;
      const double x = x0 * (1.0 + e);
      pmw::PowRef ref;
      (void)pmw::pow_pos_ref(x0, y, ref);
      const __float128 ex0 = powq((__float128)x0, (__float128)y);
      const double a = fabs((double)((((__float128)ref.h + ref.l) - ex0) / ex0));
      if (a > pp_err) pp_err = a;
      if (fabs(x - x0) <= 0x1p-10 * x0) {
        double sh, sl;
        pmw::pow_pos_near_dd(x, y, ref, sh, sl);
        const __float128 ex = powq((__float128)x, (__float128)y);
        const double b = fabs((double)((((__float128)sh + sl) - ex) / ex));
        if (b > near_err) near_err = b;
      }
    }
    printf("%lld %.6g %.6g\n", n, pp_err, near_err);
    return 0;
  }
  if (argc >= 7 && argv[5][0] == 'q') {  // pow_pos_near == pow_pos, no quadmath
    const double emax = atof(argv[6]);
    // optional: the reference's reciprocal scaled by (1 + rxerr), as an
    // approximate reciprocal would leave it
    const double rxerr = argc >= 8 ? atof(argv[7]) : 0.0;
    long long diff = 0, inrange = 0;
    for (long long i = 0; i < n; ++i) {
      const double u = (double)(next(&s) >> 11) * 0x1p-53;
      const double x0 = exp(log(lo) + u * (log(hi) - log(lo)));
      const double v = (double)(next(&s) >> 11) * 0x1p-53;
      double e = exp(log(1e-15) + v * (log(emax) - log(1e-15)));
      if (next(&s) & 1) e = -e;
      const double x = x0 * (1.0 + e);
      pmw::PowRef ref;
      (void)pmw::pow_pos_ref(x0, y, ref);
      ref.rx *= 1.0 + ((i & 1) ? rxerr : -rxerr);
      diff += pmw::pow_pos_near(x, y, ref) != pmw::pow_pos(x, y);
      inrange += fabs(x - x0) <= 0x1p-10 * x0;
    }
    printf("%lld %lld %lld\n", n, diff, inrange);
    return diff != 0;
  }
  if (argc >= 7) {  // pow_pos_near
    const double emax = atof(argv[6]);
    long long near_cr = 0, near_pp = 0, pp_cr = 0, fallback = 0;
    for (long long i = 0; i < n; ++i) {
      const double u = (double)(next(&s) >> 11) * 0x1p-53;
      const double x0 = exp(log(lo) + u * (log(hi) - log(lo)));
      const double v = (double)(next(&s) >> 11) * 0x1p-53;
      double e = exp(log(1e-15) + v * (log(emax) - log(1e-15)));
      if (next(&s) & 1) e = -e;
      if ((next(&s) & 15) == 0) e = 0.0;
      const double x = x0 * (1.0 + e);
      pmw::PowRef ref;
      (void)pmw::pow_pos_ref(x0, y, ref);
      const double near = pmw::pow_pos_near(x, y, ref);
      const double pp = pmw::pow_pos(x, y);
      const double cr = (double)powq((__float128)x, (__float128)y);
      near_cr += near != cr;
      near_pp += near != pp;
      pp_cr += pp != cr;
      fallback += !(fabs(x - x0) <= 0x1p-10 * x0);
    }
    printf("%lld %lld %lld %lld %lld\n", n, near_cr, near_pp, pp_cr, fallback);
    return 0;
  }
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

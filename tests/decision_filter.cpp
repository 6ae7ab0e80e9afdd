// Host check of removes_sq (the squared decision filter of the by-set
// search) against removes(s01 / sqrt(q), r_crit): ratios within a few hundred
// ulps of r_crit on both sides, random wide-range inputs, and the r_crit
// values for which the filter must never decide.  Prints the counts; exit
// status 1 on any disagreement.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>

#include "pc_math.hpp"

using namespace hpc;

int main() {
  std::mt19937_64 g(12345);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  long long n = 0, bad = 0, fast_rm = 0, fast_keep = 0, band = 0;
  auto check = [&](double s01, double q, double rc) {
    double lo, hi;
    square_bounds(rc, lo, hi);
    const bool want = removes(s01 / std::sqrt(q), rc);
    const bool got = removes_sq(s01, q, rc, lo, hi);
    const double s2 = s01 * s01;
    if (s2 <= lo * q) ++fast_rm; else if (s2 >= hi * q) ++fast_keep; else ++band;
    ++n;
    if (got != want) {
      if (bad < 10) std::printf("MISMATCH s01=%a q=%a r_crit=%a want=%d got=%d\n", s01, q, rc, want, got);
      ++bad;
    }
  };
  std::vector<double> rcs = {1e-90, 1e-30, 1e-8, 3e-3, 0.0123456789, 0.1, 0.5, 0.77, 0.999, 0.9999998,
                             std::nextafter(0.9999999, 0.0), 0.9999999, -1.0, 0.0, INFINITY};
  for (int k = 0; k < 200; ++k) rcs.push_back(u01(g));
  for (double rc : rcs) {
    for (int t = 0; t < 4000; ++t) {
      // q across many decades, a ratio near r_crit, both signs
      const double q = std::ldexp(0.5 + 0.5 * u01(g), -(int)(u01(g) * 80));
      const double sq = std::sqrt(q);
      const double x = std::isfinite(rc) && rc > 0 ? rc : u01(g);
      double s01 = x * sq;
      const int steps = (int)(u01(g) * 800) - 400;
      for (int s = 0; s < std::abs(steps); ++s) s01 = std::nextafter(s01, steps > 0 ? 2.0 : 0.0);
      if (t & 1) s01 = -s01;
      check(s01, q, rc);
      // wide range
      check((u01(g) * 2 - 1) * std::ldexp(1.0, -(int)(u01(g) * 60)), q, rc);
    }
  }
  std::printf("cases %lld mismatches %lld (filter removed %lld, kept %lld, band %lld)\n", n, bad,
              fast_rm, fast_keep, band);
  return bad != 0 || band == 0 || fast_rm == 0 || fast_keep == 0;
}

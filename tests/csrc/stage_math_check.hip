// Host check of fs_log (fsim_phys.h) against long-double libm:
// prints the largest error in units of the double result's ulp, and the
// fraction of arguments not rounded to the nearest double, over random
// arguments in the ranges the stage kernel uses and beyond.
#include <math.h>
#include <stdio.h>

#include <random>

#include "fsim_phys.h"

static double ulps(double got, long double ref) {
  const double r = (double)ref;
  const double u = nextafter(fabs(r), INFINITY) - fabs(r);
  return (double)(fabsl((long double)got - ref) / (long double)u);
}

int main() {
  std::mt19937_64 g(12345);
  struct Range {
    const char* fn;
    double lo, hi;
  };
  const Range ranges[] = {{"log", 0.234, 157.0}, {"log", 0.70, 1.42}, {"log", 1e-300, 1e300}};
  for (const Range& rg : ranges) {
    std::uniform_real_distribution<double> U(log(rg.lo), log(rg.hi));
    double worst = 0.0, arg = 0.0;
    long wrong = 0;
    const long n = 2000000;
    for (long i = 0; i < n; i++) {
      const double x = exp(U(g));
      const double got = fs_log(x);
      const long double ref = logl((long double)x);
      const double e = ulps(got, ref);
      if (got != (double)ref) wrong++;
      if (e > worst) {
        worst = e;
        arg = x;
      }
    }
    printf("%s [%g, %g] max_ulp %.4f at %.17g not_nearest %.6f\n", rg.fn, rg.lo, rg.hi, worst,
           arg, (double)wrong / n);
  }
  return 0;
}

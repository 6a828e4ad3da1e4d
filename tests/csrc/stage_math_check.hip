// Host check of the stage kernel's short arithmetic (fsim_phys.h), compiled
// for the host from the same header:
//   log, exp, div, rcp   largest error in units of the double result's ulp
//                        against long-double libm / long-double arithmetic,
//                        and the fraction of results not rounded to nearest
//   fit_*                the Horner-form exponent of each fit against the
//                        reference's power-sum order: largest absolute
//                        difference (= relative difference of exp(.)), as
//                        "ulp" of 1.0 (units of 2^-52)
// over random arguments in the ranges the stage kernel uses and beyond.  The
// device's division starts from v_rcp_f64; the host version starts from a
// float-precision reciprocal, a worse seed for the same two Newton steps.
#include <math.h>
#include <stdio.h>

#include <random>

#include "fsim_phys.h"

static double ulps(double got, long double ref) {
  const double r = (double)ref;
  const double u = nextafter(fabs(r), INFINITY) - fabs(r);
  return (double)(fabsl((long double)got - ref) / (long double)u);
}

struct Stat {
  double worst = 0.0, arg = 0.0;
  long wrong = 0, n = 0;
  void add(double e, bool nearest, double x) {
    n++;
    if (!nearest) wrong++;
    if (e > worst) {
      worst = e;
      arg = x;
    }
  }
  void print(const char* fn, double lo, double hi) const {
    printf("%s [%g, %g] max_ulp %.4f at %.17g not_nearest %.6f\n", fn, lo, hi, worst, arg,
           (double)wrong / n);
  }
};

int main() {
  std::mt19937_64 g(12345);
  const long n = 1000000;
  {  // log on log-uniform arguments
    const double rg[][2] = {{0.234, 157.0}, {0.70, 1.42}, {1e-300, 1e300}};
    for (auto& r : rg) {
      std::uniform_real_distribution<double> U(log(r[0]), log(r[1]));
      Stat st;
      for (long i = 0; i < n; i++) {
        const double x = exp(U(g));
        const double got = fs_log(x);
        const long double ref = logl((long double)x);
        st.add(ulps(got, ref), got == (double)ref, x);
      }
      st.print("log", r[0], r[1]);
    }
  }
  {  // exp: SG exponents [-100, 0], fit exponents [-200, 80], and wider
    const double rg[][2] = {{-100.0, 0.0}, {-200.0, 80.0}, {-700.0, 700.0}};
    for (auto& r : rg) {
      std::uniform_real_distribution<double> U(r[0], r[1]);
      Stat st;
      for (long i = 0; i < n; i++) {
        const double x = U(g);
        const double got = fs_exp(x);
        const long double ref = expl((long double)x);
        st.add(ulps(got, ref), got == (double)ref, x);
      }
      st.print("exp", r[0], r[1]);
    }
  }
  {  // div and rcp: magnitudes of the kernel's operands (densities, energies,
     // velocities x spacing, 1 - exp(-a))
    std::uniform_real_distribution<double> E(-30.0, 30.0), M(1.0, 2.0);
    Stat sd, sr;
    for (long i = 0; i < n; i++) {
      const double a = M(g) * ldexp(1.0, (int)E(g)), b = M(g) * ldexp(1.0, (int)E(g));
      const double q = fs_div(a, b);
      const long double qr = (long double)a / (long double)b;
      sd.add(ulps(q, qr), q == (double)qr, b);
      const double r = fs_rcp(b);
      const long double rr = 1.0L / (long double)b;
      sr.add(ulps(r, rr), r == (double)rr, b);
    }
    sd.print("div", 1e-9, 1e9);
    sr.print("rcp", 1e-9, 1e9);
  }
  {  // Horner-form fit exponents against the power-sum order, x = log(T ph0)
     // over each fit's own clamp interval (both branches of split fits)
    const double l5119 = (double)(float)log((double)5.119f);
    const double l5667 = (double)(float)log((double)5.667f);
    const double l75 = (double)(float)log((double)7.5f);
    FsPhys K = {};
    K.log5119 = l5119;
    struct Fit {
      const char* name;
      double lo, hi;
    };
    const Fit fits[] = {{"fit_mue", 0.234, 157.0}, {"fit_De", 0.234, 157.0},
                        {"fit_kex", 0.9611, 157.0}, {"fit_kir", 1.37, 157.0},
                        {"fit_nu", 0.234, 157.0},   {"fit_ksc", 0.5, 16.0},
                        {"fit_ksi", 0.5, 16.0}};
    for (int f = 0; f < 7; f++) {
      std::uniform_real_distribution<double> U(log(fits[f].lo), log(fits[f].hi));
      Stat st;
      for (long i = 0; i < n / 4; i++) {
        const double x = U(g);
        double a = 0, b = 0;
        switch (f) {
          case 0: a = mue_poly_h(x, l5119); b = mue_poly(K, x); break;
          case 1: a = de_poly_h(x); b = de_poly(x); break;
          case 2: a = kex_poly_h(x, l5667); b = kex_poly_ref(x, l5667); break;
          case 3: a = kir_poly_h(x, l75); b = kir_poly_ref(x, l75); break;
          case 4: a = nu_poly_h(x); b = nu_poly_ref(x); break;
          case 5: a = ksc_poly_h(x); b = ksc_poly_ref(x); break;
          case 6: a = ksi_poly_h(x); b = ksi_poly_ref(x); break;
        }
        st.add(fabs(a - b) / ldexp(1.0, -52), a == b, x);
      }
      st.print(fits[f].name, fits[f].lo, fits[f].hi);
    }
  }
  return 0;
}

// Host check of the FP64 exp, log, reciprocal, sine and cosine that the per-event code uses
// (morefit_amd/csrc/kernel_runtime.hip, from mf_exp to the MF_SINCOS_END marker), against the C
// library to 2 ulp.  tests/test_fp64_math.py cuts that block out of the runtime into
// fp64_math.inc and builds this file with g++; the hardware reciprocal approximation and the
// frexp builtins become their exact host counterparts, so mf_rcp is checked from an exact start.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#define __device__
#define __forceinline__ static inline
static inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
static inline double __builtin_amdgcn_frexp_mant(double x) { int e; return std::frexp(x, &e); }
static inline int __builtin_amdgcn_frexp_exp(double x) { int e; std::frexp(x, &e); return e; }
using std::fma; using std::rint; using std::floor; using std::fmin; using std::fmax; using std::ldexp;
using std::exp2; using std::log2; using std::sqrt;
#include "fp64_math.inc"

static double ulps(double got, double ref) {
  if (std::isnan(ref)) return std::isnan(got) ? 0 : 1e300;
  if (std::isinf(ref)) return got == ref ? 0 : 1e300;
  if (ref == 0) return got == 0 ? 0 : 1e300;
  const double u = std::nextafter(std::fabs(ref), INFINITY) - std::fabs(ref);
  return std::fabs(got - ref) / u;
}
static int fails = 0;
static void check(const char* what, double worst, double bar) {
  std::printf("%-34s max %.2f ulp\n", what, worst);
  if (!(worst <= bar)) { std::printf("  FAIL (bar %.1f ulp)\n", bar); fails++; }
}

int main() {
  std::mt19937_64 g(11);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  double w = 0;
  for (int i = 0; i < 1000000; i++) { const double x = 745.0 * u(g); w = std::fmax(w, ulps(mf_exp(x), std::exp(x))); }
  check("exp on [-745, 745]", w, 2.0);
  w = 0;
  for (int i = 0; i < 1000000; i++) { const double x = std::ldexp(1.0 + std::fabs(u(g)), (int)(1000 * u(g))); w = std::fmax(w, ulps(mf_log(x), std::log(x))); }
  for (double x = 0.5; x < 2.0; x += 1e-6) w = std::fmax(w, ulps(mf_log(x), std::log(x)));
  check("log on 2^[-1000, 1000]", w, 2.0);
  w = 0;
  for (int i = 0; i < 1000000; i++) { const double x = std::ldexp(u(g), (int)(1000 * u(g))); if (x != 0) w = std::fmax(w, ulps(mf_rcp(x), 1.0 / x)); }
  check("rcp on +-2^[-1000, 1000]", w, 1.0);
  const double ranges[] = {1e-8, 1.0, 3.5, 100.0, 1e6, 1e12};
  for (double R : ranges) {
    double ws = 0, wc = 0;
    for (int i = 0; i < 1000000; i++) {
      const double x = R * u(g);
      ws = std::fmax(ws, ulps(mf_sin(x), std::sin(x)));
      wc = std::fmax(wc, ulps(mf_cos(x), std::cos(x)));
    }
    char b[64];
    std::snprintf(b, sizeof b, "sin on [-%g, %g]", R, R); check(b, ws, 2.0);
    std::snprintf(b, sizeof b, "cos on [-%g, %g]", R, R); check(b, wc, 2.0);
  }
  w = 0;  // next to multiples of pi/2, where the reduction cancels most digits
  for (int k = -20000; k <= 20000; k++) {
    double x = k * 1.5707963267948966;
    for (int d = 0; d < 3; d++, x = std::nextafter(x, INFINITY))
      if (x != 0) w = std::fmax(w, std::fmax(ulps(mf_sin(x), std::sin(x)), ulps(mf_cos(x), std::cos(x))));
  }
  check("sin, cos next to k pi/2", w, 2.0);
  const double sp[] = {0.0, 1e-300, -0.75, 710.0, -746.0, 800.0, -800.0};
  for (double x : sp) w = std::fmax(w, ulps(mf_exp(x), std::exp(x)));
  check("exp at special arguments", w, 2.0);
  const bool specials = std::isnan(mf_sin(INFINITY)) && std::isnan(mf_cos(-INFINITY)) && std::isnan(mf_sin(NAN)) &&
                        std::isnan(mf_exp(NAN)) && mf_exp(INFINITY) == INFINITY && mf_exp(-INFINITY) == 0.0 &&
                        mf_log(0.0) == -INFINITY && std::isnan(mf_log(-1.0)) && mf_log(INFINITY) == INFINITY &&
                        mf_cos(0.0) == 1.0 && mf_sin(0.0) == 0.0;
  std::printf("special values %s\n", specials ? "ok" : "FAIL");
  if (!specials) fails++;
  std::printf("%s\n", fails ? "FAILED" : "all ok");
  return fails ? 1 : 0;
}

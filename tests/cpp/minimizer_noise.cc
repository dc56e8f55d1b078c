// minimizer_noise.cc -- CPU check of the minimiser (morefit/minimizer.hh) on the mass
// likelihood when the objective's value carries FP32-like rounding: the oracle's exact FP64
// -2 lnL, gradient and Hessian, with a coherent value perturbation of amplitude
// amp * n * 2^-24 that is piecewise constant in theta (a hash of theta's bits), as the
// rounding of FP32 kernel parameters makes it.  The same data are fitted without the
// perturbation; both fits must converge (status 0) to the same minimum.
//
// usage: minimizer_noise <data.f64> <n> <amp> <mu> <sigma> <f> <lambda> [nofloor]
// (nofloor: the perturbed objective reports no rounding floor, as an FP64 objective would)
// prints: status_exact hesse_ok status_noisy max|dtheta|/sigma calls_exact calls_noisy
// (max|dtheta|/sigma is nan when HESSE fails on the exact fit, e.g. at a parameter limit)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "morefit/minimizer.hh"

extern "C" int orc_nll(int model, const double* data, long n, long stride, const double* theta, const double* range, int mode,
                       int nthreads, double* out);

namespace {

const double kRange[2] = {5.0, 5.6};

struct MassFcn : morefit::MinimizerFcn {
  const std::vector<double>* data;
  double amp;  // perturbation amplitude in -2 lnL units (0: exact objective)
  bool fp32_noise_floor;

  double perturbation(const double* x) const {
    if (amp == 0.0) return 0.0;
    uint64_t h = 1469598103934665603ULL;
    for (int k = 0; k < 4; k++) {
      // theta rounded to FP32, as the kernel parameters are: constant between FP32 neighbours
      float xf = (float)x[k];
      uint32_t b;
      std::memcpy(&b, &xf, 4);
      for (int j = 0; j < 4; j++) {
        h ^= (b >> (8 * j)) & 0xffu;
        h *= 1099511628211ULL;
      }
    }
    return amp * ((double)(h >> 11) * std::ldexp(1.0, -53) * 2.0 - 1.0);
  }
  double sums(const double* x, int mode, double* out) const {
    if (orc_nll(0, data->data(), (long)data->size(), (long)data->size(), x, kRange, mode, 4, out) != 0) return NAN;
    return -2.0 * out[0] + perturbation(x);
  }
  double value(const double* x) override {
    double o[1];
    return sums(x, 0, o);
  }
  double value_grad(const double* x, double* g) override {
    double o[5];
    double f = sums(x, 1, o);
    for (int k = 0; k < 4; k++) g[k] = -2.0 * o[1 + k];
    return f;
  }
  bool hessian(const double* x, double* h) override {
    double o[15];
    if (!std::isfinite(sums(x, 2, o))) return false;
    int idx = 5;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j <= i; j++, idx++) h[i * 4 + j] = h[j * 4 + i] = -2.0 * o[idx];
    return true;
  }
  double value_noise() const override { return fp32_noise_floor ? morefit::fp32_sum_noise((double)data->size()) : 0.0; }
};

std::vector<morefit::MinimizerParam> params(const double* start) {
  const char* names[4] = {"mu", "sigma", "f", "lambda"};
  const double lo[4] = {5.0, 0.005, 0.0, -10.0}, hi[4] = {5.6, 0.13, 1.0, 10.0}, step[4] = {0.01, 0.001, 0.01, 0.01};
  std::vector<morefit::MinimizerParam> p(4);
  for (int k = 0; k < 4; k++) {
    p[k].name = names[k];
    p[k].value = start[k];
    p[k].step = step[k];
    p[k].lo = lo[k];
    p[k].hi = hi[k];
    p[k].has_lo = p[k].has_hi = true;
  }
  return p;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 8 && !(argc == 9 && std::strcmp(argv[8], "nofloor") == 0)) {
    std::fprintf(stderr, "usage: %s data.f64 n amp mu sigma f lambda [nofloor]\n", argv[0]);
    return 2;
  }
  const long n = std::atol(argv[2]);
  const double amp_units = std::atof(argv[3]);
  double start[4];
  for (int k = 0; k < 4; k++) start[k] = std::atof(argv[4 + k]);
  std::vector<double> data(n);
  FILE* fp = std::fopen(argv[1], "rb");
  if (!fp || std::fread(data.data(), sizeof(double), n, fp) != (size_t)n) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  std::fclose(fp);

  MassFcn exact;
  exact.data = &data;
  exact.amp = 0.0;
  exact.fp32_noise_floor = false;
  morefit::Migrad m0(params(start));
  m0.minimize(exact);
  const bool hesse_ok = m0.hesse(exact);

  MassFcn noisy;
  noisy.data = &data;
  noisy.amp = amp_units * (double)n * std::ldexp(1.0, -24);
  noisy.fp32_noise_floor = argc == 8;
  morefit::Migrad m1(params(start));
  m1.print_level = std::getenv("MOREFIT_PRINT_LEVEL") ? std::atoi(std::getenv("MOREFIT_PRINT_LEVEL")) : 0;
  m1.minimize(noisy);

  double worst = hesse_ok ? 0.0 : NAN;
  for (int k = 0; k < 4 && hesse_ok; k++) worst = std::max(worst, std::fabs(m1.x[k] - m0.x[k]) / m0.errors[k]);
  std::printf("%d %d %d %.6g %d %d\n", m0.status, (int)hesse_ok, m1.status, worst, m0.ncalls, m1.ncalls);
  return 0;
}

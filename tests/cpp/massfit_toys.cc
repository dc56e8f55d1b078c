// massfit_toys.cc -- a program written against the morefit C++ API with the calls of the
// reference's mass-fit benchmark (benchmarking/morefit_massfit.cc:22-107): compute options,
// a backend chosen by typedefs, the Gaussian + exponential SumPDF with the benchmark's
// dimension and parameters, a Xoshiro128pp seeded with 229387429, generator_options with the
// randomization on the accelerator, fitter_options, and toystudy<...>::toy.
//
// Built with -DMF_HIP_BACKEND the typedefs select HIPBackend / HIPBlock (the drop-in switch the
// reference makes between LLVMBackend and OpenCLBackend), otherwise CPUBackend / CPUBlock;
// -DMF_KERNEL_FLOAT selects FP32 kernels.
//
// usage: massfit_toys NRUNS NEVENTS ANALYTIC [BATCHED [NTHREADS [HOST]]]
//   ANALYTIC 0: analytic_gradient = analytic_hessian = false (the benchmark's setting), 1: true
//   BATCHED 1 (HIP only): after the study, the same toys (same seed, same generator calls)
//   fitted together by toystudy::toy_batched
//   NTHREADS: llvm_nthreads of the CPU backend (threads and generator streams)
//   HOST 1: generator_options::on_host (proposals and accept-reject on the host)
// Output: one JSON object {"sequential": {...}, "batched": {...}} with per-toy values,
// errors, status and event counts.
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "morefit.hh"

#ifdef MF_KERNEL_FLOAT
typedef float kernelT;
#else
typedef double kernelT;
#endif
typedef double evalT;

#ifdef MF_HIP_BACKEND
typedef morefit::HIPBackend backendT;
typedef morefit::HIPBlock<kernelT, evalT> blockT;
#else
typedef morefit::CPUBackend backendT;
typedef morefit::CPUBlock<kernelT, evalT> blockT;
#endif

template <typename S>
static void dump(const char* name, const S& toy, bool last) {
  std::printf("\"%s\": {\"values\": [", name);
  for (size_t i = 0; i < toy.values.size(); i++) {
    std::printf("%s[", i ? ", " : "");
    for (size_t j = 0; j < toy.values[i].size(); j++) std::printf("%s%.17g", j ? ", " : "", toy.values[i][j]);
    std::printf("]");
  }
  std::printf("], \"errors\": [");
  for (size_t i = 0; i < toy.errors.size(); i++) {
    std::printf("%s[", i ? ", " : "");
    for (size_t j = 0; j < toy.errors[i].size(); j++) std::printf("%s%.17g", j ? ", " : "", toy.errors[i][j]);
    std::printf("]");
  }
  std::printf("], \"status\": [");
  for (size_t i = 0; i < toy.status.size(); i++) std::printf("%s%d", i ? ", " : "", toy.status[i]);
  std::printf("], \"events\": [");
  for (size_t i = 0; i < toy.toy_events.size(); i++) std::printf("%s%zu", i ? ", " : "", toy.toy_events[i]);
  std::printf("], \"generate_ms\": %.3f, \"fit_ms\": %.3f}%s\n", toy.generate_ms, toy.fit_ms, last ? "" : ",");
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s NRUNS NEVENTS ANALYTIC [BATCHED [NTHREADS [HOST]]]\n", argv[0]);
    return 2;
  }
  const unsigned nruns = (unsigned)std::atoi(argv[1]), nevents = (unsigned)std::atoi(argv[2]);
  const bool analytic = std::atoi(argv[3]) != 0;
  const bool batched = argc > 4 && std::atoi(argv[4]) != 0;
  const unsigned nthreads = argc > 5 ? (unsigned)std::atoi(argv[5]) : 1;
  const bool on_host = argc > 6 && std::atoi(argv[6]) != 0;

  morefit::compute_options compute_opts;
  compute_opts.llvm_nthreads = nthreads;
  compute_opts.print_level = 0;
  backendT backend(&compute_opts);

  morefit::dimension<evalT> m("m", "#it{m} [GeV/#it{c}^{2}]", 5.0, 7.0, false);
  morefit::parameter<evalT> mb("mb", "m(B^{+})", 5.28, 5.0, 6.0, 0.01, false);
  morefit::parameter<evalT> sigma("sigma", "\\sigma(B^{+})", 0.06, 0.005, 0.130, 0.001, false);
  morefit::parameter<evalT> fsig("fsig", "f_{\\mathrm{sig}}", 0.3, 0.0, 1.0, 0.01, false);
  morefit::parameter<evalT> alpha("alpha", "\\alpha_{\\mathrm{bkg}}", -1.0, -10.0, 10.0, 0.01, false);
  morefit::GaussianPDF<kernelT, evalT> gaus(&m, &mb, &sigma);
  morefit::ExponentialPDF<kernelT, evalT> exp(&m, &alpha);
  morefit::SumPDF<kernelT, evalT> sum(&gaus, &exp, &fsig);
  std::vector<morefit::parameter<evalT>*> params({&mb, &sigma, &fsig, &alpha});

  morefit::Xoshiro128pp rnd;
  rnd.setSeed(229387429ULL);
  morefit::generator_options gen_opts;
  gen_opts.rndtype = on_host ? morefit::generator_options::randomization_type::on_host
                            : morefit::generator_options::randomization_type::on_accelerator;
  gen_opts.print_level = 0;
  morefit::fitter_options fit_opts;
  fit_opts.minuit_printlevel = 0;
  fit_opts.minimizer = morefit::fitter_options::minimizer_type::Minuit2;
  fit_opts.optimize_dimensions = false;
  fit_opts.optimize_parameters = true;
  fit_opts.analytic_gradient = analytic;
  fit_opts.analytic_hessian = analytic;
  fit_opts.kahan_on_accelerator = true;
  fit_opts.print_level = 0;
  fit_opts.analytic_fisher = false;

  morefit::toystudy<kernelT, evalT, backendT, blockT> toy(&fit_opts, &gen_opts, &backend, &rnd);
  toy.toy(&sum, params, {&m}, nruns, nevents, false);
  std::printf("{");
#ifdef MF_HIP_BACKEND
  dump("sequential", toy, !batched);
  if (batched) {
    morefit::Xoshiro128pp rnd2;
    rnd2.setSeed(229387429ULL);
    morefit::toystudy<kernelT, evalT, backendT, blockT> toy2(&fit_opts, &gen_opts, &backend, &rnd2);
    toy2.toy_batched(&sum, params, {&m}, nruns, nevents, false);
    dump("batched", toy2, true);
  }
#else
  dump("sequential", toy, true);
#endif
  std::printf("}\n");
  return 0;
}

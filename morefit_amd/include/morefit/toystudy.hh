/**
 * @file toystudy.hh
 * Toy studies behind the reference's interface (toystudy.hh:41-111):
 * toystudy<kernelT, evalT, backendT, computeT>::toy(pdf, params, dimensions, nruns, nevents)
 * generates `nruns` datasets at the parameters' start values and fits each one, one after
 * another, with the generator and the fitter of the given backend; the kernels of both are
 * built for the first toy and reused (toystudy.hh:74-77).  Every toy starts from the start
 * values; extended densities draw a Poisson number of events (toystudy.hh:67-72).
 *
 * Results of the last study are kept per toy (fitted values and errors of all parameters in
 * `params` order, fit status, pulls (value - start) / error of the floating parameters); the
 * reference only histograms the pulls with ROOT.  With keep_datasets the generated datasets are
 * kept as well.
 *
 * toy_batched (HIP backend only, no reference counterpart) runs the same study -- the same
 * generator calls, hence the same datasets -- and fits all toys together in the batched engine
 * (one shared launch per MIGRAD step for every toy, batch.hh), with the analytic gradient and
 * Hessian kernels whatever the fitter options say.
 */
#ifndef MOREFIT_TOYSTUDY_HH
#define MOREFIT_TOYSTUDY_HH

#include <chrono>
#include <cmath>
#include <iostream>
#include <type_traits>
#include <vector>

#include "batch.hh"
#include "compute.hh"
#include "compute_hip.hh"
#include "eventvector.hh"
#include "fitter.hh"
#include "generator.hh"
#include "parametervector.hh"
#include "pdf.hh"
#include "utils.hh"

namespace morefit {

template <typename kernelT, typename evalT, typename backendT, typename computeT, typename seedT = uint32_t>
class toystudy {
  fitter_options* fitter_opts_;
  generator_options* generator_opts_;
  backendT* backend_;
  RandomGenerator* rnd_;

 public:
  std::vector<std::vector<double>> values, errors, pulls;  // per toy; pulls of floating parameters only
  std::vector<int> status;
  std::vector<size_t> toy_events;
  bool keep_datasets = false;
  std::vector<EventVector<kernelT, evalT>> datasets;
  double generate_ms = 0.0, fit_ms = 0.0;

  toystudy(fitter_options* fitter_opts, generator_options* generator_opts, backendT* backend, RandomGenerator* rnd)
      : fitter_opts_(fitter_opts), generator_opts_(generator_opts), backend_(backend), rnd_(rnd) {}

  bool toy(PDF<kernelT, evalT>* pdf, std::vector<parameter<evalT>*> params, std::vector<dimension<evalT>*> dimensions, unsigned int nruns,
           unsigned int nevents, bool draw_pulls = false) {
    generator_opts_->print();
    generator<kernelT, evalT, backendT, computeT, seedT> gen(generator_opts_, backend_, rnd_);
    if (fitter_opts_->print_level > 0) fitter_opts_->print();
    fitter<kernelT, evalT, backendT, computeT> fit(fitter_opts_, backend_);
    clear();
    const int pad = backend_->required_padding();
    for (unsigned int i = 0; i < nruns; i++) {
      if (fitter_opts_->print_level > 0) std::cout << "toy no " << i << std::endl;
      reset(params);
      const unsigned int n = pdf->is_extended() ? rnd_->poisson(nevents) : nevents;
      EventVector<kernelT, evalT> result(dimensions, n, pad > 0, pad);
      auto t0 = std::chrono::steady_clock::now();
      gen.generate(n, pdf, params, result, i > 0);
      auto t1 = std::chrono::steady_clock::now();
      fit.fit(pdf, params, &result, i > 0);
      auto t2 = std::chrono::steady_clock::now();
      generate_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
      fit_ms += std::chrono::duration<double, std::milli>(t2 - t1).count();
      record(params, fit.status_, n);
      if (keep_datasets) datasets.push_back(std::move(result));
    }
    if (draw_pulls) print_pulls(params);
    return true;
  }

  // the same toys fitted together by the batched engine (HIP backend)
  template <typename B = backendT>
  typename std::enable_if<std::is_same<B, HIPBackend>::value, bool>::type toy_batched(PDF<kernelT, evalT>* pdf,
                                                                                       std::vector<parameter<evalT>*> params,
                                                                                       std::vector<dimension<evalT>*> dimensions,
                                                                                       unsigned int nruns, unsigned int nevents,
                                                                                       bool draw_pulls = false) {
    generator<kernelT, evalT, backendT, computeT, seedT> gen(generator_opts_, backend_, rnd_);
    clear();
    std::vector<EventVector<kernelT, evalT>> data;
    std::vector<long long> sizes;
    auto t0 = std::chrono::steady_clock::now();
    for (unsigned int i = 0; i < nruns; i++) {
      reset(params);
      const unsigned int n = pdf->is_extended() ? rnd_->poisson(nevents) : nevents;
      data.emplace_back(dimensions, n);
      gen.generate(n, pdf, params, data.back(), i > 0);
      sizes.push_back((long long)n);
    }
    auto t1 = std::chrono::steady_clock::now();
    reset(params);
    HipModel<kernelT, evalT> model(backend_, pdf, params);
    HipBatch<kernelT, evalT> batch(&model, sizes);
    const size_t nd = dimensions.size();
    for (unsigned int i = 0; i < nruns; i++) {
      const size_t n = data[i].nevents(), np = data[i].nevents_padded();
      std::vector<double> cols(nd * n);
      for (size_t d = 0; d < nd; d++)
        for (size_t k = 0; k < n; k++) cols[d * n + k] = (double)data[i].get_data()[d * np + k];
      batch.upload((int)i, cols.data(), (long long)n);
    }
    std::vector<int> segs;
    std::vector<double> start;
    for (unsigned int i = 0; i < nruns; i++) {
      segs.push_back((int)i);
      for (int k : model.theta_index) start.push_back(params[k]->get_start_value());
    }
    BatchFitter<kernelT, evalT> bf(batch);
    auto res = bf.fit(segs, start);
    auto t2 = std::chrono::steady_clock::now();
    generate_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    fit_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    for (unsigned int i = 0; i < nruns; i++) {
      std::vector<double> v, e, pl;
      size_t f = 0;
      for (size_t j = 0; j < params.size(); j++) {
        if (params[j]->is_constant()) {
          v.push_back(params[j]->get_value());
          e.push_back(0.0);
          continue;
        }
        v.push_back(res[i].values.at(f));
        e.push_back(res[i].errors.at(f));
        pl.push_back((res[i].values.at(f) - params[j]->get_start_value()) / res[i].errors.at(f));
        f++;
      }
      values.push_back(v);
      errors.push_back(e);
      pulls.push_back(pl);
      status.push_back(res[i].status);
      toy_events.push_back((size_t)sizes[i]);
      if (keep_datasets) datasets.push_back(std::move(data[i]));
    }
    if (draw_pulls) print_pulls(params);
    return true;
  }

  // mean and width of each floating parameter's pulls (the reference draws them with ROOT)
  void print_pulls(const std::vector<parameter<evalT>*>& params) const {
    size_t f = 0;
    for (auto* p : params) {
      if (p->is_constant()) continue;
      std::vector<double> v;
      for (auto& row : pulls) v.push_back(row.at(f));
      double mean = 0.0, rms = 0.0;
      mean_rms(v, mean, rms);
      std::cout << "pull(" << p->get_name() << "): mean " << mean << " width " << rms << " over " << v.size() << " toys" << std::endl;
      f++;
    }
  }

 private:
  void clear() {
    values.clear();
    errors.clear();
    pulls.clear();
    status.clear();
    toy_events.clear();
    datasets.clear();
    generate_ms = fit_ms = 0.0;
  }
  static void reset(std::vector<parameter<evalT>*>& params) {
    for (auto* p : params) p->set_value(p->get_start_value());
  }
  void record(const std::vector<parameter<evalT>*>& params, int st, size_t n) {
    std::vector<double> v, e, pl;
    for (auto* p : params) {
      v.push_back(p->get_value());
      e.push_back(p->is_constant() ? 0.0 : p->get_error());
      if (!p->is_constant()) pl.push_back((p->get_value() - p->get_start_value()) / p->get_error());
    }
    values.push_back(v);
    errors.push_back(e);
    pulls.push_back(pl);
    status.push_back(st);
    toy_events.push_back(n);
  }
};

}  // namespace morefit
#endif

/**
 * @file fitter.hh
 * Unbinned maximum-likelihood fitter over any ComputeBackend/ComputeBlock pair, with the
 * reference's interface and semantics (fitter.hh:38-1476):
 *
 *  make_kernels  log p normalised, fixed parameters substituted, simplified, optionally
 *                weighted (w * log p); first and second derivatives; parameter-only terms
 *                hoisted into the parameter buffer [floating theta..., hoisted...]; one
 *                value, one value+gradient and one Hessian kernel (fitter.hh:321-658)
 *  likelihood    -2 sum log p (+ extended term)                         (fitter.hh:1152-1220)
 *  likelihood_and_gradient  -2 sums, zeros for fixed parameters          (fitter.hh:1054-1151)
 *  hessian       -2 sum d2 log p into h[i + j*n] (+2N/nu^2 for yields)  (fitter.hh:1221-1326)
 *  fit           MIGRAD (strategy 2, max 4000 calls, tolerance 0.01, errordef 1) then
 *                HESSE; values and errors written back into the parameters (fitter.hh:660-1002).
 *                analytic_gradient / analytic_hessian = false build no gradient / Hessian
 *                kernel and MIGRAD and HESSE differentiate numerically (fitter.hh:796-814);
 *                kahan_on_accelerator = false stores per-event outputs and sums them on the
 *                host with Kahan's algorithm (fitter.hh:1110-1124)
 *
 * The minimiser is this package's Minuit-algorithm MIGRAD (minimizer.hh); the vendored
 * Minuit2 of the reference is not part of this build.  The Hessian kernel here writes the
 * Hessian of the floating parameters into their positions in the full parameter matrix
 * (the reference indexes it by floating index, which only agrees when nothing is fixed).
 */
#ifndef MOREFIT_FITTER_HH
#define MOREFIT_FITTER_HH

#include <chrono>
#include <cmath>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "compute.hh"
#include "eventvector.hh"
#include "minimizer.hh"
#include "parametervector.hh"
#include "pdf.hh"
#include "utils.hh"

namespace morefit {

struct fitter_options {
  enum minimizer_type { TMinuit, Minuit2 };
  minimizer_type minimizer = Minuit2;
  bool analytic_gradient = true;
  bool analytic_hessian = true;
  bool analytic_fisher = false;
  bool optimize_parameters = true;
  bool optimize_dimensions = false;
  bool kahan_on_accelerator = true;
  int print_level = 0;
  bool correct_weighted_fit = true;
  double minuit_strategy = 2;
  int minuit_printlevel = 0;
  double minuit_maxiterations = 4000;
  double minuit_maxcalls = 4000;
  double minuit_tolerance = 0.1;     // TMinuit only in the reference; MIGRAD here uses 0.01 like Minuit2's default
  double minuit2_tolerance = 0.01;
  double minuit_errordef = 1.0;
  float buffering_cost_threshold = 0.0f;  // hoist every parameter-only operation
  // optimize_dimensions: observable-only subexpressions costing more than this become extra
  // data columns (a column is read at every call, so cheap terms stay in the kernel)
  float event_cost_threshold = 10.0f;
  bool postrun_hesse = true;
  void print() const {
    const int w = 40;
    std::cout << "FITTER OPTIONS\n"
              << std::setw(w) << std::left << "  analytic gradient " << (analytic_gradient ? "YES" : "NO") << "\n"
              << std::setw(w) << std::left << "  analytic hessian " << (analytic_hessian ? "YES" : "NO") << "\n"
              << std::setw(w) << std::left << "  strategy " << minuit_strategy << "\n"
              << std::setw(w) << std::left << "  postrun hesse " << (postrun_hesse ? "YES" : "NO") << std::endl;
  }
};

// blocks that store observables as offsets (FP32 backends) expose SetCentres
template <typename B, typename = void> struct has_set_centres : std::false_type {};
template <typename B>
struct has_set_centres<B, std::void_t<decltype(std::declval<B&>().SetCentres(std::vector<double>()))>> : std::true_type {};

template <typename kernelT, typename evalT, typename backendT, typename computeT>
class fitter {
 public:
  typedef GraphPtr<kernelT, evalT> Ptr;
  fitter_options* opts_;
  backendT* backend_;
  computeT block_, grad_block_, hessian_block_, fisher_block_, precompute_block_;
  // optimize_dimensions: columns of the precomputed data buffer (observables, then event terms)
  std::vector<std::string> event_buffer_names_;
  std::vector<Ptr> event_buffer_expressions_;
  std::vector<parameter<evalT>*> params_;
  PDF<kernelT, evalT>* pdf_ = nullptr;
  std::vector<Ptr> buffer_expressions_, grad_buffer_expressions_, hessian_buffer_expressions_, fisher_buffer_expressions_;
  std::vector<std::string> buffer_names_, grad_buffer_names_, hessian_buffer_names_, fisher_buffer_names_;
  std::vector<std::string> floating_names_;
  std::vector<double> centres_;
  size_t nevents_ = 0;
  bool weighted_ = false;
  bool kernels_ready_ = false;
  // kahan_on_accelerator = false: per-event outputs of the value, gradient, Hessian and Fisher
  // kernels copied back here and summed on the host
  int nout_[4] = {0, 0, 0, 0};
  std::vector<dimension<evalT>> res_dims_;
  std::unique_ptr<EventVector<kernelT, evalT>> res_buffer_[4];
  // results of the last fit
  int status_ = -1;
  double fval_ = NAN, edm_ = NAN;
  int ncalls_ = 0, ngrad_ = 0, nhess_ = 0;
  std::vector<double> cov_;
  double fit_ms_ = 0.0, kernel_ms_ = 0.0;
  // failure detection of the last fit: evaluations whose sums came back non-finite; at the
  // first of them, the events with non-finite per-event terms and the first such index (from
  // the block's diagnostic pass, -1: not known) and the parameter values; events of the bound
  // data with an observable outside its range or not a number
  int nonfinite_evals_ = 0;
  long long nonfinite_events_ = -1, first_nonfinite_event_ = -1, out_of_range_events_ = 0;
  std::vector<double> nonfinite_params_;

  fitter(fitter_options* options, backendT* backend)
      : opts_(options),
        backend_(backend),
        block_(backend),
        grad_block_(backend),
        hessian_block_(backend),
        fisher_block_(backend),
        precompute_block_(backend) {}

  bool centring() const {
    return std::is_same<kernelT, float>::value && has_set_centres<computeT>::value && backend_->opts_->hip_centre_observables;
  }

  // kernel graphs and compilation (fitter.hh:321-658)
  bool make_kernels(PDF<kernelT, evalT>* pdf, std::vector<parameter<evalT>*> params, const std::vector<dimension<evalT>*>& dims,
                    const std::string& weight_name) {
    std::vector<std::string> fixed;
    std::vector<double> fixed_v;
    floating_names_.clear();
    for (auto* p : params) {
      if (p->is_constant()) {
        fixed.push_back(p->get_name());
        fixed_v.push_back(p->get_value());
      } else {
        floating_names_.push_back(p->get_name());
      }
    }
    std::vector<std::string> obs;
    centres_.clear();
    for (auto* d : dims) {
      obs.push_back(d->get_name());
      centres_.push_back(centring() && d->get_name() != weight_name ? d->get_centre() : 0.0);
    }
    Ptr lp = pdf->logprob_normalised()->substitute(fixed, fixed_v)->simplify();
    if (!weight_name.empty()) lp = (Variable<kernelT, evalT>(weight_name) * std::move(lp))->simplify();
    if (centring()) lp = centre_observables(*lp, obs, centres_);
    std::vector<dimension<evalT>> in_sig;
    for (auto* d : dims) in_sig.push_back(*d);
    // event-only terms as extra data columns, filled once per dataset by a store kernel
    // (fitter.hh:339-375): the column list starts with the observables themselves
    event_buffer_names_.clear();
    event_buffer_expressions_.clear();
    if (opts_->optimize_dimensions) {
      for (auto& o : obs) {
        event_buffer_names_.push_back(o);
        event_buffer_expressions_.emplace_back(Variable<kernelT, evalT>(o));
      }
      lp = lp->optimize_buffering_constant_terms(event_buffer_names_, event_buffer_expressions_, obs, "morefit_eventbuffer_",
                                                 opts_->event_cost_threshold, false)
               ->simplify();
      std::vector<dimension<evalT>> out_sig(in_sig);
      for (size_t i = obs.size(); i < event_buffer_names_.size(); i++) out_sig.push_back(dimension<evalT>(event_buffer_names_[i], 0.0, 1.0));
      precompute_block_.MakeComputeKernel("precompute_kernel", nevents_, in_sig, out_sig, {}, event_buffer_expressions_, false);
      in_sig = out_sig;
    }
    std::vector<Ptr> d1;
    for (auto& t : floating_names_) d1.emplace_back(factor_common(*lp->diff(t)->simplify()));
    auto hoisted = [&](std::vector<Ptr>& graphs, std::vector<std::string>& names, std::vector<Ptr>& exprs, const std::string& prefix) {
      names.clear();
      exprs.clear();
      std::vector<Ptr> out;
      for (auto& g : graphs)
        out.emplace_back(opts_->optimize_parameters
                             ? g->optimize_buffering_constant_terms(names, exprs, floating_names_, prefix, opts_->buffering_cost_threshold)->simplify()
                             : g->copy());
      return out;
    };
    auto param_sig = [&](const std::vector<std::string>& names) {
      std::vector<std::string> r(floating_names_);
      r.insert(r.end(), names.begin(), names.end());
      return r;
    };
    dimension<evalT> res("res", 0.0, 1.0);
    // per-event outputs summed on the device, or stored and summed on the host
    const bool kahan = opts_->kahan_on_accelerator;
    nout_[0] = 1;
    nout_[1] = opts_->analytic_gradient ? 1 + (int)d1.size() : 0;
    nout_[2] = opts_->analytic_hessian ? (int)(d1.size() * (d1.size() + 1) / 2) : 0;
    nout_[3] = opts_->analytic_fisher ? (int)(d1.size() * (d1.size() + 1) / 2) : 0;
    {
      std::vector<Ptr> g;
      g.emplace_back(lp->copy());
      auto k = hoisted(g, buffer_names_, buffer_expressions_, "morefit_parambuffer_");
      block_.MakeComputeKernel("lh_kernel", nevents_, in_sig, {res}, param_sig(buffer_names_), k, kahan);
    }
    if (opts_->analytic_gradient) {
      std::vector<Ptr> g;
      g.emplace_back(lp->copy());
      for (auto& d : d1) g.emplace_back(d->copy());
      auto k = hoisted(g, grad_buffer_names_, grad_buffer_expressions_, "morefit_parambuffer_");
      std::vector<dimension<evalT>> out(k.size(), res);
      grad_block_.MakeComputeKernel("lh_grad_kernel", nevents_, in_sig, out, param_sig(grad_buffer_names_), k, kahan);
    }
    if (opts_->analytic_hessian) {
      std::vector<Ptr> g;
      for (size_t i = 0; i < d1.size(); i++)
        for (size_t j = 0; j <= i; j++) g.emplace_back(factor_common(*d1[i]->diff(floating_names_[j])->simplify()));
      auto k = hoisted(g, hessian_buffer_names_, hessian_buffer_expressions_, "morefit_parambuffer_");
      std::vector<dimension<evalT>> out(k.size(), res);
      hessian_block_.MakeComputeKernel("lh_hessian_kernel", nevents_, in_sig, out, param_sig(hessian_buffer_names_), k, kahan);
    }
    // Fisher information (fitter.hh:471-497): per event (w d_i log p)(w d_j log p), i >= j, from
    // the derivatives of the weighted log-density
    if (opts_->analytic_fisher) {
      std::vector<Ptr> g;
      for (size_t i = 0; i < d1.size(); i++)
        for (size_t j = 0; j <= i; j++) g.emplace_back((d1[i]->copy() * d1[j]->copy())->simplify());
      auto k = hoisted(g, fisher_buffer_names_, fisher_buffer_expressions_, "morefit_parambuffer_");
      std::vector<dimension<evalT>> out(k.size(), res);
      fisher_block_.MakeComputeKernel("lh_fisher_kernel", nevents_, in_sig, out, param_sig(fisher_buffer_names_), k, kahan);
    }
    return true;
  }

  template <typename B>
  static void apply_centres(B& b, const std::vector<double>& c) {
    if constexpr (has_set_centres<B>::value) b.SetCentres(c);
  }

  // upload a dataset (FP64 host data may feed an FP32 kernel) and bind the blocks to it
  template <typename dataT>
  void bind_data(EventVector<dataT, evalT>* data) {
    // the blocks take int event counts (reference compute.hh:120)
    if (data->nevents_padded() > (size_t)std::numeric_limits<int>::max())
      throw std::invalid_argument("fitter: " + std::to_string(data->nevents_padded()) + " events do not fit the blocks' int event counts");
    nevents_ = data->nevents();
    // events outside the observables' ranges (or not a number) are fitted as they are, as in
    // the reference; their number is reported with the fit
    out_of_range_events_ = 0;
    std::vector<std::pair<unsigned, dimension<evalT>*>> cols;
    for (auto* d : pdf_->dimensions()) {
      const int j = data->dim_index(d->get_name());
      if (j >= 0) cols.emplace_back((unsigned)j, d);
    }
    for (size_t i = 0; i < data->nevents(); i++) {
      bool in = true;
      for (auto& c : cols) {
        const double x = (double)(*data)(i, c.first);
        in = in && x >= c.second->get_min() && x <= c.second->get_max();
      }
      if (!in) out_of_range_events_++;
    }
    apply_centres(block_, centres_);
    if (opts_->optimize_dimensions) {
      // observables in, observables and event columns out; the kernels read that buffer
      apply_centres(precompute_block_, centres_);
      precompute_block_.SetupInputBuffer(data->nevents_padded() * data->ndimensions() * sizeof(kernelT));
      precompute_block_.SetNevents((int)data->nevents(), (int)data->nevents_padded());
      precompute_block_.SetupOutputBuffer(data->nevents_padded() * event_buffer_names_.size() * sizeof(kernelT));
      precompute_block_.CopyToInputBuffer(*data);
      precompute_block_.SubmitKernel();
      precompute_block_.Finish();
      block_.SetupInputBuffer(&precompute_block_, false);
    } else {
      block_.SetupInputBuffer(data->nevents_padded() * data->ndimensions() * sizeof(kernelT));
      block_.CopyToInputBuffer(*data);
    }
    block_.SetNevents((int)data->nevents(), (int)data->nevents_padded());
    block_.SetupParameterBuffer((floating_names_.size() + buffer_names_.size() + 1) * sizeof(kernelT));
    if (opts_->analytic_gradient) {
      apply_centres(grad_block_, centres_);
      grad_block_.SetupInputBuffer(&block_, true);
      grad_block_.SetNevents((int)data->nevents(), (int)data->nevents_padded());
      grad_block_.SetupParameterBuffer((floating_names_.size() + grad_buffer_names_.size() + 1) * sizeof(kernelT));
    }
    if (opts_->analytic_hessian) {
      apply_centres(hessian_block_, centres_);
      hessian_block_.SetupInputBuffer(&block_, true);
      hessian_block_.SetNevents((int)data->nevents(), (int)data->nevents_padded());
      hessian_block_.SetupParameterBuffer((floating_names_.size() + hessian_buffer_names_.size() + 1) * sizeof(kernelT));
    }
    if (opts_->analytic_fisher) {
      apply_centres(fisher_block_, centres_);
      fisher_block_.SetupInputBuffer(&block_, true);
      fisher_block_.SetNevents((int)data->nevents(), (int)data->nevents_padded());
      fisher_block_.SetupParameterBuffer((floating_names_.size() + fisher_buffer_names_.size() + 1) * sizeof(kernelT));
    }
    if (!opts_->kahan_on_accelerator) {
      // host result buffers of nout columns, the blocks' output buffers of the same layout
      // (fitter.hh:583-603)
      int most = 0;
      for (int k = 0; k < 4; k++) most = std::max(most, nout_[k]);
      res_dims_.assign(most, dimension<evalT>("res", 0.0, 1.0));
      computeT* blocks[4] = {&block_, &grad_block_, &hessian_block_, &fisher_block_};
      for (int k = 0; k < 4; k++) {
        res_buffer_[k].reset();
        if (nout_[k] == 0) continue;
        std::vector<dimension<evalT>*> dims;
        for (int j = 0; j < nout_[k]; j++) dims.push_back(&res_dims_[j]);
        res_buffer_[k].reset(new EventVector<kernelT, evalT>(dims, data->nevents(), data->is_padded(), data->padding()));
        blocks[k]->SetupOutputBuffer(res_buffer_[k]->buffer_size());
      }
    }
  }

  template <typename dataT>
  bool prepare(PDF<kernelT, evalT>* pdf, std::vector<parameter<evalT>*> params, EventVector<dataT, evalT>* data, bool refit) {
    pdf_ = pdf;
    params_ = params;
    nevents_ = data->nevents();
    weighted_ = data->has_event_weight();
    if (!refit || !kernels_ready_) make_kernels(pdf, params, data->get_dimensions(), data->event_weight_name());
    kernels_ready_ = true;
    bind_data(data);
    return true;
  }

  // parameter buffer: floating values then the hoisted expressions evaluated in FP64 on the host
  std::vector<kernelT> parameter_buffer(const std::vector<Ptr>& exprs) const {
    std::vector<std::string> names;
    std::vector<double> values;
    std::vector<kernelT> buf;
    for (auto* p : params_)
      if (!p->is_constant()) {
        names.push_back(p->get_name());
        values.push_back(p->get_value());
        buf.push_back((kernelT)p->get_value());
      }
    for (auto& e : exprs) buf.push_back((kernelT)e->eval(names, values));
    return buf;
  }
  double total_yield() const {
    double t = 0.0;
    for (auto* p : pdf_->parameters_) t += p->get_value();
    return t;
  }
  bool is_yield(const parameter<evalT>* p) const {
    for (auto* q : pdf_->parameters_)
      if (q->get_name() == p->get_name()) return true;
    return false;
  }
  // one launch of block `which` (0 value, 1 gradient, 2 Hessian, 3 Fisher): its per-output
  // sums, from the device's compensated reduction or from a host Kahan sum of the stored
  // per-event outputs (kahan_on_accelerator = false, fitter.hh:1110-1124)
  std::vector<evalT> run(int which, const std::vector<Ptr>& exprs) {
    static const char* what[4] = {"likelihood", "gradient (analytic_gradient is off)", "Hessian (analytic_hessian is off)",
                                  "Fisher matrix (analytic_fisher is off)"};
    if (!kernels_ready_ || nout_[which] == 0) throw std::runtime_error(std::string("fitter: no ") + what[which] + " kernel");
    computeT* blocks[4] = {&block_, &grad_block_, &hessian_block_, &fisher_block_};
    computeT& b = *blocks[which];
    auto t0 = std::chrono::steady_clock::now();
    b.CopyToParameterBuffer(parameter_buffer(exprs));
    b.SubmitKernel();
    b.Finish();
    std::vector<evalT> sums;
    if (opts_->kahan_on_accelerator) {
      b.CopyFromKahanBuffer(sums);
    } else {
      EventVector<kernelT, evalT>& r = *res_buffer_[which];
      b.CopyFromOutputBuffer(r);
      for (int k = 0; k < nout_[which]; k++) sums.push_back(kahan_summation<evalT, kernelT>(r, k));
    }
    kernel_ms_ += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    bool finite = true;
    for (auto v : sums) finite = finite && std::isfinite((double)v);
    if (!finite) note_nonfinite(which, b);
    return sums;
  }

  // A non-finite sum (the reference checks the host sum of the likelihood for NaN and stops
  // the program, fitter.hh:1197-1204): counted, and at the first one in a fit the offending
  // events are counted by the block (on the device for the HIP block) or, for per-event
  // outputs on the host, here; MIGRAD then treats the point as invalid.
  void note_nonfinite(int which, computeT& b) {
    if (nonfinite_evals_++ > 0) return;
    nonfinite_params_.clear();
    for (auto* p : params_) nonfinite_params_.push_back(p->get_value());
    long long count = -1, first = -1;
    if (opts_->kahan_on_accelerator) {
      if (!b.CountNonFinite(&count, &first)) count = first = -1;
    } else {
      const EventVector<kernelT, evalT>& r = *res_buffer_[which];
      count = 0;
      for (size_t i = 0; i < r.nevents(); i++) {
        bool ok = true;
        for (int k = 0; k < nout_[which]; k++) ok = ok && std::isfinite((double)r(i, (unsigned)k));
        if (!ok && count++ == 0) first = (long long)i;
      }
    }
    nonfinite_events_ = count;
    first_nonfinite_event_ = first;
    if (opts_->print_level > 0) {
      std::cout << "sum of the " << (which == 0 ? "likelihood" : which == 1 ? "gradient" : which == 2 ? "Hessian" : "Fisher")
                << " terms is not finite: " << count << " events with non-finite terms, the first at " << first << std::endl;
      for (auto* p : params_) std::cout << "parameter values: " << p->get_name() << " " << p->get_value() << std::endl;
    }
  }

  double likelihood() { return likelihood(pdf_, params_, (EventVector<kernelT, evalT>*)nullptr); }
  template <typename dataT = kernelT>
  evalT likelihood(PDF<kernelT, evalT>* pdf, std::vector<parameter<evalT>*> params, EventVector<dataT, evalT>* data) {
    auto s = run(0, buffer_expressions_);
    double r = s.at(0);
    if (pdf_->is_extended()) {
      double nu = total_yield();
      r += nevents_ * std::log(nu) - nu;
    }
    return -2.0 * r;
  }
  double likelihood_and_gradient(int npar, double* grad) { return likelihood_and_gradient(pdf_, params_, (EventVector<kernelT, evalT>*)nullptr, npar, grad); }
  template <typename dataT = kernelT>
  evalT likelihood_and_gradient(PDF<kernelT, evalT>* pdf, std::vector<parameter<evalT>*> params, EventVector<dataT, evalT>* data, int npar,
                                double* grad) {
    auto s = run(1, grad_buffer_expressions_);
    double r = -2.0 * s.at(0);
    int k = 0;
    for (size_t i = 0; i < params_.size(); i++) grad[i] = params_[i]->is_constant() ? 0.0 : -2.0 * s.at(1 + k++);
    if (pdf_->is_extended()) {
      double nu = total_yield();
      r -= 2.0 * (nevents_ * std::log(nu) - nu);
      for (size_t i = 0; i < params_.size(); i++)
        if (!params_[i]->is_constant() && is_yield(params_[i])) grad[i] -= 2.0 * (nevents_ / nu - 1.0);
    }
    return r;
  }
  bool hessian(int npar, double* h) { return hessian(pdf_, params_, (EventVector<kernelT, evalT>*)nullptr, npar, h); }
  template <typename dataT = kernelT>
  bool hessian(PDF<kernelT, evalT>* pdf, std::vector<parameter<evalT>*> params, EventVector<dataT, evalT>* data, int npar, double* h) {
    auto s = run(2, hessian_buffer_expressions_);
    const size_t N = params_.size();
    for (size_t i = 0; i < N * N; i++) h[i] = 0.0;
    std::vector<int> fl;
    for (size_t i = 0; i < N; i++)
      if (!params_[i]->is_constant()) fl.push_back((int)i);
    size_t idx = 0;
    for (size_t i = 0; i < fl.size(); i++)
      for (size_t j = 0; j <= i; j++) {
        double v = -2.0 * s.at(idx++);
        h[fl[i] + N * fl[j]] = v;
        h[fl[j] + N * fl[i]] = v;
      }
    if (pdf_->is_extended()) {
      double nu = total_yield();
      for (size_t i = 0; i < fl.size(); i++)
        for (size_t j = 0; j < fl.size(); j++)
          if (is_yield(params_[fl[i]]) && is_yield(params_[fl[j]])) h[fl[i] + N * fl[j]] += 2.0 * nevents_ / (nu * nu);
    }
    return true;
  }

  // Fisher matrix of the (weighted) log-likelihood in -2 lnL units, n x n at full-matrix
  // positions (fitter.hh:1327-1432): -2 sum (w d_i log p)(w d_j log p), plus 2N/nu^2 for pairs of
  // yields of an extended density
  bool fisher(int npar, double* f) {
    if (!opts_->analytic_fisher) throw std::runtime_error("fitter::fisher: analytic_fisher is off");
    auto s = run(3, fisher_buffer_expressions_);
    const size_t N = params_.size();
    for (size_t i = 0; i < N * N; i++) f[i] = 0.0;
    std::vector<int> fl;
    for (size_t i = 0; i < N; i++)
      if (!params_[i]->is_constant()) fl.push_back((int)i);
    size_t idx = 0;
    for (size_t i = 0; i < fl.size(); i++)
      for (size_t j = 0; j <= i; j++) {
        double v = -2.0 * s.at(idx++);
        f[fl[i] + N * fl[j]] = v;
        f[fl[j] + N * fl[i]] = v;
      }
    if (pdf_->is_extended()) {
      double nu = total_yield();
      for (size_t i = 0; i < fl.size(); i++)
        for (size_t j = 0; j < fl.size(); j++)
          if (is_yield(params_[fl[i]]) && is_yield(params_[fl[j]])) f[fl[i] + N * fl[j]] += 2.0 * nevents_ / (nu * nu);
    }
    return true;
  }

  // covariance of a weighted fit (fitter.hh:894-947): -2 H^-1 F H^-1 over the floating
  // parameters, H the Hessian and F the Fisher matrix of -2 lnL at the minimum
  std::vector<double> weighted_covariance() {
    const size_t N = params_.size();
    std::vector<double> h(N * N), f(N * N);
    hessian((int)N, h.data());
    fisher((int)N, f.data());
    std::vector<int> fl;
    for (size_t i = 0; i < N; i++)
      if (!params_[i]->is_constant()) fl.push_back((int)i);
    const int n = (int)fl.size();
    std::vector<double> hi(n * n), fm(n * n), t(n * n, 0.0), c(n * n, 0.0);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        hi[i * n + j] = h[fl[i] + N * fl[j]];
        fm[i * n + j] = f[fl[i] + N * fl[j]];
      }
    if (!la::spd_inverse(hi, n)) throw std::runtime_error("weighted fit correction: Hessian not positive-definite");
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) t[i * n + j] += hi[i * n + k] * fm[k * n + j];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        for (int k = 0; k < n; k++) c[i * n + j] += t[i * n + k] * hi[k * n + j];
        c[i * n + j] *= -2.0;
      }
    return c;
  }

  void update_parameters(const double* x) {
    for (size_t i = 0; i < params_.size(); i++) params_[i]->set_value(x[i]);
  }

  struct Fcn : MinimizerFcn {
    fitter* f;
    explicit Fcn(fitter* ff) : f(ff) {}
    double value(const double* x) override {
      f->update_parameters(x);
      return f->likelihood();
    }
    double value_grad(const double* x, double* g) override {
      f->update_parameters(x);
      return f->likelihood_and_gradient((int)f->params_.size(), g);
    }
    bool hessian(const double* x, double* h) override {
      if (!has_hessian()) return false;
      f->update_parameters(x);
      return f->hessian((int)f->params_.size(), h);
    }
    double value_noise() const override {
      return std::is_same<kernelT, float>::value ? fp32_sum_noise((double)f->nevents_) : 0.0;
    }
    // without the kernels MIGRAD and HESSE differentiate numerically (Minuit2 with an FCN that
    // has no gradient, or no Hessian function set, fitter.hh:796-814)
    bool has_gradient() const override { return f->opts_->analytic_gradient; }
    bool has_hessian() const override { return f->opts_->analytic_hessian; }
  };

  std::vector<MinimizerParam> minimizer_params(bool reset) {
    std::vector<MinimizerParam> mp;
    int idx = 0;
    for (auto* p : params_) {
      if (reset) p->set_value(p->get_start_value());
      MinimizerParam q;
      q.name = p->get_name();
      q.value = p->get_value();
      q.step = p->get_step_size();
      q.fixed = p->is_constant();
      if (!p->is_unlimited()) {
        q.lo = p->get_min();
        q.hi = p->get_max();
        q.has_lo = std::isfinite(q.lo);
        q.has_hi = std::isfinite(q.hi);
      }
      p->set_minuit_idx(q.fixed ? -1 : idx++);
      mp.push_back(q);
    }
    return mp;
  }

  // full fit (fitter.hh:660-1002): returns true when MIGRAD converged
  template <typename dataT>
  bool fit(PDF<kernelT, evalT>* pdf, std::vector<parameter<evalT>*> params, EventVector<dataT, evalT>* data, bool refit = false) {
    auto t0 = std::chrono::steady_clock::now();
    kernel_ms_ = 0.0;
    nonfinite_evals_ = 0;
    nonfinite_events_ = first_nonfinite_event_ = -1;
    nonfinite_params_.clear();
    prepare(pdf, params, data, refit);
    Migrad m(minimizer_params(true));
    m.strategy = (int)opts_->minuit_strategy;
    m.max_calls = (int)opts_->minuit_maxcalls;
    m.max_iterations = (int)opts_->minuit_maxiterations;
    m.tolerance = opts_->minuit2_tolerance;
    m.up = opts_->minuit_errordef;
    Fcn fcn(this);
    m.minimize(fcn);
    update_parameters(m.x.data());
    if (opts_->postrun_hesse) m.hesse(fcn);
    if (weighted_ && opts_->correct_weighted_fit) {
      if (!(opts_->analytic_fisher && opts_->analytic_hessian))
        throw std::runtime_error("correction for weighted fit only supported with analytic hessian and fisher matrix");
      update_parameters(m.x.data());
      m.cov = weighted_covariance();
      int k = 0;
      for (size_t i = 0; i < params_.size(); i++)
        if (!params_[i]->is_constant()) {
          m.errors[i] = std::sqrt(m.cov[(size_t)k * (m.nfree()) + k]);
          k++;
        }
    }
    for (size_t i = 0; i < params_.size(); i++)
      if (!params_[i]->is_constant()) {
        params_[i]->set_value(m.x[i]);
        params_[i]->set_error(m.errors[i]);
      }
    status_ = m.status;
    fval_ = m.fval;
    edm_ = m.edm;
    ncalls_ = m.ncalls;
    ngrad_ = m.ngrad;
    nhess_ = m.nhess;
    cov_ = m.cov;
    fit_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (opts_->print_level > 0)
      std::cout << "fit status " << status_ << " fval " << std::setprecision(12) << fval_ << " edm " << edm_ << " calls " << ncalls_
                << " time " << fit_ms_ << " ms" << std::endl;
    return status_ == 0;
  }
};

}  // namespace morefit
#endif

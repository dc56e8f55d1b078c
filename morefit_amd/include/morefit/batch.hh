/**
 * @file batch.hh
 * Many fits at once over one HipBatch: each fit runs the ordinary blocking MIGRAD/HESSE code
 * in its own fibre (ucontext); whenever a fit asks for a value, a gradient or a Hessian it
 * posts its parameters and yields.  When every live fibre has yielded, the scheduler issues
 * one shared launch per requested kind over all requesting segments (each kind's launch
 * sequence a replayed hipGraph), hands the totals back and resumes the fibres.  Converged
 * fits leave; the others keep going.  This inverts Minuit's synchronous FCN callback
 * (reference toystudy.hh:55-111 runs one fit after another, each with its own launches).
 */
#ifndef MOREFIT_BATCH_HH
#define MOREFIT_BATCH_HH

#include <ucontext.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <vector>

#include "compute_hip.hh"
#include "minimizer.hh"

namespace morefit {

struct BatchFitResult {
  std::vector<double> values, errors;  // floating parameters
  double fval = NAN, edm = NAN;
  int status = -1, ncalls = 0, iterations = 0;
  bool hesse_ok = false;
};

struct BatchFitStats {
  int ticks = 0;
  int launches[4] = {0, 0, 0, 0};
  double wall_ms = 0.0, gpu_wait_ms = 0.0, host_ms = 0.0;
};

template <typename kernelT, typename evalT = double>
class BatchFitter {
 public:
  typedef HipBatch<kernelT, evalT> Batch;
  Batch& b;
  bool use_graph = true;
  bool run_hesse = true;
  int strategy = 2;
  double tolerance = 0.01;
  size_t stack_bytes = 256 * 1024;
  BatchFitStats stats;

  explicit BatchFitter(Batch& batch) : b(batch) {}

  // fits every segment in `segs` from its start values; parameter limits from the model's
  // parameter objects.  Results in segment order of `segs`.
  std::vector<BatchFitResult> fit(const std::vector<int>& segs, const std::vector<double>& start /* per listed seg: ntheta */) {
    auto t0 = std::chrono::steady_clock::now();
    const int F = (int)segs.size(), P = b.m->ntheta;
    fibres_.clear();
    fibres_.resize(F);
    for (int i = 0; i < F; i++) {
      Fibre& f = fibres_[i];
      f.owner = this;
      f.seg = segs[i];
      f.start.assign(start.begin() + (size_t)i * P, start.begin() + (size_t)(i + 1) * P);
      f.stack.resize(stack_bytes);
      getcontext(&f.ctx);
      f.ctx.uc_stack.ss_sp = f.stack.data();
      f.ctx.uc_stack.ss_size = f.stack.size();
      f.ctx.uc_link = &main_;
      uintptr_t p = (uintptr_t)&f;
      makecontext(&f.ctx, (void (*)())&BatchFitter::trampoline, 2, (unsigned int)(p & 0xffffffffu), (unsigned int)(p >> 32));
    }
    int live = F;
    double wait_ms = 0.0;
    while (live > 0) {
      for (auto& f : fibres_) {
        if (f.done || f.pending) continue;
        current_ = &f;
        swapcontext(&main_, &f.ctx);
        current_ = nullptr;
        if (f.done) live--;
      }
      bool any = false;
      for (auto& f : fibres_) any |= f.pending != 0;
      if (!any) break;
      for (auto& f : fibres_)
        if (f.pending) b.request(f.seg, f.pending);
      auto tw = std::chrono::steady_clock::now();
      for (int k = 1; k <= 3; k++) {
        bool want = false;
        for (auto& f : fibres_) want |= f.pending == k;
        if (!want) continue;
        b.set_modes(k, false);
        b.run(k, false, use_graph);
        stats.launches[k]++;
      }
      std::fill(b.request_.begin(), b.request_.end(), 0);
      wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw).count();
      for (auto& f : fibres_) f.pending = 0;  // results are in the batch's pinned buffers
      stats.ticks++;
    }
    std::vector<BatchFitResult> out(F);
    for (int i = 0; i < F; i++) out[i] = fibres_[i].result;
    stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats.gpu_wait_ms = wait_ms;
    stats.host_ms = stats.wall_ms - wait_ms;
    fibres_.clear();
    return out;
  }

 private:
  struct Fibre;
  struct Fcn : MinimizerFcn {
    Fibre* f;
    explicit Fcn(Fibre* ff) : f(ff) {}
    void post(const double* x, int kind) {
      double* th = f->owner->b.theta(f->seg);
      for (size_t k = 0; k < f->owner->b.m->theta_index.size(); k++) th[k] = x[f->owner->b.m->theta_index[k]];
      f->pending = kind;
      swapcontext(&f->ctx, &f->owner->main_);
    }
    double value(const double* x) override {
      post(x, EVAL_LH);
      return -2.0 * f->owner->b.result(f->seg, EVAL_LH)[0];
    }
    double value_grad(const double* x, double* g) override {
      post(x, EVAL_GRAD);
      const double* r = f->owner->b.result(f->seg, EVAL_GRAD);
      const auto& ti = f->owner->b.m->theta_index;
      const size_t N = f->owner->b.m->params.size();
      for (size_t i = 0; i < N; i++) g[i] = 0.0;
      for (size_t k = 0; k < ti.size(); k++) g[ti[k]] = -2.0 * r[1 + k];
      return -2.0 * r[0];
    }
    bool hessian(const double* x, double* h) override {
      post(x, EVAL_HESS);
      const double* r = f->owner->b.result(f->seg, EVAL_HESS);
      const auto& ti = f->owner->b.m->theta_index;
      const size_t N = f->owner->b.m->params.size(), P = ti.size();
      for (size_t i = 0; i < N * N; i++) h[i] = 0.0;
      size_t idx = 1 + P;
      for (size_t i = 0; i < P; i++)
        for (size_t j = 0; j <= i; j++) {
          double v = -2.0 * r[idx++];
          h[ti[i] * N + ti[j]] = v;
          h[ti[j] * N + ti[i]] = v;
        }
      return true;
    }
  };
  struct Fibre {
    BatchFitter* owner = nullptr;
    int seg = 0;
    std::vector<double> start;
    std::vector<char> stack;
    ucontext_t ctx;
    int pending = 0;
    bool done = false;
    BatchFitResult result;
  };
  std::vector<Fibre> fibres_;
  ucontext_t main_;
  Fibre* current_ = nullptr;

  static void trampoline(unsigned int lo, unsigned int hi) {
    Fibre* f = (Fibre*)(((uintptr_t)hi << 32) | (uintptr_t)lo);
    f->owner->run_fit(*f);
    f->done = true;
    f->pending = 0;
  }

  void run_fit(Fibre& f) {
    const auto& ps = b.m->params;
    std::vector<MinimizerParam> mp;
    size_t k = 0;
    for (size_t i = 0; i < ps.size(); i++) {
      MinimizerParam q;
      q.name = ps[i]->get_name();
      q.fixed = ps[i]->is_constant();
      q.value = q.fixed ? ps[i]->get_value() : f.start[k++];
      q.step = ps[i]->get_step_size();
      if (!ps[i]->is_unlimited()) {
        q.lo = ps[i]->get_min();
        q.hi = ps[i]->get_max();
        q.has_lo = std::isfinite(q.lo);
        q.has_hi = std::isfinite(q.hi);
      }
      mp.push_back(q);
    }
    Migrad m(mp);
    m.strategy = strategy;
    m.tolerance = tolerance;
    Fcn fcn(&f);
    try {
      m.minimize(fcn);
      if (run_hesse) f.result.hesse_ok = m.hesse(fcn);
    } catch (...) {
      m.status = 4;
    }
    for (size_t i = 0; i < b.m->theta_index.size(); i++) {
      f.result.values.push_back(m.x[b.m->theta_index[i]]);
      f.result.errors.push_back(m.errors[b.m->theta_index[i]]);
    }
    f.result.fval = m.fval;
    f.result.edm = m.edm;
    f.result.status = m.status;
    f.result.ncalls = m.ncalls;
    f.result.iterations = m.iterations;
  }
};

}  // namespace morefit
#endif

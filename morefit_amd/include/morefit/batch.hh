/**
 * @file batch.hh
 * Many fits at once.  Each fit runs the ordinary blocking MIGRAD/HESSE code in its own fibre
 * (fiber.hh); when a fit asks for a value, a gradient or a Hessian it posts its parameters
 * into its segment's slot and yields.  Once every live fibre has yielded, the scheduler
 * issues one shared launch per (batch, kind) over all requesting segments, and resumes the
 * fibres with their totals.  Converged fits leave, new fits may join between ticks (the fit
 * service), the others keep going.  This inverts Minuit's synchronous FCN callback
 * (reference toystudy.hh:55-111 runs one fit after another, each with its own launches).
 */
#ifndef MOREFIT_BATCH_HH
#define MOREFIT_BATCH_HH

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <thread>
#include <vector>

#include "compute_hip.hh"
#include "fiber.hh"
#include "minimizer.hh"

namespace morefit {

struct BatchFitResult {
  std::vector<double> values, errors;  // floating parameters
  double fval = NAN, edm = NAN;
  int status = -1, ncalls = 0, iterations = 0;
  bool hesse_ok = false;
  bool thread_moved = false;  // the fit's fibre was resumed on another OS thread (never expected)
  // evaluations whose outputs came back non-finite; at the last such point, from the diagnostic
  // kernel: events with non-finite per-event terms, the first of them (-1: none), and events
  // with an observable outside its range or not a number
  int nonfinite_evals = 0;
  long long nonfinite_events = 0, first_nonfinite_event = -1, out_of_range_events = 0;
};

struct BatchFitStats {
  int ticks = 0;
  int launches[4] = {0, 0, 0, 0};
  double wall_ms = 0.0, gpu_wait_ms = 0.0, host_ms = 0.0;
  int host_threads = 1;
  double setup_ms = 0.0;  // creating the fits (fibres, stacks)
  double step_ms = 0.0;   // running the fibres between launches (all host threads)
};

template <typename kernelT, typename evalT = double>
class FitScheduler {
 public:
  typedef HipBatch<kernelT, evalT> Batch;
  struct Task {
    FitScheduler* owner = nullptr;
    Batch* b = nullptr;
    int seg = 0;
    int worker = 0;                 // host thread that runs this fit, for its whole life
    long long id = 0;
    std::vector<double> start;
    char* stack = nullptr;          // from FiberStacks, returned when the task leaves
    void* sp = nullptr;             // saved stack pointer of this fibre while it is suspended
    void** ret = nullptr;           // where the worker that resumed it saved its own
    int pending = 0;                // kind of the evaluation it waits for (0: none)
    unsigned long long posted = 0;  // launches issued before it posted that request
    bool done = false;
    BatchFitResult result;
    std::thread::id first_thread;   // OS thread the fibre first ran on
    std::vector<double> bad_theta;  // theta of the last evaluation that came back non-finite
    typename std::list<Task>::iterator self;  // its node in tasks_
  };

  bool use_graph = false;  // direct launches: on this stack cheaper than a graph replay
  bool run_hesse = true;
  int strategy = 2;
  double tolerance = 0.01;
  // MIGRAD trace level of every fit (>= 3: per-iteration lines on stderr); the environment
  // variable MOREFIT_PRINT_LEVEL sets it for fits driven through the C API
  int print_level = std::getenv("MOREFIT_PRINT_LEVEL") ? std::atoi(std::getenv("MOREFIT_PRINT_LEVEL")) : 0;
  // host threads that run the fits' MIGRAD code between launches (0: one per hardware
  // thread, at most 16); a fit stays on the thread it was assigned, so a fibre never moves
  // between OS threads.  Device work is issued by the calling thread only.
  int host_threads = 0;
  BatchFitStats stats;

  FitScheduler() {}
  FitScheduler(const FitScheduler&) = delete;
  FitScheduler& operator=(const FitScheduler&) = delete;
  // fits still suspended are abandoned (their fibres never resume); their stacks go back
  ~FitScheduler() {
    for (auto& t : tasks_) FiberStacks::get().release(t.stack);
  }

  // a new fit of segment `seg` of batch `b` from `start` (ntheta values)
  Task* add(Batch* b, int seg, const double* start, long long id = 0) {
    tasks_.emplace_back();
    Task& t = tasks_.back();
    t.self = std::prev(tasks_.end());
    t.owner = this;
    t.b = b;
    t.seg = seg;
    t.id = id;
    t.worker = (int)(added_++ % (unsigned long long)nthreads());
    if ((int)owned_.size() <= t.worker) owned_.resize(t.worker + 1);
    owned_[t.worker].push_back(&t);
    t.start.assign(start, start + b->m->ntheta);
    t.stack = FiberStacks::get().acquire();
    t.sp = fiber_prepare(t.stack, FiberStacks::kStackBytes, &FitScheduler::fiber_main, &t);
    batches_.insert(b);
    return &t;
  }
  size_t live() const { return tasks_.size(); }

  // Runs every task whose request has been served (or that has none) until it posts its next
  // request or finishes: each host thread over its own tasks, in parallel.  A thread places its
  // tasks' new requests in their batches and collects its finished tasks itself; those are
  // handed to `on_done` and removed here.  Returns the number still live.
  template <typename F>
  size_t step(F on_done) {
    const auto ts = std::chrono::steady_clock::now();
    const int W = (int)owned_.size();
    // cleared here, not by the workers: a worker with no tasks left is not run again
    finished_.resize(W);
    for (auto& f : finished_) f.clear();
    auto work = [this](int w) {
      void* here = nullptr;
      std::vector<Task*>& v = owned_[w];
      std::vector<Task*>& fin = finished_[w];
      size_t keep = 0;
      for (Task* t : v) {
        if (!t->done && (t->pending == 0 || t->posted < launches_)) {
          t->pending = 0;
          resume(*t, &here);
          if (t->pending) t->b->request(t->seg, t->pending);  // its own segment's entry
        }
        if (t->done) fin.push_back(t);
        else v[keep++] = t;
      }
      v.resize(keep);
    };
    if (nthreads() > 1) {
      // every fibre is resumed by the worker it was assigned at add(), however many are live:
      // code in a fibre may keep thread-local addresses (errno, the C++ exception globals)
      // across its switches, so it must never continue on another OS thread
      int used = 0;
      for (int w = 0; w < W; w++)
        if (!owned_[w].empty()) used = w + 1;
      if (used > 0) HostThreads::get().run(used, work);
    } else {
      for (int w = 0; w < W; w++) work(w);
    }
    for (int w = 0; w < W; w++)
      for (Task* t : finished_[w]) {
        if (t->result.nonfinite_evals > 0) diagnose(*t);
        on_done(*t);
        FiberStacks::get().release(t->stack);
        tasks_.erase(t->self);
      }
    stats.step_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
    return tasks_.size();
  }
  // one shared launch per (batch, kind) over every segment with a request, which serves every
  // request posted so far
  void launch() {
    auto tw = std::chrono::steady_clock::now();
    for (Batch* b : batches_) {
      for (int k = 1; k <= 3; k++) {
        bool want = false;
        for (int s = 0; s < b->S && !want; s++) want = b->requested(s) == k;
        if (!want) continue;
        b->run(k, false, use_graph);
        stats.launches[k]++;
      }
      std::fill(b->request_.begin(), b->request_.end(), 0);
    }
    launches_++;
    stats.ticks++;
    stats.gpu_wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw).count();
  }

  int nthreads() const {
    if (host_threads > 0) return host_threads;
    static const int hw = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    return hw;
  }

 private:
  struct Fcn : MinimizerFcn {
    Task* t;
    explicit Fcn(Task* tt) : t(tt) {}
    void post(const double* x, int kind) {
      double* th = t->b->theta(t->seg);
      const auto& ti = t->b->m->theta_index;
      for (size_t k = 0; k < ti.size(); k++) th[k] = x[ti[k]];
      t->pending = kind;
      t->posted = t->owner->launches_;
      mf_fiber_switch(&t->sp, *t->ret);
      note_thread(*t);
      const double* r = t->b->result(t->seg, kind);
      bool finite = true;
      for (int k = 0; k < t->b->m->nout[kind]; k++) finite = finite && std::isfinite(r[k]);
      if (!finite) {
        t->result.nonfinite_evals++;
        t->bad_theta.resize(ti.size());
        for (size_t k = 0; k < ti.size(); k++) t->bad_theta[k] = x[ti[k]];
      }
    }
    double value(const double* x) override {
      post(x, EVAL_LH);
      return -2.0 * t->b->result(t->seg, EVAL_LH)[0];
    }
    double value_noise() const override {
      return std::is_same<kernelT, float>::value ? fp32_sum_noise((double)t->b->sizes[t->seg]) : 0.0;
    }
    double value_grad(const double* x, double* g) override {
      post(x, EVAL_GRAD);
      const double* r = t->b->result(t->seg, EVAL_GRAD);
      const auto& ti = t->b->m->theta_index;
      const size_t N = t->b->m->params.size();
      for (size_t i = 0; i < N; i++) g[i] = 0.0;
      for (size_t k = 0; k < ti.size(); k++) g[ti[k]] = -2.0 * r[1 + k];
      return -2.0 * r[0];
    }
    bool hessian(const double* x, double* h) override {
      post(x, EVAL_HESS);
      const double* r = t->b->result(t->seg, EVAL_HESS);
      const auto& ti = t->b->m->theta_index;
      const size_t N = t->b->m->params.size(), P = ti.size();
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

  std::list<Task> tasks_;  // stable addresses: contexts point into them
  std::set<Batch*> batches_;
  unsigned long long added_ = 0;

  unsigned long long launches_ = 0;  // launch() calls so far
  // the live tasks of worker w, in the order they were added, and those that finished in the
  // last step
  std::vector<std::vector<Task*>> owned_, finished_;

  // where the sums of a finished fit came back non-finite: count the offending events at the
  // last such theta on the device (the calling thread issues all device work)
  static void diagnose(Task& t) {
    try {
      auto d = t.b->diagnose({t.seg}, t.bad_theta);
      t.result.nonfinite_events = d[0].nonfinite;
      t.result.first_nonfinite_event = d[0].first;
      t.result.out_of_range_events = d[0].outside;
    } catch (const std::exception&) {
      t.result.nonfinite_events = -1;  // the diagnostic itself failed; the fit's status stands
    }
  }
  static void resume(Task& t, void** here) {
    t.ret = here;
    mf_fiber_switch(here, t.sp);
  }
  // inside the fibre, at its start and after each switch back into it: the thread it runs on
  // must be the one it started on (BatchFitResult::thread_moved reports a violation)
  static void note_thread(Task& t) {
    if (t.first_thread == std::thread::id()) t.first_thread = std::this_thread::get_id();
    else if (t.first_thread != std::this_thread::get_id()) t.result.thread_moved = true;
  }
  // body of every fibre: the fit, then back to the worker for good
  static void fiber_main(void* arg) {
    Task* t = (Task*)arg;
    note_thread(*t);
    try {
      t->owner->run_fit(*t);
    } catch (...) {
      t->result.status = 4;
    }
    t->done = true;
    t->pending = 0;
    mf_fiber_switch(&t->sp, *t->ret);
    __builtin_trap();  // never resumed
  }

  void run_fit(Task& t) {
    const auto& ps = t.b->m->params;
    std::vector<MinimizerParam> mp;
    size_t k = 0;
    for (size_t i = 0; i < ps.size(); i++) {
      MinimizerParam q;
      q.name = ps[i]->get_name();
      q.fixed = ps[i]->is_constant();
      q.value = q.fixed ? ps[i]->get_value() : t.start[k++];
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
    m.print_level = print_level;
    m.trace_id = t.id;
    Fcn fcn(&t);
    try {
      m.minimize(fcn);
      if (run_hesse) t.result.hesse_ok = m.hesse(fcn);
    } catch (...) {
      m.status = 4;
    }
    for (size_t i = 0; i < t.b->m->theta_index.size(); i++) {
      t.result.values.push_back(m.x[t.b->m->theta_index[i]]);
      t.result.errors.push_back(m.errors[t.b->m->theta_index[i]]);
    }
    t.result.fval = m.fval;
    t.result.edm = m.edm;
    t.result.status = m.status;
    t.result.ncalls = m.ncalls;
    t.result.iterations = m.iterations;
  }
};

// fits of a fixed set of segments of one batch (a toy study)
template <typename kernelT, typename evalT = double>
class BatchFitter {
 public:
  HipBatch<kernelT, evalT>& b;
  bool use_graph = false;
  bool run_hesse = true;
  BatchFitStats stats;
  explicit BatchFitter(HipBatch<kernelT, evalT>& batch) : b(batch) {}

  // Each listed segment is fitted once: an index outside the batch, an empty slot or a
  // repeated index is refused (two fits of one segment would share its parameter slot, and the
  // later post would overwrite the earlier one before the shared launch).
  std::vector<BatchFitResult> fit(const std::vector<int>& segs, const std::vector<double>& start /* ntheta per listed seg */) {
    const int P = b.m->ntheta;
    if (start.size() < segs.size() * (size_t)P) throw std::invalid_argument("BatchFitter::fit: fewer start values than fits");
    std::vector<char> seen(b.S, 0);
    for (int s : segs) {
      if (s < 0 || s >= b.S) throw std::invalid_argument("BatchFitter::fit: segment " + std::to_string(s) + " outside 0.." + std::to_string(b.S - 1));
      if (!b.active[s]) throw std::invalid_argument("BatchFitter::fit: segment " + std::to_string(s) + " holds no dataset");
      if (seen[s]) throw std::invalid_argument("BatchFitter::fit: segment " + std::to_string(s) + " listed twice");
      seen[s] = 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    FitScheduler<kernelT, evalT> sch;
    sch.use_graph = use_graph;
    sch.run_hesse = run_hesse;
    for (size_t i = 0; i < segs.size(); i++) sch.add(&b, segs[i], &start[i * P], (long long)i);
    const double setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::vector<BatchFitResult> out(segs.size());
    while (sch.step([&](auto& t) { out[t.id] = t.result; }) > 0) sch.launch();
    stats = sch.stats;
    stats.setup_ms = setup_ms;
    stats.host_threads = segs.size() > 1 ? sch.nthreads() : 1;
    stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats.host_ms = stats.wall_ms - stats.gpu_wait_ms;
    return out;
  }
};

}  // namespace morefit
#endif

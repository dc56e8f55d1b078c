/**
 * @file serve.hh
 * Fit service with continuous batching.  Requests (a density class, an event count, a
 * seed, an arrival time) arrive while other fits run.  A request is admitted when its
 * class's pool has a free slot and the service's HBM budget covers its events (all its data
 * columns); its dataset is generated on the device into the one event arena every class
 * shares (a single allocation of the budget), and its fit joins the shared fibre
 * scheduler.  Every tick issues one launch per (class, kind) over all fits that are waiting
 * for a value, a gradient or a Hessian, so the likelihood evaluations of concurrent fits
 * share launches; a finished fit releases its slot and memory at once.  Results are
 * bitwise the same alone and in the mix: a segment's sums depend only on its own chunks.
 *
 * No reference counterpart (the reference fits one dataset at a time, fitter.hh:660-1002).
 */
#ifndef MOREFIT_SERVE_HH
#define MOREFIT_SERVE_HH

#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
#include <thread>
#include <vector>

#include "batch.hh"

namespace morefit {

struct ServeRequest {
  long long id = 0;
  int cls = 0;
  long long n = 0;
  uint64_t seed = 0;
  double arrival_s = 0.0;
  std::vector<double> start;  // the fit's start values (ntheta of its class); empty: the class's
};

struct ServeResult {
  long long id = 0;
  int cls = 0;
  long long n = 0;
  double arrival_s = 0.0, admit_s = -1.0, done_s = -1.0;
  BatchFitResult fit;
};

struct ServeStats {
  int ticks = 0, launches = 0, peak_live = 0, rejected = 0;
  double wall_s = 0.0, peak_bytes = 0.0, generate_ms = 0.0;
  double arena_bytes = 0.0;  // device memory holding the event data of all classes together
};

template <typename kernelT, typename evalT = double>
class FitServer {
 public:
  struct Class {
    HipModel<kernelT, evalT>* model;
    std::vector<double> theta_gen, start;  // generation values and fit start values
    std::unique_ptr<HipBatch<kernelT, evalT>> pool;
  };
  std::vector<Class> classes;
  size_t budget_bytes;
  ServeStats stats;
  bool use_graph = false;  // membership changes every few ticks: direct launches
  // events per generator stream, by request size: each stream draws its events one after
  // another (about four proposals per event for the mass density), so short streams keep a
  // small request's generation latency low, while longer ones need fewer streams seeded by
  // jump-ahead and generate large requests faster (10^7 FP32 mass events: 0.34 ms at 64 per
  // stream, 0.53 ms at 16).  A request's data depend only on its size and seed: stream k starts
  // k jump()s after the seed.
  long long small_request_events = 65536;
  int events_per_stream_small = 16, events_per_stream_large = 64;

  FitServer(size_t hbm_budget_bytes, int slots_per_class) : budget_bytes(hbm_budget_bytes), slots_(slots_per_class) {}

  void add_class(HipModel<kernelT, evalT>* m, const std::vector<double>& theta_gen, const std::vector<double>& start) {
    if (!arena_) arena_ = std::make_shared<EventArena<kernelT>>(m->be->device(), (long long)(budget_bytes / sizeof(kernelT)));
    Class c;
    c.model = m;
    c.theta_gen = theta_gen;
    c.start = start;
    c.pool.reset(new HipBatch<kernelT, evalT>(m, slots_, arena_));
    classes.push_back(std::move(c));
    stats.arena_bytes = (double)arena_->device_bytes();
  }
  // bytes of the shared arena a request of n events of class c occupies (all its columns:
  // observables and precomputed event terms)
  static double request_bytes(const Class& c, long long n) {
    return (double)ArenaAllocator::round(n * c.model->ncol) * sizeof(kernelT);
  }

  // Serves `reqs` (sorted by arrival) in real time; returns one result per request.
  std::vector<ServeResult> run(const std::vector<ServeRequest>& reqs) {
    std::vector<ServeResult> res(reqs.size());
    for (size_t i = 0; i < reqs.size(); i++) {
      res[i].id = reqs[i].id;
      res[i].cls = reqs[i].cls;
      res[i].n = reqs[i].n;
      res[i].arrival_s = reqs[i].arrival_s;
    }
    FitScheduler<kernelT, evalT> sch;
    sch.use_graph = use_graph;
    std::deque<size_t> queue;
    size_t next = 0, finished = 0;
    double used = 0.0;
    std::vector<int> slot_of(reqs.size(), -1);
    const auto t0 = std::chrono::steady_clock::now();
    auto now = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    while (finished < reqs.size()) {
      const double t = now();
      while (next < reqs.size() && reqs[next].arrival_s <= t) queue.push_back(next++);
      // admission in arrival order, bounded by slots and the HBM budget; the data of all
      // requests admitted in this tick are generated in one launch per class
      std::vector<std::vector<size_t>> admitted(classes.size());
      while (!queue.empty()) {
        size_t i = queue.front();
        const ServeRequest& r = reqs[i];
        Class& c = classes.at(r.cls);
        const double need = request_bytes(c, r.n);
        if (need > (double)budget_bytes) {  // can never fit: reject
          queue.pop_front();
          stats.rejected++;
          res[i].fit.status = -2;
          res[i].done_s = now();
          finished++;
          continue;
        }
        if (used + need > (double)budget_bytes || !c.pool->fits(r.n)) break;
        int slot = c.pool->add_segment(r.n);
        if (slot < 0) break;
        queue.pop_front();
        used += need;
        stats.peak_bytes = std::max(stats.peak_bytes, used);
        slot_of[i] = slot;
        admitted[r.cls].push_back(i);
      }
      for (size_t ci = 0; ci < classes.size(); ci++) {
        if (admitted[ci].empty()) continue;
        Class& c = classes[ci];
        auto tgen = std::chrono::steady_clock::now();
        for (int large = 0; large < 2; large++) {
          std::vector<double> tg;
          std::vector<uint64_t> seeds;
          std::vector<int> slots;
          for (size_t i : admitted[ci]) {
            if ((reqs[i].n > small_request_events) != (large == 1)) continue;
            tg.insert(tg.end(), c.theta_gen.begin(), c.theta_gen.end());
            seeds.push_back(reqs[i].seed);
            slots.push_back(slot_of[i]);
          }
          if (!slots.empty()) c.pool->generate(tg, seeds, large ? events_per_stream_large : events_per_stream_small, slots);
        }
        stats.generate_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tgen).count();
        for (size_t i : admitted[ci]) {
          res[i].admit_s = now();
          const std::vector<double>& st = reqs[i].start.empty() ? c.start : reqs[i].start;
          if ((int)st.size() < c.model->ntheta) throw std::invalid_argument("FitServer: request start values too short");
          sch.add(c.pool.get(), slot_of[i], st.data(), (long long)i);
        }
      }
      if (sch.live() == 0) {
        if (next < reqs.size() && queue.empty()) {
          double wait = reqs[next].arrival_s - now();
          if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
        }
        continue;
      }
      stats.peak_live = std::max(stats.peak_live, (int)sch.live());
      sch.step([&](auto& task) {
        size_t i = (size_t)task.id;
        res[i].fit = task.result;
        res[i].done_s = now();
        Class& c = classes.at(reqs[i].cls);
        c.pool->remove_segment(slot_of[i]);
        used -= request_bytes(c, reqs[i].n);
        finished++;
      });
      if (sch.live() > 0) sch.launch();
    }
    stats.ticks = sch.stats.ticks;
    stats.launches = sch.stats.launches[1] + sch.stats.launches[2] + sch.stats.launches[3];
    stats.wall_s = now();
    return res;
  }

 private:
  int slots_;
  std::shared_ptr<EventArena<kernelT>> arena_;
};

}  // namespace morefit
#endif

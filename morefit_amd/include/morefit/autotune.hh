/**
 * @file autotune.hh
 * Launch-shape autotuner for the evaluation kernels of the MI355X backend.
 *
 * A launch shape is the work-group size, the events per thread (chunk = block x events per
 * thread) and the requested unroll (events per thread whose loads are issued together; the
 * code generator may lower it per kernel to fit the register budget).  The tuner builds the
 * density's module and a batch of the workload's segment sizes for each candidate, fills the
 * batch on the device from the generating values, and times back-to-back launches of one
 * evaluation kind (per-chunk kernel and final pass, HIP events on the device stream).  The
 * backend's options are set to each candidate while its model and batch exist and are left at
 * the fastest, so models built afterwards use it.
 *
 * No counterpart in the reference (its OpenCL work-group size is a fixed option,
 * compute.hh:33); the candidates span the shapes that matter on gfx950: 4-8 waves per
 * work-group and 8-64 events per thread for large workloads (unroll 16 too from 10^8 events),
 * 2-8 waves and 2-16 events per thread below a few million events (launch_candidates).
 */
#ifndef MOREFIT_AUTOTUNE_HH
#define MOREFIT_AUTOTUNE_HH

#include <cmath>
#include <vector>

#include "compute_hip.hh"

namespace morefit {

struct LaunchShape {
  int block = 256;
  int events_per_thread = 32;
  int unroll = 8;
};

inline std::vector<LaunchShape> default_launch_candidates() {
  std::vector<LaunchShape> c;
  for (int block : {256, 512})
    for (int ept : {8, 16, 32, 64})
      for (int unroll : {4, 8}) c.push_back({block, ept, unroll});
  return c;
}

// Candidates for a workload of `total_events` events.  From a few million events up the
// default set (4-8 waves per work-group, 8-64 events per thread) already gives every CU
// several work-groups; below, smaller chunks are needed for that (10^6 events at 256 x 16 are
// 245 work-groups on 256 CUs), so the set moves to 2-8 waves and 2-16 events per thread (an
// unroll above the events per thread is lowered to them).  From 10^8 events (beyond the
// last-level cache, every work-group one of many rounds) 16 events' loads in flight per thread
// are tried as well: there the kernel streams, and deeper loads keep more bytes in flight
// (FP32 mass gradient over 2 GB: 6.64 -> 6.78 TB/s at 512 x 64).
inline std::vector<LaunchShape> launch_candidates(long long total_events) {
  if (total_events >= 100000000) {
    auto c = default_launch_candidates();
    for (int block : {256, 512})
      for (int ept : {32, 64}) c.push_back({block, ept, 16});
    return c;
  }
  if (total_events >= 4000000) return default_launch_candidates();
  std::vector<LaunchShape> c;
  for (int block : {128, 256, 512})
    for (int ept : {2, 4, 8, 16})
      for (int unroll : {4, 8}) {
        const LaunchShape s{block, ept, std::min(unroll, ept)};
        bool seen = false;
        for (auto& q : c) seen = seen || (q.block == s.block && q.events_per_thread == s.events_per_thread && q.unroll == s.unroll);
        if (!seen) c.push_back(s);
      }
  return c;
}

// Returns the time per launch (us) of each candidate; *best receives the fastest, which is
// also left in the backend's options.
template <typename kernelT, typename evalT = double>
std::vector<double> autotune_launch_shape(HIPBackend* be, PDF<kernelT, evalT>* pdf, const std::vector<parameter<evalT>*>& params,
                                          const std::vector<long long>& sizes, int kind, const std::vector<double>& theta_gen,
                                          const std::vector<LaunchShape>& candidates, int iters, LaunchShape* best) {
  compute_options& o = *be->opts_;
  const compute_options saved = o;
  std::vector<double> us(candidates.size(), NAN);
  double best_us = INFINITY;
  LaunchShape win = {saved.hip_block_size, saved.hip_events_per_thread, saved.hip_unroll};
  for (size_t c = 0; c < candidates.size(); c++) {
    o.hip_block_size = candidates[c].block;
    o.hip_events_per_thread = candidates[c].events_per_thread;
    o.hip_unroll = candidates[c].unroll;
    HipModel<kernelT, evalT> m(be, pdf, params, kind == EVAL_HESS);
    if (m.nout[kind] == 0) throw std::runtime_error("autotune: evaluation kind not compiled");
    HipBatch<kernelT, evalT> b(&m, sizes);
    std::vector<double> tg;
    std::vector<uint64_t> seeds;
    for (int s = 0; s < b.S; s++) {
      tg.insert(tg.end(), theta_gen.begin(), theta_gen.end());
      seeds.push_back(1000 + (uint64_t)s);
    }
    b.generate(tg, seeds);
    b.set_modes(kind, true);
    b.prep(kind);
    b.launch_kernel(kind);  // warm-up
    hip_check(mf_sync(b.dev), "autotune");
    hip_check(mf_timer_begin(b.dev), "autotune timer");
    for (int i = 0; i < iters; i++) b.launch_kernel(kind);
    float ms = 0.0f;
    hip_check(mf_timer_end(b.dev, &ms), "autotune timer");
    us[c] = 1e3 * (double)ms / iters;
    if (us[c] < best_us) {
      best_us = us[c];
      win = candidates[c];
    }
  }
  o = saved;
  o.hip_block_size = win.block;
  o.hip_events_per_thread = win.events_per_thread;
  o.hip_unroll = win.unroll;
  if (best) *best = win;
  return us;
}

}  // namespace morefit
#endif

// fiber_check.cc -- CPU check of the host fibres the batched fits run in (morefit/fiber.hh):
// many fibres pinned to the workers of the process-wide thread pool, each yielding back to its
// worker many times; floating-point state and callee-saved registers survive the switches;
// exceptions thrown and caught inside a fibre; a deep recursion on the fibre stack; stacks
// reused by a second round.  Prints "ok <sum>" or "FAIL <reason>" and exits non-zero on failure.
#include <cfenv>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <thread>
#include <list>
#include <stdexcept>
#include <vector>

#include "morefit/fiber.hh"

using namespace morefit;

namespace {

struct Task {
  char* stack = nullptr;
  void* sp = nullptr;
  void** ret = nullptr;
  int id = 0, worker = 0, yields = 0, left = 0;
  bool done = false;
  double acc = 0.0;
  int round_mode = 0;
};

int depth_sum(int d) {
  volatile char pad[256];
  pad[0] = (char)d;
  return d == 0 ? pad[0] : d + depth_sum(d - 1) + (pad[0] - (char)d);
}

void yield(Task* t) { mf_fiber_switch(&t->sp, *t->ret); }

void body(void* arg) {
  Task* t = (Task*)arg;
  // a rounding mode of its own: the MXCSR is saved per fibre
  std::fesetround(t->id % 2 ? FE_UPWARD : FE_DOWNWARD);
  double x = 1.0 + t->id;
  for (int k = 0; k < t->left; k++) {
    const double before = x;
    x = std::sqrt(x * x + 1.0);
    t->acc += x - before;
    if (k % 7 == 3) {
      try {
        throw std::runtime_error("inside a fibre");
      } catch (const std::exception&) {
        t->acc += 1.0;
      }
    }
    if (k == 5) t->acc += depth_sum(200) - 200 * 201 / 2;
    t->yields++;
    yield(t);
    t->round_mode = std::fegetround();
  }
  t->done = true;
  yield(t);
  __builtin_trap();
}

}  // namespace

int main() {
  const int T = 8, N = 3000, K = 20;
  double total = 0.0;
  std::vector<char*> first_round;
  for (int round = 0; round < 2; round++) {
    std::list<Task> tasks;
    for (int i = 0; i < N; i++) {
      tasks.emplace_back();
      Task& t = tasks.back();
      t.id = i;
      t.worker = i % T;
      t.left = K + i % 5;
      t.stack = FiberStacks::get().acquire();
      if (round == 0) first_round.push_back(t.stack);
      t.sp = fiber_prepare(t.stack, FiberStacks::kStackBytes, &body, &t);
    }
    std::vector<std::vector<Task*>> groups(T);
    for (auto& t : tasks) groups[t.worker].push_back(&t);
    bool live = true;
    while (live) {
      HostThreads::get().run(T, [&](int w) {
        void* here = nullptr;
        std::fesetround(FE_TONEAREST);
        for (Task* t : groups[w])
          if (!t->done) {
            t->ret = &here;
            mf_fiber_switch(&here, t->sp);
            if (std::fegetround() != FE_TONEAREST) {
              std::printf("FAIL worker rounding mode changed by a fibre\n");
              std::exit(1);
            }
          }
      });
      live = false;
      for (auto& t : tasks) live |= !t.done;
    }
    for (auto& t : tasks) {
      if (t.yields != t.left) {
        std::printf("FAIL task %d yielded %d times, expected %d\n", t.id, t.yields, t.left);
        return 1;
      }
      const int want = t.id % 2 ? FE_UPWARD : FE_DOWNWARD;
      if (t.round_mode != want) {
        std::printf("FAIL task %d lost its rounding mode\n", t.id);
        return 1;
      }
      total += t.acc;
      FiberStacks::get().release(t.stack);
    }
  }
  // the thread pool alone: runs of 1-16 threads back to back (workers still polling) and after
  // pauses longer than the polling window (workers blocked): every run calls fn(w) once for each
  // w < T, worker w is always the same OS thread, and the caller sees every write of the run
  {
    const int R = 6000;
    std::vector<int> last(HostThreads::kMaxThreads, -1), calls(HostThreads::kMaxThreads, 0);
    std::vector<std::thread::id> ids(HostThreads::kMaxThreads);
    for (int r = 0; r < R; r++) {
      const int Tr = 1 + (r * 7) % HostThreads::kMaxThreads;
      if (r % 1000 == 999) std::this_thread::sleep_for(std::chrono::milliseconds(2));
      HostThreads::get().run(Tr, [&](int w) {
        if (ids[w] == std::thread::id()) ids[w] = std::this_thread::get_id();
        else if (ids[w] != std::this_thread::get_id()) last[w] = -2;  // moved: reported below
        if (last[w] != -2) last[w] = r;
        calls[w]++;
      });
      for (int w = 0; w < HostThreads::kMaxThreads; w++)
        if ((w < Tr && last[w] != r) || last[w] == -2) {
          std::printf("FAIL pool run %d (T=%d): worker %d saw %d\n", r, Tr, w, last[w]);
          return 1;
        }
    }
    int sum = 0;
    for (int c : calls) sum += c;
    int want = 0;
    for (int r = 0; r < R; r++) want += 1 + (r * 7) % HostThreads::kMaxThreads;
    if (sum != want) {
      std::printf("FAIL pool calls %d, expected %d\n", sum, want);
      return 1;
    }
  }
  // the second round ran on stacks of the first (the pool reuses them)
  std::printf("ok %.6f %zu\n", total, first_round.size());
  return 0;
}

/**
 * @file fiber.hh
 * Host fibres for the batched fits: a context switch of a few instructions, stacks from a
 * process-wide pool, and a process-wide pool of worker threads.
 *
 * A toy study or the fit service runs thousands of MIGRAD fits, each in its own fibre, and
 * switches into every live fibre once per launch.  ucontext's swapcontext saves and restores
 * the signal mask with a system call on every switch, and fresh 128 KB stacks per fit cost an
 * allocation and first-touch page faults per fit; both dominated the host time of a study.
 * Here a switch saves the callee-saved registers of the x86-64 System V ABI (rbx, rbp,
 * r12-r15, the MXCSR and x87 control words) on the current stack and swaps stack pointers;
 * stacks (with a guard page below each) are mapped in slabs once and reused by later fits, so
 * their pages stay resident.  No counterpart in the reference (its toy study fits one toy
 * after another, toystudy.hh:55-111).
 */
#ifndef MOREFIT_FIBER_HH
#define MOREFIT_FIBER_HH

#include <sys/mman.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <vector>

#if !defined(__x86_64__)
#error "morefit fibres: x86-64 hosts only"
#endif

// void mf_fiber_switch(void** save_sp, void* next_sp): pushes the callee-saved state on the
// current stack, stores the stack pointer in *save_sp, loads next_sp and pops the state saved
// there, returning into the other fibre.  mf_fiber_start is where a fresh fibre's first
// switch returns to: it calls r13(r12) and never returns.  Weak and hidden: every translation
// unit that includes this header may emit them.
asm(R"(
  .text
  .weak mf_fiber_switch
  .hidden mf_fiber_switch
  .type mf_fiber_switch, @function
mf_fiber_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  subq $8, %rsp
  stmxcsr (%rsp)
  fnstcw 4(%rsp)
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  ldmxcsr (%rsp)
  fldcw 4(%rsp)
  addq $8, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
  .size mf_fiber_switch, .-mf_fiber_switch

  .weak mf_fiber_start
  .hidden mf_fiber_start
  .type mf_fiber_start, @function
mf_fiber_start:
  movq %r12, %rdi
  callq *%r13
  ud2
  .size mf_fiber_start, .-mf_fiber_start
)");
extern "C" void mf_fiber_switch(void** save_sp, void* next_sp);
extern "C" void mf_fiber_start();

namespace morefit {

// Stack pointer of a fresh fibre on [lo, lo + bytes) whose first switch calls fn(arg); fn must
// never return (it switches away for the last time instead).
inline void* fiber_prepare(char* lo, size_t bytes, void (*fn)(void*), void* arg) {
  uintptr_t top = ((uintptr_t)lo + bytes) & ~(uintptr_t)15;
  top -= 16;
  uint64_t* sp = (uint64_t*)(top - 64);  // 16-aligned; after the 64 bytes are popped, rsp = top
  uint32_t mxcsr = 0;
  uint16_t fpucw = 0;
  asm volatile("stmxcsr %0" : "=m"(mxcsr));
  asm volatile("fnstcw %0" : "=m"(fpucw));
  sp[0] = (uint64_t)mxcsr | ((uint64_t)fpucw << 32);
  sp[1] = 0;                               // r15
  sp[2] = 0;                               // r14
  sp[3] = (uint64_t)(uintptr_t)fn;         // r13
  sp[4] = (uint64_t)(uintptr_t)arg;        // r12
  sp[5] = 0;                               // rbx
  sp[6] = 0;                               // rbp
  sp[7] = (uint64_t)(uintptr_t)&mf_fiber_start;  // return address of the first switch
  return sp;
}

// Process-wide pool of fibre stacks of kStackBytes, each above a PROT_NONE guard page (an
// overflow faults instead of running into a neighbour).  Slabs are never unmapped: a stack's
// pages stay resident for the next fit that takes it.
class FiberStacks {
 public:
  static constexpr size_t kStackBytes = 128 * 1024;
  static constexpr size_t kPerSlab = 64;
  static FiberStacks& get() {
    static FiberStacks pool;
    return pool;
  }
  char* acquire() {
    std::lock_guard<std::mutex> lk(mu_);
    if (free_.empty()) grow();
    char* s = free_.back();
    free_.pop_back();
    return s;
  }
  void release(char* s) {
    if (!s) return;
    std::lock_guard<std::mutex> lk(mu_);
    free_.push_back(s);
  }

 private:
  FiberStacks() = default;
  void grow() {
    const size_t page = (size_t)sysconf(_SC_PAGESIZE);
    const size_t stride = kStackBytes + page;
    void* base = mmap(nullptr, stride * kPerSlab, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (base == MAP_FAILED) throw std::runtime_error("fibre stacks: mmap failed");
    for (size_t i = 0; i < kPerSlab; i++) {
      char* guard = (char*)base + i * stride;
      if (mprotect(guard, page, PROT_NONE) != 0) throw std::runtime_error("fibre stacks: mprotect failed");
      free_.push_back(guard + page);
    }
  }
  std::mutex mu_;
  std::vector<char*> free_;
};

// Process-wide pool of host threads: run(T, fn) calls fn(w) for w = 0..T-1, each on worker
// thread w (so work keyed by w always runs on the same OS thread), and returns when all have
// finished.  Threads are started on first use, up to the largest T asked for, and stopped at
// process exit.  One run at a time (callers serialise on run_mu_).
class HostThreads {
 public:
  static constexpr int kMaxThreads = 16;
  // A worker that has done its part polls for the next run for kPollUs before it blocks, and
  // the caller polls for the end of a run likewise: the ticks of a batched fit come every few
  // hundred microseconds, and a blocked thread took tens of microseconds to wake (a run of 16
  // threads over empty work: ~40-60 us through the condition variables alone).
  static constexpr int kPollUs = 300;
  static HostThreads& get() {
    static HostThreads pool;
    return pool;
  }
  void run(int T, const std::function<void(int)>& fn) {
    if (T <= 0) return;
    if (T > kMaxThreads) throw std::invalid_argument("HostThreads::run: too many threads");
    std::lock_guard<std::mutex> serial(run_mu_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      while ((int)threads_.size() < T) {
        const int w = (int)threads_.size();
        threads_.emplace_back([this, w] { worker(w); });
      }
      fn_ = &fn;
      active_ = T;
      running_.store(T, std::memory_order_relaxed);
      generation_.fetch_add(1, std::memory_order_release);
    }
    cv_go_.notify_all();
    auto finished = [this] { return running_.load(std::memory_order_acquire) == 0; };
    if (!poll(finished)) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_done_.wait(lk, finished);
    }
    std::lock_guard<std::mutex> lk(mu_);
    fn_ = nullptr;
  }
  ~HostThreads() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stopping_.store(true, std::memory_order_release);
    }
    cv_go_.notify_all();
    for (auto& t : threads_) t.join();
  }

 private:
  HostThreads() = default;
  // polls `ready` for up to kPollUs (a pause between polls, a yield every 64); false when it
  // never held
  template <class P>
  static bool poll(P ready) {
    const auto end = std::chrono::steady_clock::now() + std::chrono::microseconds(kPollUs);
    for (unsigned i = 1;; i++) {
      if (ready()) return true;
      if ((i & 63) == 0) {
        if (std::chrono::steady_clock::now() > end) return false;
        std::this_thread::yield();
      } else {
        __builtin_ia32_pause();
      }
    }
  }
  void worker(int w) {
    unsigned long long seen = 0;
    auto go = [&] { return stopping_.load(std::memory_order_acquire) || generation_.load(std::memory_order_acquire) != seen; };
    for (;;) {
      const std::function<void(int)>* fn = nullptr;
      poll(go);
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_go_.wait(lk, go);
        if (stopping_.load(std::memory_order_acquire)) return;
        seen = generation_.load(std::memory_order_acquire);
        if (w >= active_) continue;  // not part of this run
        fn = fn_;
      }
      (*fn)(w);
      if (running_.fetch_sub(1, std::memory_order_acq_rel) == 1) {
        std::lock_guard<std::mutex> lk(mu_);  // the caller checks under mu_ before it blocks
        cv_done_.notify_one();
      }
    }
  }
  std::mutex run_mu_, mu_;
  std::condition_variable cv_go_, cv_done_;
  std::vector<std::thread> threads_;
  const std::function<void(int)>* fn_ = nullptr;
  std::atomic<unsigned long long> generation_{0};
  std::atomic<int> running_{0};
  std::atomic<bool> stopping_{false};
  int active_ = 0;
};

}  // namespace morefit
#endif

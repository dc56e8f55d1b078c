// launch_floor.hip -- calibration of the per-call floor of a one-segment evaluation on this device:
// the host round trip of a kernel launch whose completion is polled in mapped memory (as the
// engine's one-segment calls complete), of two dependent launches (per-chunk kernel + final
// kernel), of a launch followed by hipStreamSynchronize, and the host cost of the launch call
// itself, with an 8-byte and with a 1.1 KB argument block (the size of mf_eval_args).
// Output: one line per pass (two passes).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>

#define CK(x)                                                                      \
  do {                                                                             \
    hipError_t e = (x);                                                            \
    if (e != hipSuccess) {                                                         \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e), __LINE__); \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

// stores v into the mapped flag after a system-scope fence, as the engine's last kernel does
__global__ void k_flag(int* flag, int v) {
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(flag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
// an argument block of the size of mf_eval_args
struct BigArgs {
  int* flag;
  int v;
  unsigned int words[128];
  double terms[64];
};
__global__ void k_flag_big(BigArgs b) {
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(b.flag, b.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
// writes nothing when given a null pointer (it always is here)
__global__ void k_nop(int* p) {
  if (p != nullptr && threadIdx.x == 0) p[0] = 1;
}

int main() {
  int *flag, *flag_d;
  CK(hipHostMalloc(&flag, 64, hipHostMallocMapped));
  CK(hipHostGetDevicePointer((void**)&flag_d, flag, 0));
  *flag = 0;
  hipStream_t s;
  CK(hipStreamCreate(&s));
  // bounded wait for the flag (seconds at most)
  auto poll = [&](int v) {
    for (long spin = 0; spin < 2000000000L; spin++)
      if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == v) return true;
    return false;
  };
  using clk = std::chrono::steady_clock;
  auto us = [](clk::time_point a, clk::time_point b, int n) { return std::chrono::duration<double, std::micro>(b - a).count() / n; };
  const int it = 3000;
  int v = 0;
  for (int pass = 0; pass < 2; pass++) {
    auto t0 = clk::now();
    for (int i = 0; i < it; i++) {
      k_flag<<<1, 64, 0, s>>>(flag_d, ++v);
      if (!poll(v)) return std::printf("flag not set\n"), 1;
    }
    const double one = us(t0, clk::now(), it);
    t0 = clk::now();
    for (int i = 0; i < it; i++) {
      k_nop<<<1, 64, 0, s>>>(nullptr);
      k_flag<<<1, 64, 0, s>>>(flag_d, ++v);
      if (!poll(v)) return std::printf("flag not set\n"), 1;
    }
    const double two = us(t0, clk::now(), it);
    t0 = clk::now();
    for (int i = 0; i < it; i++) {
      k_nop<<<1, 64, 0, s>>>(nullptr);
      CK(hipStreamSynchronize(s));
    }
    const double sync = us(t0, clk::now(), it);
    t0 = clk::now();
    for (int i = 0; i < it; i++) k_nop<<<1, 64, 0, s>>>(nullptr);
    const double api = us(t0, clk::now(), it);
    CK(hipStreamSynchronize(s));
    BigArgs b{};
    b.flag = flag_d;
    t0 = clk::now();
    for (int i = 0; i < it; i++) {
      b.v = ++v;
      k_flag_big<<<1, 64, 0, s>>>(b);
      if (!poll(v)) return std::printf("flag not set\n"), 1;
    }
    const double big = us(t0, clk::now(), it);
    t0 = clk::now();
    for (int i = 0; i < it; i++) k_flag_big<<<1, 64, 0, s>>>(b);
    const double big_api = us(t0, clk::now(), it);
    CK(hipStreamSynchronize(s));
    std::printf("launch+poll %.2f us, two launches+poll %.2f us, launch+hipStreamSynchronize %.2f us, launch call %.2f us; "
                "1.1 KB arguments: launch+poll %.2f us, launch call %.2f us\n",
                one, two, sync, api, big, big_api);
  }
  CK(hipStreamDestroy(s));
  CK(hipHostFree(flag));
  return 0;
}

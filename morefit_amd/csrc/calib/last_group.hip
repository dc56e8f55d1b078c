// last_group.hip -- calibration of in-kernel segment reduction ("the last work-group to count
// reduces"): per launch of G work-groups of 512 threads that each write a 6-value partial and
// then (by mode) do nothing more; count on one device-scope counter between two fences (the
// runtime's fused_final scheme); count with a relaxed agent-scope atomic; count with an acq_rel
// agent-scope atomic; count on one of eight counters (work-group mod 8) and then on a second-level
// counter.  The work-group that counts last reduces all partials (512 threads, 8 loads in flight
// each) and resets the counters.  G = 256, 1221 (the headline workload's chunks) and 2442.  The
// relaxed and acq_rel modes do not order the partials' stores before the count: they time the
// counting only (the printed check uses the fenced mode).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#define CK(x)                                                                      \
  do {                                                                             \
    hipError_t e = (x);                                                            \
    if (e != hipSuccess) {                                                         \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e), __LINE__); \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

constexpr int NO = 6, GMAX = 4096;

template <int MODE>
__global__ void __launch_bounds__(512) k_count(double* part, unsigned* cnt, double* out) {
  __shared__ int last;
  __shared__ double red[512];
  const int g = blockIdx.x, G = gridDim.x;
  if (threadIdx.x < NO) part[g * NO + threadIdx.x] = (double)(g + threadIdx.x);
  __syncthreads();
  if (MODE == 0) return;
  if (threadIdx.x == 0) {
    unsigned prev = 0, n = (unsigned)G;
    unsigned* c = cnt;
    if (MODE == 1) {
      __threadfence();
      prev = atomicAdd(c, 1u);
      __threadfence();
    } else if (MODE == 2) {
      prev = __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (MODE == 3) {
      prev = __hip_atomic_fetch_add(c, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      c = cnt + 1 + (g & 7);
      n = (unsigned)((G - (g & 7) + 7) / 8);  // work-groups of this residue
      __threadfence();
      prev = atomicAdd(c, 1u);
      if (prev == n - 1) {  // last of its group: reset the group counter, count on the top one
        *c = 0u;
        prev = atomicAdd(cnt, 1u);
        n = 8;
        c = cnt;
      } else {
        prev = 0;
        n = 2;  // not last
      }
      __threadfence();
    }
    last = prev == n - 1;
    if (last) *c = 0u;
  }
  __syncthreads();
  if (last) {
    // output k = t % 8 (< NO), chunk rows r = t / 8, r + 64, ... with 8 loads in flight
    const int t = threadIdx.x, k = t & 7, r = t >> 3;
    double s = 0.0;
    if (k < NO) {
      for (int j0 = r; j0 < G; j0 += 8 * 64) {
        double b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) b[u] = (j0 + u * 64 < G) ? part[(j0 + u * 64) * NO + k] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++) s += b[u];
      }
    }
    red[t] = s;
    __syncthreads();
    if (t < NO) {
      double tot = 0.0;
      for (int rr = 0; rr < 64; rr++) tot += red[rr * 8 + t];
      out[t] = tot;
    }
  }
}

template <typename F>
float us_per_launch(F f, int iters) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1.0f;
  for (int i = 0; i < 20; i++) f();
  (void)hipEventRecord(a, 0);
  for (int i = 0; i < iters; i++) f();
  (void)hipEventRecord(b, 0);
  (void)hipEventSynchronize(b);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms * 1e3f / iters;
}

int main() {
  double *part, *out;
  unsigned* cnt;
  CK(hipMalloc(&part, GMAX * NO * sizeof(double)));
  CK(hipMalloc(&out, 64 * sizeof(double)));
  CK(hipMalloc(&cnt, 64 * sizeof(unsigned)));
  CK(hipMemset(cnt, 0, 64 * sizeof(unsigned)));
  CK(hipDeviceSynchronize());
  for (int G : {256, 1221, 2442}) {
    std::printf("G %d: partials only %.2f us", G, us_per_launch([&] { k_count<0><<<G, 512>>>(part, cnt, out); }, 500));
    std::printf(", fence+atomic+fence %.2f", us_per_launch([&] { k_count<1><<<G, 512>>>(part, cnt, out); }, 200));
    std::printf(", relaxed %.2f", us_per_launch([&] { k_count<2><<<G, 512>>>(part, cnt, out); }, 200));
    std::printf(", acq_rel %.2f", us_per_launch([&] { k_count<3><<<G, 512>>>(part, cnt, out); }, 200));
    std::printf(", 8+1 counters %.2f us\n", us_per_launch([&] { k_count<4><<<G, 512>>>(part, cnt, out); }, 200));
  }
  CK(hipDeviceSynchronize());
  std::vector<double> o(NO);
  CK(hipMemcpy(o.data(), out, NO * sizeof(double), hipMemcpyDeviceToHost));
  std::printf("check: out[0] %g (expect %g)\n", o[0], 2442.0 * 2441.0 / 2.0);
  CK(hipFree(part));
  CK(hipFree(out));
  CK(hipFree(cnt));
  return 0;
}

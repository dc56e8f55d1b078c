// chunk_flags.hip -- calibration of the whole call (launch + polled completion, host clock) for
// four ways to reduce a segment of 1221 chunks (the headline workload's) after a per-chunk pass
// of realistic weight (512 threads x 16 float events, exp2 + log2 + reciprocal + 6 sums per event):
//  mode 0: per-chunk kernel + one-work-group final kernel writing host totals and the flag (the
//          engine's scheme);
//  mode 1: one kernel; every work-group publishes its partials behind a per-chunk epoch flag
//          (release fence + agent-scope store); work-group G-1, after its own chunk, polls all
//          flags (bounded spin; a timeout is reported in the totals), then reduces and signals;
//  mode 2: one kernel; every work-group counts on one device-scope counter between two fences;
//          the one that counts last reduces and signals;
//  mode 3: one kernel; every work-group writes its partials through to memory (agent-scope stores),
//          waits for their acknowledgement and counts with relaxed atomics on one of 16 counters on
//          separate 256-byte lines, the last of each counting on a top counter (no release fence,
//          whose L2 write-back serialises); the last overall takes an acquire fence, reduces and
//          signals.
// Also the per-chunk kernel alone, back to back (HIP events).  All three totals are printed.
#include <hip/hip_runtime.h>

#include <chrono>
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

constexpr int NO = 6, G = 1221, NT = 512, EPT = 16, NEV = G * NT * EPT;
constexpr long SPIN_LIMIT = 1L << 20;

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// one work-group of NTR threads reduces G x NO partials; thread 0 writes totals + flag
template <int NTR>
__device__ void reduce_all(const double* part, double* out, int* flag, int v, int err) {
  constexpr int KP = 8, R = NTR / KP, NW = NTR / 64;
  __shared__ double s_fin[NTR];
  const int t = threadIdx.x, q = t % KP, r = t / KP, lane = t & 63, wave = t >> 6;
  double acc = 0.0, er = 0.0;
  if (q < NO) {
    for (int j0 = r; j0 < G; j0 += 16 * R) {
      double b[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int j = j0 + u * R;
        b[u] = j < G ? part[j * NO + q] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 16; u++) {
        double s, e;
        two_sum(acc, b[u], s, e);
        acc = s;
        er += e;
      }
    }
  }
  double vv = acc + er;
#pragma unroll
  for (int off = 32; off >= KP; off >>= 1) vv += __shfl_xor(vv, off, 64);
  if (lane < KP) s_fin[wave * KP + lane] = vv;
  __syncthreads();
  if (t < NO) {
    double tot = 0.0;
    for (int w = 0; w < NW; w++) tot += s_fin[w * KP + t];
    out[t] = tot;
  }
  if (t == NO) out[NO] = (double)err;
  __syncthreads();
  if (t == 0) {
    __threadfence_system();
    __hip_atomic_store(flag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

template <int MODE>
__global__ void __launch_bounds__(NT) k_chunk(const float* data, double* part, int* cflags, unsigned* cnt, double* out,
                                              int* flag, int v, float a, float b, float c) {
  __shared__ double s_all[NO * NT];
  __shared__ int s_last;
  const int g = blockIdx.x, t = threadIdx.x;
  float acc[NO] = {0, 0, 0, 0, 0, 0};
  const float* col = data + (long)g * NT * EPT;
  float x[EPT];
#pragma unroll
  for (int u = 0; u < EPT; u++) x[u] = __builtin_nontemporal_load(col + t + u * NT);
#pragma unroll
  for (int u = 0; u < EPT; u++) {
    const float e = exp2f(x[u] * a + b);
    const float l = log2f(e + c);
    const float r = __frcp_rn(e + c);
    acc[0] += l;
    acc[1] += e * r;
    acc[2] += e * r * x[u];
    acc[3] += r;
    acc[4] += r * x[u];
    acc[5] += r * x[u] * x[u];
  }
#pragma unroll
  for (int k = 0; k < NO; k++) s_all[k * NT + t] = acc[k];
  __syncthreads();
  {
    const int j = t & 31;
    for (int k = t >> 5; k < NO; k += NT / 32) {
      double s = 0.0;
      for (int gg = 0; gg < NT / 32; gg++) s += s_all[k * NT + gg * 32 + j];
#pragma unroll
      for (int off = 16; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (j == 0) {
        if (MODE == 3) __hip_atomic_store(part + g * NO + k, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else part[g * NO + k] = s;
      }
    }
  }
  if (MODE == 0) return;
  if (MODE == 3) {
    __builtin_amdgcn_s_waitcnt(0);  // the partial stores acknowledged
    __syncthreads();
    if (t == 0) {
      constexpr int C = 16, LINE = 64;
      const int q = g % C;
      const unsigned n1 = (unsigned)((G - q + C - 1) / C);
      int last = 0;
      unsigned prev = __hip_atomic_fetch_add(cnt + (1 + q) * LINE, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (prev == n1 - 1) {
        __hip_atomic_store(cnt + (1 + q) * LINE, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        prev = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == C - 1) {
          __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          last = 1;
        }
      }
      s_last = last;
    }
    __syncthreads();
    if (s_last) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      reduce_all<NT>(part, out, flag, v, 0);
    }
    return;
  }
  __syncthreads();
  if (MODE == 1) {
    if (t == 0) {
      __threadfence();  // release: partials before the flag
      __hip_atomic_store(cflags + g, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (g != G - 1) return;
    int bad = 0;
    for (int j = t; j < G; j += NT) {
      long spins = 0;
      while (__hip_atomic_load(cflags + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != v && spins < SPIN_LIMIT) {
        __builtin_amdgcn_s_sleep(1);
        spins++;
      }
      bad |= spins >= SPIN_LIMIT;
    }
    bad = __syncthreads_or(bad);
    if (t == 0) __threadfence();  // acquire
    __syncthreads();
    reduce_all<NT>(part, out, flag, v, bad);
  } else {
    if (t == 0) {
      __threadfence();
      const unsigned prev = atomicAdd(cnt, 1u);
      __threadfence();
      s_last = prev == (unsigned)(G - 1);
      if (s_last) *cnt = 0u;
    }
    __syncthreads();
    if (s_last) reduce_all<NT>(part, out, flag, v, 0);
  }
}

__global__ void __launch_bounds__(1024) k_final(const double* part, double* out, int* flag, int v) {
  reduce_all<1024>(part, out, flag, v, 0);
}

int main() {
  float* data;
  double *part, *hout, *hout_d;
  int *cflags, *flag, *flag_d;
  unsigned* cnt;
  CK(hipMalloc(&data, (size_t)NEV * sizeof(float)));
  std::vector<float> h(NEV);
  for (int i = 0; i < NEV; i++) h[i] = 5.0f + 0.6f * (float)((i * 2654435761u) % 1000003u) / 1000003.0f;
  CK(hipMemcpy(data, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  CK(hipMalloc(&part, G * NO * sizeof(double)));
  CK(hipMalloc(&cflags, G * sizeof(int)));
  CK(hipMemset(cflags, 0, G * sizeof(int)));
  CK(hipMalloc(&cnt, 17 * 64 * 4));
  CK(hipMemset(cnt, 0, 17 * 64 * 4));
  CK(hipHostMalloc(&hout, 64 * sizeof(double), hipHostMallocMapped));
  CK(hipHostGetDevicePointer((void**)&hout_d, hout, 0));
  CK(hipHostMalloc(&flag, 64, hipHostMallocMapped));
  CK(hipHostGetDevicePointer((void**)&flag_d, flag, 0));
  *flag = 0;
  CK(hipDeviceSynchronize());
  hipStream_t s;
  CK(hipStreamCreate(&s));
  auto poll = [&](int v) {
    for (long spin = 0; spin < 2000000000L; spin++)
      if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == v) return true;
    return false;
  };
  using clk = std::chrono::steady_clock;
  const float a = -1.5f, b = 7.0f, c = 0.25f;
  int v = 0;
  const int it = 3000;
  double tot[4][NO + 1];
  for (int pass = 0; pass < 2; pass++) {
    for (int mode = 0; mode < 4; mode++) {
      clk::time_point t0;
      for (int i = -200; i < it; i++) {
        if (i == 0) t0 = clk::now();
        ++v;
        if (mode == 0) {
          k_chunk<0><<<G, NT, 0, s>>>(data, part, cflags, cnt, hout_d, flag_d, v, a, b, c);
          k_final<<<1, 1024, 0, s>>>(part, hout_d, flag_d, v);
        } else if (mode == 1) {
          k_chunk<1><<<G, NT, 0, s>>>(data, part, cflags, cnt, hout_d, flag_d, v, a, b, c);
        } else if (mode == 2) {
          k_chunk<2><<<G, NT, 0, s>>>(data, part, cflags, cnt, hout_d, flag_d, v, a, b, c);
        } else {
          k_chunk<3><<<G, NT, 0, s>>>(data, part, cflags, cnt, hout_d, flag_d, v, a, b, c);
        }
        if (!poll(v)) return std::printf("flag not set (mode %d)\n", mode), 1;
      }
      const double us = std::chrono::duration<double, std::micro>(clk::now() - t0).count() / it;
      for (int k = 0; k <= NO; k++) tot[mode][k] = hout[k];
      std::printf("mode %d: %.2f us per call (err flag %g)\n", mode, us, hout[NO]);
    }
    // kernel alone (events), mode 0 chunk kernel only
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    CK(hipEventRecord(e0, s));
    for (int i = 0; i < 500; i++) k_chunk<0><<<G, NT, 0, s>>>(data, part, cflags, cnt, hout_d, flag_d, 0, a, b, c);
    CK(hipEventRecord(e1, s));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("chunk kernel alone (back to back) %.2f us\n", ms * 1e3 / 500);
  }
  for (int k = 0; k < NO; k++) std::printf("total %d: %.17g %.17g %.17g %.17g\n", k, tot[0][k], tot[1][k], tot[2][k], tot[3][k]);
  CK(hipStreamDestroy(s));
  return 0;
}

// issue_rate.hip -- calibration of the VALU issue cost of FP32 transcendentals on this device:
// per wave and loop iteration, NV independent FMA chains and NT chains of hardware
// transcendentals (v_rcp_f32, or v_exp_f32 / v_log_f32 in turn so the values stay bounded),
// 8 independent accumulators each, at full occupancy (256 CUs x 32 work-groups of 256
// threads).  The time per wave-iteration, in SIMD cycles at 2.4 GHz, gives the cycles one
// wave64 instruction of each kind occupies its SIMD: 4 for an FMA, and for the transcendentals
// 8 (half rate) on gfx950.  The NV + NT rows show whether the two kinds overlap (they add up:
// one issue port).  The FP64 rows do the same with FP64 FMAs and v_rcp_f64.  Output: one line per
// configuration.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(x)                                                                      \
  do {                                                                             \
    hipError_t e = (x);                                                            \
    if (e != hipSuccess) {                                                         \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e), __LINE__); \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

// KIND 0: v_rcp_f32 chains; 1: v_exp_f32 / v_log_f32 alternating; 2: FMAs in place of both;
// 3: the FP64 kernel k_mix64
template <int NV, int NT, int KIND>
__global__ void __launch_bounds__(256) k_mix(float* out, int iters, float c, float d) {
  float a[8], b[8];
  for (int i = 0; i < 8; i++) {
    a[i] = threadIdx.x * 1e-3f + i;
    b[i] = 0.5f + 0.01f * i + threadIdx.x * 1e-5f;
  }
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int v = 0; v < NV; v++) a[v % 8] = fmaf(a[v % 8], c, d);
#pragma unroll
    for (int t = 0; t < NT; t++) {
      if (KIND == 0) b[t % 8] = __builtin_amdgcn_rcpf(b[t % 8]);
      else if (KIND == 1) b[t % 8] = (t & 1) ? __builtin_amdgcn_logf(b[t % 8]) : __builtin_amdgcn_exp2f(b[t % 8]);
      else b[t % 8] = fmaf(b[t % 8], c, d);
    }
  }
  float s = 0.f;
  for (int i = 0; i < 8; i++) s += a[i] + b[i];
  out[blockIdx.x * 256 + threadIdx.x] = s;  // grid * 256 floats
}

// FP64: NV FMA chains and NT v_rcp_f64 chains (the FP64 kernels' one hardware transcendental)
template <int NV, int NT>
__global__ void __launch_bounds__(256) k_mix64(float* out, int iters, double c, double d) {
  double a[8], b[8];
  for (int i = 0; i < 8; i++) {
    a[i] = threadIdx.x * 1e-3 + i;
    b[i] = 0.5 + 0.01 * i + threadIdx.x * 1e-5;
  }
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int v = 0; v < NV; v++) a[v % 8] = fma(a[v % 8], c, d);
#pragma unroll
    for (int t = 0; t < NT; t++) b[t % 8] = __builtin_amdgcn_rcp(b[t % 8]);
  }
  double s = 0.0;
  for (int i = 0; i < 8; i++) s += a[i] + b[i];
  out[blockIdx.x * 256 + threadIdx.x] = (float)s;  // grid * 256 floats
}

template <int NV, int NT, int KIND>
int run(float* out, int grid, int iters, const char* label) {
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  auto launch = [&] {
    if (KIND == 3) k_mix64<NV, NT><<<grid, 256>>>(out, iters, 0.999, 0.001);
    else k_mix<NV, NT, KIND><<<grid, 256>>>(out, iters, 0.999f, 0.001f);
  };
  launch();
  CK(hipGetLastError());
  CK(hipEventRecord(e0));
  for (int r = 0; r < 5; r++) launch();
  CK(hipEventRecord(e1));
  CK(hipEventSynchronize(e1));
  float ms = 0;
  CK(hipEventElapsedTime(&ms, e0, e1));
  const double wave_iters = (double)grid * 4 * iters * 5;  // four wave64 per work-group
  const double cycles = ms * 1e-3 / wave_iters * 1024 * 2.4e9;  // 1024 SIMDs at 2.4 GHz
  std::printf("%2d FMA + %2d %-8s per iteration: %.3f ms per launch, %6.2f SIMD cycles per wave-iteration\n", NV, NT, label, ms / 5,
              cycles);
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  return 0;
}

int main() {
  const int grid = 256 * 32, iters = 2000;
  float* out = nullptr;
  CK(hipMalloc(&out, (size_t)grid * 256 * sizeof(float)));
  int rc = 0;
  rc |= run<16, 0, 2>(out, grid, iters, "-");
  rc |= run<32, 0, 2>(out, grid, iters, "-");
  rc |= run<0, 16, 0>(out, grid, iters, "rcp");
  rc |= run<0, 16, 1>(out, grid, iters, "exp/log");
  rc |= run<16, 16, 0>(out, grid, iters, "rcp");
  rc |= run<16, 16, 1>(out, grid, iters, "exp/log");
  rc |= run<32, 8, 0>(out, grid, iters, "rcp");
  rc |= run<32, 8, 1>(out, grid, iters, "exp/log");
  rc |= run<16, 0, 3>(out, grid, iters, "- (FP64)");
  rc |= run<0, 16, 3>(out, grid, iters, "rcp_f64");
  rc |= run<16, 16, 3>(out, grid, iters, "rcp_f64");
  CK(hipFree(out));
  return rc;
}

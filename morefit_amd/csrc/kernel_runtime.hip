// kernel_runtime.hip -- device runtime that libmorefit_hip prepends to every generated
// likelihood module before compiling it with hipRTC for gfx950.
//
// The generated source defines, before this runtime:
//   MF_T            float or double (storage and per-event arithmetic type)
//   MF_FP32         1 for float kernels
//   MF_NOBS         observables per event (SoA columns)
//   MF_NTHETA       floating parameters (FP64 on the host, first kernel-parameter slots)
//   MF_NHOIST       hoisted parameter-only terms (FP64 on the device prep kernel)
//   MF_BLOCK        work-group size of the evaluation kernels (multiple of 64)
//   MF_ACC          0: FP64 accumulation per event, 1: compensated (Kahan) FP32 per lane
//   MF_FAST_MATH    1: hardware exp2/log2/rcp approximations in FP32
//   MF_GEN_LO<d>, MF_GEN_W<d>  generator proposal box per observable (stored coordinates)
// and, after it, functor structs { static constexpr int nout; static void eval(x, p, o); },
// mf_hoist(theta, buf) and the kernel instantiations (MF_EVAL_KERNEL etc. below).  The
// unroll of each evaluation kernel (events per thread whose loads are issued together) is an
// argument of MF_EVAL_KERNEL, chosen per kernel by the host so that the kernel's registers
// stay in the 256 architectural VGPRs (see mf_module_load).
//
// Semantics it implements (reference compute_opencl.hh:646-995): per event the log density
// and its derivatives; per output the compensated sum over the events of each segment.
// Differences by design: the sum is fused into the per-event kernel (no N x outputs
// round trip through memory, no ceil(log32 N) extra passes); partial sums are indexed by
// (segment, chunk) and a second small kernel reduces each segment in a fixed order, so a
// result is bitwise the same whatever else shares the launch, with no fences or atomics.

#define MF_NPARAM (MF_NTHETA + MF_NHOIST)
#define MF_GEN_MAX_TRIES (1u << 20)
#define MF_NWAVE (MF_BLOCK / 64)

// ------------------------------------------------------------------ math
#if MF_FP32
#if MF_FAST_MATH
__device__ __forceinline__ float mf_exp(float x) { return __expf(x); }
__device__ __forceinline__ float mf_log(float x) { return __logf(x); }
__device__ __forceinline__ float mf_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float mf_exp2(float x) { return __builtin_amdgcn_exp2f(x); }  // v_exp_f32
__device__ __forceinline__ float mf_log2(float x) { return __builtin_amdgcn_logf(x); }   // v_log_f32
#else
__device__ __forceinline__ float mf_exp(float x) { return expf(x); }
__device__ __forceinline__ float mf_log(float x) { return logf(x); }
__device__ __forceinline__ float mf_rcp(float x) { return __builtin_amdgcn_rcpf(x); }  // 1 ulp
__device__ __forceinline__ float mf_exp2(float x) { return exp2f(x); }
__device__ __forceinline__ float mf_log2(float x) { return log2f(x); }
#endif
__device__ __forceinline__ float mf_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ float mf_sin(float x) { return sinf(x); }
__device__ __forceinline__ float mf_cos(float x) { return cosf(x); }
__device__ __forceinline__ float mf_tan(float x) { return tanf(x); }
__device__ __forceinline__ float mf_asin(float x) { return asinf(x); }
__device__ __forceinline__ float mf_acos(float x) { return acosf(x); }
__device__ __forceinline__ float mf_atan(float x) { return atanf(x); }
__device__ __forceinline__ float mf_erf(float x) { return erff(x); }
__device__ __forceinline__ float mf_pow(float x, float y) { return powf(x, y); }
#else
__device__ __forceinline__ double mf_exp(double x) { return exp(x); }
__device__ __forceinline__ double mf_log(double x) { return log(x); }
__device__ __forceinline__ double mf_exp2(double x) { return exp2(x); }
__device__ __forceinline__ double mf_log2(double x) { return log2(x); }
__device__ __forceinline__ double mf_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ double mf_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double mf_sin(double x) { return sin(x); }
__device__ __forceinline__ double mf_cos(double x) { return cos(x); }
__device__ __forceinline__ double mf_tan(double x) { return tan(x); }
__device__ __forceinline__ double mf_asin(double x) { return asin(x); }
__device__ __forceinline__ double mf_acos(double x) { return acos(x); }
__device__ __forceinline__ double mf_atan(double x) { return atan(x); }
__device__ __forceinline__ double mf_erf(double x) { return erf(x); }
__device__ __forceinline__ double mf_pow(double x, double y) { return pow(x, y); }
#endif
// FP64 versions for the hoisted terms (always double)
__device__ __forceinline__ double mfd_exp(double x) { return exp(x); }
__device__ __forceinline__ double mfd_exp2(double x) { return exp2(x); }
__device__ __forceinline__ double mfd_log2(double x) { return log2(x); }
__device__ __forceinline__ double mfd_log(double x) { return log(x); }
__device__ __forceinline__ double mfd_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ double mfd_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double mfd_sin(double x) { return sin(x); }
__device__ __forceinline__ double mfd_cos(double x) { return cos(x); }
__device__ __forceinline__ double mfd_tan(double x) { return tan(x); }
__device__ __forceinline__ double mfd_asin(double x) { return asin(x); }
__device__ __forceinline__ double mfd_acos(double x) { return acos(x); }
__device__ __forceinline__ double mfd_atan(double x) { return atan(x); }
__device__ __forceinline__ double mfd_erf(double x) { return erf(x); }
__device__ __forceinline__ double mfd_pow(double x, double y) { return pow(x, y); }

// ------------------------------------------------------------------ accumulation
// Per-lane accumulator of NO outputs.  MF_ACC == 1: Kahan in FP32 (4 VALU ops per output
// and event), converted to FP64 once per chunk; otherwise FP64 sums.  With MF_ACC == 0 every
// event's value is added in FP64 (1 convert + 1 FP64 add); with MF_ACC == 2 the event loop
// first sums the U events of one unrolled group in FP32 (U - 1 adds) and adds that group sum
// in FP64, about 1 + 2/U VALU ops per output and event.
template <int NO>
struct mf_acc {
#if MF_FP32 && MF_ACC == 1
  float s[NO], c[NO];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < NO; k++) { s[k] = 0.0f; c[k] = 0.0f; }
  }
  __device__ __forceinline__ void add(const float* o) {
#pragma unroll
    for (int k = 0; k < NO; k++) {
      float y = o[k] - c[k];
      float t = s[k] + y;
      c[k] = (t - s[k]) - y;
      s[k] = t;
    }
  }
  __device__ __forceinline__ void get(double* v) const {
#pragma unroll
    for (int k = 0; k < NO; k++) v[k] = (double)s[k] - (double)c[k];
  }
#else
  double s[NO];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < NO; k++) s[k] = 0.0;
  }
  __device__ __forceinline__ void add(const MF_T* o) {
#pragma unroll
    for (int k = 0; k < NO; k++) s[k] += (double)o[k];
  }
  __device__ __forceinline__ void get(double* v) const {
#pragma unroll
    for (int k = 0; k < NO; k++) v[k] = s[k];
  }
#endif
};

// butterfly over the 64 lanes of a wavefront: every lane ends with the same total and the
// order of additions is fixed, so the result is deterministic
__device__ __forceinline__ double mf_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// error-free transformation a + b = s + e (Knuth two-sum)
__device__ __forceinline__ void mf_two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// FP64 factor of output k's total for segment `seg`: the parameter-only factor the code
// generator split off the per-event term (functor scale(k) >= 0), else 1
template <class F>
__device__ __forceinline__ double mf_output_scale(const mf_eval_args& a, int seg, int k) {
  const int si = F::scale(k);
  return (si >= 0 && a.hoist64) ? a.hoist64[(long long)seg * MF_NHOIST + si] : 1.0;
}

// ------------------------------------------------------------------ fused evaluation
// Grid-stride over the chunk table (a persistent grid when gridDim < nchunks).  Per chunk:
// coalesced SoA loads, per-event outputs from the generated functor F, per-lane
// accumulation, wave64 butterfly + LDS reduction to one FP64 partial per output, written to
// partials[chunk][k].  No fences or atomics: the segment totals come from the separate
// final kernel below, launched right after on the same stream.
template <class F, int U>
__device__ __forceinline__ void mf_eval_impl(const mf_eval_args& a) {
  constexpr int NO = F::nout;
  __shared__ double s_red[MF_NWAVE][NO];
  const MF_T* __restrict__ data = (const MF_T*)a.data;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
    const mf_chunk ch = a.chunks[c];
    const mf_segment sg = a.segs[ch.seg];
    if (sg.mode == 0) continue;
    const MF_T* __restrict__ P = (const MF_T*)a.params + sg.param_off;
    const MF_T* __restrict__ col = data + sg.data_off;
    mf_acc<NO> acc;
    acc.zero();
    int i = ch.begin + (int)threadIdx.x;
    // U events per thread per step: their loads are issued together, then computed
    for (; i + (U - 1) * MF_BLOCK < ch.end; i += U * MF_BLOCK) {
      MF_T x[U][MF_NOBS];
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int d = 0; d < MF_NOBS; d++) x[u][d] = __builtin_nontemporal_load(col + d * sg.stride + i + u * MF_BLOCK);
#if MF_FP32 && MF_ACC == 2
      MF_T g[NO];
      F::eval(x[0], P, g);
#pragma unroll
      for (int u = 1; u < U; u++) {
        MF_T o[NO];
        F::eval(x[u], P, o);
#pragma unroll
        for (int k = 0; k < NO; k++) g[k] += o[k];
      }
      acc.add(g);
#else
#pragma unroll
      for (int u = 0; u < U; u++) {
        MF_T o[NO];
        F::eval(x[u], P, o);
        acc.add(o);
      }
#endif
    }
    for (; i < ch.end; i += MF_BLOCK) {
      MF_T x[MF_NOBS];
#pragma unroll
      for (int d = 0; d < MF_NOBS; d++) x[d] = __builtin_nontemporal_load(col + d * sg.stride + i);
      MF_T o[NO];
      F::eval(x, P, o);
      acc.add(o);
    }
    double v[NO];
    acc.get(v);
#pragma unroll
    for (int k = 0; k < NO; k++) {
      double w = mf_wave_sum(v[k]);
      if (lane == 0) s_red[wave][k] = w;
    }
    __syncthreads();
    if (threadIdx.x < NO) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < MF_NWAVE; w++) t += s_red[w][threadIdx.x];
      a.partials[(long long)c * NO + threadIdx.x] = t;
      // a segment of one chunk is complete here: its totals need no second pass
      if (sg.chunk_end - sg.chunk_begin == 1) a.results[sg.out_off + threadIdx.x] = t * mf_output_scale<F>(a, ch.seg, threadIdx.x);
    }
    __syncthreads();
  }
}

// Segment totals of segments with more than one chunk: one work-group of MF_FINAL_BLOCK
// threads per segment.  Thread t sums chunks t, t + MF_FINAL_BLOCK, ... with compensated
// (two-sum) FP64 additions, MF_FINAL_KB outputs at a time; the lanes are combined by the fixed
// wave64 butterfly and the waves in wave order through LDS.  The order of every addition
// depends only on the segment's own chunk count, so a segment's totals are bitwise the same
// whatever else shares the launch.
#define MF_FINAL_BLOCK 256
#define MF_FINAL_KB 8
template <class F>
__device__ __forceinline__ void mf_final_impl(const mf_eval_args& a) {
  constexpr int NO = F::nout;
  __shared__ double s_red[MF_FINAL_BLOCK / 64][MF_FINAL_KB];
  const int s = blockIdx.x;
  if (s >= a.nseg) return;
  const mf_segment sg = a.segs[s];
  const int nch = sg.chunk_end - sg.chunk_begin;
  if (sg.mode == 0 || nch <= 1) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const double* __restrict__ part = a.partials + (long long)sg.chunk_begin * NO;
  for (int k0 = 0; k0 < NO; k0 += MF_FINAL_KB) {
    double acc[MF_FINAL_KB], err[MF_FINAL_KB];
#pragma unroll
    for (int kk = 0; kk < MF_FINAL_KB; kk++) acc[kk] = err[kk] = 0.0;
#pragma unroll 4
    for (int j = t; j < nch; j += MF_FINAL_BLOCK) {
#pragma unroll
      for (int kk = 0; kk < MF_FINAL_KB; kk++) {
        if (k0 + kk < NO) {
          double u, e;
          mf_two_sum(acc[kk], part[(long long)j * NO + k0 + kk], u, e);
          acc[kk] = u;
          err[kk] += e;
        }
      }
    }
#pragma unroll
    for (int kk = 0; kk < MF_FINAL_KB; kk++) {
      const double w = mf_wave_sum(acc[kk] + err[kk]);
      if (lane == 0) s_red[wave][kk] = w;
    }
    __syncthreads();
    if (t < MF_FINAL_KB && k0 + t < NO) {
      double tot = 0.0;
#pragma unroll
      for (int w = 0; w < MF_FINAL_BLOCK / 64; w++) tot += s_red[w][t];
      a.results[sg.out_off + k0 + t] = tot * mf_output_scale<F>(a, s, k0 + t);
    }
    __syncthreads();
  }
}

// per-event outputs written to memory (the reference's kahan_on_accelerator == false path
// and its per-event precompute kernels): store[k * store_stride + data_off + i]
template <class F>
__device__ __forceinline__ void mf_store_impl(const mf_eval_args& a) {
  constexpr int NO = F::nout;
  const MF_T* __restrict__ data = (const MF_T*)a.data;
  MF_T* __restrict__ out = (MF_T*)a.store;
  for (int c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
    const mf_chunk ch = a.chunks[c];
    const mf_segment sg = a.segs[ch.seg];
    if (sg.mode == 0) continue;
    const MF_T* __restrict__ P = (const MF_T*)a.params + sg.param_off;
    const MF_T* __restrict__ col = data + sg.data_off;
    for (int i = ch.begin + (int)threadIdx.x; i < ch.end; i += MF_BLOCK) {
      MF_T x[MF_NOBS];
#pragma unroll
      for (int d = 0; d < MF_NOBS; d++) x[d] = col[d * sg.stride + i];
      MF_T o[NO];
      F::eval(x, P, o);
#pragma unroll
      for (int k = 0; k < NO; k++) out[k * a.store_stride + sg.data_off + i] = o[k];
    }
  }
}

#define MF_EVAL_KERNEL(NAME, FUNCTOR, UNROLL)                                                                    \
  extern "C" __global__ void __launch_bounds__(MF_BLOCK) NAME(mf_eval_args a) { mf_eval_impl<FUNCTOR, UNROLL>(a); } \
  extern "C" __global__ void __launch_bounds__(MF_FINAL_BLOCK) NAME##_final(mf_eval_args a) { mf_final_impl<FUNCTOR>(a); }
#define MF_STORE_KERNEL(NAME, FUNCTOR) \
  extern "C" __global__ void __launch_bounds__(MF_BLOCK) NAME(mf_eval_args a) { mf_store_impl<FUNCTOR>(a); }

// ------------------------------------------------------------------ parameter preparation
// One thread per segment: hoisted parameter-only terms in FP64 from the segment's floating
// parameters, then [theta, hoisted] rounded once to MF_T into the kernel parameter block.
// theta may live in mapped host memory (read once, before anything is stored); with seg_src
// the segment's table entry is staged from there into the device table the evaluation kernels
// read; with hoist64 the hoisted terms are also kept in FP64 (output scales, mf_eval_args).
__device__ void mf_hoist(const double* th, double* buf);

#define MF_PREP_KERNEL                                                                                  \
  extern "C" __global__ void mf_prep(const double* theta, MF_T* params, double* hoist64, const int* seg_list, \
                                     int nseg, const mf_segment* seg_src, mf_segment* seg_dst) {        \
    const int i = blockIdx.x * blockDim.x + threadIdx.x;                                                \
    if (i >= nseg) return;                                                                              \
    const int s = seg_list ? seg_list[i] : i;                                                           \
    double th[MF_NTHETA > 0 ? MF_NTHETA : 1];                                                           \
    for (int k = 0; k < MF_NTHETA; k++) th[k] = theta[(long long)s * MF_NTHETA + k];                    \
    if (seg_src) seg_dst[s] = seg_src[s];                                                               \
    MF_T* out = params + (long long)s * MF_NPARAM;                                                      \
    double hb[MF_NHOIST > 0 ? MF_NHOIST : 1];                                                           \
    mf_hoist(th, hb);                                                                                   \
    for (int k = 0; k < MF_NTHETA; k++) out[k] = (MF_T)th[k];                                           \
    for (int j = 0; j < MF_NHOIST; j++) out[MF_NTHETA + j] = (MF_T)hb[j];                               \
    if (hoist64)                                                                                        \
      for (int j = 0; j < MF_NHOIST; j++) hoist64[(long long)s * MF_NHOIST + j] = hb[j];                \
  }

// ------------------------------------------------------------------ generator
// xoshiro128++ (Blackman & Vigna 2019), the stream the reference seeds per work item
__device__ __forceinline__ unsigned int mf_rotl32(unsigned int x, int k) { return (x << k) | (x >> (32 - k)); }
__device__ __forceinline__ unsigned int mf_xoshiro_next(unsigned int* s) {
  const unsigned int r = mf_rotl32(s[0] + s[3], 7) + s[0];
  const unsigned int t = s[1] << 9;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = mf_rotl32(s[3], 11);
  return r;
}
// uniform in [0,1): r * 2^-32 formed in double and then rounded to MF_T (compute_opencl.hh:438-449)
__device__ __forceinline__ MF_T mf_uniform(unsigned int* s) { return (MF_T)(mf_xoshiro_next(s) * 2.3283064365386963e-10); }

// Accept-reject (compute_opencl.hh:540-575): propose each observable uniformly in the box,
// accept when maxprob * u < p(x).  Each stream fills its own fixed event range, so the
// output does not depend on scheduling.  L is the log-density functor (nout == 1, its value
// the log density itself: no output scale).
template <class L>
__device__ __forceinline__ void mf_generate_impl(const mf_gen_args& g) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.nstreams) return;
  const mf_gen_stream st = g.streams[k];
  const mf_segment sg = g.segs[st.seg];
  unsigned int s[4];
#pragma unroll
  for (int j = 0; j < 4; j++) s[j] = g.states[4 * k + j];
  const MF_T* __restrict__ P = (const MF_T*)g.params + sg.param_off;
  const MF_T mp = (MF_T)g.maxprob[st.seg];
  MF_T* __restrict__ out = (MF_T*)g.data + sg.data_off;
  unsigned long long props = 0;
  bool failed = false;
  for (int i = st.begin; i < st.end; i++) {
    MF_T x[MF_NOBS];
    // bounded: a density that is NaN or far above maxprob's scale must not hang the stream
    unsigned int tries = 0;
    for (;;) {
      props++;
      MF_GEN_PROPOSE(x, s);
      MF_T lp[1];
      L::eval(x, P, lp);
      const MF_T p = mf_exp(lp[0]);
      const MF_T u = mf_uniform(s);
      if (mp * u < p) break;
      if (++tries >= MF_GEN_MAX_TRIES) {
        failed = true;
        break;
      }
    }
    if (failed) {
      for (int j = i; j < st.end; j++)
        for (int d = 0; d < MF_NOBS; d++) out[d * sg.stride + j] = (MF_T)(0.0 / 0.0);
      break;
    }
#pragma unroll
    for (int d = 0; d < MF_NOBS; d++) out[d * sg.stride + i] = x[d];
  }
#pragma unroll
  for (int j = 0; j < 4; j++) g.states[4 * k + j] = s[j];
  if (g.proposals) atomicAdd(g.proposals + st.seg, props);
}

#define MF_GENERATE_KERNEL(FUNCTOR) \
  extern "C" __global__ void mf_generate(mf_gen_args g) { mf_generate_impl<FUNCTOR>(g); }

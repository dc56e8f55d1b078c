// kernel_runtime.hip -- device runtime that libmorefit_hip prepends to every generated
// likelihood module before compiling it with hipRTC for gfx950.
//
// The generated source defines, before this runtime:
//   MF_T            float or double (storage and per-event arithmetic type)
//   MF_FP32         1 for float kernels
//   MF_NOBS         observables per event (SoA columns the generator writes and uploads fill)
//   MF_NCOL         columns the evaluation kernels read: the observables, then event-only terms
//                   the precompute kernel fills (MF_NCOL > MF_NOBS; mf_event_terms)
//   MF_NTHETA       floating parameters (FP64 on the host, first kernel-parameter slots)
//   MF_NHOIST       hoisted parameter-only terms (FP64 on the device prep kernel)
//   MF_BLOCK        work-group size of the evaluation kernels (multiple of 64)
//   MF_ACC          0: FP64 accumulation per event, 1: compensated (Kahan) FP32 per lane
//   MF_FAST_MATH    1: hardware exp2/log2/rcp/sqrt/sin/cos approximations in FP32
//   MF_GEN_LO<d>, MF_GEN_W<d>  generator proposal box per observable (stored coordinates)
// and, after it, functor structs { static constexpr int nout; static void eval(x, p, o); },
// mf_hoist_term(k, theta) and the kernel instantiations (MF_EVAL_KERNEL etc. below).  The
// unroll of each evaluation kernel (events per thread whose loads are issued together) is an
// argument of MF_EVAL_KERNEL, chosen per kernel by the host so that the kernel's registers
// stay in the 256 architectural VGPRs (see mf_module_load).
//
// Semantics it implements (reference compute_opencl.hh:646-995): per event the log density
// and its derivatives; per output the compensated sum over the events of each segment.
// Differences by design: the sum is fused into the per-event kernel (no N x outputs
// round trip through memory, no ceil(log32 N) extra passes); partial sums are indexed by
// (segment, chunk) and each segment is reduced in a fixed order, so a result is bitwise the
// same whatever else shares the launch.  The reduction runs either in the work-group that
// completes a segment's last chunk (a per-segment counter tells it, mf_eval_args.seg_done) or
// in a second small kernel launched after the first (no fences or atomics).

#ifndef MF_NCOL
#define MF_NCOL MF_NOBS
#endif
#ifndef MF_NMOM
#define MF_NMOM 0  // data moments per segment (mf_eval_args.moments)
#endif
// read by the host at module load: launches of the evaluation kernels need the moments
extern "C" __device__ __attribute__((used)) int mf_nmom = MF_NMOM;
#define MF_NPARAM (MF_NTHETA + MF_NHOIST)
#define MF_GEN_MAX_TRIES (1u << 20)
#define MF_NWAVE (MF_BLOCK / 64)
#define MF_CONST __attribute__((address_space(4)))  // constant memory: read-only for the kernel's lifetime
#ifndef MF_PREFETCH
#define MF_PREFETCH 1       // evaluation kernels load the next event group before computing this one
#endif
#ifndef MF_PREFETCH_MAX
#define MF_PREFETCH_MAX 32  // ... when a group holds at most this many values per thread
#endif

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
#if MF_FAST_MATH
__device__ __forceinline__ float mf_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }  // v_sqrt_f32, 1 ulp
// The hardware sine / cosine (v_sin_f32 / v_cos_f32 take the argument in turns and are valid
// for |turns| <= 256): one multiply and one half-rate instruction instead of the library's
// range reduction and polynomials.  Beyond 256 turns (1608 rad) the fraction of the turn is
// passed, which loses nothing the FP32 argument still holds; below, the turns are passed as
// they are, keeping the relative accuracy of small negative angles that a fraction would lose.
__device__ __forceinline__ float mf_turns(float x) {
  const float t = x * 0.159154943f;  // 1 / (2 pi)
  return fabsf(t) <= 256.0f ? t : __builtin_amdgcn_fractf(t);
}
__device__ __forceinline__ float mf_sin(float x) { return __builtin_amdgcn_sinf(mf_turns(x)); }
__device__ __forceinline__ float mf_cos(float x) { return __builtin_amdgcn_cosf(mf_turns(x)); }
#else
__device__ __forceinline__ float mf_sqrt(float x) { return sqrtf(x); }  // correctly rounded
__device__ __forceinline__ float mf_sin(float x) { return sinf(x); }
__device__ __forceinline__ float mf_cos(float x) { return cosf(x); }
#endif
__device__ __forceinline__ float mf_tan(float x) { return tanf(x); }
__device__ __forceinline__ float mf_asin(float x) { return asinf(x); }
__device__ __forceinline__ float mf_acos(float x) { return acosf(x); }
__device__ __forceinline__ float mf_atan(float x) { return atanf(x); }
__device__ __forceinline__ float mf_erf(float x) { return erff(x); }
__device__ __forceinline__ float mf_pow(float x, float y) { return powf(x, y); }
#else
// FP64 exp, log, reciprocal, sine and cosine of the per-event code, written out: the library's
// exp, log and reciprocal take 42, 98 and 12 VALU instructions on gfx950 and were most of an FP64 event's work.  Each is
// within about 2 ulp and branch-free (a branch to the library for special arguments, inlined at
// every unrolled event, doubled the kernels' registers): special arguments are handled by
// clamping and selects.
// exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2 / 2 (ln2 split so that k * ln2_hi is exact
// for |k| < 2^11), e^r by its Taylor polynomial to degree 13 (truncation < 1e-17), times 2^k.
// The argument is clamped to [-746, 710], where exp underflows to 0 / overflows to inf.
__device__ __forceinline__ double mf_exp(double x) {
  const double xc = fmin(fmax(x, -746.0), 710.0);
  const double k = rint(xc * 0x1.71547652b82fep+0);
  double r = fma(-k, 0x1.62e42fefa3800p-1, xc);
  r = fma(-k, 0x1.ef35793c76730p-45, r);
  double p = 1.0 / 6227020800.0;  // 1/13!
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  const double v = ldexp(p, (int)k);
  return x == x ? v : x;
}
// reciprocal: the hardware approximation refined by two Newton steps; where those give NaN
// (zero, infinite or NaN argument) the hardware value itself (inf, 0, NaN)
__device__ __forceinline__ double mf_rcp(double x) {
  const double r0 = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r0, 1.0);
  double r = fma(r0, e, r0);
  e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  return r == r ? r : r0;
}
// log: x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m - 1) / (m + 1),
// |s| <= 0.1716, by its series 2 s + 2 s^3 / 3 + ... + 2 s^21 / 21 (truncation < 3e-17 of it);
// log x = e ln2_hi + (e ln2_lo + log m).  Zero, negative, infinite, NaN: -inf, NaN, inf, NaN.
__device__ __forceinline__ double mf_log(double x) {
  double m = __builtin_amdgcn_frexp_mant(x);
  int e = __builtin_amdgcn_frexp_exp(x);
  const bool low = m < 0x1.6a09e667f3bcdp-1;
  m = low ? 2.0 * m : m;
  e = low ? e - 1 : e;
  const double s = (m - 1.0) * mf_rcp(m + 1.0);
  const double z = s * s;
  double p = 2.0 / 21.0;
  p = fma(p, z, 2.0 / 19.0);
  p = fma(p, z, 2.0 / 17.0);
  p = fma(p, z, 2.0 / 15.0);
  p = fma(p, z, 2.0 / 13.0);
  p = fma(p, z, 2.0 / 11.0);
  p = fma(p, z, 2.0 / 9.0);
  p = fma(p, z, 2.0 / 7.0);
  p = fma(p, z, 2.0 / 5.0);
  p = fma(p, z, 2.0 / 3.0);
  const double lm = fma(s * z, p, 2.0 * s);
  const double de = (double)e;
  const double v = fma(de, 0x1.62e42fefa3800p-1, fma(de, 0x1.ef35793c76730p-45, lm));
  return (x > 0.0 && x < __builtin_inf()) ? v : (x == 0.0 ? -__builtin_inf() : (x > 0.0 ? x : __builtin_nan("")));
}
__device__ __forceinline__ double mf_exp2(double x) { return exp2(x); }
__device__ __forceinline__ double mf_log2(double x) { return log2(x); }
__device__ __forceinline__ double mf_sqrt(double x) { return sqrt(x); }
// sine and cosine: Cody-Waite reduction x = k pi/2 + r, pi/2 split into three doubles (the split
// is within 6e-50 of pi/2; x - k P1 is exact, being a multiple of 2^-53 below 1 in magnitude), so
// |r| <= pi/4 + rounding for |x| < 2^50, beyond which the reduction loses accuracy; sin r and cos r
// by their Taylor polynomials to degree 17 and 16 (truncation < 1e-18 relative at |r| = 0.8);
// k mod 4 selects the function and the sign.  Infinite and NaN arguments give NaN.  The two share
// every instruction up to the final selects, so a density with sin(a) and cos(a) pays for one.
// MF_SINCOS_BEGIN
__device__ __forceinline__ void mf_sincos_parts(double x, double& s, double& c, int& q) {
  const double k = rint(x * 0x1.45f306dc9c883p-1);
  double r = fma(-k, 0x1.921fb54442d18p+0, x);
  r = fma(-k, 0x1.1a62633145c07p-54, r);
  r = fma(-k, -0x1.f1976b7ed8fbcp-110, r);
  const double z = r * r;
  double ps = 1.0 / 355687428096000.0;  // 1/17!
  ps = fma(ps, z, -1.0 / 1307674368000.0);
  ps = fma(ps, z, 1.0 / 6227020800.0);
  ps = fma(ps, z, -1.0 / 39916800.0);
  ps = fma(ps, z, 1.0 / 362880.0);
  ps = fma(ps, z, -1.0 / 5040.0);
  ps = fma(ps, z, 1.0 / 120.0);
  ps = fma(ps, z, -1.0 / 6.0);
  s = fma(r * z, ps, r);
  double pc = 1.0 / 20922789888000.0;  // 1/16!
  pc = fma(pc, z, -1.0 / 87178291200.0);
  pc = fma(pc, z, 1.0 / 479001600.0);
  pc = fma(pc, z, -1.0 / 3628800.0);
  pc = fma(pc, z, 1.0 / 40320.0);
  pc = fma(pc, z, -1.0 / 720.0);
  pc = fma(pc, z, 1.0 / 24.0);
  c = fma(z * z, pc, fma(-0.5, z, 1.0));
  const double kq = k - 4.0 * floor(0.25 * k);  // k mod 4, exact
  q = kq == kq ? (int)kq : 0;
}
__device__ __forceinline__ double mf_sin(double x) {
  double s, c;
  int q;
  mf_sincos_parts(x, s, c, q);
  const double v = (q & 1) ? c : s;
  return (q & 2) ? -v : v;
}
__device__ __forceinline__ double mf_cos(double x) {
  double s, c;
  int q;
  mf_sincos_parts(x, s, c, q);
  const double v = (q & 1) ? s : c;
  return ((q + 1) & 2) ? -v : v;
}
// MF_SINCOS_END
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

// ------------------------------------------------------------------ event loads
// Raw buffer resource over `n` elements from `base` (gfx9 descriptor word 3: 32-bit data
// format, no swizzle, no stride).  Loads through it take a per-lane byte offset (VGPR) and a
// uniform byte offset (SGPR), so walking a chunk changes scalar registers only.  The range
// check is a safety net: the evaluation kernels load in-range events only.
typedef __amdgpu_buffer_rsrc_t mf_rsrc_t;
__device__ __forceinline__ mf_rsrc_t mf_make_rsrc(const MF_T* base, int n) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, n * (int)sizeof(MF_T), 0x00020000);
}
// event loads with the cache policy MF_LOAD_AUX (buffer instruction aux bits; default 2, the nt
// policy: every event is read once per launch)
#ifndef MF_LOAD_AUX
#define MF_LOAD_AUX 2
#endif
__device__ __forceinline__ MF_T mf_load_nt(mf_rsrc_t r, int voff, int soff) {
#if MF_FP32
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, MF_LOAD_AUX));
#else
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, MF_LOAD_AUX));
#endif
}

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

// The argument block as it lies in the kernarg segment (the evaluation kernels' only explicit
// argument, at offset 0): parameters passed by value are read from there with scalar loads.
__device__ __forceinline__ const MF_CONST char* mf_args_block() { return (const MF_CONST char*)__builtin_amdgcn_kernarg_segment_ptr(); }
__device__ __forceinline__ const MF_CONST MF_T* mf_params_of(const mf_eval_args& a) {
  return a.inline_params ? (const MF_CONST MF_T*)(mf_args_block() + __builtin_offsetof(mf_eval_args, pinline))
                         : (const MF_CONST MF_T*)a.params;
}

// Totals of segment s (its F::nout outputs) from the segment sums S of the F::nacc per-event
// terms, formed by the generated F::combine in FP64 with the segment's hoisted terms
template <class F>
__device__ __forceinline__ void mf_write_totals(const mf_eval_args& a, int s, const mf_segment& sg, const double* S) {
  double out[F::nout];
  const double* h = a.inline_params ? (const double*)(mf_args_block() + __builtin_offsetof(mf_eval_args, hinline))
                    : a.hoist64     ? a.hoist64 + (long long)s * MF_NHOIST
                                    : nullptr;
  const double* M = MF_NMOM > 0 && a.moments ? a.moments + (long long)s * (MF_NMOM > 0 ? MF_NMOM : 1) : nullptr;
  F::combine(S, h, (double)sg.n, M, out);
#pragma unroll
  for (int k = 0; k < F::nout; k++) a.results[sg.out_off + k] = out[k];
}

// ------------------------------------------------------------------ segment totals
// Totals of segment s from its chunk partials part[j][k] (j < nch chunks, k < NO outputs), by
// the NT threads of one work-group.  Outputs go in passes of KP (NO rounded up to a power of
// two, at most NT); thread t = r * KP + q takes output k0 + q and sums chunks r, r + R, r + 2R,
// ... (R = NT / KP) with compensated (two-sum) FP64 additions -- consecutive threads read
// consecutive addresses, and each thread holds one sum.  The R sums of an output are then
// combined in a fixed tree: the wave64 butterfly over the lanes of one output, and the waves in
// wave order through LDS.  Every addition's order depends only on the segment's chunk count
// and NT, so a segment's totals are bitwise the same whatever else shares the launch.
#define MF_FINAL_BLOCK MF_FINAL_THREADS
// threadIdx.x through a volatile register move: what the segment reduction derives from the
// thread index is then computed where it is used, not hoisted above the event loop of the
// evaluation kernel, where it would hold registers for the whole loop
__device__ __forceinline__ int mf_thread_index() {
  int t;
  asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((int)threadIdx.x));
  return t;
}
__host__ __device__ constexpr int mf_pow2_ceil(int v) { return v <= 1 ? 1 : 2 * mf_pow2_ceil((v + 1) / 2); }
template <class F, int NT, int UNR>
__device__ __forceinline__ void mf_segment_totals(const mf_eval_args& a, int s, const mf_segment& sg) {
  constexpr int NO = F::nacc;
  __shared__ double s_tot[NO];
  constexpr int KP = mf_pow2_ceil(NO) < NT ? mf_pow2_ceil(NO) : NT;
  constexpr int R = NT / KP;
  constexpr int NW = NT / 64;
  __shared__ double s_fin[NT];
  const int nch = sg.chunk_end - sg.chunk_begin;
  const int t = mf_thread_index(), q = t % KP, r = t / KP, lane = t & 63, wave = t >> 6;
  const double* __restrict__ part = a.partials + (long long)sg.chunk_begin * NO;
  for (int k0 = 0; k0 < NO; k0 += KP) {
    const int k = k0 + q;
    double acc = 0.0, err = 0.0;
    if (k < NO) {
      // batches of UNR partials: all loads issued before the additions, so a thread's ~nch / R
      // loads take one memory latency, not one each (a runtime trip count below the unroll
      // left them to a one-at-a-time remainder loop); past the end a batch holds zeros, which
      // the two-sum adds exactly, so the order and the bits are those of the plain loop
      for (int j0 = r; j0 < nch; j0 += UNR * R) {
        double b[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
          const int j = j0 + u * R;
          b[u] = j < nch ? part[(long long)j * NO + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
          double s2, e;
          mf_two_sum(acc, b[u], s2, e);
          acc = s2;
          err += e;
        }
      }
    }
    double v = acc + err;
    if constexpr (KP < 64) {
      // lanes of one output differ by multiples of KP: butterfly over those offsets
#pragma unroll
      for (int off = 32; off >= KP; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane < KP) s_fin[wave * KP + lane] = v;
      __syncthreads();
      if (t < KP && k < NO) {
        double tot = 0.0;
#pragma unroll
        for (int w = 0; w < NW; w++) tot += s_fin[w * KP + t];
        s_tot[k] = tot;
      }
    } else {
      // each lane of a wave is its own output: combine the R threads of output q in order
      s_fin[t] = v;
      __syncthreads();
      if (t < KP && k < NO) {
        double tot = 0.0;
#pragma unroll
        for (int rr = 0; rr < R; rr++) tot += s_fin[rr * KP + t];
        s_tot[k] = tot;
      }
    }
    __syncthreads();
  }
  if (t == 0) mf_write_totals<F>(a, s, sg, s_tot);
  __syncthreads();
}

// ------------------------------------------------------------------ fused evaluation
// Grid-stride over the chunk table (a persistent grid when gridDim < nchunks).  Per chunk:
// coalesced SoA loads, per-event outputs from the generated functor F, per-lane
// accumulation, wave64 butterfly + LDS reduction to one FP64 partial per output, written to
// partials[chunk][k].  A segment of one chunk gets its totals right there.  For longer
// segments, with a.seg_done each work-group publishes its partials (device-scope release
// fence) and counts its chunk on the segment's counter; the one that counts last (acquire
// fence) reduces the segment and sets the counter back to zero for the next launch.  Nothing
// waits on another work-group.  Without a.seg_done the final kernel below does the reduction.
// The chunk and segment tables and the kernel parameters are read through constant-address-
// space pointers (this kernel never writes them): uniform loads then go through the scalar
// cache into SGPRs, which the compiler cannot otherwise prove safe once the loop holds the
// counter atomics and fences.
template <class F, int U, bool FUSED, bool PF>
__device__ __forceinline__ void mf_eval_impl(const mf_eval_args& a) {
  constexpr int NO = F::nacc;
  __shared__ double s_red[MF_NWAVE][NO];
  __shared__ double s_tot[NO];
  const MF_T* __restrict__ data = (const MF_T*)a.data;
  const MF_CONST mf_chunk* chunks = (const MF_CONST mf_chunk*)a.chunks;
  const MF_CONST mf_segment* segs = (const MF_CONST mf_segment*)a.segs;
  const MF_CONST MF_T* params = mf_params_of(a);
  unsigned int* seg_done = a.seg_done;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
    const mf_chunk ch = chunks[c];
    const mf_segment sg = segs[ch.seg];
    if (sg.mode == 0) continue;
    const MF_T* __restrict__ P = (const MF_T*)(params + (a.inline_params ? 0 : sg.param_off));
    const MF_T* __restrict__ col = data + sg.data_off;
    mf_acc<NO> acc;
    acc.zero();
    // one buffer resource per column based at the chunk's first event: a load's address is the
    // lane's byte offset (one VGPR for the whole loop) plus a uniform offset in an SGPR, so the
    // event loop spends no vector instructions on addresses; the trip count of the unrolled
    // part is uniform (whole groups of U * MF_BLOCK events), the rest goes event by event
    const int len = ch.end - ch.begin;
    mf_rsrc_t rs[MF_NCOL];
#pragma unroll
    for (int d = 0; d < MF_NCOL; d++) rs[d] = mf_make_rsrc(col + d * sg.stride + ch.begin, len);
    const int lane_off = (int)threadIdx.x * (int)sizeof(MF_T);
    // U events per thread per step: their loads are issued together, then computed.  With
    // MF_PREFETCH the loads of the next group are issued before the current group is computed
    // (U * MF_NCOL more registers), so a wave waits on memory once per chunk, not per group
    const int ngroups = len / (U * MF_BLOCK);
    auto load_group = [&](MF_T (&dst)[U][MF_NCOL], int g) {
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int d = 0; d < MF_NCOL; d++) dst[u][d] = mf_load_nt(rs[d], lane_off, (g * U + u) * MF_BLOCK * (int)sizeof(MF_T));
    };
    auto compute_group = [&](const MF_T (&x)[U][MF_NCOL]) {
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
    };
    if constexpr (PF && MF_PREFETCH && U * MF_NCOL <= MF_PREFETCH_MAX) {
      // two register sets in turn: group g + 1 loads into one while group g is computed from
      // the other (no copies between them).  Every load sits on an unconditional path to the
      // group computed next, so the compiler's wait counts stay exact (a load issued under a
      // condition made it wait for all loads in flight, the prefetched ones included)
      MF_T xa[U][MF_NCOL], xb[U][MF_NCOL];
      if (ngroups > 0) {
        load_group(xa, 0);
        int gi = 0;
        for (; gi + 2 < ngroups; gi += 2) {
          load_group(xb, gi + 1);
          compute_group(xa);
          load_group(xa, gi + 2);
          compute_group(xb);
        }
        if (gi + 1 < ngroups) {
          load_group(xb, gi + 1);
          compute_group(xa);
          compute_group(xb);
        } else {
          compute_group(xa);
        }
      }
    } else {
      for (int gi = 0; gi < ngroups; gi++) {
        MF_T x[U][MF_NCOL];
        load_group(x, gi);
        compute_group(x);
      }
    }
    // the rest of the chunk (fewer than U * MF_BLOCK events) as one more group: its loads are
    // issued together, so the tail waits on memory once, not once per event.  A lane whose event
    // lies past the chunk's end loads the chunk's last event instead (every load stays in range)
    // and its outputs are replaced by zeros; of the U slots only the first nu (uniform) hold
    // events of the chunk and are computed
    const int rest = len - ngroups * U * MF_BLOCK;
    if constexpr (U == 1) {
      // one event per lane at most: the plain loop (no selects, fewest registers)
      for (int i = ngroups * MF_BLOCK + (int)threadIdx.x; i < len; i += MF_BLOCK) {
        MF_T x[MF_NCOL];
#pragma unroll
        for (int d = 0; d < MF_NCOL; d++) x[d] = mf_load_nt(rs[d], i * (int)sizeof(MF_T), 0);
        MF_T o[NO];
        F::eval(x, P, o);
        acc.add(o);
      }
    } else if (rest > 0) {
      const int first = ngroups * U * MF_BLOCK + (int)threadIdx.x;
      const int nu = (rest + MF_BLOCK - 1) / MF_BLOCK;
      MF_T x[U][MF_NCOL];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int i = min(first + u * MF_BLOCK, len - 1);
#pragma unroll
        for (int d = 0; d < MF_NCOL; d++) x[u][d] = mf_load_nt(rs[d], i * (int)sizeof(MF_T), 0);
      }
#if MF_FP32 && MF_ACC == 2
      MF_T g[NO];
#pragma unroll
      for (int k = 0; k < NO; k++) g[k] = 0.0f;
#endif
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (u < nu) {
          const bool in = first + u * MF_BLOCK < len;
          MF_T o[NO];
          F::eval(x[u], P, o);
#if MF_FP32 && MF_ACC == 2
#pragma unroll
          for (int k = 0; k < NO; k++) g[k] += in ? o[k] : (MF_T)0;
#else
#pragma unroll
          for (int k = 0; k < NO; k++) o[k] = in ? o[k] : (MF_T)0;
          acc.add(o);
#endif
        }
      }
#if MF_FP32 && MF_ACC == 2
      acc.add(g);
#endif
    }
    double v[NO];
    acc.get(v);
    if constexpr (NO * MF_BLOCK * 8 <= 32768) {
      // through LDS: every thread stores its NO sums (transposed), then 32 lanes per output
      // add a column of MF_BLOCK / 32 values each and a 32-lane butterfly joins them: about
      // MF_BLOCK / 32 + 5 additions per output in 32 threads instead of 6 in every thread
      __shared__ double s_all[NO * MF_BLOCK];
      // the thread index through a register move the compiler cannot hoist: computed here, the
      // LDS addresses do not sit in registers across the event loop
      const int ti = mf_thread_index();
#pragma unroll
      for (int k = 0; k < NO; k++) s_all[k * MF_BLOCK + ti] = v[k];
      __syncthreads();
      // 32-lane group ti / 32 takes outputs ti / 32, ti / 32 + MF_BLOCK / 32, ... (both halves of a
      // wave run the loop body for their own output; the butterfly stays inside a half)
      const int j = ti & 31;
      for (int k = ti >> 5; k < NO; k += MF_BLOCK / 32) {
        double t = 0.0;
#pragma unroll 4
        for (int g = 0; g < MF_BLOCK / 32; g++) t += s_all[k * MF_BLOCK + g * 32 + j];
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        if (j == 0) {
          a.partials[(long long)c * NO + k] = t;
          s_tot[k] = t;
        }
      }
      __syncthreads();
    } else {
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
        s_tot[threadIdx.x] = t;
      }
      __syncthreads();
    }
    // a segment of one chunk is complete here: its totals need no second pass
    if (threadIdx.x == 0 && sg.chunk_end - sg.chunk_begin == 1) mf_write_totals<F>(a, ch.seg, sg, s_tot);
    const int nch = sg.chunk_end - sg.chunk_begin;
    if (FUSED && nch > 1 && seg_done) {
      __shared__ int s_last;
      if (threadIdx.x == 0) {
        __threadfence();  // release: this work-group's partials before its count
        const unsigned int prev = atomicAdd(seg_done + ch.seg, 1u);
        __threadfence();  // acquire: every other chunk's partials after the last count
        s_last = prev == (unsigned int)(nch - 1);
      }
      __syncthreads();
      if (s_last) {
        __threadfence();
        mf_segment_totals<F, MF_BLOCK, 8>(a, ch.seg, sg);
        if (threadIdx.x == 0) seg_done[ch.seg] = 0u;
      }
    }
  }
  // a launch of one chunk (one work-group) published its totals in the loop: now the flag the
  // host polls, as the final kernel does.  Kept after the loop: a fence inside it cost the loop
  // its scalar loads of the tables (FP32 mass gradient: 62 -> 74 VGPRs, +5 % time)
  if (a.done_flag && a.nchunks == 1 && threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(a.done_flag, a.done_value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Segment totals of segments with more than one chunk as a kernel of its own (used when
// mf_eval_args.seg_done is NULL): one work-group of MF_FINAL_BLOCK threads per segment.
template <class F>
__device__ __forceinline__ void mf_final_impl(const mf_eval_args& a) {
  const int s = blockIdx.x;
  if (s < a.nseg) {
    const mf_segment sg = a.segs[s];
    if (sg.mode != 0 && sg.chunk_end - sg.chunk_begin > 1) mf_segment_totals<F, MF_FINAL_BLOCK, 16>(a, s, sg);
  }
  // one segment (checked by the host): this work-group wrote every total; publish them to the
  // host, then the flag it polls
  if (a.done_flag) {
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence_system();
      __hip_atomic_store(a.done_flag, a.done_value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// per-event outputs written to memory (the reference's kahan_on_accelerator == false path
// and its per-event precompute kernels): store[k * store_stride + data_off + i]
template <class F>
__device__ __forceinline__ void mf_store_impl(const mf_eval_args& a) {
  constexpr int NO = F::nacc;
  const MF_T* __restrict__ data = (const MF_T*)a.data;
  MF_T* __restrict__ out = (MF_T*)a.store;
  for (int c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
    const mf_chunk ch = a.chunks[c];
    const mf_segment sg = a.segs[ch.seg];
    if (sg.mode == 0) continue;
    const MF_T* __restrict__ P = (const MF_T*)a.params + sg.param_off;
    const MF_T* __restrict__ col = data + sg.data_off;
    for (int i = ch.begin + (int)threadIdx.x; i < ch.end; i += MF_BLOCK) {
      MF_T x[MF_NCOL];
#pragma unroll
      for (int d = 0; d < MF_NCOL; d++) x[d] = col[d * sg.stride + i];
      MF_T o[NO];
      F::eval(x, P, o);
#pragma unroll
      for (int k = 0; k < NO; k++) out[k * a.store_stride + sg.data_off + i] = o[k];
    }
  }
}

// ------------------------------------------------------------------ diagnostics
// For sums that came back non-finite: per active segment, the events whose per-event terms are
// not all finite, the smallest such event index, and the events with an observable outside the
// density's range or not a number.  a.store points at unsigned int [nseg][3] = {count, first,
// outside}, set by the host to {0, 0xffffffff, 0}.  One ballot per wave and event step, one
// atomic per wave and counter.  Finiteness is a test of the exponent bits, which no
// floating-point mode folds away.
__device__ __forceinline__ bool mf_finite(float v) { return (__float_as_uint(v) & 0x7f800000u) != 0x7f800000u; }
__device__ __forceinline__ bool mf_finite(double v) {
  return ((unsigned long long)__double_as_longlong(v) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull;
}
template <class F, class FM>
__device__ __forceinline__ void mf_check_impl(const mf_eval_args& a) {
  constexpr int NO = F::nacc, NM = FM::nacc;
  const MF_T* __restrict__ data = (const MF_T*)a.data;
  unsigned int* __restrict__ out = (unsigned int*)a.store;
  const MF_CONST MF_T* params = mf_params_of(a);
#ifdef MF_OBS_LO
  const MF_T lo[MF_NOBS] = MF_OBS_LO, hi[MF_NOBS] = MF_OBS_HI;  // stored coordinates
#endif
  for (int c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
    const mf_chunk ch = a.chunks[c];
    const mf_segment sg = a.segs[ch.seg];
    if (sg.mode == 0) continue;
    const MF_T* __restrict__ P = (const MF_T*)(params + (a.inline_params ? 0 : sg.param_off));
    const MF_T* __restrict__ col = data + sg.data_off;
    // the trip count is the same for every lane, so every lane reaches the ballots
    for (int i0 = ch.begin; i0 < ch.end; i0 += MF_BLOCK) {
      const int i = i0 + (int)threadIdx.x;
      bool bad = false, outside = false;
      if (i < ch.end) {
        MF_T x[MF_NCOL];
#pragma unroll
        for (int d = 0; d < MF_NCOL; d++) x[d] = col[d * sg.stride + i];
#ifdef MF_OBS_LO
#pragma unroll
        for (int d = 0; d < MF_NOBS; d++) outside = outside || !(x[d] >= lo[d] && x[d] <= hi[d]);
#endif
        MF_T o[NO];
        F::eval(x, P, o);
#pragma unroll
        for (int k = 0; k < NO; k++) bad = bad || !mf_finite(o[k]);
        // the observable-only terms summed through the data moments belong to the event too
        if constexpr (NM > 0) {
          MF_T om[NM > 0 ? NM : 1];
          FM::eval(x, P, om);
#pragma unroll
          for (int k = 0; k < NM; k++) bad = bad || !mf_finite(om[k]);
        }
      }
      const unsigned long long mb = __ballot(bad), mo = __ballot(outside);
      if ((threadIdx.x & 63) == 0) {
        unsigned int* o3 = out + 3 * ch.seg;
        if (mb) {
          atomicAdd(o3, (unsigned int)__popcll(mb));
          atomicMin(o3 + 1, (unsigned int)(i + __ffsll((unsigned long long)mb) - 1));
        }
        if (mo) atomicAdd(o3 + 2, (unsigned int)__popcll(mo));
      }
    }
  }
}
// a module without data moments
struct mf_no_moments {
  static constexpr int nacc = 0;
  __device__ __forceinline__ static void eval(const MF_T*, const MF_T*, MF_T*) {}
};
#define MF_CHECK_KERNEL(NAME, FUNCTOR) \
  extern "C" __global__ void __launch_bounds__(MF_BLOCK) NAME(mf_eval_args a) { mf_check_impl<FUNCTOR, mf_no_moments>(a); }
#define MF_CHECK_KERNEL_MOM(NAME, FUNCTOR, MOMENTS) \
  extern "C" __global__ void __launch_bounds__(MF_BLOCK) NAME(mf_eval_args a) { mf_check_impl<FUNCTOR, MOMENTS>(a); }

// ------------------------------------------------------------------ event columns
// Event-only terms of the density (generated mf_event_terms: FP64 from the stored observables)
// into columns MF_NOBS..MF_NCOL-1 of every segment whose mode is set, rounded once to MF_T.
// Runs once per upload or generation; the evaluation kernels then read the terms instead of
// recomputing them at every call (the reference's optimize_dimensions, fitter.hh:339-375).
#if MF_NCOL > MF_NOBS
__device__ __forceinline__ void mf_event_terms(const double* x, double* buf);
__device__ __forceinline__ void mf_precompute_impl(const mf_eval_args& a) {
  MF_T* __restrict__ data = (MF_T*)a.data;  // this kernel writes the event columns of the data buffer
  for (int c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
    const mf_chunk ch = a.chunks[c];
    const mf_segment sg = a.segs[ch.seg];
    if (sg.mode == 0) continue;
    MF_T* __restrict__ col = data + sg.data_off;
    for (int i = ch.begin + (int)threadIdx.x; i < ch.end; i += MF_BLOCK) {
      double xd[MF_NOBS];
#pragma unroll
      for (int d = 0; d < MF_NOBS; d++) xd[d] = (double)col[d * sg.stride + i];
      double o[MF_NCOL - MF_NOBS];
      mf_event_terms(xd, o);
#pragma unroll
      for (int k = 0; k < MF_NCOL - MF_NOBS; k++) col[(MF_NOBS + k) * sg.stride + i] = (MF_T)o[k];
    }
  }
}
#define MF_PRECOMPUTE_KERNEL \
  extern "C" __global__ void __launch_bounds__(MF_BLOCK) mf_precompute(mf_eval_args a) { mf_precompute_impl(a); }
#endif

// NAME##_nacc: the per-chunk partials the kernel writes (F::nacc doubles per chunk), read by the
// host at module load to bound the launches that bring no partials buffer of their own
#define MF_EVAL_KERNEL(NAME, FUNCTOR, UNROLL, FUSED, PREFETCH)                                                \
  extern "C" __global__ void __launch_bounds__(MF_BLOCK) NAME(mf_eval_args a) {                                  \
    mf_eval_impl<FUNCTOR, UNROLL, FUSED, PREFETCH>(a);                                                           \
  }                                                                                                               \
  extern "C" __global__ void __launch_bounds__(MF_FINAL_BLOCK) NAME##_final(mf_eval_args a) { mf_final_impl<FUNCTOR>(a); } \
  extern "C" __device__ __attribute__((used)) int NAME##_nacc = FUNCTOR::nacc;
#define MF_STORE_KERNEL(NAME, FUNCTOR) \
  extern "C" __global__ void __launch_bounds__(MF_BLOCK) NAME(mf_eval_args a) { mf_store_impl<FUNCTOR>(a); }

// ------------------------------------------------------------------ parameter preparation
// Hoisted parameter-only terms in FP64 from a segment's floating parameters, then [theta,
// hoisted] rounded once to MF_T into the kernel parameter block.  Term-parallel: work-group
// (j, b) of 64 lanes handles segments b * 64 + lane and hoisted terms j, j + MF_PREP_SPLIT, ...
// (the term index is uniform over the work-group: no divergence), so a segment's terms --
// erf, exp and reciprocals with long FP64 dependency chains -- are computed side by side
// rather than one after another.  Work-group j == 0 also stores theta and stages the segment's
// table entry (seg_src -> seg_dst) for the evaluation kernels.  theta may live in mapped host
// memory (read in full before anything is stored); with hoist64 the hoisted terms are also
// kept in FP64 (output scales, mf_eval_args).
// hoisted term k in FP64 from the floating parameters (generated, after this runtime)
__device__ double mf_hoist_term(int k, const double* th);

__device__ __forceinline__ void mf_prep_part(const double* th_src, int s, int j, MF_T* params, double* hoist64, const mf_segment* seg_src,
                                             mf_segment* seg_dst) {
  double th[MF_NTHETA > 0 ? MF_NTHETA : 1];
  for (int k = 0; k < MF_NTHETA; k++) th[k] = th_src[k];
  MF_T* out = params + (long long)s * MF_NPARAM;
  if (j == 0) {
    if (seg_src) seg_dst[s] = seg_src[s];
    for (int k = 0; k < MF_NTHETA; k++) out[k] = (MF_T)th[k];
  }
  for (int k = j; k < MF_NHOIST; k += MF_PREP_SPLIT) {
    const double h = mf_hoist_term(k, th);
    out[MF_NTHETA + k] = (MF_T)h;
    if (hoist64) hoist64[(long long)s * MF_NHOIST + k] = h;
  }
}

#define MF_PREP_KERNEL                                                                                             \
  extern "C" __global__ void mf_prep(const double* theta, MF_T* params, double* hoist64, const int* seg_list, int nseg, \
                                     const mf_segment* seg_src, mf_segment* seg_dst) {                             \
    const int i = blockIdx.y * blockDim.x + threadIdx.x;                                                           \
    if (i >= nseg) return;                                                                                         \
    const int s = seg_list ? seg_list[i] : i;                                                                      \
    mf_prep_part(theta + (long long)s * MF_NTHETA, s, blockIdx.x, params, hoist64, seg_src, seg_dst);              \
  }                                                                                                                \
  extern "C" __global__ void mf_prep_inline(mf_prep_inline_args g) {                                               \
    const int s = blockIdx.y * blockDim.x + threadIdx.x;                                                           \
    if (s >= g.nseg) return;                                                                                       \
    mf_prep_part(g.theta + s * MF_NTHETA, s, blockIdx.x, (MF_T*)g.params, g.hoist64, g.seg_src, g.seg_dst);        \
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
  const bool active = k < g.nstreams;  // every lane stays to the end: the wave reduces the proposal counts
  int seg = -1;
  unsigned long long props = 0;
  if (active) {
    const mf_gen_stream st = g.streams[k];
    seg = st.seg;
    const mf_segment sg = g.segs[st.seg];
    unsigned int s[4];
#pragma unroll
    for (int j = 0; j < 4; j++) s[j] = g.states[4 * k + j];
    const MF_T* __restrict__ P = (const MF_T*)g.params + sg.param_off;
    const MF_T mp = (MF_T)g.maxprob[st.seg];
    MF_T* __restrict__ out = (MF_T*)g.data + sg.data_off;
    // One proposal per iteration, accepted events stored in order: a lane goes on to its next
    // event without waiting for the other lanes of its wave to accept their current one (with a
    // per-event inner loop the wave ran, per event, the most tries of any of its 64 lanes).  The
    // stream's draws and accepted events are the same as one event after another.
    int i = st.begin;
    unsigned int tries = 0;  // bounded: a density that is NaN or far above maxprob's scale must not hang the stream
    while (i < st.end) {
      MF_T x[MF_NOBS];
      props++;
      MF_GEN_PROPOSE(x, s);
      MF_T lp[1];
      L::eval(x, P, lp);
      const MF_T p = mf_exp(lp[0]);
      const MF_T u = mf_uniform(s);
      if (mp * u < p) {
#pragma unroll
        for (int d = 0; d < MF_NOBS; d++) out[d * sg.stride + i] = x[d];
        i++;
        tries = 0;
      } else if (++tries >= MF_GEN_MAX_TRIES) {
        for (int j = i; j < st.end; j++)
          for (int d = 0; d < MF_NOBS; d++) out[d * sg.stride + j] = (MF_T)(0.0 / 0.0);
        break;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) g.states[4 * k + j] = s[j];
  }
  // proposal counts: one atomic per wave when its streams belong to one segment (the common case),
  // else one per stream; device-scope atomics on a few counters serialise (~10 ns each), and one per
  // stream had cost milliseconds at 10^5-10^6 streams
  if (g.proposals) {
    const int lane = threadIdx.x & 63;
    if (__ballot(active)) {  // lane 0 holds the wave's lowest stream, so it is active
      const int s0 = __shfl(seg, 0);
      if (__ballot(active && seg != s0) == 0) {
        unsigned long long v = props;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) atomicAdd(g.proposals + s0, v);
      } else if (active) {
        atomicAdd(g.proposals + seg, props);
      }
    }
  }
}

#define MF_GENERATE_KERNEL(FUNCTOR) \
  extern "C" __global__ void mf_generate(mf_gen_args g) { mf_generate_impl<FUNCTOR>(g); }

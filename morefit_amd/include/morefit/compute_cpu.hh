/**
 * @file compute_cpu.hh
 * Host CPU backend with the semantics of the reference LLVM backend (compute_llvm.hh:212-941):
 * the same generated functor text as the GPU modules, compiled here by the host C++ compiler
 * into a shared object and loaded with dlopen (the reference JIT-compiles C with Clang/ORC);
 * contiguous per-thread event slices with the last thread taking the remainder, an in-loop
 * Kahan sum per output in evalT, and a Kahan sum of the thread partials (compute_llvm.hh:
 * 556-709, 785-850).  Generation (MakeGenerateKernel) runs one xoshiro128++ stream per thread
 * over the same contiguous slices, accept-reject as the reference's generate kernel
 * (compute_llvm.hh:362-525, 751-784).  It runs the whole graph -> code -> fit pipeline where
 * no GPU exists.
 */
#ifndef MOREFIT_COMPUTE_CPU_HH
#define MOREFIT_COMPUTE_CPU_HH

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "codegen.hh"
#include "compute.hh"
#include "utils.hh"

namespace morefit {

class CPUBackend : public ComputeBackend {
 public:
  explicit CPUBackend(compute_options* opts) : ComputeBackend(opts) {}
  int required_padding() const override { return 0; }
};

inline std::string cpu_cache_dir() {
  const char* env = std::getenv("MOREFIT_KCACHE");
  if (env && *env) return env;
  const char* tmp = std::getenv("TMPDIR");
  return std::string(tmp && *tmp ? tmp : "/tmp") + "/morefit_cpu_kcache";
}

inline uint64_t cpu_fnv1a(const std::string& s) {
  uint64_t h = 1469598103934665603ULL;
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

// FP64 SIMD code of the host backend: 4 events per AVX2 vector, the transcendentals from
// glibc's vector math library (libmvec, its AVX2 variants _ZGVdN4v_*), the interleaved
// 4-lane Kahan sums of the reference's vectorised LLVM kernels (compute_llvm.hh:556-568,
// 656-709)
inline std::string cpu_vector_prelude() {
  std::ostringstream s;
  s << "#include <immintrin.h>\n#include <cstring>\n"
    << "typedef double mf_v4d __attribute__((vector_size(32)));\n"
    << "static inline mf_v4d mf_bc(mf_v4d v) { return v; }\n"
    << "static inline mf_v4d mf_bc(double v) { return mf_v4d{v, v, v, v}; }\n";
  const char* fns[] = {"exp", "log", "exp2", "log2", "sin", "cos", "tan", "asin", "acos", "atan", "erf"};
  for (auto* f : fns)
    s << "extern \"C\" __m256d _ZGVdN4v_" << f << "(__m256d);\n"
      << "static inline mf_v4d mf_" << f << "(mf_v4d x) { return (mf_v4d)_ZGVdN4v_" << f << "((__m256d)x); }\n";
  s << "extern \"C\" __m256d _ZGVdN4vv_pow(__m256d, __m256d);\n"
    << "static inline mf_v4d mf_pow(mf_v4d x, mf_v4d y) { return (mf_v4d)_ZGVdN4vv_pow((__m256d)x, (__m256d)y); }\n"
    << "static inline mf_v4d mf_pow(mf_v4d x, double y) { return mf_pow(x, mf_bc(y)); }\n"
    << "static inline mf_v4d mf_pow(double x, mf_v4d y) { return mf_pow(mf_bc(x), y); }\n"
    << "static inline mf_v4d mf_sqrt(mf_v4d x) { return (mf_v4d)_mm256_sqrt_pd((__m256d)x); }\n"
    << "static inline mf_v4d mf_rcp(mf_v4d x) { return 1.0 / x; }\n";
  return s.str();
}

// host translation unit around generated functors; `vector_functor` (FP64 only): the same
// outputs as a SIMD functor mf_v_block, which the sum then runs 4 events at a time
inline std::string cpu_source(const ModuleSource& ms, const std::string& kernel_fn, const std::string& vector_functor = "") {
  std::ostringstream s;
  s << "#include <cmath>\n#include <cstddef>\n"
    << "#define __device__\n#define __forceinline__ inline\n"
    << "typedef " << (ms.fp32 ? "float" : "double") << " MF_T;\n";
  const char* fns[] = {"exp", "log", "exp2", "log2", "sqrt", "sin", "cos", "tan", "asin", "acos", "atan", "erf"};
  for (auto* f : fns) {
    s << "static inline MF_T mf_" << f << "(MF_T x) { return std::" << f << "(x); }\n";
    s << "static inline double mfd_" << f << "(double x) { return std::" << f << "(x); }\n";
  }
  s << "static inline MF_T mf_rcp(MF_T x) { return (MF_T)1 / x; }\n"
    << "static inline double mfd_rcp(double x) { return 1.0 / x; }\n"
    << "static inline MF_T mf_pow(MF_T x, MF_T y) { return std::pow(x, y); }\n"
    << "static inline double mfd_pow(double x, double y) { return std::pow(x, y); }\n"
    << ms.functors << "\n";
  if (!vector_functor.empty()) {
    s << cpu_vector_prelude() << vector_functor << "\n";
    // 4 interleaved lane sums with in-loop Kahan per output; events of the tail go to the lane
    // of their position in the slice; the lanes' sums then added with Kahan (the lanes'
    // carries dropped, as in the reference's final pass)
    s << "extern \"C\" void " << kernel_fn << "_sum(const MF_T* data, long long stride, long long from, long long to, const MF_T* p, double* out) {\n"
      << "  const int NO = mf_f_block::nout;\n  mf_v4d sum[NO], c[NO];\n"
      << "  for (int k = 0; k < NO; k++) { sum[k] = mf_bc(0.0); c[k] = mf_bc(0.0); }\n"
      << "  long long i = from;\n"
      << "  for (; i + 4 <= to; i += 4) {\n    mf_v4d x[" << ms.nobs << "];\n"
      << "    for (int d = 0; d < " << ms.nobs << "; d++) std::memcpy(&x[d], data + d * stride + i, sizeof(mf_v4d));\n"
      << "    mf_v4d o[NO];\n    mf_v_block::eval(x, p, o);\n"
      << "    for (int k = 0; k < NO; k++) { const mf_v4d y = o[k] - c[k]; const mf_v4d t = sum[k] + y; c[k] = (t - sum[k]) - y; sum[k] = t; }\n"
      << "  }\n"
      << "  double ls[NO][4], lc[NO][4];\n"
      << "  for (int k = 0; k < NO; k++) for (int j = 0; j < 4; j++) { ls[k][j] = sum[k][j]; lc[k][j] = c[k][j]; }\n"
      << "  for (; i < to; i++) {\n    const int j = (int)((i - from) & 3);\n    MF_T x[" << ms.nobs << "];\n"
      << "    for (int d = 0; d < " << ms.nobs << "; d++) x[d] = data[d * stride + i];\n"
      << "    MF_T o[NO];\n    mf_f_block::eval(x, p, o);\n"
      << "    for (int k = 0; k < NO; k++) { double y = (double)o[k] - lc[k][j]; volatile double t = ls[k][j] + y; lc[k][j] = (t - ls[k][j]) - y; ls[k][j] = t; }\n"
      << "  }\n"
      << "  for (int k = 0; k < NO; k++) {\n    double s = 0.0, cc = 0.0;\n"
      << "    for (int j = 0; j < 4; j++) { double y = ls[k][j] - cc; volatile double t = s + y; cc = (t - s) - y; s = t; }\n"
      << "    out[k] = s;\n  }\n}\n";
    s << "extern \"C\" void " << kernel_fn << "_store(const MF_T* data, long long stride, long long from, long long to, const MF_T* p, MF_T* out, long long ostride) {\n"
      << "  const int NO = mf_f_block::nout;\n"
      << "  for (long long i = from; i < to; i++) {\n    MF_T x[" << ms.nobs << "];\n"
      << "    for (int d = 0; d < " << ms.nobs << "; d++) x[d] = data[d * stride + i];\n"
      << "    MF_T o[NO];\n    mf_f_block::eval(x, p, o);\n    for (int k = 0; k < NO; k++) out[k * ostride + i] = o[k];\n  }\n}\n";
    return s.str();
  }
  // sums: in-loop Kahan per output in double (evalT), events [from, to)
  s << "extern \"C\" void " << kernel_fn << "_sum(const MF_T* data, long long stride, long long from, long long to, const MF_T* p, double* out) {\n"
    << "  const int NO = mf_f_block::nout;\n  double sum[NO], c[NO];\n  for (int k = 0; k < NO; k++) { sum[k] = 0.0; c[k] = 0.0; }\n"
    << "  for (long long i = from; i < to; i++) {\n    MF_T x[" << ms.nobs << "];\n"
    << "    for (int d = 0; d < " << ms.nobs << "; d++) x[d] = data[d * stride + i];\n"
    << "    MF_T o[NO];\n    mf_f_block::eval(x, p, o);\n"
    << "    for (int k = 0; k < NO; k++) { double y = (double)o[k] - c[k]; volatile double t = sum[k] + y; c[k] = (t - sum[k]) - y; sum[k] = t; }\n"
    << "  }\n  for (int k = 0; k < NO; k++) out[k] = sum[k];\n}\n";
  s << "extern \"C\" void " << kernel_fn << "_store(const MF_T* data, long long stride, long long from, long long to, const MF_T* p, MF_T* out, long long ostride) {\n"
    << "  const int NO = mf_f_block::nout;\n"
    << "  for (long long i = from; i < to; i++) {\n    MF_T x[" << ms.nobs << "];\n"
    << "    for (int d = 0; d < " << ms.nobs << "; d++) x[d] = data[d * stride + i];\n"
    << "    MF_T o[NO];\n    mf_f_block::eval(x, p, o);\n    for (int k = 0; k < NO; k++) out[k * ostride + i] = o[k];\n  }\n}\n";
  return s.str();
}

// host translation unit of a generate kernel: the density p (one functor output) and
// accept-reject over events [from, to) of one xoshiro128++ stream (the reference's generate
// kernel, compute_llvm.hh:380-500): every observable proposed uniformly in its range, accepted
// when maxprob * u < p; after 2^20 rejections in a row (a density that is NaN, or far above
// the envelope) the remaining events of the slice are NaN
inline std::string cpu_generate_source(const ModuleSource& ms) {
  std::ostringstream s;
  s << cpu_source(ms, "mf_cpu_unused")
    << "#include <cstdint>\n"
    << "static inline MF_T mf_xoshiro_u(uint32_t* s) {\n"
    << "  const uint32_t r = ((s[0] + s[3]) << 7 | (s[0] + s[3]) >> 25) + s[0];\n"
    << "  const uint32_t t = s[1] << 9;\n"
    << "  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = (s[3] << 11) | (s[3] >> 21);\n"
    << "  return (MF_T)(r * 2.3283064365386963e-10);\n}\n"
    << "extern \"C\" void mf_cpu_gen(long long stride, long long from, long long to, uint32_t* st, MF_T* out, double maxprob,\n"
    << "                            const double* lo, const double* width) {\n"
    << "  const MF_T mp = (MF_T)maxprob;\n"
    << "  uint32_t s[4] = {st[0], st[1], st[2], st[3]};\n"
    << "  const MF_T* p = nullptr;\n"
    << "  for (long long i = from; i < to; i++) {\n"
    << "    unsigned long tries = 0;\n"
    << "    for (;;) {\n"
    << "      MF_T x[" << ms.nobs << "];\n"
    << "      for (int d = 0; d < " << ms.nobs << "; d++) x[d] = (MF_T)(lo[d] + width[d] * (double)mf_xoshiro_u(s));\n"
    << "      MF_T o[1];\n      mf_f_block::eval(x, p, o);\n"
    << "      if (mp * mf_xoshiro_u(s) < o[0]) {\n"
    << "        for (int d = 0; d < " << ms.nobs << "; d++) out[d * stride + i] = x[d];\n"
    << "        break;\n      }\n"
    << "      if (++tries >= (1ul << 20)) {\n"
    << "        for (long long j = i; j < to; j++) for (int d = 0; d < " << ms.nobs << "; d++) out[d * stride + j] = (MF_T)NAN;\n"
    << "        i = to;\n        break;\n      }\n"
    << "    }\n  }\n"
    << "  for (int j = 0; j < 4; j++) st[j] = s[j];\n}\n";
  return s.str();
}

// compiles (or reuses) a shared object for `src`; returns the dlopen handle
inline void* cpu_compile(const std::string& src, const compute_options& o) {
  // one compilation at a time per process (the temporary files are named by process)
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  std::string key = src + "\n" + o.cpu_compiler + " " + o.cpu_flags;
  char name[64];
  std::snprintf(name, sizeof(name), "/cpu_%016llx", (unsigned long long)cpu_fnv1a(key));
  std::string dir = cpu_cache_dir();
  ::mkdir(dir.c_str(), 0755);
  std::string base = dir + name, so = base + ".so";
  if (::access(so.c_str(), R_OK) != 0) {
    std::string cc = base + "." + std::to_string(::getpid()) + ".cc";
    {
      std::ofstream f(cc);
      f << src;
    }
    std::string tmpso = base + "." + std::to_string(::getpid()) + ".so";
    std::string cmd = o.cpu_compiler + " " + o.cpu_flags + " -std=c++17 -fPIC -shared " + cc + " -o " + tmpso + " -lmvec -lm 2>&1";
    FILE* pp = ::popen(cmd.c_str(), "r");
    std::string log;
    char buf[512];
    while (pp && std::fgets(buf, sizeof(buf), pp)) log += buf;
    int rc = pp ? ::pclose(pp) : -1;
    ::unlink(cc.c_str());
    if (rc != 0) throw std::runtime_error("CPU backend: compilation failed:\n" + log);
    ::rename(tmpso.c_str(), so.c_str());
  }
  void* h = ::dlopen(so.c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!h) throw std::runtime_error(std::string("CPU backend: dlopen failed: ") + ::dlerror());
  return h;
}

template <typename kernelT, typename evalT, typename seedT = uint32_t>
class CPUBlock : public ComputeBlock<kernelT, evalT, seedT> {
  typedef void (*sum_fn)(const kernelT*, long long, long long, long long, const kernelT*, double*);
  typedef void (*store_fn)(const kernelT*, long long, long long, long long, const kernelT*, kernelT*, long long);
  typedef void (*gen_fn)(long long, long long, long long, uint32_t*, kernelT*, double, const double*, const double*);
  struct State {
    CPUBackend* be = nullptr;
    std::shared_ptr<std::vector<kernelT>> input, output;
    std::vector<kernelT> params;
    std::vector<double> sums;
    std::vector<double> centres;
    void* handle = nullptr;
    sum_fn fsum = nullptr;
    store_fn fstore = nullptr;
    gen_fn fgen = nullptr;
    bool generate = false;
    double maxprob = 1.0;
    std::vector<double> gen_lo, gen_width;
    std::vector<uint32_t> seeds;  // 4 words per stream
    bool kahan = true;
    int nout = 0;
    long long nevents = 0, stride = 0;
    ~State() {
      if (handle) ::dlclose(handle);
    }
  };
  std::shared_ptr<State> s_;

 public:
  explicit CPUBlock(CPUBackend* backend) : s_(std::make_shared<State>()) { s_->be = backend; }
  void SetCentres(const std::vector<double>& c) { s_->centres = c; }
  bool SetupInputBuffer(unsigned long int nbytes) override {
    s_->input = std::make_shared<std::vector<kernelT>>(nbytes / sizeof(kernelT));
    return true;
  }
  bool SetupInputBuffer(ComputeBlock<kernelT, evalT>* other, bool use_data_in = false) override {
    auto* o = dynamic_cast<CPUBlock<kernelT, evalT>*>(other);
    if (!o) throw std::runtime_error("CPUBlock::SetupInputBuffer: source block is not a CPUBlock");
    s_->input = use_data_in ? o->s_->input : o->s_->output;
    return true;
  }
  bool SetupSeedBuffer(unsigned long int) override { return true; }
  bool SetupParameterBuffer(unsigned long int) override { return true; }
  bool SetupOutputBuffer(unsigned long int nbytes) override {
    s_->output = std::make_shared<std::vector<kernelT>>(nbytes / sizeof(kernelT));
    return true;
  }
  bool SetNevents(int nevents, int nevents_padded = -1) override {
    s_->nevents = nevents;
    s_->stride = nevents_padded > 0 ? nevents_padded : nevents;
    return true;
  }
  template <typename srcT>
  bool CopyToInputBufferAny(const EventVector<srcT, evalT>& data) {
    const size_t nd = data.ndimensions(), np = data.nevents_padded();
    s_->input = std::make_shared<std::vector<kernelT>>(nd * np);
    for (size_t d = 0; d < nd; d++) {
      double c = d < s_->centres.size() ? s_->centres[d] : 0.0;
      for (size_t i = 0; i < np; i++) (*s_->input)[d * np + i] = (kernelT)((double)data.get_data()[d * np + i] - c);
    }
    return true;
  }
  bool CopyToInputBuffer(const EventVector<kernelT, evalT>& data) override { return CopyToInputBufferAny(data); }
  template <typename srcT, typename std::enable_if<!std::is_same<srcT, kernelT>::value, int>::type = 0>
  bool CopyToInputBuffer(const EventVector<srcT, evalT>& data) { return CopyToInputBufferAny(data); }
  // one stream per thread: row t holds the 4 state words of thread t's xoshiro128++
  bool CopyToSeedBuffer(const EventVector<seedT, evalT>& data) override {
    s_->seeds.assign(4 * data.nevents(), 0u);
    for (size_t k = 0; k < data.nevents(); k++)
      for (unsigned j = 0; j < 4; j++) s_->seeds[4 * k + j] = (uint32_t)data(k, j);
    return true;
  }
  bool CopyToParameterBuffer(const std::vector<kernelT>& params) override {
    s_->params = params;
    if (s_->params.empty()) s_->params.push_back(kernelT(0));
    return true;
  }
  bool CopyFromOutputBuffer(EventVector<kernelT, evalT>& data) override {
    if (!s_->output) throw std::runtime_error("CPUBlock::CopyFromOutputBuffer: no output buffer");
    std::copy(s_->output->begin(), s_->output->begin() + std::min(s_->output->size(), data.buffer_size() / sizeof(kernelT)), data.get_data());
    return true;
  }
  bool CopyFromKahanBuffer(std::vector<evalT>& sums) override {
    sums.assign(s_->sums.begin(), s_->sums.end());
    return true;
  }
  bool ProvidesKahan() override { return true; }
  bool MakeComputeKernel(std::string name, unsigned int nevents, std::vector<dimension<evalT>> input_signature,
                         std::vector<dimension<evalT>> output_signature, const std::vector<std::string>& params,
                         const std::vector<std::unique_ptr<ComputeGraphNode<kernelT, evalT>>>& graphs, bool kahan_summation = false) override {
    ModuleSource ms;
    ms.fp32 = std::is_same<kernelT, float>::value;
    ms.nobs = (int)input_signature.size();
    std::vector<std::string> obs;
    for (auto& d : input_signature) obs.push_back(d.get_name());
    std::vector<const ComputeGraphNode<kernelT, evalT>*> outs;
    for (auto& g : graphs) outs.push_back(g.get());
    ms.functors = emit_functor("mf_f_block", outs, obs, params, ms.fp32, &ms.statements);
    // FP64 sums run 4 events per AVX2 vector when vectorisation is on (the reference's
    // llvm_vectorization, width 4)
    const compute_options& o = *s_->be->opts_;
    std::string vf;
    if (!ms.fp32 && o.llvm_vectorization && kahan_summation)
      vf = emit_functor("mf_v_block", outs, obs, params, false, nullptr, nullptr, nullptr, "mf_v4d");
    std::string src = cpu_source(ms, "mf_cpu", vf);
    if (s_->be->opts_->print_kernel) std::cout << src << std::endl;
    if (s_->handle) ::dlclose(s_->handle);
    s_->handle = cpu_compile(src, *s_->be->opts_);
    s_->fsum = (sum_fn)::dlsym(s_->handle, "mf_cpu_sum");
    s_->fstore = (store_fn)::dlsym(s_->handle, "mf_cpu_store");
    if (!s_->fsum || !s_->fstore) throw std::runtime_error("CPU backend: symbols missing");
    s_->kahan = kahan_summation;
    s_->generate = false;
    s_->nout = (int)graphs.size();
    if (s_->nevents == 0) s_->nevents = s_->stride = nevents;
    return true;
  }
  bool MakeGenerateKernel(std::string name, unsigned int nevents, std::vector<dimension<evalT>> input_signature,
                          std::vector<dimension<evalT>> output_signature, const std::vector<std::string>& params,
                          const std::vector<std::unique_ptr<ComputeGraphNode<kernelT, evalT>>>& graphs, evalT maxprob = 1.0) override {
    if (graphs.size() != 1) throw std::runtime_error("CPUBlock::MakeGenerateKernel: expects the density as one graph");
    ModuleSource ms;
    ms.fp32 = std::is_same<kernelT, float>::value;
    ms.nobs = (int)output_signature.size();
    std::vector<std::string> obs;
    s_->gen_lo.clear();
    s_->gen_width.clear();
    for (auto& d : output_signature) {
      obs.push_back(d.get_name());
      s_->gen_lo.push_back(d.get_min());
      s_->gen_width.push_back(d.get_max() - d.get_min());
    }
    std::vector<const ComputeGraphNode<kernelT, evalT>*> outs{graphs[0].get()};
    ms.functors = emit_functor("mf_f_block", outs, obs, params, ms.fp32, &ms.statements);
    std::string src = cpu_generate_source(ms);
    if (s_->be->opts_->print_kernel) std::cout << src << std::endl;
    if (s_->handle) ::dlclose(s_->handle);
    s_->handle = cpu_compile(src, *s_->be->opts_);
    s_->fgen = (gen_fn)::dlsym(s_->handle, "mf_cpu_gen");
    if (!s_->fgen) throw std::runtime_error("CPU backend: generate symbol missing");
    s_->fsum = nullptr;
    s_->fstore = nullptr;
    s_->generate = true;
    s_->maxprob = maxprob;
    s_->nout = ms.nobs;
    if (s_->nevents == 0) s_->nevents = s_->stride = nevents;
    return true;
  }
  bool SubmitKernel() override {
    State& s = *s_;
    if (s.generate) {
      // thread t owns events [t * n / nt, ...) (the last thread the remainder) and stream t
      const unsigned nt = (unsigned)(s.seeds.size() / 4);
      if (nt == 0) throw std::runtime_error("CPUBlock::SubmitKernel: generator without seeds");
      if (!s.output || (long long)s.output->size() < s.stride * s.nout) s.output = std::make_shared<std::vector<kernelT>>(s.stride * s.nout);
      const long long size = s.nevents / nt;
      auto slice = [&](unsigned t) {
        long long from = t * size, to = (t == nt - 1) ? s.nevents : from + size;
        s.fgen(s.stride, from, to, &s.seeds[4 * t], s.output->data(), s.maxprob, s.gen_lo.data(), s.gen_width.data());
      };
      if (nt == 1) {
        slice(0);
      } else {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; t++) th.emplace_back(slice, t);
        for (auto& t : th) t.join();
      }
      return true;
    }
    if (!s.input) throw std::runtime_error("CPUBlock::SubmitKernel: no input buffer");
    const kernelT* data = s.input->data();
    const kernelT* p = s.params.empty() ? nullptr : s.params.data();
    unsigned int nt = std::max(1u, s.be->opts_->llvm_nthreads);
    if ((long long)nt > s.nevents) nt = (unsigned)std::max<long long>(1, s.nevents);
    if (!s.kahan) {
      if (!s.output || (long long)s.output->size() < s.stride * s.nout) s.output = std::make_shared<std::vector<kernelT>>(s.stride * s.nout);
      s.fstore(data, s.stride, 0, s.nevents, p, s.output->data(), s.stride);
      return true;
    }
    std::vector<std::vector<double>> part(nt, std::vector<double>(s.nout, 0.0));
    const long long size = s.nevents / nt;
    auto slice = [&](unsigned t) {
      long long from = t * size, to = (t == nt - 1) ? s.nevents : from + size;
      s.fsum(data, s.stride, from, to, p, part[t].data());
    };
    if (nt == 1) {
      slice(0);
    } else {
      std::vector<std::thread> th;
      for (unsigned t = 0; t < nt; t++) th.emplace_back(slice, t);
      for (auto& t : th) t.join();
    }
    s.sums.assign(s.nout, 0.0);
    for (int k = 0; k < s.nout; k++) {
      std::vector<double> v;
      for (unsigned t = 0; t < nt; t++) v.push_back(part[t][k]);
      s.sums[k] = kahan_summation<double, double>(v);
    }
    return true;
  }
  bool Finish() override { return true; }
  // the per-event outputs of the block's functor over its events, counted on the host
  bool CountNonFinite(long long* count, long long* first) override {
    State& s = *s_;
    if (!s.fstore || s.generate || !s.input || s.nout == 0) return false;
    std::vector<kernelT> o((size_t)s.nout * (size_t)std::max<long long>(1, s.nevents));
    s.fstore(s.input->data(), s.stride, 0, s.nevents, s.params.empty() ? nullptr : s.params.data(), o.data(), s.nevents);
    *count = 0;
    *first = -1;
    for (long long i = 0; i < s.nevents; i++) {
      bool ok = true;
      for (int k = 0; k < s.nout; k++) ok = ok && std::isfinite((double)o[(size_t)k * s.nevents + i]);
      if (!ok) {
        if (*first < 0) *first = i;
        (*count)++;
      }
    }
    return true;
  }
};

}  // namespace morefit
#endif

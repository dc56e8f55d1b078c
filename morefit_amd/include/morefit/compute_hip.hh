/**
 * @file compute_hip.hh
 * MI355X backend of the MoreFit compute boundary, over the C ABI of libmorefit_hip
 * (include/morefit_hip.h).
 *
 *  HIPBackend   : ComputeBackend   (reference OpenCLBackend, compute_opencl.hh:30-206)
 *  HIPBlock     : ComputeBlock     (reference OpenCLBlock, compute_opencl.hh:208-1012): the
 *                 drop-in block the generic fitter/generator drive exactly as they drive the
 *                 reference backends (host-evaluated hoisted terms, one launch per call).
 *  HipModel     : one compiled likelihood module for a PDF and its floating parameters:
 *                 value, gradient and Hessian kernels sharing one kernel-parameter layout, a
 *                 device prep kernel computing the hoisted terms in FP64, and the generator.
 *  HipBatch     : many datasets ("segments": toys, requests, or one big dataset) resident
 *                 in one HBM arena with a static chunk table; per launch each segment is
 *                 evaluated or skipped, the launch sequence of a kind is a captured hipGraph.
 *
 * FP32 kernels (kernelT = float) store observables as offsets from the range centre and
 * receive parameter-only terms computed in FP64 and rounded once; per-lane sums are
 * compensated and block/segment totals are FP64.
 */
#ifndef MOREFIT_COMPUTE_HIP_HH
#define MOREFIT_COMPUTE_HIP_HH

#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "codegen.hh"
#include "compute.hh"
#include "hostprog.hh"
#include "morefit_hip.h"
#include "pdf.hh"
#include "random.hh"

namespace morefit {

inline void hip_check(int rc, const char* what) {
  if (rc != MF_OK) throw std::runtime_error(std::string(what) + ": " + mf_last_error());
}

// Fit every evaluation kernel into the 256 architectural VGPRs: a kernel that allocates AGPRs or
// scratch first drops the prefetch of its next event batch (one register set instead of two),
// then has its unroll halved until it fits.  A kernel that
// needs AGPRs even at unroll 1 is an error: mf_module_load refuses such modules, because the
// compiler was found to leave loop-carried AGPR copies unset in lanes outside a divergent
// branch.  Each pass compiles through the code-object cache, so the load that follows finds
// the final code object there.  `res` receives the resources of the final module.
inline void fit_register_budget(ModuleSource& ms, const std::string& options, std::vector<mf_kernel_resources>* res = nullptr) {
  constexpr int kMaxKernels = 64;
  mf_kernel_resources r[kMaxKernels];
  for (;;) {
    const std::string src = ms.device_source();
    int n = 0;
    hip_check(mf_module_resources(src.c_str(), options.c_str(), r, kMaxKernels, &n), "mf_module_resources");
    n = std::min(n, kMaxKernels);
    bool changed = false;
    for (auto& e : ms.evals)
      for (int k = 0; k < n; k++) {
        if (e.name != r[k].name || (r[k].agpr_count == 0 && r[k].private_segment_bytes == 0)) continue;
        if (e.prefetch) {
          e.prefetch = false;  // one register set for the event batch
          changed = true;
        } else if (e.unroll > 1) {
          e.unroll /= 2;
          changed = true;
        } else if (e.fused) {
          e.fused = false;  // the segment reduction moves to the kernel's _final kernel
          changed = true;
        } else if (r[k].agpr_count > 0) {
          throw std::runtime_error("kernel " + e.name + " needs " + std::to_string(r[k].vgpr_count) + " VGPRs + " +
                                   std::to_string(r[k].agpr_count) + " AGPRs at unroll 1; more than the 256 architectural "
                                   "VGPRs is not supported");
        }
      }
    if (!changed) {
      if (res) res->assign(r, r + n);
      return;
    }
  }
}

// ---------------------------------------------------------------------------- backend
class HIPBackend : public ComputeBackend {
  mf_device* dev_ = nullptr;

 public:
  explicit HIPBackend(compute_options* opts) : ComputeBackend(opts) {
    hip_check(mf_device_open(opts->hip_device, &dev_), "HIPBackend: mf_device_open");
  }
  ~HIPBackend() override { mf_device_close(dev_); }
  HIPBackend(const HIPBackend&) = delete;
  HIPBackend& operator=(const HIPBackend&) = delete;
  // tails are masked in the kernels, so events never need padding
  int required_padding() const override { return 0; }
  mf_device* device() const { return dev_; }
  mf_device_props props() const {
    mf_device_props p;
    hip_check(mf_device_query(dev_, &p), "mf_device_query");
    return p;
  }
};

// device allocation owned by one object
struct DevBuf {
  mf_device* dev = nullptr;
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() {}
  DevBuf(mf_device* d, size_t n) { alloc(d, n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  void alloc(mf_device* d, size_t n) {
    if (d == dev && n <= bytes && ptr) return;
    release();
    dev = d;
    hip_check(mf_malloc(d, n, &ptr), "mf_malloc");
    bytes = n;
  }
  void release() {
    if (ptr) mf_free(dev, ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <typename T> T* as() const { return (T*)ptr; }
};

// pinned host allocation
struct HostBuf {
  mf_device* dev = nullptr;
  void* ptr = nullptr;
  void* dptr = nullptr;  // device address of mapped buffers (alloc_mapped), else null
  size_t bytes = 0;
  HostBuf() {}
  ~HostBuf() { release(); }
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  void alloc(mf_device* d, size_t n) {
    if (d == dev && n <= bytes && ptr) return;
    release();
    dev = d;
    hip_check(mf_host_alloc(d, n, &ptr), "mf_host_alloc");
    bytes = n;
  }
  // coherent pinned memory that kernels read and write in place through dptr
  void alloc_mapped(mf_device* d, size_t n) {
    if (d == dev && n <= bytes && ptr && dptr) return;
    release();
    dev = d;
    hip_check(mf_host_alloc_mapped(d, n, &ptr, &dptr), "mf_host_alloc_mapped");
    bytes = n;
  }
  void release() {
    if (ptr) mf_host_free(dev, ptr);
    ptr = dptr = nullptr;
    bytes = 0;
  }
  template <typename T> T* as() const { return (T*)ptr; }
};

// refuses an event count a segment cannot hold (32-bit event indices in the kernels)
inline long long check_segment_events(long long n, const char* who) {
  if (n < 0 || n > MF_MAX_SEGMENT_EVENTS)
    throw std::invalid_argument(std::string(who) + ": " + std::to_string(n) + " events in one segment; at most " +
                                std::to_string(MF_MAX_SEGMENT_EVENTS) + " (split the dataset into segments)");
  return n;
}

// one segment's chunk table [chunk_begin, chunk_end) for n events at `chunk` events per chunk
// Splits segment `seg` of n events into chunks of about `chunk` events.  With `resident` (the
// work-groups the whole device holds at once) a segment of at least half that many chunks is
// split into a multiple of `resident` equal chunks instead, so the last round of work-groups
// is full (e.g. 10^7 events at 8192 per chunk: 1221 chunks on a 1024-group device would run
// 1.19 rounds; 1024 chunks of 9766 run one).
inline void append_chunks(std::vector<mf_chunk>& chunks, mf_segment& sg, int seg, long long n, int chunk, int resident = 0) {
  check_segment_events(n, "append_chunks");
  if (chunk <= 0) throw std::invalid_argument("append_chunks: chunk size " + std::to_string(chunk));
  sg.chunk_begin = (int)chunks.size();
  if (resident > 0 && n > chunk) {
    const long long nch0 = (n + chunk - 1) / chunk;
    if (2 * nch0 >= resident) {
      const long long rounds = std::max(1ll, std::llround((double)nch0 / resident));
      chunk = (int)((n + rounds * resident - 1) / (rounds * resident));
    }
  }
  if (n <= 0) {
    chunks.push_back({seg, 0, 0, 0});  // an empty chunk still completes the segment (totals 0)
  } else {
    for (long long b = 0; b < n; b += chunk) chunks.push_back({seg, (int)b, (int)std::min<long long>(n, b + chunk), 0});
  }
  sg.chunk_end = (int)chunks.size();
}

// ---------------------------------------------------------------------------- HIPBlock
template <typename kernelT, typename evalT, typename seedT = uint32_t>
class HIPBlock : public ComputeBlock<kernelT, evalT, seedT> {
  struct State {
    HIPBackend* be = nullptr;
    std::shared_ptr<DevBuf> input, output;
    DevBuf params, seeds_states, streams, maxprob, segs, chunks, partials, results, seg_done;
    HostBuf results_h;
    mf_module* mod = nullptr;
    std::string kernel;
    int nout = 0;
    bool kahan = false, generate = false, tables_ready = false;
    bool fused = false;           // the evaluation kernel reduces the segment itself (mf_eval_args.seg_done)
    long long nevents = 0, nevents_padded = 0;
    int nstreams = 0;
    double maxprob_value = 1.0;
    int block = 256;              // work-group size the module was compiled for (MF_BLOCK)
    std::vector<double> centres;  // observables stored as x - centre (FP32 path), empty: raw
    ~State() {
      if (mod) mf_module_unload(mod);
    }
  };
  std::shared_ptr<State> s_;

  mf_device* dev() const { return s_->be->device(); }
  const compute_options& opts() const { return *s_->be->opts_; }

  void build_tables() {
    State& s = *s_;
    if (s.tables_ready) return;
    mf_segment sg{};
    sg.data_off = 0;
    sg.stride = s.nevents_padded > 0 ? s.nevents_padded : s.nevents;
    sg.n = (int)s.nevents;
    sg.param_off = 0;
    sg.out_off = 0;
    sg.mode = 1;
    std::vector<mf_chunk> ch;
    append_chunks(ch, sg, 0, s.nevents, s.block * opts().hip_events_per_thread);
    s.segs.alloc(dev(), sizeof(mf_segment));
    s.chunks.alloc(dev(), ch.size() * sizeof(mf_chunk));
    hip_check(mf_memcpy_h2d(dev(), s.segs.ptr, &sg, sizeof(sg)), "HIPBlock segs");
    hip_check(mf_memcpy_h2d(dev(), s.chunks.ptr, ch.data(), ch.size() * sizeof(mf_chunk)), "HIPBlock chunks");
    s.partials.alloc(dev(), ch.size() * std::max(1, s.nout) * sizeof(double));
    s.results.alloc(dev(), std::max(1, s.nout) * sizeof(double));
    s.results_h.alloc(dev(), std::max(1, s.nout) * sizeof(double));
    s.seg_done.alloc(dev(), sizeof(unsigned int));
    hip_check(mf_fill_zero(dev(), s.seg_done.ptr, s.seg_done.bytes), "HIPBlock counters");
    hip_check(mf_sync(dev()), "HIPBlock tables");
    nchunks_ = (int)ch.size();
    s.tables_ready = true;
  }
  int nchunks_ = 0;

  void compile(ModuleSource& ms) {
    State& s = *s_;
    if (s.mod) mf_module_unload(s.mod);
    s.mod = nullptr;
    fit_register_budget(ms, opts().hip_compile_options);
    s.block = ms.block;
    s.fused = !ms.evals.empty() && ms.evals[0].fused;
    std::string src = ms.device_source();
    if (opts().print_kernel) std::cout << "KERNEL SOURCE\n" << src << "END KERNEL SOURCE" << std::endl;
    hip_check(mf_module_load(dev(), src.c_str(), opts().hip_compile_options.c_str(), &s.mod), "HIPBlock: mf_module_load");
  }

 public:
  explicit HIPBlock(HIPBackend* backend) : s_(std::make_shared<State>()) { s_->be = backend; }

  // FP32 path: observables are stored as offsets from these centres (set before uploading)
  void SetCentres(const std::vector<double>& c) { s_->centres = c; }
  const std::vector<double>& Centres() const { return s_->centres; }

  bool SetupInputBuffer(unsigned long int nbytes) override {
    s_->input = std::make_shared<DevBuf>(dev(), nbytes);
    return true;
  }
  bool SetupInputBuffer(ComputeBlock<kernelT, evalT>* other, bool use_data_in = false) override {
    auto* o = dynamic_cast<HIPBlock<kernelT, evalT>*>(other);
    if (!o) throw std::runtime_error("HIPBlock::SetupInputBuffer: source block is not a HIPBlock");
    s_->input = use_data_in ? o->s_->input : o->s_->output;
    if (use_data_in && s_->centres.empty()) s_->centres = o->s_->centres;
    return s_->input != nullptr;
  }
  bool SetupSeedBuffer(unsigned long int nbytes) override { return true; }
  bool SetupParameterBuffer(unsigned long int nbytes) override {
    s_->params.alloc(dev(), std::max<unsigned long>(nbytes, sizeof(kernelT)));
    return true;
  }
  bool SetupOutputBuffer(unsigned long int nbytes) override {
    if (!s_->output || s_->output->bytes < nbytes) s_->output = std::make_shared<DevBuf>(dev(), std::max<unsigned long>(nbytes, 16));
    return true;
  }
  bool SetupKahanBuffer(unsigned long int nbytes) override { return true; }
  bool SetNevents(int nevents, int nevents_padded = -1) override {
    check_segment_events(nevents, "HIPBlock::SetNevents");
    if (nevents != s_->nevents || (nevents_padded > 0 ? nevents_padded : nevents) != s_->nevents_padded) s_->tables_ready = false;
    s_->nevents = nevents;
    s_->nevents_padded = nevents_padded > 0 ? nevents_padded : nevents;
    return true;
  }
  // uploads SoA data, converting/centring as the kernel expects
  template <typename srcT>
  bool CopyToInputBufferAny(const EventVector<srcT, evalT>& data) {
    const size_t nd = data.ndimensions(), np = data.nevents_padded();
    if (!s_->input || s_->input->bytes < nd * np * sizeof(kernelT)) SetupInputBuffer(nd * np * sizeof(kernelT));
    std::vector<kernelT> tmp(nd * np);
    for (size_t d = 0; d < nd; d++) {
      double c = d < s_->centres.size() ? s_->centres[d] : 0.0;
      const srcT* col = data.get_data() + d * np;
      for (size_t i = 0; i < np; i++) tmp[d * np + i] = (kernelT)((double)col[i] - c);
    }
    hip_check(mf_memcpy_h2d(dev(), s_->input->ptr, tmp.data(), tmp.size() * sizeof(kernelT)), "CopyToInputBuffer");
    hip_check(mf_sync(dev()), "CopyToInputBuffer");
    return true;
  }
  bool CopyToInputBuffer(const EventVector<kernelT, evalT>& data) override { return CopyToInputBufferAny(data); }
  template <typename srcT, typename std::enable_if<!std::is_same<srcT, kernelT>::value, int>::type = 0>
  bool CopyToInputBuffer(const EventVector<srcT, evalT>& data) { return CopyToInputBufferAny(data); }

  bool CopyToSeedBuffer(const EventVector<seedT, evalT>& data) override {
    // reference layout: 4 seed columns of one row per work item; here one row per stream.
    // Stream k owns events [k * per, (k + 1) * per): per = 1, or n / max_generator_workitems + 1
    // when there are more events than work items (compute_opencl.hh:436-440, 596-602)
    State& s = *s_;
    const size_t ns = data.nevents();
    std::vector<unsigned int> st(4 * ns);
    for (size_t k = 0; k < ns; k++)
      for (int j = 0; j < 4; j++) st[4 * k + j] = (unsigned int)data(k, (unsigned)j);
    const long long n = s.nevents;
    const int maxw = opts().opencl_max_generator_workitems;
    const long long per = (maxw > 0 && n > maxw) ? n / maxw + 1 : 1;
    if (ns == 0 || (long long)ns * per < n)
      throw std::runtime_error("HIPBlock::CopyToSeedBuffer: " + std::to_string(ns) + " seeds cannot cover " + std::to_string(n) + " events");
    std::vector<mf_gen_stream> streams;
    for (size_t k = 0; k < ns; k++) {
      long long b = (long long)k * per, e = std::min(n, b + per);
      if (b >= e) break;
      streams.push_back({0, (int)b, (int)e, 0});
    }
    s.nstreams = (int)streams.size();
    s.seeds_states.alloc(dev(), std::max<size_t>(16, st.size() * 4));
    s.streams.alloc(dev(), std::max<size_t>(16, streams.size() * sizeof(mf_gen_stream)));
    hip_check(mf_memcpy_h2d(dev(), s.seeds_states.ptr, st.data(), st.size() * 4), "CopyToSeedBuffer");
    hip_check(mf_memcpy_h2d(dev(), s.streams.ptr, streams.data(), streams.size() * sizeof(mf_gen_stream)), "CopyToSeedBuffer");
    hip_check(mf_sync(dev()), "CopyToSeedBuffer");
    return true;
  }
  bool CopyToParameterBuffer(const std::vector<kernelT>& params) override {
    if (s_->params.bytes < params.size() * sizeof(kernelT)) SetupParameterBuffer(params.size() * sizeof(kernelT));
    if (!params.empty()) hip_check(mf_memcpy_h2d(dev(), s_->params.ptr, params.data(), params.size() * sizeof(kernelT)), "CopyToParameterBuffer");
    return true;
  }
  bool CopyFromOutputBuffer(EventVector<kernelT, evalT>& data) override {
    const size_t nd = data.ndimensions(), np = data.nevents_padded();
    if (!s_->output || s_->output->bytes < nd * np * sizeof(kernelT)) throw std::runtime_error("CopyFromOutputBuffer: output buffer too small");
    hip_check(mf_memcpy_d2h(dev(), data.get_data(), s_->output->ptr, nd * np * sizeof(kernelT)), "CopyFromOutputBuffer");
    hip_check(mf_sync(dev()), "CopyFromOutputBuffer");
    if (s_->generate && !s_->centres.empty())
      for (size_t d = 0; d < nd && d < s_->centres.size(); d++)
        for (size_t i = 0; i < np; i++) data.get_data()[d * np + i] = (kernelT)((double)data.get_data()[d * np + i] + s_->centres[d]);
    return true;
  }
  bool CopyFromKahanBuffer(std::vector<evalT>& sums) override {
    sums.resize(s_->nout);
    const double* r = s_->results_h.template as<double>();
    for (int k = 0; k < s_->nout; k++) sums[k] = (evalT)r[k];
    return true;
  }
  bool ProvidesKahan() override { return true; }

  bool MakeComputeKernel(std::string name, unsigned int nevents, std::vector<dimension<evalT>> input_signature,
                         std::vector<dimension<evalT>> output_signature, const std::vector<std::string>& params,
                         const std::vector<std::unique_ptr<ComputeGraphNode<kernelT, evalT>>>& graphs, bool kahan_summation = false) override {
    State& s = *s_;
    ModuleSource ms;
    ms.fp32 = std::is_same<kernelT, float>::value;
    ms.nobs = (int)input_signature.size();
    ms.ntheta = (int)params.size();
    ms.nhoist = 0;
    ms.block = opts().hip_block_size;
    ms.acc = opts().hip_accumulate;
    ms.fast_math = opts().hip_fast_math;
    std::vector<std::string> obs;
    for (auto& d : input_signature) obs.push_back(d.get_name());
    std::vector<const ComputeGraphNode<kernelT, evalT>*> outs;
    for (auto& g : graphs) outs.push_back(g.get());
    ms.functors = emit_functor("mf_f_block", outs, obs, params, ms.fp32, &ms.statements);
    ms.hoist = empty_hoist();
    s.kernel = name;
    if (kahan_summation) {
      ms.evals.push_back({name, "mf_f_block", std::max(1, opts().hip_unroll), opts().hip_fused_final, opts().hip_prefetch});
      ms.kernels.push_back("MF_CHECK_KERNEL(" + name + "_check, mf_f_block)");
    } else {
      ms.kernels.push_back("MF_STORE_KERNEL(" + name + ", mf_f_block)");
    }
    compile(ms);
    s.nout = (int)graphs.size();
    s.kahan = kahan_summation;
    s.generate = false;
    if (s.nevents == 0) s.nevents = s.nevents_padded = nevents;
    s.tables_ready = false;
    return true;
  }

  bool MakeGenerateKernel(std::string name, unsigned int nevents, std::vector<dimension<evalT>> input_signature,
                          std::vector<dimension<evalT>> output_signature, const std::vector<std::string>& params,
                          const std::vector<std::unique_ptr<ComputeGraphNode<kernelT, evalT>>>& graphs, evalT maxprob = 1.0) override {
    State& s = *s_;
    ModuleSource ms;
    ms.fp32 = std::is_same<kernelT, float>::value;
    ms.nobs = (int)output_signature.size();
    ms.ntheta = (int)params.size();
    ms.block = opts().hip_block_size;
    std::vector<std::string> obs;
    for (auto& d : output_signature) {
      obs.push_back(d.get_name());
      ms.gen_lo.push_back(d.get_min());
      ms.gen_width.push_back(d.get_max() - d.get_min());
    }
    auto lp = Log<kernelT, evalT>(graphs.at(0)->copy())->simplify();
    ms.functors = emit_functor<kernelT, evalT>("mf_f_gen", {lp.get()}, obs, params, ms.fp32, &ms.statements);
    ms.hoist = empty_hoist();
    ms.kernels.push_back("MF_GENERATE_KERNEL(mf_f_gen)");
    compile(ms);
    s.generate = true;
    s.kernel = "mf_generate";
    s.maxprob_value = maxprob;
    s.nevents = nevents;
    if (s.nevents_padded < (long long)nevents) s.nevents_padded = nevents;
    s.centres.clear();
    return true;
  }

  bool SubmitKernel() override {
    State& s = *s_;
    if (!s.mod) throw std::runtime_error("HIPBlock::SubmitKernel: no kernel");
    if (s.generate) {
      mf_segment sg{};
      sg.stride = s.nevents_padded;
      sg.n = (int)s.nevents;
      sg.mode = 1;
      s.segs.alloc(dev(), sizeof(mf_segment));
      s.maxprob.alloc(dev(), sizeof(double));
      hip_check(mf_memcpy_h2d(dev(), s.segs.ptr, &sg, sizeof(sg)), "generate segs");
      hip_check(mf_memcpy_h2d(dev(), s.maxprob.ptr, &s.maxprob_value, sizeof(double)), "generate maxprob");
      if (!s.output) throw std::runtime_error("HIPBlock::SubmitKernel: generator needs an output buffer");
      if (s.params.bytes == 0) SetupParameterBuffer(sizeof(kernelT));
      mf_gen_args g{};
      g.data = s.output->ptr;
      g.segs = s.segs.template as<mf_segment>();
      g.streams = s.streams.template as<mf_gen_stream>();
      g.nstreams = s.nstreams;
      g.states = s.seeds_states.template as<unsigned int>();
      g.params = s.params.ptr;
      g.maxprob = s.maxprob.template as<double>();
      g.proposals = nullptr;
      const int block = 64;
      hip_check(mf_launch_generate(s.mod, &g, (s.nstreams + block - 1) / block, block), "mf_launch_generate");
      return true;
    }
    build_tables();
    if (!s.input) throw std::runtime_error("HIPBlock::SubmitKernel: no input buffer");
    mf_eval_args a{};
    a.data = s.input->ptr;
    a.segs = s.segs.template as<mf_segment>();
    a.chunks = s.chunks.template as<mf_chunk>();
    a.nchunks = nchunks_;
    a.nseg = 1;
    if (s.params.bytes == 0) SetupParameterBuffer(sizeof(kernelT));
    a.params = s.params.ptr;
    a.partials = s.partials.template as<double>();
    a.results = s.results.template as<double>();
    if (nchunks_ <= 1) a.flags |= MF_EVAL_SKIP_FINAL;  // one chunk: the per-chunk kernel writes the totals
    else if (s.kahan && s.fused) a.seg_done = s.seg_done.template as<unsigned int>();
    if (!s.kahan) {
      if (!s.output) throw std::runtime_error("HIPBlock::SubmitKernel: store kernel needs an output buffer");
      a.store = s.output->ptr;
      a.store_stride = s.nevents_padded;
    }
    const int grid = opts().hip_grid > 0 ? std::min(opts().hip_grid, nchunks_) : nchunks_;
    hip_check(mf_launch_eval(s.mod, s.kernel.c_str(), &a, grid, s.block), "mf_launch_eval");
    if (s.kahan) hip_check(mf_memcpy_d2h(dev(), s.results_h.ptr, s.results.ptr, s.nout * sizeof(double)), "kahan results");
    return true;
  }
  bool Finish() override {
    hip_check(mf_sync(dev()), "HIPBlock::Finish");
    return true;
  }
  // the summing kernel's diagnostic twin (NAME_check) over the block's events
  bool CountNonFinite(long long* count, long long* first) override {
    State& s = *s_;
    if (!s.mod || !s.kahan || s.generate || !s.input || s.params.bytes == 0) return false;
    build_tables();
    unsigned int c[3] = {0u, 0xffffffffu, 0u};
    DevBuf out(dev(), sizeof(c));
    hip_check(mf_memcpy_h2d(dev(), out.ptr, c, sizeof(c)), "CountNonFinite");
    mf_eval_args a{};
    a.data = s.input->ptr;
    a.segs = s.segs.template as<mf_segment>();
    a.chunks = s.chunks.template as<mf_chunk>();
    a.nchunks = nchunks_;
    a.nseg = 1;
    a.params = s.params.ptr;
    a.store = out.ptr;
    hip_check(mf_launch_check(s.mod, (s.kernel + "_check").c_str(), &a, std::max(1, nchunks_), s.block), "mf_launch_check");
    hip_check(mf_memcpy_d2h(dev(), c, out.ptr, sizeof(c)), "CountNonFinite");
    hip_check(mf_sync(dev()), "CountNonFinite");
    *count = c[0];
    *first = c[1] == 0xffffffffu ? -1 : (long long)c[1];
    return true;
  }
  mf_module* module() const { return s_->mod; }
};

// ---------------------------------------------------------------------------- engine
enum EvalKind { EVAL_NONE = 0, EVAL_LH = 1, EVAL_GRAD = 2, EVAL_HESS = 3 };

// graphs of one density for a fixed set of floating parameters (fixed ones substituted)
template <typename kernelT, typename evalT = double>
struct LikelihoodGraphs {
  typedef GraphPtr<kernelT, evalT> Ptr;
  std::vector<std::string> obs, theta;
  // the graphs' observables are offsets from `centres`; stored data are such offsets (FP32), or
  // are stored as given and the kernels compute the offsets per event (shift = -centres, FP64)
  std::vector<double> centres, lo, hi, shift;
  std::vector<double> stored_centres() const { return shift.empty() ? centres : std::vector<double>(centres.size(), 0.0); }
  // A kernel kind (value; value + gradient; value + gradient + Hessian): the per-event terms the
  // kernel accumulates, and how the kind's outputs follow from their segment sums (FP64)
  struct KindForm {
    std::vector<Ptr> acc;
    CombineForm form;
    bool linear = false;  // acc are the distinct monomials of the outputs' linear forms
    bool pivot = false;   // the monomials evaluate the density around a pivot exponential (exp_pivot)
  };
  KindForm lh, grad, hess;
  Ptr gen;                                   // log p for the generator, hoisted but otherwise as written
  std::vector<std::string> hoist_names;
  std::vector<Ptr> hoist_exprs;
  // event-only terms kept as extra data columns (in terms of the stored observables)
  std::vector<std::string> ev_names;
  std::vector<Ptr> ev_exprs;
  // data moments: observable-only terms whose sums over a segment are computed once per dataset
  // (mf_eval_mom) and enter the outputs' combination (CombineTerm::mom); var^k of one observable
  std::vector<std::string> mom_keys;
  std::vector<Ptr> mom_exprs;
  size_t nodes = 0;

  // base2: rewrite exp/log into the hardware's exp2/log2 (FP32 kernels), see to_base2;
  // event_threshold >= 0: subexpressions of the kernel outputs that depend on the observables
  // only and cost more than it become precomputed data columns (the reference's
  // optimize_dimensions, fitter.hh:339-375); the generator's log density keeps its form
  // linear: 0 every output summed as it is (its parameter-only factor split off as an FP64
  // scale), 1 a kind's outputs as linear forms over shared monomials when that emits fewer
  // operations (linear_expand), 2 always linear forms where they exist; pivot: a density that
  // is a sum of exponential terms is evaluated around one of them in the linear forms (one
  // exponential fewer per event, exp_pivot)
  // moments: with the pivot, the pivot exponent's share of log p is summed from data moments
  // and the other exponentials' arguments are taken in Horner form (pivot_moments)
  LikelihoodGraphs(PDF<kernelT, evalT>* pdf, const std::vector<parameter<evalT>*>& params, bool centre, bool with_hessian,
                   bool base2 = false, double hoist_threshold = 0.0, double event_threshold = -1.0, int linear = 1,
                   bool pivot = true, bool centre_in_kernel = false, bool moments = true) {
    std::vector<std::string> fixed;
    std::vector<double> fixed_v;
    std::map<std::string, Interval> box;  // observable ranges (as stored) and parameter limits
    for (auto* p : params) {
      if (p->is_constant()) {
        fixed.push_back(p->get_name());
        fixed_v.push_back(p->get_value());
      } else {
        theta.push_back(p->get_name());
        box[p->get_name()] = Interval{(double)p->get_min(), (double)p->get_max()};
      }
    }
    for (auto* d : pdf->dimensions()) {
      obs.push_back(d->get_name());
      lo.push_back(d->get_min());
      hi.push_back(d->get_max());
      centres.push_back(centre ? d->get_centre() : 0.0);
      box[obs.back()] = Interval{lo.back() - centres.back(), hi.back() - centres.back()};
      if (centre && centre_in_kernel) shift.push_back(-centres.back());
    }
    Ptr lp = pdf->logprob_normalised()->substitute(fixed, fixed_v)->simplify();
    if (centre) lp = centre_observables(*lp, obs, centres);
    lp = double_angle(*lp)->simplify();
    // event columns come out of log p itself, before differentiation: the derivatives then
    // treat them as opaque variables (products of columns stay products, not new columns)
    Ptr lpe = event_threshold >= 0.0
                  ? lp->optimize_buffering_constant_terms(ev_names, ev_exprs, obs, "morefit_eventbuffer_", (float)event_threshold, false)->simplify()
                  : lp->copy();
    std::vector<Ptr> d1;
    for (auto& t : theta) d1.emplace_back(factor_common(*lpe->diff(t)->simplify()));
    std::vector<Ptr> d2;
    if (with_hessian)
      for (size_t i = 0; i < theta.size(); i++)
        for (size_t j = 0; j <= i; j++) d2.emplace_back(factor_common(*d1[i]->diff(theta[j])->simplify()));
    const std::string prefix = "morefit_parambuffer_";
    const float thr = (float)hoist_threshold;
    auto hoist_index = [&](Ptr e) {
      const std::string nm = ComputeGraphNode<kernelT, evalT>::hoist(std::move(e), hoist_names, hoist_exprs, prefix);
      return (int)(std::find(hoist_names.begin(), hoist_names.end(), nm) - hoist_names.begin());
    };
    // a combine term's parameter-only coefficient: constants stay in mult, the rest is hoisted
    auto coefficient = [&](Ptr c, CombineTerm& t) {
      c = c->simplify();
      if (c->is_constant()) {
        t.mult = c->value();
        return;
      }
      if (c->kind() == NodeKind::Prod) {
        std::vector<Ptr> rest;
        for (auto& f : c->children_) {
          if (f->is_constant()) t.mult *= f->value();
          else rest.emplace_back(f->copy());
        }
        if (rest.empty()) return;
        c = rest.size() == 1 ? std::move(rest[0]) : make_node<kernelT, evalT>(NodeKind::Prod, std::move(rest));
      }
      t.hoist = hoist_index(std::move(c));
    };
    // every output summed as it is, its parameter-only factor applied to the total
    auto build_scaled = [&](const std::vector<Ptr>& outs) {
      KindForm kf;
      for (size_t k = 0; k < outs.size(); k++) {
        Ptr sc;
        Ptr ev = split_output_scale(*outs[k], theta, sc);
        CombineTerm t;
        t.acc = (int)k;
        if (sc) coefficient(std::move(sc), t);
        kf.form.push_back({t});
        kf.acc.emplace_back(ev->optimize_buffering_constant_terms(hoist_names, hoist_exprs, theta, prefix, thr)->simplify());
      }
      return kf;
    };
    // index of the data moment var^k (registered on first use)
    auto moment_index = [&](const std::string& var, int k) {
      const std::string key = var + "^" + std::to_string(k);
      auto it = std::find(mom_keys.begin(), mom_keys.end(), key);
      if (it != mom_keys.end()) return (int)(it - mom_keys.begin());
      Ptr v = make_node<kernelT, evalT>(NodeKind::Variable, {}, 0.0, var);
      if (k > 1) {
        std::vector<Ptr> ch;
        ch.emplace_back(std::move(v));
        ch.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, (double)k));
        v = make_node<kernelT, evalT>(NodeKind::Pow, std::move(ch));
      }
      mom_keys.push_back(key);
      mom_exprs.emplace_back(std::move(v));
      return (int)mom_keys.size() - 1;
    };
    // With the exponential pivot, log p = a_k + log q and the other terms enter as
    // E(a_i - a_k).  When the pivot exponent a_k is a polynomial in one observable x (the
    // exponential density: lambda x), (1) its share of every output, coefficient * sum_events a_k,
    // is a combination of the segment's event count and moments sum x^j, so the kernels sum log q
    // alone; (2) each a_i - a_k, then a polynomial in x, is written in Horner form in the shift
    // t = x + s the monomials use most (the Gaussian's m - mu): for f Gauss + (1 - f) Exp the
    // kernels no longer compute lambda x per event, and (m - mu)^2 / 2 sigma^2 - lambda m costs two
    // multiply-adds.  Everything is rewritten or nothing (false).
    std::function<bool(std::vector<Ptr>&, CombineForm&, std::vector<std::map<int, Ptr>>&, const std::string&, const ComputeGraphNode<kernelT, evalT>&)>
        pivot_moments = [&](std::vector<Ptr>& monos, CombineForm& form, std::vector<std::map<int, Ptr>>& cexpr, const std::string& pv,
                            const ComputeGraphNode<kernelT, evalT>& pe) -> bool {
      std::string var;
      int nvar = 0;
      for (const auto& o : obs)
        if (pe.variable_in_tree(o)) {
          var = o;
          nvar++;
        }
      for (const auto& e : ev_names)
        if (pe.variable_in_tree(e)) nvar++;
      std::vector<Ptr> pa;
      if (nvar != 1 || !poly_in(pe, var, theta, pa, 4)) return false;
      // the log p monomial: a_k (as pv) + log q
      int i0 = -1;
      for (size_t i = 0; i < monos.size(); i++) {
        const auto& m = *monos[i];
        if (m.kind() != NodeKind::Sum || m.children_.size() != 2) continue;
        for (const auto& c : m.children_)
          if (c->kind() == NodeKind::Variable && c->name() == pv) i0 = (int)i;
      }
      if (i0 < 0) return false;
      // the shift of x used most in the monomials (Horner variable t = x + s), else x itself
      std::map<std::string, int> count;
      std::map<std::string, const ComputeGraphNode<kernelT, evalT>*> repr;
      std::function<void(const ComputeGraphNode<kernelT, evalT>&)> collect = [&](const ComputeGraphNode<kernelT, evalT>& n) {
        if (linear_shift(n, theta) && n.variable_in_tree(var)) {
          count[n.get_kernel()]++;
          repr.emplace(n.get_kernel(), &n);
          return;
        }
        for (const auto& c : n.children_) collect(*c);
      };
      for (const auto& m : monos) collect(*m);
      Ptr tnode = make_node<kernelT, evalT>(NodeKind::Variable, {}, 0.0, var), snode;
      int most = 0;
      for (const auto& [key, c] : count)
        if (c > most) {
          most = c;
          tnode = repr[key]->copy();
        }
      if (most > 0) {
        std::vector<Ptr> sp;
        for (const auto& c : tnode->children_)
          if (!c->contains_other_variables(theta)) sp.emplace_back(c->copy());
        snode = sp.size() == 1 ? std::move(sp[0]) : make_node<kernelT, evalT>(NodeKind::Sum, std::move(sp));
      }
      // E(a_i - pv) -> E(Horner(a_i - a_k)) in every monomial
      bool ok = true;
      std::function<Ptr(const ComputeGraphNode<kernelT, evalT>&)> rewrite = [&](const ComputeGraphNode<kernelT, evalT>& n) -> Ptr {
        if ((n.kind() == NodeKind::Exp || n.kind() == NodeKind::Exp2) && n.children_[0]->variable_in_tree(pv)) {
          Ptr e = n.children_[0]->replace_variable(pv, pe)->simplify();
          std::vector<Ptr> c;
          if (!poly_in(*e, var, theta, c, 4)) {
            ok = false;
            return n.copy();
          }
          std::vector<Ptr> arg;
          arg.emplace_back(horner_in_shift(c, *tnode, snode.get()));
          return make_node<kernelT, evalT>(n.kind(), std::move(arg));
        }
        std::vector<Ptr> ch;
        for (const auto& c : n.children_) ch.emplace_back(rewrite(*c));
        return make_node<kernelT, evalT>(n.kind(), std::move(ch), n.value(), n.name(), n.comparator());
      };
      std::vector<Ptr> out;
      for (size_t i = 0; i < monos.size(); i++) {
        if ((int)i == i0) {
          for (const auto& c : monos[i]->children_)
            if (!(c->kind() == NodeKind::Variable && c->name() == pv)) out.emplace_back(rewrite(*c));
        } else {
          out.emplace_back(rewrite(*monos[i]));
        }
        if (!ok || out.back()->variable_in_tree(pv)) return false;
      }
      // each output's coefficient c of log p: c * a_k = c a_0 n + sum_j c a_j sum x^j
      CombineForm f2(form);
      for (size_t k = 0; k < f2.size(); k++) {
        auto it = cexpr[k].find(i0);
        if (it == cexpr[k].end()) continue;
        for (size_t j = 0; j < pa.size(); j++) {
          if (!pa[j]) continue;
          std::vector<Ptr> pr;
          pr.emplace_back(it->second->copy());
          pr.emplace_back(pa[j]->copy());
          CombineTerm t;
          if (j > 0) t.mom = moment_index(var, (int)j);
          coefficient(make_node<kernelT, evalT>(NodeKind::Prod, std::move(pr)), t);
          if (t.hoist < 0 && t.mult == 0.0) continue;
          f2[k].push_back(t);
        }
      }
      monos = std::move(out);
      form = std::move(f2);
      return true;
    };
    // outputs as linear forms over the kind's distinct monomials (false: expansion too large)
    // (canon: one shift per event variable, linear_canonical_shifts)
    auto build_linear = [&](const std::vector<Ptr>& outs, KindForm& kf, bool canon) {
      std::map<std::string, int> index;
      std::vector<Ptr> monos;
      std::vector<std::vector<LinearTerm<kernelT, evalT>>> all(outs.size());
      std::vector<std::map<int, Ptr>> cexpr(outs.size());  // each output's coefficient of each monomial, unhoisted
      for (size_t k = 0; k < outs.size(); k++)
        if (!linear_expand(*outs[k], theta, all[k], 64)) return false;
      if (canon && !linear_canonical_shifts(all, theta, 64)) return false;
      for (size_t k = 0; k < outs.size(); k++) {
        std::vector<LinearTerm<kernelT, evalT>>& terms = all[k];
        std::vector<std::string> keys;
        std::map<std::string, std::vector<Ptr>> coefs;
        for (auto& t : terms) {
          std::string key;
          Ptr mono = linear_monomial(t.factors, key);
          if (!key.empty() && !index.count(key)) {
            index[key] = (int)monos.size();
            monos.emplace_back(std::move(mono));
          }
          if (!coefs.count(key)) keys.push_back(key);
          coefs[key].emplace_back(t.coef ? std::move(t.coef) : make_node<kernelT, evalT>(NodeKind::Constant, {}, 1.0));
        }
        std::vector<CombineTerm> ft;
        for (const auto& key : keys) {
          auto& cl = coefs[key];
          Ptr c = cl.size() == 1 ? std::move(cl[0]) : make_node<kernelT, evalT>(NodeKind::Sum, std::move(cl));
          CombineTerm t;
          t.acc = key.empty() ? -1 : index[key];
          if (t.acc >= 0 && moments) cexpr[k][t.acc] = c->copy();
          coefficient(std::move(c), t);
          if (t.hoist < 0 && t.mult == 0.0) continue;
          ft.push_back(t);
        }
        kf.form.push_back(ft);
      }
      const std::string pv = "morefit_pivot";
      Ptr pe;
      // E(a_i - a_k) up to 2^40 in FP32 (a Hessian monomial's product of two stays below
      // 2^127), 2^400 in FP64
      if (pivot) kf.pivot = exp_pivot(monos, theta, box, pv, pe, std::is_same<kernelT, float>::value ? 40.0 : 400.0);
      const bool pv_gone = kf.pivot && moments && pivot_moments(monos, kf.form, cexpr, pv, *pe);
      if (kf.pivot && !pv_gone) pe = pe->optimize_buffering_constant_terms(hoist_names, hoist_exprs, theta, prefix, thr)->simplify();
      for (auto& m : monos) {
        Ptr h = m->optimize_buffering_constant_terms(hoist_names, hoist_exprs, theta, prefix, thr)->simplify();
        // the pivot exponent substituted after simplification: a_i - a_k stays a difference
        // of the two exponents as computed, sharing a_k with log p = a_k + log q
        kf.acc.emplace_back(kf.pivot && !pv_gone ? h->replace_variable(pv, *pe) : std::move(h));
      }
      kf.linear = true;
      return true;
    };
    // issue cost of the statements the code generator emits for a kind's per-event terms
    // (after CSE, the cheaper of its two product orders, as emit_functor keeps), in full-rate
    // operations
    auto statements = [&](const std::vector<Ptr>& acc) {
      double best = INFINITY;
      for (bool search : {true, false}) {
        KernelEmitter em(std::is_same<kernelT, float>::value);
        em.search_order = search;
        std::vector<std::string> cols(obs);
        cols.insert(cols.end(), ev_names.begin(), ev_names.end());
        for (size_t d = 0; d < cols.size(); d++) em.bindings[cols[d]] = "x[" + std::to_string(d) + "]";
        std::vector<std::string> pn(theta);
        pn.insert(pn.end(), hoist_names.begin(), hoist_names.end());
        for (size_t k = 0; k < pn.size(); k++) em.bindings[pn[k]] = "p[" + std::to_string(k) + "]";
        for (auto& g : acc) em.count_products(*g);
        for (auto& g : acc) em.emit(*g);
        best = std::min(best, em.cost);
      }
      return best;
    };
    // canon: linear forms may take one shift per event variable (value and gradient kinds only:
    // in second derivatives the shift differences enter squared, and the FP32 Hessian sums
    // lost accuracy to the cancellation between the expanded terms)
    auto make_kind = [&](std::vector<Ptr> outs, bool allow_linear, bool canon) {
      if (base2)
        for (auto& g : outs) g = to_base2(*g)->simplify();
      if (linear == 0 || !allow_linear) return build_scaled(outs);
      std::vector<std::string> names0(hoist_names);
      std::vector<Ptr> exprs0;
      for (auto& e : hoist_exprs) exprs0.emplace_back(e->copy());
      const size_t nmom0 = mom_exprs.size();
      auto restore = [&] {
        hoist_names = names0;
        hoist_exprs.clear();
        for (auto& e : exprs0) hoist_exprs.emplace_back(e->copy());
        mom_keys.resize(nmom0);
        mom_exprs.resize(nmom0);
      };
      // candidates: linear forms with one shift per event variable (c = 1) or as expanded
      // (c = 0), and every output summed as it is; the one with the lowest emitted cost wins
      // (ties: the earlier).  Accumulating a term costs about 1.25 operations per event (an FP32
      // add, and an FP64 convert + add per unrolled group).
      double cost[2] = {INFINITY, INFINITY};
      for (int c = canon ? 1 : 0; c >= 0; c--) {
        KindForm b;
        if (build_linear(outs, b, c == 1)) cost[c] = statements(b.acc) + 1.25 * b.acc.size();
        restore();
      }
      const int cb = cost[1] <= cost[0] ? 1 : 0;
      if (std::isfinite(cost[cb]) && linear == 2) {
        KindForm b;
        build_linear(outs, b, cb == 1);
        return b;
      }
      KindForm a = build_scaled(outs);
      const double cost_a = statements(a.acc) + 1.25 * a.acc.size();
      if (!std::isfinite(cost[cb]) || cost_a <= cost[cb]) return a;
      restore();
      KindForm b;
      build_linear(outs, b, cb == 1);
      return b;
    };
    gen = lp->optimize_buffering_constant_terms(hoist_names, hoist_exprs, theta, prefix, thr)->simplify();
    {
      std::vector<Ptr> o;
      o.emplace_back(lpe->copy());
      lh = make_kind(std::move(o), true, true);
    }
    {
      std::vector<Ptr> o;
      o.emplace_back(lpe->copy());
      for (auto& g : d1) o.emplace_back(g->copy());
      grad = make_kind(std::move(o), true, true);
    }
    if (with_hessian) {
      std::vector<Ptr> o;
      o.emplace_back(lpe->copy());
      for (auto& g : d1) o.emplace_back(g->copy());
      for (auto& g : d2) o.emplace_back(g->copy());
      // FP64 Hessians keep per-event outputs: second derivatives cancel within each event
      // (E''/p - (E'/p)^2 for a dominant component), and a linear form moves that cancellation
      // to segment sums of 10^6 terms, costing FP64 digits (1e-12 of the reference) where f or
      // lambda sit near their limits; the Hessian is evaluated once per fit
      hess = make_kind(std::move(o), std::is_same<kernelT, float>::value, false);
    }
    for (auto& g : hess.acc) nodes += g->node_count();
  }
  std::vector<std::string> columns() const {
    std::vector<std::string> r(obs);
    r.insert(r.end(), ev_names.begin(), ev_names.end());
    return r;
  }
  std::vector<std::string> param_names() const {
    std::vector<std::string> r(theta);
    r.insert(r.end(), hoist_names.begin(), hoist_names.end());
    return r;
  }
  static std::vector<const ComputeGraphNode<kernelT, evalT>*> raw(const std::vector<Ptr>& v) {
    std::vector<const ComputeGraphNode<kernelT, evalT>*> r;
    for (auto& g : v) r.push_back(g.get());
    return r;
  }
  ModuleSource module(const compute_options& o, bool with_hessian) const {
    ModuleSource ms;
    ms.fp32 = std::is_same<kernelT, float>::value;
    ms.nobs = (int)obs.size();
    ms.ntheta = (int)theta.size();
    ms.nhoist = (int)hoist_names.size();
    ms.block = o.hip_block_size;
    ms.acc = o.hip_accumulate;
    ms.fast_math = o.hip_fast_math;
    const auto sc = stored_centres();
    for (size_t d = 0; d < obs.size(); d++) {
      ms.gen_lo.push_back(lo[d] - sc[d]);
      ms.gen_width.push_back(hi[d] - lo[d]);
      ms.obs_lo.push_back(lo[d] - sc[d]);
      ms.obs_hi.push_back(hi[d] - sc[d]);
    }
    auto pn = param_names();
    const auto cols = columns();
    ms.ncol = (int)cols.size();
    const std::vector<double>* sh = shift.empty() ? nullptr : &shift;
    ms.functors = emit_functor("mf_f_lh", raw(lh.acc), cols, pn, ms.fp32, &ms.statements, &lh.form, sh) +
                  emit_functor("mf_f_grad", raw(grad.acc), cols, pn, ms.fp32, &ms.statements, &grad.form, sh);
    if (with_hessian) ms.functors += emit_functor("mf_f_hess", raw(hess.acc), cols, pn, ms.fp32, &ms.statements, &hess.form, sh);
    ms.nmom = (int)mom_exprs.size();
    if (ms.nmom) ms.functors += emit_functor("mf_f_mom", raw(mom_exprs), cols, pn, ms.fp32, &ms.statements, nullptr, sh);
    ms.nacc[EVAL_LH] = (int)lh.acc.size();
    ms.nacc[EVAL_GRAD] = (int)grad.acc.size();
    ms.nacc[EVAL_HESS] = with_hessian ? (int)hess.acc.size() : 0;
    ms.functors += emit_functor<kernelT, evalT>("mf_f_gen", {gen.get()}, obs, pn, ms.fp32, &ms.statements, nullptr, sh);
    ms.hoist = emit_hoist(raw(hoist_exprs), theta);
    if (!ev_exprs.empty()) ms.event_terms = emit_fp64_function("mf_event_terms", "x", raw(ev_exprs), obs, sh);
    const int u = std::max(1, o.hip_unroll);
    const bool f = o.hip_fused_final;
    const bool pf = o.hip_prefetch;
    ms.evals = {{"mf_eval_lh", "mf_f_lh", u, f, pf}, {"mf_eval_grad", "mf_f_grad", u, f, pf}};
    if (with_hessian) ms.evals.push_back({"mf_eval_hess", "mf_f_hess", u, f, pf});
    if (ms.nmom) ms.evals.push_back({"mf_eval_mom", "mf_f_mom", u, false, pf});
    ms.kernels = {"MF_PREP_KERNEL", "MF_GENERATE_KERNEL(mf_f_gen)",
                  ms.nmom ? "MF_CHECK_KERNEL_MOM(mf_check_lh, mf_f_lh, mf_f_mom)" : "MF_CHECK_KERNEL(mf_check_lh, mf_f_lh)"};
    if (!ev_exprs.empty()) ms.kernels.push_back("MF_PRECOMPUTE_KERNEL");
    return ms;
  }
};

template <typename kernelT, typename evalT = double>
class HipModel {
 public:
  HIPBackend* be;
  PDF<kernelT, evalT>* pdf;
  std::vector<parameter<evalT>*> params;
  std::vector<int> theta_index;  // position in params of each floating parameter
  std::vector<std::string> obs;
  std::vector<double> centres, lo, hi;
  // the hoisted terms as a host program over theta: one-segment calls pass [theta, hoisted] by
  // value instead of running the prep kernel (empty when the block does not fit the arguments)
  HostProgram<kernelT, evalT> hoist_prog;
  bool host_params = false;
  int nobs = 0, ntheta = 0, nhoist = 0, nparam = 0;
  int ncol = 0;  // data columns per event: observables, then precomputed event terms
  int nout[4] = {0, 0, 0, 0};
  int nacc[4] = {0, 0, 0, 0};  // per-event terms each kind's kernel accumulates (partials per chunk)
  int nmom = 0;                // data moments per segment (mf_eval_mom; 0: none)
  bool fp32 = std::is_same<kernelT, float>::value;
  bool with_hessian = true;
  std::string source;
  size_t statements = 0, graph_nodes = 0;
  double codegen_ms = 0.0, compile_ms = 0.0;
  mf_module* mod = nullptr;
  std::vector<std::pair<std::string, int>> unrolls;  // evaluation kernel -> unroll
  std::vector<std::string> fused;                    // evaluation kernels that reduce multi-chunk segments themselves
  std::vector<mf_kernel_resources> resources;        // every kernel of the module
  int block = 256;                                   // work-group size compiled into the module
  int grid_opt = 0;                                  // hip_grid at construction (0: one group per chunk)
  int resident = 0;                                  // gradient work-groups the device holds at once (0: no balancing)

  // generated module with the unroll of each evaluation kernel fitted to the register budget
  static ModuleSource make_source(PDF<kernelT, evalT>* pdf, const std::vector<parameter<evalT>*>& params, const compute_options& o,
                                  bool with_hessian, LikelihoodGraphs<kernelT, evalT>** keep = nullptr,
                                  std::vector<mf_kernel_resources>* res = nullptr, double* compile_ms = nullptr) {
    const bool f32 = std::is_same<kernelT, float>::value;
    std::unique_ptr<LikelihoodGraphs<kernelT, evalT>> g(
        new LikelihoodGraphs<kernelT, evalT>(pdf, params, o.hip_centre_observables, with_hessian,
                                             o.hip_base2 >= 2 || (f32 && o.hip_base2 == 1), 0.0,
                                             o.hip_event_columns ? o.hip_event_cost_threshold : -1.0, o.hip_linear_outputs,
                                             o.hip_exp_pivot, !f32));
    ModuleSource ms = g->module(o, with_hessian);
    auto t0 = std::chrono::steady_clock::now();
    fit_register_budget(ms, o.hip_compile_options, res);
    if (compile_ms) *compile_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (keep) *keep = g.release();
    return ms;
  }
  // unroll and register count of an evaluation kernel ("mf_eval_lh", ...) in the loaded module
  int kernel_unroll(const std::string& name) const {
    for (auto& e : unrolls)
      if (e.first == name) return e.second;
    return 0;
  }
  bool kernel_fused(const std::string& name) const {
    for (auto& e : fused)
      if (e == name) return true;
    return false;
  }
  int kernel_vgprs(const std::string& name) const {
    for (auto& r : resources)
      if (name == r.name) return r.vgpr_count;
    return 0;
  }

  HipModel(HIPBackend* backend, PDF<kernelT, evalT>* p, const std::vector<parameter<evalT>*>& ps, bool hessian = true)
      : be(backend), pdf(p), params(ps), with_hessian(hessian) {
    auto t0 = std::chrono::steady_clock::now();
    LikelihoodGraphs<kernelT, evalT>* g = nullptr;
    double fit_ms = 0.0;
    ModuleSource ms = make_source(p, ps, *be->opts_, hessian, &g, &resources, &fit_ms);
    std::unique_ptr<LikelihoodGraphs<kernelT, evalT>> keep(g);
    for (auto& e : ms.evals) {
      unrolls.push_back({e.name, e.unroll});
      if (e.fused) fused.push_back(e.name);
    }
    block = ms.block;
    grid_opt = be->opts_->hip_grid;
    for (size_t i = 0; i < ps.size(); i++)
      if (!ps[i]->is_constant()) theta_index.push_back((int)i);
    obs = g->obs;
    centres = g->stored_centres();
    lo = g->lo;
    hi = g->hi;
    nobs = ms.nobs;
    ncol = std::max(ms.ncol, ms.nobs);
    ntheta = ms.ntheta;
    nhoist = ms.nhoist;
    host_params = be->opts_->hip_host_params && (size_t)(ms.ntheta + ms.nhoist) * sizeof(kernelT) <= MF_INLINE_PARAM_WORDS * 4u &&
                  ms.nhoist <= MF_INLINE_HOIST;
    if (host_params) hoist_prog = HostProgram<kernelT, evalT>(LikelihoodGraphs<kernelT, evalT>::raw(g->hoist_exprs), g->theta);
    nparam = ntheta + nhoist;
    nout[EVAL_LH] = 1;
    nout[EVAL_GRAD] = 1 + ntheta;
    nout[EVAL_HESS] = hessian ? 1 + ntheta + ntheta * (ntheta + 1) / 2 : 0;
    for (int k = 1; k <= 3; k++) nacc[k] = nout[k] ? ms.nacc[k] : 0;
    nmom = ms.nmom;
    statements = ms.statements;
    graph_nodes = g->nodes;
    source = ms.device_source();
    // codegen_ms: graphs and source; compile_ms: hipRTC passes of the register fit (0 when the
    // code objects come from the cache) plus the module load
    codegen_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() - fit_ms;
    if (be->opts_->print_kernel) std::cout << source << std::endl;
    auto t1 = std::chrono::steady_clock::now();
    hip_check(mf_module_load(be->device(), source.c_str(), be->opts_->hip_compile_options.c_str(), &mod), "HipModel: mf_module_load");
    if (be->opts_->hip_balance_chunks && grid_opt == 0) {
      int per_cu = 0;
      hip_check(mf_module_occupancy(mod, "mf_eval_grad", block, &per_cu), "HipModel: mf_module_occupancy");
      resident = per_cu * be->props().compute_units;
    }
    compile_ms = fit_ms + std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
  }
  ~HipModel() { mf_module_unload(mod); }
  HipModel(const HipModel&) = delete;
  HipModel& operator=(const HipModel&) = delete;

  // accept-reject envelope at the given floating parameter values (PDF::get_max)
  double maxprob(const double* theta) const {
    std::vector<evalT> saved;
    for (size_t k = 0; k < theta_index.size(); k++) {
      saved.push_back(params[theta_index[k]]->get_value());
      params[theta_index[k]]->set_value(theta[k]);
    }
    double m = pdf->get_max();
    for (size_t k = 0; k < theta_index.size(); k++) params[theta_index[k]]->set_value(saved[k]);
    return m;
  }
};

// xoshiro128++ stream states: `n` streams separated by jump() from the seeded state
inline void xoshiro_streams(uint64_t seed, int n, unsigned int* out) {
  Xoshiro128pp r;
  r.setSeed(seed);
  for (int k = 0; k < n; k++) {
    for (int j = 0; j < 4; j++) out[4 * k + j] = r.getSeed(j);
    r.jump();
  }
}

// first-fit allocator over an arena of `capacity` elements (offsets aligned to 64 elements)
class ArenaAllocator {
  std::vector<std::pair<long long, long long>> free_;  // (offset, length), sorted by offset

 public:
  long long capacity = 0;
  explicit ArenaAllocator(long long cap = 0) { reset(cap); }
  void reset(long long cap) {
    capacity = cap;
    free_.clear();
    if (cap > 0) free_.push_back({0, cap});
  }
  static long long round(long long n) { return (std::max<long long>(n, 1) + 63) / 64 * 64; }
  long long alloc(long long n) {
    n = round(n);
    for (size_t i = 0; i < free_.size(); i++)
      if (free_[i].second >= n) {
        long long off = free_[i].first;
        free_[i].first += n;
        free_[i].second -= n;
        if (free_[i].second == 0) free_.erase(free_.begin() + i);
        return off;
      }
    return -1;
  }
  void release(long long off, long long n) {
    n = round(n);
    auto it = free_.begin();
    while (it != free_.end() && it->first < off) ++it;
    it = free_.insert(it, {off, n});
    size_t i = it - free_.begin();
    if (i + 1 < free_.size() && free_[i].first + free_[i].second == free_[i + 1].first) {
      free_[i].second += free_[i + 1].second;
      free_.erase(free_.begin() + i + 1);
    }
    if (i > 0 && free_[i - 1].first + free_[i - 1].second == free_[i].first) {
      free_[i - 1].second += free_[i].second;
      free_.erase(free_.begin() + i);
    }
  }
  long long largest_free() const {
    long long m = 0;
    for (auto& f : free_) m = std::max(m, f.second);
    return m;
  }
};

// Device storage for event columns with a first-fit allocator over it.  One batch owns its
// arena, or several batches share one (the fit service's classes: one allocation that bounds
// the event data of every class together).
template <typename kernelT>
struct EventArena {
  DevBuf buf;
  ArenaAllocator alloc;
  EventArena(mf_device* dev, long long elems) : alloc(elems) { buf.alloc(dev, std::max<long long>(1, elems) * sizeof(kernelT)); }
  long long capacity() const { return alloc.capacity; }
  size_t device_bytes() const { return buf.bytes; }
};

// Segments of one model resident in one HBM arena.  Built either with fixed sizes (a toy
// study, one big dataset) or with a slot capacity and an arena budget, in which case
// segments are added and removed while other segments keep being evaluated (the fit
// service).  Chunk tables are rebuilt when membership changes; a changed table drops the
// captured graphs, which are recaptured on the next graph launch.
template <typename kernelT, typename evalT = double>
class HipBatch {
 public:
  HipModel<kernelT, evalT>* m;
  mf_device* dev;
  int S = 0;                             // slot capacity
  std::vector<long long> sizes, offsets; // per slot (events, element offset)
  std::vector<char> active;
  std::shared_ptr<EventArena<kernelT>> arena_;  // event columns (owned, or shared with other batches)
  long long total_elems = 0;             // arena capacity in elements
  int chunk = 0, nchunks = 0, chunk_cap = 0;
  int max_seg_chunks = 0;                // most chunks of any active segment (<= 1: no final pass)
  bool fused_[4] = {false, false, false, false};  // per kind: the kernel reduces multi-chunk segments itself
  bool segs_stale_[4] = {true, true, true, true};  // per kind: device segment table differs from the mapped one
  bool dirty = true;
  std::vector<mf_segment> segs_h;
  DevBuf params_d, hoist64_d, segs_d[4], chunks_d, partials_d, seg_done_d, moments_d;
  // generator scratch, kept between calls (grows as needed)
  DevBuf gen_powers_, gen_streams_, gen_states_, gen_base_, gen_maxprob_, gen_segs_, gen_props_, gen_slots_;
  int gen_levels_ = 0;
  // pinned staging of the tables copied to the device (a copy from pageable memory costs the
  // runtime ~0.1 ms of staging per call on MI355X): the chunk table, and the generator's tables
  // and proposal counts
  HostBuf chunks_stage_, gen_stage_;
  // staging in mapped host memory, read and written by the kernels in place: theta and the
  // per-kind segment tables (modes) in, totals out -- a call's graph holds kernels only
  HostBuf theta_h, segs_hk[4], results_h[4], done_h, segs_pre_h;
  int done_seq_ = 0;
  mf_graph* graphs[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<int> request_;
  DevBuf check_segs_, check_out_;  // diagnostic pass: its segment table and counters

  // fixed segments, packed back to back
  HipBatch(HipModel<kernelT, evalT>* model, const std::vector<long long>& seg_sizes) : m(model), dev(model->be->device()) {
    long long tot = 0;
    for (auto n : seg_sizes) tot += ArenaAllocator::round(n * m->ncol);
    init((int)seg_sizes.size(), tot);
    for (size_t s = 0; s < seg_sizes.size(); s++)
      if (add_segment(seg_sizes[s]) != (int)s) throw std::runtime_error("HipBatch: packing failed");
    rebuild();
  }
  // `capacity` slots in an arena of `arena_bytes`; segments come and go via add/remove
  HipBatch(HipModel<kernelT, evalT>* model, int capacity, size_t arena_bytes) : m(model), dev(model->be->device()) {
    init(capacity, (long long)(arena_bytes / sizeof(kernelT)));
  }
  // `capacity` slots whose events live in an arena shared with other batches of the same
  // element type (models of other densities); segments come and go via add/remove
  HipBatch(HipModel<kernelT, evalT>* model, int capacity, std::shared_ptr<EventArena<kernelT>> shared)
      : m(model), dev(model->be->device()) {
    init(capacity, shared->capacity(), shared);
  }
  void* data_ptr() const { return arena_->buf.ptr; }
  kernelT* data_elems() const { return arena_->buf.template as<kernelT>(); }
  ~HipBatch() { drop_graphs(); }
  HipBatch(const HipBatch&) = delete;
  HipBatch& operator=(const HipBatch&) = delete;

  void init(int capacity, long long elems, std::shared_ptr<EventArena<kernelT>> shared = nullptr) {
    const compute_options& o = *m->be->opts_;
    S = capacity;
    chunk = m->block * o.hip_events_per_thread;
    arena_ = shared ? shared : std::make_shared<EventArena<kernelT>>(dev, elems);
    total_elems = elems;
    sizes.assign(S, 0);
    offsets.assign(S, 0);
    active.assign(S, 0);
    segs_h.assign(S, mf_segment{});
    theta_h.alloc_mapped(dev, std::max(1, S * m->ntheta) * sizeof(double));
    params_d.alloc(dev, std::max(1, S * m->nparam) * sizeof(kernelT));
    hoist64_d.alloc(dev, (size_t)std::max(1, S) * std::max(1, m->nhoist) * sizeof(double));
    // ceil per segment adds at most one chunk each; balancing (append_chunks) rounds a segment
    // of nch0 >= resident/2 chunks to the nearest multiple of resident, at most doubling it
    chunk_cap = (int)(2 * (elems / m->ncol / chunk)) + 2 * S + 1 + m->resident;
    chunks_d.alloc(dev, (size_t)chunk_cap * sizeof(mf_chunk));
    const int nmax = std::max(std::max(1, m->nmom), std::max(m->nacc[EVAL_LH], std::max(m->nacc[EVAL_GRAD], m->nacc[EVAL_HESS])));
    partials_d.alloc(dev, (size_t)chunk_cap * nmax * sizeof(double));
    if (m->nmom) {
      moments_d.alloc(dev, (size_t)std::max(1, S) * m->nmom * sizeof(double));
      hip_check(mf_fill_zero(dev, moments_d.ptr, moments_d.bytes), "HipBatch moments");
    }
    seg_done_d.alloc(dev, (size_t)std::max(1, S) * sizeof(unsigned int));
    hip_check(mf_fill_zero(dev, seg_done_d.ptr, seg_done_d.bytes), "HipBatch counters");
    for (int k = 1; k <= 3; k++) {
      if (m->nout[k] == 0) continue;
      segs_d[k].alloc(dev, std::max(1, S) * sizeof(mf_segment));
      segs_hk[k].alloc_mapped(dev, std::max(1, S) * sizeof(mf_segment));
      results_h[k].alloc_mapped(dev, (size_t)std::max(1, S) * m->nout[k] * sizeof(double));
      std::memset(results_h[k].ptr, 0, results_h[k].bytes);
    }
    std::memset(theta_h.ptr, 0, theta_h.bytes);
    done_h.alloc_mapped(dev, 64);
    if (m->ncol > m->nobs || m->nmom) segs_pre_h.alloc_mapped(dev, std::max(1, S) * sizeof(mf_segment));
    std::memset(done_h.ptr, 0, done_h.bytes);
    request_.assign(S, 0);
    for (int k = 1; k <= 3; k++) fused_[k] = m->nout[k] > 0 && m->kernel_fused(kernel_name(k));
    hip_check(mf_sync(dev), "HipBatch init");
  }

  // ---------------------------------------------------------------- membership
  // returns the slot, or -1 when no slot or no contiguous arena space is left
  int add_segment(long long n) {
    check_segment_events(n, "HipBatch::add_segment");
    int slot = -1;
    for (int s = 0; s < S && slot < 0; s++)
      if (!active[s]) slot = s;
    if (slot < 0) return -1;
    long long off = arena_->alloc.alloc(n * m->ncol);
    if (off < 0) return -1;
    active[slot] = 1;
    sizes[slot] = n;
    offsets[slot] = off;
    dirty = true;
    return slot;
  }
  void remove_segment(int s) {
    if (s < 0 || s >= S || !active[s]) return;
    arena_->alloc.release(offsets[s], sizes[s] * m->ncol);
    active[s] = 0;
    sizes[s] = 0;
    request_[s] = 0;
    dirty = true;
  }
  bool fits(long long n) const {
    bool slot = false;
    for (int s = 0; s < S && !slot; s++) slot = !active[s];
    return slot && arena_->alloc.largest_free() >= ArenaAllocator::round(n * m->ncol);
  }
  int active_count() const {
    int c = 0;
    for (auto a : active) c += a;
    return c;
  }
  void drop_graphs() {
    for (auto*& g : graphs)
      if (g) {
        mf_graph_destroy(g);
        g = nullptr;
      }
  }
  // chunk table over the active slots, in slot order
  void rebuild() {
    if (!dirty) return;
    std::vector<mf_chunk> ch;
    max_seg_chunks = 0;
    for (int s = 0; s < S; s++) {
      mf_segment sg{};
      sg.data_off = offsets[s];
      sg.stride = sizes[s];
      sg.n = (int)sizes[s];
      sg.param_off = s * m->nparam;
      sg.mode = 0;
      if (active[s]) append_chunks(ch, sg, s, sizes[s], chunk, m->resident);
      max_seg_chunks = std::max(max_seg_chunks, sg.chunk_end - sg.chunk_begin);
      segs_h[s] = sg;
    }
    if ((int)ch.size() > chunk_cap) throw std::runtime_error("HipBatch: chunk table over capacity");
    nchunks = (int)ch.size();
    if (nchunks) {
      // a previous copy out of the staging buffer may still be queued: drain the stream first
      // (layout changes only)
      if (chunks_stage_.ptr) hip_check(mf_sync(dev), "HipBatch chunks");
      chunks_stage_.alloc(dev, ch.size() * sizeof(mf_chunk));
      std::memcpy(chunks_stage_.ptr, ch.data(), ch.size() * sizeof(mf_chunk));
      hip_check(mf_memcpy_h2d(dev, chunks_d.ptr, chunks_stage_.ptr, ch.size() * sizeof(mf_chunk)), "HipBatch chunks");
    }
    for (int k = 1; k <= 3; k++) {
      if (m->nout[k] == 0) continue;
      mf_segment* t = segs_hk[k].template as<mf_segment>();
      for (int s = 0; s < S; s++) {
        t[s] = segs_h[s];
        t[s].out_off = s * m->nout[k];
        t[s].mode = 0;
      }
    }
    hip_check(mf_sync(dev), "HipBatch rebuild");
    drop_graphs();
    for (int k = 0; k < 4; k++) segs_stale_[k] = true;
    dirty = false;
  }

  double* theta(int s) { return theta_h.template as<double>() + (size_t)s * m->ntheta; }
  const double* result(int s, int kind) const { return results_h[kind].template as<double>() + (size_t)s * m->nout[kind]; }
  long long data_bytes() const {
    long long b = 0;
    for (int s = 0; s < S; s++) b += active[s] ? sizes[s] * m->ncol : 0;
    return b * (long long)sizeof(kernelT);
  }

  // host FP64 columns (SoA, col_stride elements apart) -> stored type, centred
  void upload(int s, const double* cols, long long col_stride) {
    const long long n = sizes[s];
    std::vector<kernelT> tmp((size_t)n * m->nobs);
    for (int d = 0; d < m->nobs; d++)
      for (long long i = 0; i < n; i++) tmp[d * n + i] = (kernelT)(cols[d * col_stride + i] - m->centres[d]);
    hip_check(mf_memcpy_h2d(dev, data_elems() + offsets[s], tmp.data(), tmp.size() * sizeof(kernelT)), "HipBatch upload");
    precompute({s});
    hip_check(mf_sync(dev), "HipBatch upload");
  }
  // after new data in the listed slots: their event columns (precomputed event-only terms) from
  // their observables, and their data moments (mf_eval_mom: per-chunk sums and the final pass,
  // totals into moments_d[slot]); the segment table of these launches is read from mapped host
  // memory
  void precompute(const std::vector<int>& slots) {
    if ((m->ncol == m->nobs && m->nmom == 0) || slots.empty()) return;
    rebuild();
    if (nchunks == 0) return;
    mf_segment* t = segs_pre_h.template as<mf_segment>();
    for (int s = 0; s < S; s++) {
      t[s] = segs_h[s];
      t[s].mode = 0;
      t[s].out_off = s * m->nmom;
    }
    for (int s : slots) t[s].mode = active[s] ? 1 : 0;
    mf_eval_args a{};
    a.data = data_ptr();
    a.segs = (const mf_segment*)segs_pre_h.dptr;
    a.chunks = chunks_d.template as<mf_chunk>();
    a.nchunks = nchunks;
    a.nseg = S;
    a.params = params_d.ptr;
    a.partials = partials_d.template as<double>();
    if (m->ncol > m->nobs) {
      a.flags = MF_EVAL_SKIP_FINAL;
      hip_check(mf_launch_eval(m->mod, "mf_precompute", &a, grid(), m->block), "mf_precompute");
    }
    if (m->nmom) {
      a.flags = 0;
      a.results = moments_d.template as<double>();
      hip_check(mf_launch_eval(m->mod, "mf_eval_mom", &a, grid(), m->block), "mf_eval_mom");
    }
    hip_check(mf_sync(dev), "HipBatch precompute");
  }
  // stored columns of segment s back as FP64 observables (centre added back)
  void download(int s, double* cols) {
    const long long n = sizes[s];
    std::vector<kernelT> tmp((size_t)n * m->nobs);
    hip_check(mf_memcpy_d2h(dev, tmp.data(), data_elems() + offsets[s], tmp.size() * sizeof(kernelT)), "HipBatch download");
    hip_check(mf_sync(dev), "HipBatch download");
    for (int d = 0; d < m->nobs; d++)
      for (long long i = 0; i < n; i++) cols[d * n + i] = (double)tmp[d * n + i] + m->centres[d];
  }

  // Kernel parameters from theta, and the segment table of `kind` staged to the device.  Direct
  // launches pass a small theta by value and copy the table only when its modes changed since
  // the last copy, so the kernel reads nothing from host memory; a launch for graph capture
  // reads both from mapped host memory at each replay.
  void prep(int kind = 0, bool direct = true) {
    const bool copy = kind > 0 && (!direct || segs_stale_[kind]);
    const mf_segment* src = copy ? (const mf_segment*)segs_hk[kind].dptr : nullptr;
    mf_segment* dst = copy ? segs_d[kind].template as<mf_segment>() : nullptr;
    if (direct && (long long)S * m->ntheta <= MF_PREP_INLINE_THETA)
      hip_check(mf_launch_prep_inline(m->mod, theta_h.template as<double>(), m->ntheta, S, params_d.ptr,
                                      hoist64_d.template as<double>(), src, dst),
                "mf_launch_prep_inline");
    else
      hip_check(mf_launch_prep(m->mod, (const double*)theta_h.dptr, params_d.ptr, hoist64_d.template as<double>(), nullptr, S, src, dst),
                "mf_launch_prep");
    if (copy) segs_stale_[kind] = false;
  }

  // On-device accept-reject into the listed slots (all active slots when `slots` is empty).
  // theta_gen: ntheta generation values per listed slot; seeds: one per listed slot.  Slot
  // s is split into streams of events_per_stream events; stream k starts from the seed's
  // xoshiro128++ state advanced by k jump()s, computed on the device by jump-ahead
  // (random.hh xoshiro_jump_powers).  Returns the total number of proposals.
  unsigned long long generate(const std::vector<double>& theta_gen, const std::vector<uint64_t>& seeds, int events_per_stream = 64,
                              std::vector<int> slots = {}) {
    rebuild();
    if (slots.empty())
      for (int s = 0; s < S; s++)
        if (active[s]) slots.push_back(s);
    const int L = (int)slots.size();
    for (int i = 0; i < L; i++) std::memcpy(theta(slots[i]), &theta_gen[(size_t)i * m->ntheta], m->ntheta * sizeof(double));
    prep();
    // per listed slot: its streams' prefix sum (the stream table itself is built on the device)
    std::vector<long long> first(1, 0), counts;
    std::vector<int> slot_of;
    std::vector<unsigned int> base(4 * (size_t)S, 0u);
    std::vector<double> mp(S, 1.0);
    long long max_streams = 1;
    for (int i = 0; i < L; i++) {
      const int s = slots[i];
      mp[s] = m->maxprob(&theta_gen[(size_t)i * m->ntheta]);
      Xoshiro128pp r;
      r.setSeed(seeds[i]);
      for (int w = 0; w < 4; w++) base[4 * (size_t)s + w] = r.getSeed(w);
      const long long ns = (sizes[s] + events_per_stream - 1) / events_per_stream;
      max_streams = std::max(max_streams, ns);
      first.push_back(first.back() + ns);
      counts.push_back(sizes[s]);
      slot_of.push_back(s);
    }
    const long long nst = first.back();
    if (nst == 0) return 0;
    if (nst > (long long)std::numeric_limits<int>::max()) throw std::invalid_argument("HipBatch::generate: too many generator streams");
    int levels = 1;
    while ((1ll << levels) < max_streams) levels++;
    if (gen_levels_ < levels) {
      const int lv = std::max(levels, 24);
      auto pw = xoshiro_jump_powers(lv);
      gen_powers_.alloc(dev, pw.size() * sizeof(uint32_t));
      hip_check(mf_memcpy_h2d(dev, gen_powers_.ptr, pw.data(), pw.size() * sizeof(uint32_t)), "generate jump powers");
      gen_levels_ = lv;
    }
    gen_streams_.alloc(dev, (size_t)nst * sizeof(mf_gen_stream));
    gen_states_.alloc(dev, (size_t)nst * 4 * sizeof(unsigned int));
    gen_base_.alloc(dev, base.size() * sizeof(unsigned int));
    gen_maxprob_.alloc(dev, S * sizeof(double));
    gen_segs_.alloc(dev, S * sizeof(mf_segment));
    gen_props_.alloc(dev, S * sizeof(unsigned long long));
    // the tables in one pinned staging buffer (8-byte aligned parts), then the proposal counts:
    // per-slot stream prefix sums, event counts and slot numbers, seeds, envelopes, segments
    const size_t b_first = (L + 1) * sizeof(long long), b_cnt = L * sizeof(long long), b_slot = (L * sizeof(int) + 7) & ~(size_t)7,
                 b_base = (base.size() * sizeof(unsigned int) + 7) & ~(size_t)7, b_mp = S * sizeof(double),
                 b_sg = ((S * sizeof(mf_segment)) + 7) & ~(size_t)7, b_pr = S * sizeof(unsigned long long);
    gen_slots_.alloc(dev, b_first + b_cnt + b_slot);
    gen_stage_.alloc(dev, b_first + b_cnt + b_slot + b_base + b_mp + b_sg + b_pr);
    char* stage = (char*)gen_stage_.ptr;
    std::memcpy(stage, first.data(), b_first);
    std::memcpy(stage + b_first, counts.data(), b_cnt);
    std::memcpy(stage + b_first + b_cnt, slot_of.data(), L * sizeof(int));
    char* rest = stage + b_first + b_cnt + b_slot;
    std::memcpy(rest, base.data(), base.size() * sizeof(unsigned int));
    std::memcpy(rest + b_base, mp.data(), b_mp);
    std::memcpy(rest + b_base + b_mp, segs_h.data(), S * sizeof(mf_segment));
    unsigned long long* props_h = (unsigned long long*)(rest + b_base + b_mp + b_sg);
    char* slots_d = (char*)gen_slots_.ptr;
    hip_check(mf_memcpy_h2d(dev, slots_d, stage, b_first + b_cnt + b_slot), "generate slot table");
    hip_check(mf_memcpy_h2d(dev, gen_base_.ptr, rest, base.size() * sizeof(unsigned int)), "generate seeds");
    hip_check(mf_memcpy_h2d(dev, gen_maxprob_.ptr, rest + b_base, b_mp), "generate maxprob");
    hip_check(mf_memcpy_h2d(dev, gen_segs_.ptr, rest + b_base + b_mp, S * sizeof(mf_segment)), "generate segs");
    hip_check(mf_fill_zero(dev, gen_props_.ptr, S * sizeof(unsigned long long)), "generate proposals");
    hip_check(mf_launch_expand_streams(dev, (const long long*)slots_d, (const int*)(slots_d + b_first + b_cnt),
                                       (const long long*)(slots_d + b_first), L, events_per_stream,
                                       gen_streams_.template as<mf_gen_stream>(), (int)nst),
              "mf_launch_expand_streams");
    hip_check(mf_launch_seed_streams(dev, gen_powers_.template as<unsigned int>(), levels, gen_streams_.template as<mf_gen_stream>(),
                                     (int)nst, events_per_stream, gen_base_.template as<unsigned int>(),
                                     gen_states_.template as<unsigned int>()),
              "mf_launch_seed_streams");
    mf_gen_args g{};
    g.data = data_ptr();
    g.segs = gen_segs_.template as<mf_segment>();
    g.streams = gen_streams_.template as<mf_gen_stream>();
    g.nstreams = (int)nst;
    g.states = gen_states_.template as<unsigned int>();
    g.params = params_d.ptr;
    g.maxprob = gen_maxprob_.template as<double>();
    g.proposals = gen_props_.template as<unsigned long long>();
    const int block = 64;
    hip_check(mf_launch_generate(m->mod, &g, (int)((nst + block - 1) / block), block), "mf_launch_generate");
    hip_check(mf_memcpy_d2h(dev, props_h, gen_props_.ptr, b_pr), "generate proposals");
    hip_check(mf_sync(dev), "generate");
    precompute(slots);
    unsigned long long tot = 0;
    for (int k = 0; k < S; k++) tot += props_h[k];
    return tot;
  }

  void request(int s, int kind) { request_[s] = kind; }
  int requested(int s) const { return request_[s]; }

  static const char* kernel_name(int kind) {
    return kind == EVAL_LH ? "mf_eval_lh" : (kind == EVAL_GRAD ? "mf_eval_grad" : "mf_eval_hess");
  }
  int grid() const {
    const int g = m->grid_opt;
    return g > 0 ? std::min(g, nchunks) : nchunks;
  }
  mf_eval_args eval_args(int kind) const {
    mf_eval_args a{};
    a.data = data_ptr();
    a.segs = segs_d[kind].template as<mf_segment>();
    a.chunks = chunks_d.template as<mf_chunk>();
    a.nchunks = nchunks;
    a.nseg = S;
    a.params = params_d.ptr;
    a.partials = partials_d.template as<double>();
    a.results = (double*)results_h[kind].dptr;
    a.hoist64 = hoist64_d.template as<double>();
    a.moments = m->nmom ? moments_d.template as<double>() : nullptr;
    if (max_seg_chunks <= 1) a.flags |= MF_EVAL_SKIP_FINAL;
    else if (fused_[kind]) a.seg_done = seg_done_d.template as<unsigned int>();
    return a;
  }
  // just the evaluation kernels of `kind` (segment table and parameters already on device);
  // main_only launches the per-chunk kernel without any reduction of multi-chunk segments (for
  // timing the event loop alone)
  void launch_kernel(int kind, bool main_only = false, int done_value = 0, bool by_value = false) {
    if (nchunks == 0) return;
    mf_eval_args a = eval_args(kind);
    if (by_value) {
      // [theta, hoisted] of segment 0 in the argument block, hoisted terms computed here
      const double* th = theta(0);
      double h[MF_INLINE_HOIST];
      m->hoist_prog.run(th, h);
      kernelT* p = reinterpret_cast<kernelT*>(a.pinline);
      for (int k = 0; k < m->ntheta; k++) p[k] = (kernelT)th[k];
      for (int j = 0; j < m->nhoist; j++) {
        p[m->ntheta + j] = (kernelT)h[j];
        a.hinline[j] = h[j];
      }
      a.inline_params = 1;
    }
    if (main_only) {
      a.flags |= MF_EVAL_SKIP_FINAL;
      a.seg_done = nullptr;
    }
    if (done_value) {
      a.done_flag = (int*)done_h.dptr;
      a.done_value = done_value;
    }
    hip_check(mf_launch_eval(m->mod, kernel_name(kind), &a, grid(), m->block), "mf_launch_eval");
  }
  // the full per-call sequence of one kind: prep (theta and segment modes from mapped host
  // memory) -> per-chunk kernel -> final pass (totals into mapped host memory)
  void enqueue(int kind, bool direct = true, int done_value = 0) {
    if (by_value(kind, direct)) {
      if (segs_stale_[kind]) {  // stage the segment's table entry (its mode changed) with a copy
        hip_check(mf_memcpy_h2d(dev, segs_d[kind].ptr, segs_hk[kind].template as<mf_segment>(), sizeof(mf_segment) * (size_t)S),
                  "segment table");
        segs_stale_[kind] = false;
      }
      launch_kernel(kind, false, done_value, true);
      return;
    }
    prep(kind, direct);
    launch_kernel(kind, false, done_value);
  }
  // a one-segment direct call: its kernels take theta and the host-computed hoisted terms by
  // value (no prep kernel)
  bool by_value(int kind, bool direct) const { return direct && S == 1 && m->host_params && kind > 0; }
  // a one-segment batch whose totals the final kernel reduces, or which is one chunk: the last
  // kernel signals completion through a flag in mapped host memory, which the host polls (a
  // stream synchronisation costs several microseconds more)
  bool can_poll(int kind) const { return S == 1 && (nchunks == 1 || (nchunks > 0 && max_seg_chunks > 1 && !fused_[kind])); }
  // wait for done_value in the flag: a bounded spin, then the stream (which reports any error)
  void wait_done(int done_value) {
    const int* f = done_h.template as<int>();
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 1;; spin++) {
      if (__atomic_load_n(f, __ATOMIC_ACQUIRE) == done_value) {
        // every kQueryEvery calls a non-blocking check of the stream: without one the runtime
        // keeps every launch's bookkeeping and reclaims it all at once every ~1000 launches (two
        // stalls of 100-250 us inside launch calls every ~500 calls).  Measured at N = 10^7 (one
        // call ~20.6 us): every 64 calls the check costs ~20 us and no stall remains; every 8
        // it costs ~15 us each time (+1.9 us per call), every 256 the stalls come back in pairs
        if (++polled_ % kQueryEvery == 0) {
          int idle = 0;
          hip_check(mf_query(dev, &idle), "HipBatch run");
        }
        return;
      }
      if ((spin & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(kPollMicros)) break;
    }
    hip_check(mf_sync(dev), "HipBatch run");
    if (__atomic_load_n(f, __ATOMIC_ACQUIRE) != done_value) throw std::runtime_error("HipBatch: completion flag not set after the stream drained");
  }
  static constexpr int kPollMicros = 200;
  static constexpr unsigned kQueryEvery = 64;
  unsigned polled_ = 0;
  void set_modes(int kind, bool all) {
    mf_segment* t = segs_hk[kind].template as<mf_segment>();
    for (int s = 0; s < S; s++) {
      const int mode = active[s] && (all || request_[s] == kind) ? 1 : 0;
      if (t[s].mode != mode) {
        t[s].mode = mode;
        segs_stale_[kind] = true;
      }
    }
  }
  // evaluate `kind` on the requested (or all active) segments; as a replayed hipGraph when use_graph
  void run(int kind, bool all, bool use_graph) {
    if (m->nout[kind] == 0) throw std::runtime_error("HipBatch::run: kind not compiled");
    rebuild();
    set_modes(kind, all);
    if (use_graph) {
      if (!graphs[kind]) {
        hip_check(mf_sync(dev), "graph capture");
        hip_check(mf_graph_begin(dev), "mf_graph_begin");
        enqueue(kind, false);
        hip_check(mf_graph_end(dev, &graphs[kind]), "mf_graph_end");
      }
      hip_check(mf_graph_launch(graphs[kind]), "mf_graph_launch");
    } else if (can_poll(kind)) {
      const int v = ++done_seq_ > 0 ? done_seq_ : (done_seq_ = 1);
      enqueue(kind, true, v);
      wait_done(v);
      return;
    } else {
      enqueue(kind);
    }
    hip_check(mf_sync(dev), "HipBatch run");
  }
  // Per listed slot, at the slot's theta in `thetas` (ntheta each): the events whose per-event
  // log-density terms are not all finite, the smallest such event index (-1: none), and the
  // events with an observable outside its range or not a number -- the diagnostic kernel
  // mf_check_lh, run for fits whose sums came back non-finite (the reference's NaN check of
  // its host sum, fitter.hh:1197-1204).  Kernel parameters of every slot are recomputed from
  // the mapped theta first, as before any launch.
  struct Diagnosis {
    long long nonfinite = 0, first = -1, outside = 0;
  };
  std::vector<Diagnosis> diagnose(const std::vector<int>& slots, const std::vector<double>& thetas) {
    if (thetas.size() < slots.size() * (size_t)m->ntheta) throw std::invalid_argument("HipBatch::diagnose: fewer theta values than slots");
    rebuild();
    std::vector<Diagnosis> r(slots.size());
    std::vector<mf_segment> t(segs_h);
    for (auto& e : t) e.mode = 0;
    for (size_t i = 0; i < slots.size(); i++) {
      const int s = slots[i];
      if (s < 0 || s >= S || !active[s]) throw std::invalid_argument("HipBatch::diagnose: slot " + std::to_string(s) + " holds no dataset");
      t[s].mode = 1;
      std::memcpy(theta(s), &thetas[i * (size_t)m->ntheta], m->ntheta * sizeof(double));
    }
    if (slots.empty() || nchunks == 0) return r;
    std::vector<unsigned int> c(3 * (size_t)S, 0u);
    for (int s = 0; s < S; s++) c[3 * (size_t)s + 1] = 0xffffffffu;
    check_segs_.alloc(dev, t.size() * sizeof(mf_segment));
    check_out_.alloc(dev, c.size() * sizeof(unsigned int));
    hip_check(mf_memcpy_h2d(dev, check_segs_.ptr, t.data(), t.size() * sizeof(mf_segment)), "diagnose segments");
    hip_check(mf_memcpy_h2d(dev, check_out_.ptr, c.data(), c.size() * sizeof(unsigned int)), "diagnose counters");
    hip_check(mf_launch_prep(m->mod, (const double*)theta_h.dptr, params_d.ptr, hoist64_d.template as<double>(), nullptr, S, nullptr, nullptr),
              "mf_launch_prep");
    mf_eval_args a{};
    a.data = data_ptr();
    a.segs = check_segs_.template as<mf_segment>();
    a.chunks = chunks_d.template as<mf_chunk>();
    a.nchunks = nchunks;
    a.nseg = S;
    a.params = params_d.ptr;
    a.store = check_out_.ptr;
    hip_check(mf_launch_check(m->mod, "mf_check_lh", &a, grid(), m->block), "mf_launch_check");
    hip_check(mf_memcpy_d2h(dev, c.data(), check_out_.ptr, c.size() * sizeof(unsigned int)), "diagnose counters");
    hip_check(mf_sync(dev), "diagnose");
    for (size_t i = 0; i < slots.size(); i++) {
      const size_t s = (size_t)slots[i];
      r[i].nonfinite = c[3 * s];
      r[i].first = c[3 * s + 1] == 0xffffffffu ? -1 : (long long)c[3 * s + 1];
      r[i].outside = c[3 * s + 2];
    }
    return r;
  }

  // one launch per kind that any segment requested; requests are cleared
  int run_requested(bool use_graph) {
    int launches = 0;
    for (int k = 1; k <= 3; k++) {
      bool any = false;
      for (int s = 0; s < S && !any; s++) any = request_[s] == k;
      if (!any) continue;
      run(k, false, use_graph);
      launches++;
    }
    std::fill(request_.begin(), request_.end(), 0);
    return launches;
  }
};

}  // namespace morefit
#endif

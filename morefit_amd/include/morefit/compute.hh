/**
 * @file compute.hh
 * The compute boundary: compute_options, ComputeBackend and the abstract ComputeBlock.
 * Same interface as the reference compute.hh:30-149, so fitter/generator/toystudy stay
 * templated on <backendT, computeT> and any backend is a drop-in.  Added fields: hip_*
 * (device, launch shape, precision knobs of the MI355X backend) and cpu_* (the host-compiled
 * CPU backend that follows the reference LLVM backend's semantics).
 */
#ifndef MOREFIT_COMPUTE_HH
#define MOREFIT_COMPUTE_HH

#include <cstdint>
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "eventvector.hh"
#include "graph.hh"
#include "parametervector.hh"

namespace morefit {

struct compute_options {
  // general
  bool print_kernel = false;
  int print_level = 0;
  // reference OpenCL options (kept for source compatibility)
  unsigned int opencl_platform = 0;
  unsigned int opencl_device = 0;
  unsigned int opencl_workgroup_size = 256;
  int opencl_max_generator_workitems = 8192;
  int opencl_max_compute_workitems = -1;
  unsigned int opencl_kahan_reduction_factor = 32;
  // reference LLVM options (llvm_nthreads also drives the CPU backend's thread count)
  unsigned int llvm_nthreads = 1;
  bool llvm_print_intermediate = false;
  bool llvm_vectorization = true;
  int llvm_vectorization_width = 4;
  std::string llvm_triple = "x86_64-pc-linux-gnu";
  std::string llvm_cpu = "x86-64-v3";
  std::string llvm_tunecpu = "x86-64";
  std::string llvm_optimization = "-O3";
  // MI355X backend
  int hip_device = 0;
  int hip_block_size = 256;          // work-group size of the evaluation kernels
  int hip_events_per_thread = 32;    // chunk = block * events_per_thread events
  int hip_grid = 0;                  // 0: one work-group per chunk, >0: persistent grid of that many
  int hip_unroll = 8;                // events per thread whose loads are issued together
  int hip_base2 = 1;                 // exp/log -> exp2/log2 rewrite: 0 never, 1 FP32 kernels, 2 FP32 and FP64
  bool hip_balance_chunks = false;
  // event-only subexpressions costing more than hip_event_cost_threshold (cost units of the
  // graph: 1 per add/multiply, 20 per transcendental) become data columns computed once per
  // upload (the reference's optimize_dimensions); each column adds sizeof(T) bytes per event
  bool hip_event_columns = false;
  // a kind's outputs as linear forms over shared per-event monomials, combined per segment in
  // FP64: 0 never, 1 when the generated code is shorter, 2 always (where the expansion exists)
  int hip_linear_outputs = 1;
  // in the linear forms, a density that is a sum of exponential terms is evaluated around the
  // term whose exponent bounds allow it: exp, log and rcp per event instead of two exp, log and
  // rcp for f Gauss + (1 - f) Exp
  bool hip_exp_pivot = true;
  // one-segment calls compute the hoisted terms on the host and pass them with theta in the
  // kernels' argument block (no prep kernel); batches prepare them on the device
  bool hip_host_params = true;
  double hip_event_cost_threshold = 10.0;
  // segment totals in the work-group that completes a segment (no final kernel).  Off: each
  // work-group's device-scope release fence writes back its XCD's L2, and on MI355X those
  // write-backs cost more than the final kernel's launch (measured 36.7 vs 19.0 us per
  // gradient call at N = 1e7)
  bool hip_fused_final = false;   // large segments: chunk count a multiple of the resident work-groups (measured: no gain)
  bool hip_fast_math = true;         // FP32: hardware exp2/log2/rcp/sqrt/sin/cos (error measured against the oracle)
  int hip_accumulate = 2;            // FP32 per-lane sums: 0 FP64 per event, 1 Kahan FP32, 2 FP32 per unrolled group + FP64
  // observables as offsets from the range centre: FP32 kernels store them so, FP64 kernels
  // compute them per event from the data as given
  bool hip_centre_observables = true;
  // evaluation kernels issue the loads of their next event batch before computing the current
  // one (two register sets; dropped per kernel where the registers do not fit)
  bool hip_prefetch = true;
  int hip_generator_events_per_stream = 64;
  std::string hip_compile_options;   // extra hipRTC options
  // host-compiled CPU backend
  std::string cpu_compiler = "g++";
  std::string cpu_flags = "-O3 -march=native -fno-math-errno";

  void print() const {
    const int w = 40;
    std::cout << "COMPUTE OPTIONS\n"
              << std::setw(w) << std::left << "  print level " << print_level << "\n"
              << std::setw(w) << std::left << "  hip device " << hip_device << "\n"
              << std::setw(w) << std::left << "  hip block size " << hip_block_size << "\n"
              << std::setw(w) << std::left << "  hip events per thread " << hip_events_per_thread << "\n"
              << std::setw(w) << std::left << "  hip grid " << hip_grid << "\n"
              << std::setw(w) << std::left << "  hip fast math " << (hip_fast_math ? "YES" : "NO") << "\n"
              << std::setw(w) << std::left << "  hip accumulate "
              << (hip_accumulate == 1 ? "FP32 compensated" : hip_accumulate == 2 ? "FP32 groups + FP64" : "FP64") << "\n"
              << std::setw(w) << std::left << "  cpu threads " << llvm_nthreads << std::endl;
  }
};

class ComputeBackend {
 public:
  compute_options* opts_;
  virtual ~ComputeBackend() = default;
  // events must be padded to a multiple of this (0: no padding; both backends here mask tails)
  virtual int required_padding() const { return 0; }

 protected:
  explicit ComputeBackend(compute_options* opts) : opts_(opts) {}
};

template <typename kernelT, typename evalT, typename seedT = uint32_t>
class ComputeBlock {
 public:
  virtual ~ComputeBlock() = default;
  virtual bool SetupInputBuffer(unsigned long int nbytes) = 0;
  virtual bool SetupSeedBuffer(unsigned long int nbytes) = 0;
  // alias the input (use_data_in) or output buffer of another block as this block's input
  virtual bool SetupInputBuffer(ComputeBlock<kernelT, evalT>* input, bool use_data_in = false) = 0;
  virtual bool SetupParameterBuffer(unsigned long int nbytes) = 0;
  virtual bool SetupOutputBuffer(unsigned long int nbytes) = 0;
  virtual bool SetupKahanBuffer(unsigned long int nbytes) { return false; }
  virtual bool SetNevents(int nevents, int nevents_padded = -1) { return false; }
  virtual bool CopyToInputBuffer(const EventVector<kernelT, evalT>& data) = 0;
  virtual bool CopyToSeedBuffer(const EventVector<seedT, evalT>& data) = 0;
  virtual bool CopyToParameterBuffer(const std::vector<kernelT>& params) = 0;
  virtual bool CopyFromOutputBuffer(EventVector<kernelT, evalT>& data) = 0;
  virtual bool CopyFromKahanBuffer(std::vector<evalT>& kahan_sums) { return false; }
  virtual bool MakeGenerateKernel(std::string name, unsigned int nevents, std::vector<dimension<evalT>> input_signature,
                                  std::vector<dimension<evalT>> output_signature, const std::vector<std::string>& params,
                                  const std::vector<std::unique_ptr<ComputeGraphNode<kernelT, evalT>>>& graphs, evalT maxprob = 1.0) {
    return true;
  }
  virtual bool MakeComputeKernel(std::string name, unsigned int nevents, std::vector<dimension<evalT>> input_signature,
                                 std::vector<dimension<evalT>> output_signature, const std::vector<std::string>& params,
                                 const std::vector<std::unique_ptr<ComputeGraphNode<kernelT, evalT>>>& graphs,
                                 bool kahan_summation = false) = 0;
  virtual bool SubmitKernel() = 0;
  virtual bool Finish() = 0;
  virtual bool ProvidesKahan() { return false; }
  // events whose per-event outputs of the current summing kernel are not all finite at the
  // current parameter buffer, and the smallest such index (-1: none); false when the block
  // cannot tell (no counterpart in the reference, which prints its per-event buffer instead)
  virtual bool CountNonFinite(long long* count, long long* first) { return false; }
};

}  // namespace morefit
#endif

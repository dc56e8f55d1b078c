/**
 * @file generator.hh
 * Accept-reject event generation behind the reference's generator interface
 * (generator.hh:30-175): generator_options, and generator<kernelT, evalT, backendT, computeT>
 * whose generate() fills an EventVector with events of a PDF at the current parameter values.
 *
 *  on_accelerator  the seeds of per-stream xoshiro128++ generators are drawn from the caller's
 *                  Xoshiro128pp by jump() (generator.hh:96-132) and the block's generate kernel
 *                  proposes and accepts on the device (or in the CPU backend's threads).  A
 *                  backend that runs one stream per host thread (CPUBackend, the reference's
 *                  LLVMBackend branch) takes llvm_nthreads seeds; any other backend takes one
 *                  seed per work item, each work item owning nevents / max_generator_workitems
 *                  + 1 consecutive events when there are more events than
 *                  opencl_max_generator_workitems (generator.hh:114-132, compute_opencl.hh:
 *                  436-440).  The caller's generator advances by one jump() per seed, so
 *                  successive calls (toys) draw independent streams.
 *  on_host         proposals and the accept-reject uniforms from the caller's RandomGenerator,
 *                  in the reference's order (all proposals of a chunk dimension by dimension,
 *                  then one uniform per proposal); the block evaluates the density of a chunk
 *                  (generator.hh:176-262).
 *
 * regen = true reuses the kernel built by an earlier call (so the density keeps the parameter
 * values and the envelope of that call), as the reference's toy study does for toys after the
 * first (toystudy.hh:74-77).
 */
#ifndef MOREFIT_GENERATOR_HH
#define MOREFIT_GENERATOR_HH

#include <chrono>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "compute.hh"
#include "compute_cpu.hh"
#include "eventvector.hh"
#include "parametervector.hh"
#include "pdf.hh"
#include "random.hh"

namespace morefit {

struct generator_options {
  enum randomization_type { on_accelerator, on_host };
  randomization_type rndtype = on_accelerator;
  double chunksize_multiplier = 1.0;
  int print_level = 2;
  void print() const {
    if (print_level <= 0) return;
    const int w = 40;
    std::cout << "GENERATOR OPTIONS\n"
              << std::setw(w) << std::left << "  randomization type " << (rndtype == on_accelerator ? "ON ACCELERATOR" : "ON HOST") << "\n"
              << std::setw(w) << std::left << "  chunksize multiplier " << chunksize_multiplier << "\n"
              << std::setw(w) << std::left << "  print level " << print_level << std::endl;
  }
};

template <typename kernelT, typename evalT, typename backendT, typename computeT, typename seedT = uint32_t>
class generator {
  generator_options* opts_;
  backendT* backend_;
  computeT block_;
  RandomGenerator* rnd_;
  std::vector<dimension<evalT>> seed_dims_;
  double last_ms_ = 0.0;

  // the density at the current values of all parameters, as the kernel's only graph
  static std::vector<GraphPtr<kernelT, evalT>> density(PDF<kernelT, evalT>* prob, const std::vector<parameter<evalT>*>& params) {
    std::vector<std::string> names;
    std::vector<double> values;
    for (auto* p : params) {
      names.push_back(p->get_name());
      values.push_back(p->get_value());
    }
    std::vector<GraphPtr<kernelT, evalT>> g;
    g.emplace_back(prob->prob_normalised()->substitute(names, values)->simplify());
    return g;
  }

 public:
  generator(generator_options* opts, backendT* backend, RandomGenerator* rnd) : opts_(opts), backend_(backend), block_(backend), rnd_(rnd) {
    for (int i = 0; i < 4; i++) seed_dims_.push_back(dimension<evalT>("rnd_seed" + std::to_string(i)));
  }
  double last_ms() const { return last_ms_; }

  bool generate(unsigned int nevents, PDF<kernelT, evalT>* prob, std::vector<parameter<evalT>*> params, EventVector<kernelT, evalT>& result,
                bool regen = false) {
    const auto t0 = std::chrono::steady_clock::now();
    if (nevents > (unsigned)std::numeric_limits<int>::max())
      throw std::invalid_argument("generator: " + std::to_string(nevents) + " events do not fit the blocks' int event counts");
    const evalT maxprob = prob->get_max();
    const int pad = backend_->required_padding();
    if (opts_->rndtype == generator_options::on_accelerator) {
      result.resize(nevents);
      if (pad > 0 && nevents % pad != 0) result.set_padding(true, pad);
      auto* xoshiro = dynamic_cast<Xoshiro128pp*>(rnd_);
      if (!xoshiro) throw std::runtime_error("generator: generation on the accelerator needs a Xoshiro128pp random generator");
      std::vector<dimension<evalT>*> sd;
      for (auto& d : seed_dims_) sd.push_back(&d);
      EventVector<seedT, evalT> seeds(sd);
      size_t nseeds;
      if (dynamic_cast<CPUBackend*>(backend_) != nullptr) {
        nseeds = std::max(1u, backend_->opts_->llvm_nthreads);  // one stream per host thread
      } else {
        nseeds = result.nevents_padded();
        const int maxw = backend_->opts_->opencl_max_generator_workitems;
        if (maxw > 0 && (long long)nseeds > maxw) {
          const size_t per = result.nevents_padded() / (size_t)maxw + 1;
          nseeds = result.nevents_padded() / per + 1;
        }
      }
      seeds.resize(nseeds);
      for (size_t i = 0; i < nseeds; i++) {
        for (unsigned j = 0; j < 4; j++) seeds(i, j) = (seedT)xoshiro->getSeed((int)j);
        xoshiro->jump();
      }
      block_.SetupSeedBuffer(seeds.buffer_size());
      block_.SetupOutputBuffer(result.buffer_size());
      block_.SetNevents((int)result.nevents(), (int)result.nevents_padded());
      if (!regen) {
        auto g = density(prob, params);
        block_.MakeGenerateKernel("gen_kernel", nevents, seeds.copy_dimensions(), result.copy_dimensions(), {}, g, maxprob);
      }
      block_.Finish();
      block_.CopyToSeedBuffer(seeds);
      block_.SubmitKernel();
      block_.Finish();
      block_.CopyFromOutputBuffer(result);
    } else {
      result.resize(nevents);
      unsigned int chunk = std::max(1u, (unsigned int)(nevents * opts_->chunksize_multiplier));
      if (pad > 0 && chunk % pad != 0) chunk = (chunk / pad + 1) * pad;
      if (pad > 0) result.set_padding(true, pad);
      EventVector<kernelT, evalT> props(result.get_dimensions(), chunk, pad > 0, pad);
      dimension<evalT> res("res", 0.0, maxprob);
      EventVector<kernelT, evalT> probs({&res}, chunk, pad > 0, pad);
      block_.SetupInputBuffer(props.buffer_size());
      block_.SetupOutputBuffer(probs.buffer_size());
      block_.SetNevents((int)probs.nevents(), (int)probs.nevents_padded());
      if (!regen) {
        auto g = density(prob, params);
        block_.MakeComputeKernel("gen_kernel", chunk, props.copy_dimensions(), probs.copy_dimensions(), {}, g, false);
      }
      block_.Finish();
      const auto dims = result.get_dimensions();
      size_t accepted = 0;
      while (accepted < nevents) {
        for (unsigned d = 0; d < dims.size(); d++) {
          const evalT lo = dims[d]->get_min(), width = dims[d]->get_max() - dims[d]->get_min();
          for (unsigned j = 0; j < chunk; j++) props(j, d) = (kernelT)(lo + width * rnd_->random());
        }
        block_.CopyToInputBuffer(props);
        block_.SubmitKernel();
        block_.Finish();
        block_.CopyFromOutputBuffer(probs);
        for (unsigned i = 0; i < chunk && accepted < nevents; i++)
          if (maxprob * rnd_->random() < probs(i, 0u)) {
            for (unsigned d = 0; d < dims.size(); d++) result(accepted, d) = props(i, d);
            accepted++;
          }
      }
    }
    last_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (opts_->print_level > 1) std::cout << "generation of " << nevents << " events took " << last_ms_ << " ms" << std::endl;
    return true;
  }
};

}  // namespace morefit
#endif

/**
 * @file codegen.hh
 * Turns graphs into the source of one likelihood module: functor structs (one per kernel
 * kind, all outputs of a kind emitted together with CSE), the FP64 hoisted-term function
 * mf_hoist_term, the generator's proposal box, and the kernel instantiation lines.  The same
 * functor text compiles for gfx950 through hipRTC (with the device runtime prepended by
 * libmorefit_hip) and for the host CPU backend (with host definitions of the mf_* wrappers).
 *
 * Reference counterpart: the string building in compute_opencl.hh:646-879 and
 * compute_llvm.hh:535-750 (one expression per output, unsuffixed decimal constants).
 */
#ifndef MOREFIT_CODEGEN_HH
#define MOREFIT_CODEGEN_HH

#include <cstdio>
#include <sstream>
#include <string>
#include <vector>

#include "graph.hh"

namespace morefit {

// one MF_EVAL_KERNEL instantiation: kernel name, functor, events per thread per load batch,
// whether the kernel reduces multi-chunk segments itself (else its _final kernel does), and
// whether it loads the next batch while computing the current one (two register sets)
struct EvalKernelSpec {
  std::string name, functor;
  int unroll = 1;
  bool fused = true;
  bool prefetch = true;
};

struct ModuleSource {
  bool fp32 = false;
  int nobs = 1;                            // observables (generator output, uploaded columns)
  int ncol = 0;                            // columns the evaluation kernels read: observables, then
                                           // precomputed event terms (0: same as nobs)
  int ntheta = 0;
  int nhoist = 0;
  int block = 256;
  int acc = 1;
  bool fast_math = false;
  std::vector<double> gen_lo, gen_width;  // proposal box per observable, in stored coordinates
  std::vector<double> obs_lo, obs_hi;      // ranges of the observables in stored coordinates, for the
                                           // diagnostic kernel (empty: no range test)
  std::string functors;                    // functor structs
  std::string hoist;                       // mf_hoist_term definition
  std::string event_terms;                 // mf_event_terms definition (ncol > nobs)
  std::vector<EvalKernelSpec> evals;       // evaluation kernels (unroll chosen per kernel)
  std::vector<std::string> kernels;        // other instantiation lines
  size_t statements = 0;                   // SSA statements emitted (all functors)
  int nacc[4] = {0, 0, 0, 0};              // per evaluation kind: per-event terms accumulated
  int nmom = 0;                            // data moments per segment (mf_eval_mom's outputs)

  std::string prelude() const {
    std::ostringstream s;
    s << "#define MF_T " << (fp32 ? "float" : "double") << "\n"
      << "#define MF_FP32 " << (fp32 ? 1 : 0) << "\n"
      << "#define MF_NOBS " << nobs << "\n"
      << "#define MF_NCOL " << (ncol > nobs ? ncol : nobs) << "\n"
      << "#define MF_NTHETA " << ntheta << "\n"
      << "#define MF_NHOIST " << nhoist << "\n"
      << "#define MF_BLOCK " << block << "\n"
      << "#define MF_ACC " << acc << "\n"
      << "#define MF_FAST_MATH " << (fast_math ? 1 : 0) << "\n"
      << "#define MF_NMOM " << nmom << "\n";
    s << "#define MF_GEN_PROPOSE(x, s) do {";
    for (int d = 0; d < nobs; d++) {
      double lo = d < (int)gen_lo.size() ? gen_lo[d] : 0.0, w = d < (int)gen_width.size() ? gen_width[d] : 1.0;
      char buf[160];
      if (fp32) std::snprintf(buf, sizeof(buf), " x[%d] = %af + %af * mf_uniform(s);", d, (double)(float)lo, (double)(float)w);
      else std::snprintf(buf, sizeof(buf), " x[%d] = %a + %a * mf_uniform(s);", d, lo, w);
      s << buf;
    }
    s << " } while (0)\n";
    if ((int)obs_lo.size() == nobs && (int)obs_hi.size() == nobs && nobs > 0) {
      // rounded to the kernel type as the uploaded observables are, so an event at a boundary
      // compares equal to it
      auto list = [&](const std::vector<double>& v) {
        std::string r = "{";
        for (int d = 0; d < nobs; d++) {
          char buf[64];
          if (fp32) std::snprintf(buf, sizeof(buf), "%s%af", d ? ", " : "", (double)(float)v[d]);
          else std::snprintf(buf, sizeof(buf), "%s%a", d ? ", " : "", v[d]);
          r += buf;
        }
        return r + "}";
      };
      s << "#define MF_OBS_LO " << list(obs_lo) << "\n#define MF_OBS_HI " << list(obs_hi) << "\n";
    }
    return s.str();
  }
  // generated module text: prelude, the runtime marker, then functors/hoist/kernels
  std::string device_source() const {
    std::ostringstream s;
    s << "// generated by morefit codegen.hh\n" << prelude() << "//@MF_RUNTIME@\n" << functors << "\n" << hoist << "\n" << event_terms << "\n";
    for (const auto& e : evals)
      s << "MF_EVAL_KERNEL(" << e.name << ", " << e.functor << ", " << e.unroll << ", " << (e.fused ? 1 : 0) << ", "
        << (e.prefetch ? 1 : 0) << ")\n";
    for (const auto& k : kernels) s << k << "\n";
    return s.str();
  }
};

// One term of an output's combination from the segment sums of a kind's accumulated per-event
// terms: mult * (hoist >= 0 ? h[hoist] : 1) * (mom >= 0 ? M[mom] : acc >= 0 ? S[acc] : n), h the
// segment's FP64 hoisted terms, M its data moments (sums of observable-only terms, computed once
// per dataset), n its event count.
struct CombineTerm {
  int acc = -1;
  int hoist = -1;
  double mult = 1.0;
  int mom = -1;
};
typedef std::vector<std::vector<CombineTerm>> CombineForm;  // per output

// FP64 statement of out[k] from CombineForm terms
inline std::string emit_combine_terms(const std::vector<CombineTerm>& terms) {
  if (terms.empty()) return "0.0";
  std::ostringstream s;
  for (size_t i = 0; i < terms.size(); i++) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%a", terms[i].mult);
    s << (i ? " + " : "") << buf;
    if (terms[i].hoist >= 0) s << " * h[" << terms[i].hoist << "]";
    s << " * "
      << (terms[i].mom >= 0   ? "M[" + std::to_string(terms[i].mom) + "]"
          : terms[i].acc >= 0 ? "S[" + std::to_string(terms[i].acc) + "]"
                              : std::string("n"));
  }
  return s.str();
}

// struct NAME { nacc; nout; eval(x, p, o); combine(S, h, n, M, out) }: eval computes the nacc
// per-event terms `acc` (their sums over a segment's events are accumulated) from the columns
// x[d] and kernel parameters p[k]; combine forms the nout outputs from those sums S, the
// segment's FP64 hoisted terms h, its event count n and its data moments M, in FP64 (form;
// null: out[k] = S[k]).
// shift (optional, one entry per observable): the graphs' variable d is x[d] + shift[d],
// computed once per event (offsets from the range centre for data stored as given).
inline std::string emit_shift_prologue(KernelEmitter& em, const std::vector<std::string>& observables, const std::vector<double>* shift,
                                       const std::string& type) {
  std::string pro;
  for (size_t d = 0; d < observables.size(); d++) {
    const std::string xd = "x[" + std::to_string(d) + "]";
    if (shift && d < shift->size() && (*shift)[d] != 0.0) {
      const std::string nm = "xs" + std::to_string(d);
      pro += "    const " + type + " " + nm + " = " + xd + " + " + em.literal((*shift)[d]) + ";\n";
      em.bindings[observables[d]] = nm;
      em.statements++;
      em.cost += 1.0;
    } else {
      em.bindings[observables[d]] = xd;
    }
  }
  return pro;
}

// vector_type (host CPU backend): the per-event values are of that SIMD type, one lane per
// event -- x and o point to vectors, parameters stay scalar, and every temporary goes through
// mf_bc (a broadcast of scalar subexpressions, the identity on vectors)
template <typename kernelT, typename evalT>
std::string emit_functor(const std::string& name, const std::vector<const ComputeGraphNode<kernelT, evalT>*>& acc,
                         const std::vector<std::string>& observables, const std::vector<std::string>& param_names,
                         bool fp32, size_t* statements = nullptr, const CombineForm* form = nullptr,
                         const std::vector<double>* shift = nullptr, const std::string& vector_type = "") {
  const std::string vt = vector_type.empty() ? "MF_T" : vector_type;
  // the product chains in the searched operand order and in plain frequency order: the one with
  // the lower emitted cost is kept
  auto run = [&](KernelEmitter& em, std::string& pro, std::vector<std::string>& res) {
    if (!vector_type.empty()) {
      em.T = vector_type;
      em.wrap = "mf_bc";
    }
    pro = emit_shift_prologue(em, observables, shift, vector_type.empty() ? (fp32 ? "float" : "double") : vector_type);
    for (size_t k = 0; k < param_names.size(); k++) em.bindings[param_names[k]] = "p[" + std::to_string(k) + "]";
    for (auto* g : acc) em.count_products(*g);
    for (auto* g : acc) res.push_back(em.emit(*g));
  };
  KernelEmitter em(fp32), em2(fp32);
  em2.search_order = false;
  std::string pro, pro2;
  std::vector<std::string> res, res2;
  run(em, pro, res);
  run(em2, pro2, res2);
  if (em2.cost < em.cost) {
    std::swap(em, em2);
    std::swap(pro, pro2);
    std::swap(res, res2);
  }
  const size_t nout = form ? form->size() : acc.size();
  std::ostringstream s;
  s << "struct " << name << " {\n  static constexpr int nacc = " << acc.size() << ";\n  static constexpr int nout = " << nout << ";\n"
    << "  __device__ __forceinline__ static void eval(const " << vt << "* __restrict__ x, const MF_T* __restrict__ p, " << vt
    << "* __restrict__ o) {\n"
    << "    (void)x; (void)p;\n"
    << pro << em.body.str();
  for (size_t k = 0; k < res.size(); k++) s << "    o[" << k << "] = " << (vector_type.empty() ? res[k] : "mf_bc(" + res[k] + ")") << ";\n";
  s << "  }\n  __device__ __forceinline__ static void combine(const double* S, const double* h, double n, const double* M, "
       "double* out) {\n"
    << "    (void)S; (void)h; (void)n; (void)M;\n";
  for (size_t k = 0; k < nout; k++) {
    if (form) s << "    out[" << k << "] = " << emit_combine_terms((*form)[k]) << ";\n";
    else s << "    out[" << k << "] = S[" << k << "];\n";
  }
  s << "  }\n};\n";
  if (statements) *statements += (size_t)em.statements;
  return s.str();
}

// void NAME(const double* ARG, double* buf): FP64 values of `exprs`, whose variables are
// var_names bound to ARG[k] (+ shift[k], when given)
template <typename kernelT, typename evalT>
std::string emit_fp64_function(const std::string& name, const std::string& arg, const std::vector<const ComputeGraphNode<kernelT, evalT>*>& exprs,
                               const std::vector<std::string>& var_names, const std::vector<double>* shift = nullptr) {
  KernelEmitter em(false);
  em.prefix = "mfd_";
  em.tmp = "h";
  std::string pro;
  for (size_t k = 0; k < var_names.size(); k++) {
    const std::string a = arg + "[" + std::to_string(k) + "]";
    if (shift && k < shift->size() && (*shift)[k] != 0.0) {
      const std::string nm = arg + "s" + std::to_string(k);
      pro += "  const double " + nm + " = " + a + " + " + em.literal((*shift)[k]) + ";\n";
      em.bindings[var_names[k]] = nm;
    } else {
      em.bindings[var_names[k]] = a;
    }
  }
  std::vector<std::string> res;
  for (auto* g : exprs) res.push_back(em.emit(*g));
  std::ostringstream s;
  s << "__device__ __forceinline__ void " << name << "(const double* " << arg << ", double* buf) {\n  (void)" << arg << "; (void)buf;\n"
    << pro << em.body.str();
  for (size_t k = 0; k < res.size(); k++) s << "  buf[" << k << "] = " << res[k] << ";\n";
  s << "}\n";
  return s.str();
}
// double mf_hoist_term(int k, const double* th): FP64 value of hoisted parameter-only term k,
// each term computed on its own (the prep kernels evaluate the terms side by side)
template <typename kernelT, typename evalT>
std::string emit_hoist(const std::vector<const ComputeGraphNode<kernelT, evalT>*>& exprs, const std::vector<std::string>& theta_names) {
  std::ostringstream s;
  s << "__device__ double mf_hoist_term(int k, const double* th) {\n  (void)th;\n  switch (k) {\n";
  for (size_t k = 0; k < exprs.size(); k++) {
    KernelEmitter em(false);
    em.prefix = "mfd_";
    em.tmp = "h";
    for (size_t j = 0; j < theta_names.size(); j++) em.bindings[theta_names[j]] = "th[" + std::to_string(j) + "]";
    const std::string r = em.emit(*exprs[k]);
    s << "    case " << k << ": {\n" << em.body.str() << "      return " << r << ";\n    }\n";
  }
  s << "  }\n  return 0.0;\n}\n";
  return s.str();
}
// the hoist of a module without hoisted terms
inline std::string empty_hoist() { return "__device__ double mf_hoist_term(int k, const double* th) { (void)k; (void)th; return 0.0; }\n"; }

}  // namespace morefit
#endif

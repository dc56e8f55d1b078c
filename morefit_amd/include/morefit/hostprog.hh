/**
 * @file hostprog.hh
 * Parameter-only expressions (the hoisted terms of a density) as a flat FP64 program that the
 * host runs per call.
 *
 * A one-fit evaluation computes its hoisted terms on the host and passes them with theta in the
 * kernels' argument block. This saves the parameter-preparation kernel and the kernel boundary
 * after it, about 5 us of a 30 us call on MI355X. The reference likewise evaluates its parameter
 * buffer on the host before each launch (fitter.hh:1063-1072), by walking the graphs. Here the
 * graphs are flattened once into ops with shared subexpressions computed once: an op list in
 * evaluation order, n-ary sums and products accumulated left to right from 0 and 1 as
 * ComputeGraphNode::eval does, and the same libm functions (apply_unary).
 */
#ifndef MOREFIT_HOSTPROG_HH
#define MOREFIT_HOSTPROG_HH

#include <cmath>
#include <string>
#include <unordered_map>
#include <vector>

#include "graph.hh"

namespace morefit {

template <typename kernelT, typename evalT>
class HostProgram {
 public:
  typedef ComputeGraphNode<kernelT, evalT> Node;
  HostProgram() = default;
  // expressions over the variables `vars` (bound by position at run time)
  HostProgram(const std::vector<const Node*>& exprs, const std::vector<std::string>& vars) {
    std::unordered_multimap<uint64_t, std::pair<const Node*, int>> memo;
    for (auto* e : exprs) out_.push_back(add(*e, vars, memo));
  }
  size_t outputs() const { return out_.size(); }
  size_t ops() const { return ops_.size(); }
  // out[k] = expression k with vars[j] = x[j]
  void run(const double* x, double* out) const {
    double stack[kStack];
    std::vector<double> heap;
    double* r = stack;
    if (ops_.size() > kStack) {
      heap.resize(ops_.size());
      r = heap.data();
    }
    for (size_t i = 0; i < ops_.size(); i++) {
      const Op& o = ops_[i];
      switch (o.kind) {
        case NodeKind::Constant: r[i] = o.v; break;
        case NodeKind::Variable: r[i] = o.var >= 0 ? x[o.var] : NAN; break;
        case NodeKind::Sum: {
          double s = 0.0;
          for (int a : o.a) s += r[a];
          r[i] = s;
          break;
        }
        case NodeKind::Prod: {
          double p = 1.0;
          for (int a : o.a) p *= r[a];
          r[i] = p;
          break;
        }
        case NodeKind::Pow: r[i] = std::pow(r[o.a[0]], r[o.a[1]]); break;
        case NodeKind::Conditional: r[i] = Node::cmp_holds(o.cmp, r[o.a[0]]) ? r[o.a[1]] : r[o.a[2]]; break;
        default: r[i] = Node::apply_unary(o.kind, r[o.a[0]]); break;
      }
    }
    for (size_t k = 0; k < out_.size(); k++) out[k] = r[out_[k]];
  }

 private:
  static constexpr size_t kStack = 512;
  struct Op {
    NodeKind kind = NodeKind::Constant;
    int cmp = 0;
    int var = -1;
    double v = 0.0;
    std::vector<int> a;
  };
  std::vector<Op> ops_;
  std::vector<int> out_;

  int add(const Node& n, const std::vector<std::string>& vars, std::unordered_multimap<uint64_t, std::pair<const Node*, int>>& memo) {
    const uint64_t h = structural_hash(n);
    auto range = memo.equal_range(h);
    for (auto it = range.first; it != range.second; ++it)
      if (it->second.first->equals(n)) return it->second.second;
    Op o;
    o.kind = n.kind();
    o.cmp = n.comparator();
    if (n.kind() == NodeKind::Constant) {
      o.v = n.value();
    } else if (n.kind() == NodeKind::Variable) {
      for (size_t j = 0; j < vars.size(); j++)
        if (vars[j] == n.name()) o.var = (int)j;
    } else {
      for (const auto& c : n.children_) o.a.push_back(add(*c, vars, memo));
    }
    ops_.push_back(std::move(o));
    const int idx = (int)ops_.size() - 1;
    memo.insert({h, {&n, idx}});
    return idx;
  }
};

}  // namespace morefit
#endif

/**
 * @file graph.hh
 * Symbolic compute graph of MoreFit densities: construction, exact derivatives,
 * simplification, parameter substitution, hoisting of parameter-only subexpressions,
 * host evaluation and code emission.
 *
 * API parity with the reference graph.hh: the node class names (ConstantNode, VariableNode,
 * SumNode, ProdNode, NegNode, InvNode, ExpNode, LogNode, PowNode, SqrtNode, SinNode, CosNode,
 * TanNode, AsinNode, AcosNode, AtanNode, ErfNode, ConditionalNode), the virtual interface of
 * ComputeGraphNode (graph.hh:93-230: copy, get_kernel, print, diff, substitute, simplify,
 * cost, eval, variable_in_tree, contains_other_variables, optimize_buffering_constant_terms,
 * operator==, find_in_list) and the builder functions/operators (graph.hh:2491-3010) are
 * kept, so PDF code written against the reference compiles unchanged.
 *
 * Design differences (MI355X-first):
 *  - one node representation, dispatched on a kind tag, instead of one class per operation;
 *  - correct derivative rules for asin/acos/tan (the reference omits the square root in
 *    asin/acos and uses 1+x^2 for tan, graph.hh:1554-1565, 1758-1769, 1862-1873);
 *  - emit_kernel(): all outputs of a kernel are emitted together as SSA statements with
 *    structural common-subexpression elimination (the reference emits one independent
 *    expression string per output, graph.hh:245-2290), and constants are emitted as exact
 *    hexadecimal literals (the reference prints 16 digits, graph.hh:252);
 *  - centre_observables(): rewrites x -> x' + c so an FP32 kernel works on offsets from the
 *    range centre, with parameter-only parts moved into hoisted FP64 terms.
 */
#ifndef MOREFIT_GRAPH_HH
#define MOREFIT_GRAPH_HH

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace morefit {

enum class NodeKind {
  Constant, Variable, Sum, Prod, Neg, Inv, Exp, Log, Pow, Sqrt,
  Sin, Cos, Tan, Asin, Acos, Atan, Erf, Conditional,
  Exp2, Log2  // base-2 forms, introduced only by to_base2() for FP32 kernels
};

template <typename kernelT, typename evalT> class ComputeGraphNode;
template <typename kernelT, typename evalT> using GraphPtr = std::unique_ptr<ComputeGraphNode<kernelT, evalT>>;

template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> make_node(NodeKind kind, std::vector<GraphPtr<kernelT, evalT>> children,
                                   double value = 0.0, const std::string& name = "", int cmp = 0);

template <typename T> inline std::string type_name_str() {
  if constexpr (std::is_same_v<T, float>) return "float";
  else if constexpr (std::is_same_v<T, double>) return "double";
  else if constexpr (std::is_same_v<T, long double>) return "long double";
  else return "unknown";
}

inline std::string hex_double(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%a", v);
  return buf;
}

template <typename kernelT, typename evalT>
class ComputeGraphNode {
 public:
  typedef GraphPtr<kernelT, evalT> Ptr;
  enum comparator_type { Equal, Unequal, Larger, Smaller, LargerEqual, SmallerEqual };  // all against zero
  std::vector<Ptr> children_;

 protected:
  NodeKind kind_;
  double value_ = 0.0;   // Constant
  std::string name_;     // Variable
  int cmp_ = 0;          // Conditional

 public:
  ComputeGraphNode(NodeKind kind, std::vector<Ptr> children, double value = 0.0, const std::string& name = "", int cmp = 0)
      : children_(std::move(children)), kind_(kind), value_(value), name_(name), cmp_(cmp) {}
  virtual ~ComputeGraphNode() = default;

  NodeKind kind() const { return kind_; }
  double value() const { return value_; }
  const std::string& name() const { return name_; }
  int comparator() const { return cmp_; }
  bool is_constant() const { return kind_ == NodeKind::Constant; }
  bool is_variable() const { return kind_ == NodeKind::Variable; }
  const ComputeGraphNode& child(size_t i) const { return *children_.at(i); }

  std::string kernelT_str() const { return type_name_str<kernelT>(); }
  std::string evalT_str() const { return type_name_str<evalT>(); }

  // ---------------------------------------------------------------- copy
  std::vector<Ptr> copy_children() const {
    std::vector<Ptr> r;
    r.reserve(children_.size());
    for (const auto& c : children_) r.emplace_back(c->copy());
    return r;
  }
  virtual Ptr copy() const { return make_node<kernelT, evalT>(kind_, copy_children(), value_, name_, cmp_); }

  virtual void rename_variable(const std::string& variable, const std::string& new_name) {
    if (kind_ == NodeKind::Variable && name_ == variable) name_ = new_name;
    for (auto& c : children_) c->rename_variable(variable, new_name);
  }

  // ---------------------------------------------------------------- queries
  virtual bool variable_in_tree(const std::string& variable) const {
    if (kind_ == NodeKind::Variable) return name_ == variable;
    for (const auto& c : children_)
      if (c->variable_in_tree(variable)) return true;
    return false;
  }
  // true if the tree holds a variable that is not in `variables`
  virtual bool contains_other_variables(const std::vector<std::string>& variables) const {
    if (kind_ == NodeKind::Variable) return std::find(variables.begin(), variables.end(), name_) == variables.end();
    for (const auto& c : children_)
      if (c->contains_other_variables(variables)) return true;
    return false;
  }
  // true if the tree holds any of `variables`
  bool contains_any(const std::vector<std::string>& variables) const {
    if (kind_ == NodeKind::Variable) return std::find(variables.begin(), variables.end(), name_) != variables.end();
    for (const auto& c : children_)
      if (c->contains_any(variables)) return true;
    return false;
  }
  void collect_variables(std::set<std::string>& out) const {
    if (kind_ == NodeKind::Variable) out.insert(name_);
    for (const auto& c : children_) c->collect_variables(out);
  }
  size_t node_count() const {
    size_t n = 1;
    for (const auto& c : children_) n += c->node_count();
    return n;
  }

  // structural equality
  bool equals(const ComputeGraphNode& o) const {
    if (kind_ != o.kind_ || children_.size() != o.children_.size()) return false;
    if (kind_ == NodeKind::Constant && !(value_ == o.value_ || (std::isnan(value_) && std::isnan(o.value_)))) return false;
    if (kind_ == NodeKind::Variable && name_ != o.name_) return false;
    if (kind_ == NodeKind::Conditional && cmp_ != o.cmp_) return false;
    for (size_t i = 0; i < children_.size(); i++)
      if (!children_[i]->equals(*o.children_[i])) return false;
    return true;
  }
  virtual bool operator==(const Ptr& rhs) const { return rhs && equals(*rhs); }
  virtual int find_in_list(std::vector<Ptr>& list) const {
    for (size_t i = 0; i < list.size(); i++)
      if (equals(*list[i])) return (int)i;
    return -1;
  }

  // ---------------------------------------------------------------- cost heuristic
  virtual double cost() const {
    double c = 0.0;
    for (const auto& ch : children_) c += ch->cost();
    switch (kind_) {
      case NodeKind::Constant:
      case NodeKind::Variable: return 0.0;
      case NodeKind::Sum:
      case NodeKind::Prod: return c + (double)(children_.size() - 1);
      case NodeKind::Neg: return c + 0.5;
      case NodeKind::Inv:
      case NodeKind::Sqrt: return c + 4.0;
      case NodeKind::Pow: return c + 30.0;
      case NodeKind::Conditional: return c + 1.0;
      default: return c + 20.0;  // transcendental
    }
  }

  // ---------------------------------------------------------------- host evaluation
  static int cmp_holds(int cmp, double v) {
    switch (cmp) {
      case Equal: return v == 0.0;
      case Unequal: return v != 0.0;
      case Larger: return v > 0.0;
      case Smaller: return v < 0.0;
      case LargerEqual: return v >= 0.0;
      default: return v <= 0.0;
    }
  }
  static double apply_unary(NodeKind k, double a) {
    switch (k) {
      case NodeKind::Neg: return -a;
      case NodeKind::Inv: return 1.0 / a;
      case NodeKind::Exp: return std::exp(a);
      case NodeKind::Log: return std::log(a);
      case NodeKind::Exp2: return std::exp2(a);
      case NodeKind::Log2: return std::log2(a);
      case NodeKind::Sqrt: return std::sqrt(a);
      case NodeKind::Sin: return std::sin(a);
      case NodeKind::Cos: return std::cos(a);
      case NodeKind::Tan: return std::tan(a);
      case NodeKind::Asin: return std::asin(a);
      case NodeKind::Acos: return std::acos(a);
      case NodeKind::Atan: return std::atan(a);
      case NodeKind::Erf: return std::erf(a);
      default: return NAN;
    }
  }
  // evaluates with the named variables bound; unbound variables give NaN (graph.hh:2224-2234)
  virtual evalT eval(const std::vector<std::string>& variables, const std::vector<double>& values) const {
    switch (kind_) {
      case NodeKind::Constant: return value_;
      case NodeKind::Variable: {
        for (size_t i = 0; i < variables.size(); i++)
          if (variables[i] == name_) return values.at(i);
        return NAN;
      }
      case NodeKind::Sum: {
        double s = 0.0;
        for (const auto& c : children_) s += c->eval(variables, values);
        return s;
      }
      case NodeKind::Prod: {
        double p = 1.0;
        for (const auto& c : children_) p *= c->eval(variables, values);
        return p;
      }
      case NodeKind::Pow: return std::pow((double)children_[0]->eval(variables, values), (double)children_[1]->eval(variables, values));
      case NodeKind::Conditional:
        return cmp_holds(cmp_, children_[0]->eval(variables, values)) ? children_[1]->eval(variables, values)
                                                                       : children_[2]->eval(variables, values);
      default: return apply_unary(kind_, children_[0]->eval(variables, values));
    }
  }

  // ---------------------------------------------------------------- substitution
  virtual Ptr substitute(const std::vector<std::string>& variables, const std::vector<double>& values) const {
    if (kind_ == NodeKind::Variable) {
      for (size_t i = 0; i < variables.size(); i++)
        if (variables[i] == name_) return make_node<kernelT, evalT>(NodeKind::Constant, {}, values.at(i));
      return copy();
    }
    std::vector<Ptr> ch;
    for (const auto& c : children_) ch.emplace_back(c->substitute(variables, values));
    return make_node<kernelT, evalT>(kind_, std::move(ch), value_, name_, cmp_);
  }
  // replaces every occurrence of the variable `variable` by a copy of `replacement`
  Ptr replace_variable(const std::string& variable, const ComputeGraphNode& replacement) const {
    if (kind_ == NodeKind::Variable) return name_ == variable ? replacement.copy() : copy();
    std::vector<Ptr> ch;
    for (const auto& c : children_) ch.emplace_back(c->replace_variable(variable, replacement));
    return make_node<kernelT, evalT>(kind_, std::move(ch), value_, name_, cmp_);
  }

  // ---------------------------------------------------------------- derivatives
  static Ptr K(double c) { return make_node<kernelT, evalT>(NodeKind::Constant, {}, c); }
  static Ptr U(NodeKind k, Ptr a) {
    std::vector<Ptr> v;
    v.emplace_back(std::move(a));
    return make_node<kernelT, evalT>(k, std::move(v));
  }
  static Ptr B(NodeKind k, Ptr a, Ptr b) {
    std::vector<Ptr> v;
    v.emplace_back(std::move(a));
    v.emplace_back(std::move(b));
    return make_node<kernelT, evalT>(k, std::move(v));
  }
  static Ptr P3(Ptr a, Ptr b, Ptr c) {
    std::vector<Ptr> v;
    v.emplace_back(std::move(a));
    v.emplace_back(std::move(b));
    v.emplace_back(std::move(c));
    return make_node<kernelT, evalT>(NodeKind::Prod, std::move(v));
  }

  virtual Ptr diff(std::string variable) const {
    if (!variable_in_tree(variable)) return K(0.0);
    const ComputeGraphNode* a = children_.empty() ? nullptr : children_[0].get();
    switch (kind_) {
      case NodeKind::Variable: return K(name_ == variable ? 1.0 : 0.0);
      case NodeKind::Constant: return K(0.0);
      case NodeKind::Sum: {
        std::vector<Ptr> terms;
        for (const auto& c : children_)
          if (c->variable_in_tree(variable)) terms.emplace_back(c->diff(variable));
        return make_node<kernelT, evalT>(NodeKind::Sum, std::move(terms));
      }
      case NodeKind::Prod: {  // product rule over all factors
        std::vector<Ptr> terms;
        for (size_t i = 0; i < children_.size(); i++) {
          if (!children_[i]->variable_in_tree(variable)) continue;
          std::vector<Ptr> f;
          for (size_t j = 0; j < children_.size(); j++) f.emplace_back(i == j ? children_[j]->diff(variable) : children_[j]->copy());
          terms.emplace_back(make_node<kernelT, evalT>(NodeKind::Prod, std::move(f)));
        }
        return make_node<kernelT, evalT>(NodeKind::Sum, std::move(terms));
      }
      case NodeKind::Neg: return U(NodeKind::Neg, a->diff(variable));
      case NodeKind::Inv:  // -(a') / a^2
        return U(NodeKind::Neg, P3(a->diff(variable), U(NodeKind::Inv, a->copy()), U(NodeKind::Inv, a->copy())));
      case NodeKind::Exp: return B(NodeKind::Prod, copy(), a->diff(variable));
      case NodeKind::Log: return B(NodeKind::Prod, a->diff(variable), U(NodeKind::Inv, a->copy()));
      case NodeKind::Sqrt: return P3(K(0.5), a->diff(variable), U(NodeKind::Inv, copy()));
      case NodeKind::Sin: return B(NodeKind::Prod, U(NodeKind::Cos, a->copy()), a->diff(variable));
      case NodeKind::Cos: return U(NodeKind::Neg, B(NodeKind::Prod, U(NodeKind::Sin, a->copy()), a->diff(variable)));
      case NodeKind::Tan: {  // (1 + tan^2 a) a'
        std::vector<Ptr> s;
        s.emplace_back(K(1.0));
        s.emplace_back(B(NodeKind::Prod, copy(), copy()));
        return B(NodeKind::Prod, make_node<kernelT, evalT>(NodeKind::Sum, std::move(s)), a->diff(variable));
      }
      case NodeKind::Asin:
      case NodeKind::Acos: {  // +-a' / sqrt(1 - a^2)
        std::vector<Ptr> s;
        s.emplace_back(K(1.0));
        s.emplace_back(U(NodeKind::Neg, B(NodeKind::Prod, a->copy(), a->copy())));
        Ptr r = B(NodeKind::Prod, a->diff(variable), U(NodeKind::Inv, U(NodeKind::Sqrt, make_node<kernelT, evalT>(NodeKind::Sum, std::move(s)))));
        return kind_ == NodeKind::Asin ? std::move(r) : U(NodeKind::Neg, std::move(r));
      }
      case NodeKind::Atan: {  // a' / (1 + a^2)
        std::vector<Ptr> s;
        s.emplace_back(K(1.0));
        s.emplace_back(B(NodeKind::Prod, a->copy(), a->copy()));
        return B(NodeKind::Prod, a->diff(variable), U(NodeKind::Inv, make_node<kernelT, evalT>(NodeKind::Sum, std::move(s))));
      }
      case NodeKind::Erf:  // 2/sqrt(pi) exp(-a^2) a'
        return P3(K(1.1283791670955126), U(NodeKind::Exp, U(NodeKind::Neg, B(NodeKind::Prod, a->copy(), a->copy()))), a->diff(variable));
      case NodeKind::Exp2:  // ln2 2^a a'
        return P3(K(0.6931471805599453), copy(), a->diff(variable));
      case NodeKind::Log2:  // a' / (a ln2)
        return P3(K(1.4426950408889634), a->diff(variable), U(NodeKind::Inv, a->copy()));
      case NodeKind::Pow: {
        const auto& e = children_[1];
        if (!e->variable_in_tree(variable)) {  // e * a^(e-1) * a'
          std::vector<Ptr> s;
          s.emplace_back(e->copy());
          s.emplace_back(K(-1.0));
          return P3(e->copy(), B(NodeKind::Pow, a->copy(), make_node<kernelT, evalT>(NodeKind::Sum, std::move(s))), a->diff(variable));
        }
        // a^e (e' log a + e a'/a)
        std::vector<Ptr> s;
        s.emplace_back(B(NodeKind::Prod, e->diff(variable), U(NodeKind::Log, a->copy())));
        s.emplace_back(P3(e->copy(), a->diff(variable), U(NodeKind::Inv, a->copy())));
        return B(NodeKind::Prod, copy(), make_node<kernelT, evalT>(NodeKind::Sum, std::move(s)));
      }
      case NodeKind::Conditional: {
        std::vector<Ptr> ch;
        ch.emplace_back(children_[0]->copy());
        ch.emplace_back(children_[1]->diff(variable));
        ch.emplace_back(children_[2]->diff(variable));
        return make_node<kernelT, evalT>(NodeKind::Conditional, std::move(ch), 0.0, "", cmp_);
      }
    }
    return K(0.0);
  }

  // ---------------------------------------------------------------- simplification
  virtual Ptr simplify() const {
    std::vector<Ptr> ch;
    for (const auto& c : children_) ch.emplace_back(c->simplify());
    auto all_const = [&ch]() {
      for (const auto& c : ch)
        if (!c->is_constant()) return false;
      return true;
    };
    switch (kind_) {
      case NodeKind::Constant:
      case NodeKind::Variable: return copy();
      case NodeKind::Sum: {
        std::vector<Ptr> flat;
        double csum = 0.0;
        bool has_c = false;
        std::vector<Ptr> stack;
        for (auto& c : ch) stack.emplace_back(std::move(c));
        for (size_t i = 0; i < stack.size(); i++) {
          Ptr& c = stack[i];
          if (c->kind() == NodeKind::Sum) {
            for (auto& g : c->children_) flat.emplace_back(std::move(g));
          } else {
            flat.emplace_back(std::move(c));
          }
        }
        std::vector<Ptr> out;
        for (auto& c : flat) {
          if (c->is_constant()) { csum += c->value(); has_c = true; }
          else out.emplace_back(std::move(c));
        }
        if (has_c && csum != 0.0) out.emplace_back(K(csum));
        if (out.empty()) return K(has_c ? csum : 0.0);
        if (out.size() == 1) return std::move(out[0]);
        return make_node<kernelT, evalT>(NodeKind::Sum, std::move(out));
      }
      case NodeKind::Prod: {
        // flatten nested products, also those under a negation (the sign is collected)
        double cprod = 1.0;
        std::vector<Ptr> out;
        bool negate = false;
        std::vector<Ptr> work;
        for (auto& c : ch) work.emplace_back(std::move(c));
        for (size_t i = 0; i < work.size(); i++) {
          Ptr c = std::move(work[i]);
          while (c->kind() == NodeKind::Neg) {
            negate = !negate;
            c = std::move(c->children_[0]);
          }
          if (c->kind() == NodeKind::Prod) {
            for (auto& g : c->children_) work.emplace_back(std::move(g));
          } else if (c->is_constant()) {
            cprod *= c->value();
          } else {
            out.emplace_back(std::move(c));
          }
        }
        if (negate) cprod = -cprod;
        if (cprod == 0.0) return K(0.0);
        if (out.empty()) return K(cprod);
        if (cprod == 1.0 && out.size() == 1) return std::move(out[0]);
        if (cprod == -1.0 && out.size() == 1) return U(NodeKind::Neg, std::move(out[0]));
        if (cprod != 1.0) out.insert(out.begin(), K(cprod));
        return make_node<kernelT, evalT>(NodeKind::Prod, std::move(out));
      }
      case NodeKind::Neg: {
        Ptr& a = ch[0];
        if (a->is_constant()) return K(-a->value());
        if (a->kind() == NodeKind::Neg) return std::move(a->children_[0]);
        if (a->kind() == NodeKind::Prod && a->children_[0]->is_constant()) {
          Ptr r = a->copy();
          r->children_[0] = K(-a->children_[0]->value());
          if (r->children_[0]->value() == 1.0) {
            r->children_.erase(r->children_.begin());
            if (r->children_.size() == 1) return std::move(r->children_[0]);
          }
          return r;
        }
        return U(NodeKind::Neg, std::move(a));
      }
      case NodeKind::Inv: {
        Ptr& a = ch[0];
        if (a->is_constant()) return K(1.0 / a->value());
        if (a->kind() == NodeKind::Inv) return std::move(a->children_[0]);
        if (a->kind() == NodeKind::Neg) return U(NodeKind::Neg, U(NodeKind::Inv, std::move(a->children_[0])));
        return U(NodeKind::Inv, std::move(a));
      }
      case NodeKind::Exp:
        if (ch[0]->is_constant()) return K(std::exp(ch[0]->value()));
        if (ch[0]->kind() == NodeKind::Log) return std::move(ch[0]->children_[0]);
        return U(kind_, std::move(ch[0]));
      case NodeKind::Log:
        if (ch[0]->is_constant()) return K(std::log(ch[0]->value()));
        if (ch[0]->kind() == NodeKind::Exp) return std::move(ch[0]->children_[0]);
        return U(kind_, std::move(ch[0]));
      case NodeKind::Pow: {
        if (all_const()) return K(std::pow(ch[0]->value(), ch[1]->value()));
        if (ch[1]->is_constant()) {
          if (ch[1]->value() == 0.0) return K(1.0);
          if (ch[1]->value() == 1.0) return std::move(ch[0]);
        }
        return B(NodeKind::Pow, std::move(ch[0]), std::move(ch[1]));
      }
      case NodeKind::Conditional: {
        if (ch[0]->is_constant()) return cmp_holds(cmp_, ch[0]->value()) ? std::move(ch[1]) : std::move(ch[2]);
        if (ch[1]->equals(*ch[2])) return std::move(ch[1]);
        return make_node<kernelT, evalT>(NodeKind::Conditional, std::move(ch), 0.0, "", cmp_);
      }
      default:
        if (ch[0]->is_constant()) return K(apply_unary(kind_, ch[0]->value()));
        return U(kind_, std::move(ch[0]));
    }
  }

  // ---------------------------------------------------------------- hoisting
  // Replaces subtrees that depend only on `variables` (the floating parameters) and
  // constants by new variables prefix+k whose expressions are appended to
  // bufferexpressions (deduplicated structurally).  Parameter-only operands of a Sum or
  // Prod that also has event-dependent operands are grouped and hoisted together.
  static std::string hoist(Ptr expr, std::vector<std::string>& names, std::vector<Ptr>& exprs, const std::string& prefix) {
    for (size_t i = 0; i < exprs.size(); i++)
      if (exprs[i]->equals(*expr)) return names[i];
    std::string n = prefix + std::to_string(exprs.size());
    names.push_back(n);
    exprs.emplace_back(std::move(expr));
    return n;
  }
  // group_always: a group of such operands is hoisted even below the threshold when it has
  // the kind of its parent (free for per-segment parameter terms; event columns pass false,
  // since every column costs memory bandwidth per event)
  virtual Ptr optimize_buffering_constant_terms(std::vector<std::string>& buffernames, std::vector<Ptr>& bufferexpressions,
                                                const std::vector<std::string>& variables,
                                                const std::string& prefix = "morefit_buffer_",
                                                float buffering_cost_threshold = 2.0, bool group_always = true) {
    if (kind_ == NodeKind::Constant || kind_ == NodeKind::Variable) return copy();
    if (!contains_other_variables(variables)) {
      if (cost() > buffering_cost_threshold)
        return make_node<kernelT, evalT>(NodeKind::Variable, {}, 0.0, hoist(copy(), buffernames, bufferexpressions, prefix));
      return copy();
    }
    std::vector<Ptr> ch;
    if (kind_ == NodeKind::Sum || kind_ == NodeKind::Prod) {
      // parameter-only operands (constants join them when at least one parameter term exists)
      bool has_param = false;
      for (auto& c : children_) has_param |= !c->contains_other_variables(variables) && !c->is_constant();
      std::vector<Ptr> group;
      for (auto& c : children_) {
        if (has_param && !c->contains_other_variables(variables)) group.emplace_back(c->copy());
        else ch.emplace_back(c->optimize_buffering_constant_terms(buffernames, bufferexpressions, variables, prefix, buffering_cost_threshold, group_always));
      }
      if (!group.empty()) {
        Ptr g = group.size() == 1 ? std::move(group[0]) : make_node<kernelT, evalT>(kind_, std::move(group));
        if (g->cost() > buffering_cost_threshold || (group_always && g->kind() == kind_))
          ch.emplace_back(make_node<kernelT, evalT>(NodeKind::Variable, {}, 0.0, hoist(std::move(g), buffernames, bufferexpressions, prefix)));
        else
          ch.emplace_back(std::move(g));
      }
      return make_node<kernelT, evalT>(kind_, std::move(ch));
    }
    for (auto& c : children_)
      ch.emplace_back(c->optimize_buffering_constant_terms(buffernames, bufferexpressions, variables, prefix, buffering_cost_threshold, group_always));
    return make_node<kernelT, evalT>(kind_, std::move(ch), value_, name_, cmp_);
  }

  // ---------------------------------------------------------------- printing
  static const char* kind_str(NodeKind k) {
    switch (k) {
      case NodeKind::Constant: return "Constant";
      case NodeKind::Variable: return "Variable";
      case NodeKind::Sum: return "Sum";
      case NodeKind::Prod: return "Prod";
      case NodeKind::Neg: return "Neg";
      case NodeKind::Inv: return "Inv";
      case NodeKind::Exp: return "exp";
      case NodeKind::Log: return "log";
      case NodeKind::Pow: return "pow";
      case NodeKind::Sqrt: return "sqrt";
      case NodeKind::Sin: return "sin";
      case NodeKind::Cos: return "cos";
      case NodeKind::Tan: return "tan";
      case NodeKind::Asin: return "asin";
      case NodeKind::Acos: return "acos";
      case NodeKind::Atan: return "atan";
      case NodeKind::Erf: return "erf";
      case NodeKind::Exp2: return "exp2";
      case NodeKind::Log2: return "log2";
      case NodeKind::Conditional: return "Conditional";
    }
    return "?";
  }
  static const char* cmp_str(int c) {
    static const char* s[] = {"==", "!=", ">", "<", ">=", "<="};
    return s[c];
  }
  // single C expression (no CSE); kernels use emit_kernel() instead
  virtual std::string get_kernel() const {
    switch (kind_) {
      case NodeKind::Constant: return "(" + hex_double(value_) + ")";
      case NodeKind::Variable: return name_;
      case NodeKind::Sum:
      case NodeKind::Prod: {
        std::string r = "(";
        for (size_t i = 0; i < children_.size(); i++) r += (i ? (kind_ == NodeKind::Sum ? " + " : " * ") : "") + children_[i]->get_kernel();
        return r + ")";
      }
      case NodeKind::Neg: return "(-" + children_[0]->get_kernel() + ")";
      case NodeKind::Inv: return "(1.0/" + children_[0]->get_kernel() + ")";
      case NodeKind::Pow: return "pow(" + children_[0]->get_kernel() + ", " + children_[1]->get_kernel() + ")";
      case NodeKind::Conditional:
        return "((" + children_[0]->get_kernel() + cmp_str(cmp_) + "0.0) ? " + children_[1]->get_kernel() + " : " + children_[2]->get_kernel() + ")";
      default: return std::string(kind_str(kind_)) + "(" + children_[0]->get_kernel() + ")";
    }
  }
  virtual void print(unsigned int level = 0) const {
    std::cout << std::string(2 * level, ' ') << kind_str(kind_);
    if (kind_ == NodeKind::Constant) std::cout << " " << value_;
    if (kind_ == NodeKind::Variable) std::cout << " " << name_;
    if (kind_ == NodeKind::Conditional) std::cout << " " << cmp_str(cmp_) << "0";
    std::cout << std::endl;
    for (const auto& c : children_) c->print(level + 1);
  }
  virtual std::string draw_node() const {
    std::string label = kind_ == NodeKind::Constant ? hex_double(value_) : (kind_ == NodeKind::Variable ? name_ : kind_str(kind_));
    std::string r = "\"" + label + "\"";
    if (!children_.empty()) {
      r += " -> {";
      for (size_t i = 0; i < children_.size(); i++) r += (i ? ", " : "") + children_[i]->draw_node();
      r += "}";
    }
    return r;
  }
  void draw(std::string filename) const {
    std::ofstream f(filename);
    f << "\\documentclass[tikz,border=10pt]{standalone}\n\\usetikzlibrary{graphdrawing,graphs}\n\\usegdlibrary{trees}\n"
         "\\begin{document}\n\\begin{tikzpicture}\n\\graph [tree layout, fresh nodes]\n{\n"
      << draw_node() << "};\n\\end{tikzpicture}\n\\end{document}\n";
  }
};

// ==================================================================== node classes
#define MOREFIT_UNARY_NODE(NAME, KIND)                                                           \
  template <typename kernelT, typename evalT>                                                  \
  class NAME : public ComputeGraphNode<kernelT, evalT> {                                       \
   public:                                                                                     \
    explicit NAME(GraphPtr<kernelT, evalT> a)                                                  \
        : ComputeGraphNode<kernelT, evalT>(NodeKind::KIND, [&] {                               \
            std::vector<GraphPtr<kernelT, evalT>> v;                                           \
            v.emplace_back(std::move(a));                                                      \
            return v;                                                                          \
          }()) {}                                                                              \
    NAME(NodeKind k, std::vector<GraphPtr<kernelT, evalT>> ch) : ComputeGraphNode<kernelT, evalT>(k, std::move(ch)) {} \
  };

MOREFIT_UNARY_NODE(NegNode, Neg)
MOREFIT_UNARY_NODE(InvNode, Inv)
MOREFIT_UNARY_NODE(ExpNode, Exp)
MOREFIT_UNARY_NODE(LogNode, Log)
MOREFIT_UNARY_NODE(SqrtNode, Sqrt)
MOREFIT_UNARY_NODE(SinNode, Sin)
MOREFIT_UNARY_NODE(CosNode, Cos)
MOREFIT_UNARY_NODE(TanNode, Tan)
MOREFIT_UNARY_NODE(AsinNode, Asin)
MOREFIT_UNARY_NODE(AcosNode, Acos)
MOREFIT_UNARY_NODE(AtanNode, Atan)
MOREFIT_UNARY_NODE(ErfNode, Erf)
MOREFIT_UNARY_NODE(Exp2Node, Exp2)
MOREFIT_UNARY_NODE(Log2Node, Log2)
#undef MOREFIT_UNARY_NODE

template <typename kernelT, typename evalT>
class ConstantNode : public ComputeGraphNode<kernelT, evalT> {
 public:
  explicit ConstantNode(double c) : ComputeGraphNode<kernelT, evalT>(NodeKind::Constant, {}, c) {}
};

template <typename kernelT, typename evalT>
class VariableNode : public ComputeGraphNode<kernelT, evalT> {
 public:
  explicit VariableNode(std::string name) : ComputeGraphNode<kernelT, evalT>(NodeKind::Variable, {}, 0.0, name) {}
};

#define MOREFIT_NARY_NODE(NAME, KIND)                                                                     \
  template <typename kernelT, typename evalT>                                                           \
  class NAME : public ComputeGraphNode<kernelT, evalT> {                                                \
   public:                                                                                              \
    explicit NAME(std::vector<GraphPtr<kernelT, evalT>> ch) : ComputeGraphNode<kernelT, evalT>(NodeKind::KIND, std::move(ch)) {} \
    NAME(GraphPtr<kernelT, evalT> a, GraphPtr<kernelT, evalT> b)                                        \
        : ComputeGraphNode<kernelT, evalT>(NodeKind::KIND, [&] {                                        \
            std::vector<GraphPtr<kernelT, evalT>> v;                                                    \
            v.emplace_back(std::move(a));                                                               \
            v.emplace_back(std::move(b));                                                               \
            return v;                                                                                   \
          }()) {}                                                                                       \
  };
MOREFIT_NARY_NODE(SumNode, Sum)
MOREFIT_NARY_NODE(ProdNode, Prod)
MOREFIT_NARY_NODE(PowNode, Pow)
#undef MOREFIT_NARY_NODE

template <typename kernelT, typename evalT>
class ConditionalNode : public ComputeGraphNode<kernelT, evalT> {
 public:
  using typename ComputeGraphNode<kernelT, evalT>::comparator_type;
  ConditionalNode(int comparator, GraphPtr<kernelT, evalT> expression, GraphPtr<kernelT, evalT> a, GraphPtr<kernelT, evalT> b)
      : ComputeGraphNode<kernelT, evalT>(NodeKind::Conditional, [&] {
          std::vector<GraphPtr<kernelT, evalT>> v;
          v.emplace_back(std::move(expression));
          v.emplace_back(std::move(a));
          v.emplace_back(std::move(b));
          return v;
        }(), 0.0, "", comparator) {}
};

template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> make_node(NodeKind kind, std::vector<GraphPtr<kernelT, evalT>> ch, double value, const std::string& name, int cmp) {
  switch (kind) {
    case NodeKind::Constant: return std::make_unique<ConstantNode<kernelT, evalT>>(value);
    case NodeKind::Variable: return std::make_unique<VariableNode<kernelT, evalT>>(name);
    case NodeKind::Sum: return std::make_unique<SumNode<kernelT, evalT>>(std::move(ch));
    case NodeKind::Prod: return std::make_unique<ProdNode<kernelT, evalT>>(std::move(ch));
    case NodeKind::Pow: return std::make_unique<PowNode<kernelT, evalT>>(std::move(ch));
    case NodeKind::Conditional:
      return std::make_unique<ConditionalNode<kernelT, evalT>>(cmp, std::move(ch[0]), std::move(ch[1]), std::move(ch[2]));
    case NodeKind::Neg: return std::make_unique<NegNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Inv: return std::make_unique<InvNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Exp: return std::make_unique<ExpNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Log: return std::make_unique<LogNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Sqrt: return std::make_unique<SqrtNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Sin: return std::make_unique<SinNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Cos: return std::make_unique<CosNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Tan: return std::make_unique<TanNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Asin: return std::make_unique<AsinNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Acos: return std::make_unique<AcosNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Atan: return std::make_unique<AtanNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Erf: return std::make_unique<ErfNode<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Exp2: return std::make_unique<Exp2Node<kernelT, evalT>>(kind, std::move(ch));
    case NodeKind::Log2: return std::make_unique<Log2Node<kernelT, evalT>>(kind, std::move(ch));
  }
  return nullptr;
}

// ==================================================================== builders
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Variable(std::string name) { return std::make_unique<VariableNode<kernelT, evalT>>(name); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Constant(evalT c) { return std::make_unique<ConstantNode<kernelT, evalT>>(c); }

template <typename kernelT, typename evalT>
bool operator==(const GraphPtr<kernelT, evalT>& lhs, const GraphPtr<kernelT, evalT>& rhs) { return lhs->equals(*rhs); }

#define MOREFIT_CONDITIONAL_BUILDER(NAME, CMP)                                                                      \
  template <typename kernelT, typename evalT>                                                                   \
  GraphPtr<kernelT, evalT> NAME(GraphPtr<kernelT, evalT> e, GraphPtr<kernelT, evalT> A, GraphPtr<kernelT, evalT> B) { \
    return std::make_unique<ConditionalNode<kernelT, evalT>>(ComputeGraphNode<kernelT, evalT>::CMP, std::move(e), std::move(A), std::move(B)); \
  }
MOREFIT_CONDITIONAL_BUILDER(ConditionalEqual, Equal)
MOREFIT_CONDITIONAL_BUILDER(ConditionalUnequal, Unequal)
MOREFIT_CONDITIONAL_BUILDER(ConditionalLarger, Larger)
MOREFIT_CONDITIONAL_BUILDER(ConditionalLargerEqual, LargerEqual)
MOREFIT_CONDITIONAL_BUILDER(ConditionalSmaller, Smaller)
MOREFIT_CONDITIONAL_BUILDER(ConditionalSmallerEqual, SmallerEqual)
#undef MOREFIT_CONDITIONAL_BUILDER

template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Sum(GraphPtr<kernelT, evalT> A, GraphPtr<kernelT, evalT> B) { return std::make_unique<SumNode<kernelT, evalT>>(std::move(A), std::move(B)); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Prod(GraphPtr<kernelT, evalT> A, GraphPtr<kernelT, evalT> B) { return std::make_unique<ProdNode<kernelT, evalT>>(std::move(A), std::move(B)); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Neg(GraphPtr<kernelT, evalT> A) { return std::make_unique<NegNode<kernelT, evalT>>(std::move(A)); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Neg(std::string A) { return Neg<kernelT, evalT>(Variable<kernelT, evalT>(A)); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Neg(evalT A) { return Constant<kernelT, evalT>(-A); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Inv(GraphPtr<kernelT, evalT> A) { return std::make_unique<InvNode<kernelT, evalT>>(std::move(A)); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Inv(std::string A) { return Inv<kernelT, evalT>(Variable<kernelT, evalT>(A)); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Inv(evalT A) { return Constant<kernelT, evalT>(1.0 / A); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> Pow(GraphPtr<kernelT, evalT> A, GraphPtr<kernelT, evalT> B) { return std::make_unique<PowNode<kernelT, evalT>>(std::move(A), std::move(B)); }

#define MOREFIT_UNARY_BUILDER(NAME, NODE)                                                                                    \
  template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> NAME(GraphPtr<kernelT, evalT> A) { return std::make_unique<NODE<kernelT, evalT>>(std::move(A)); } \
  template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> NAME(std::string A) { return NAME<kernelT, evalT>(Variable<kernelT, evalT>(A)); } \
  template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> NAME(evalT A) { return NAME<kernelT, evalT>(Constant<kernelT, evalT>(A)); }
MOREFIT_UNARY_BUILDER(Sqrt, SqrtNode)
MOREFIT_UNARY_BUILDER(Exp, ExpNode)
MOREFIT_UNARY_BUILDER(Log, LogNode)
MOREFIT_UNARY_BUILDER(Sin, SinNode)
MOREFIT_UNARY_BUILDER(Cos, CosNode)
MOREFIT_UNARY_BUILDER(Tan, TanNode)
MOREFIT_UNARY_BUILDER(Asin, AsinNode)
MOREFIT_UNARY_BUILDER(Acos, AcosNode)
MOREFIT_UNARY_BUILDER(Atan, AtanNode)
MOREFIT_UNARY_BUILDER(Erf, ErfNode)
#undef MOREFIT_UNARY_BUILDER

// a-b -> Sum(a, Neg(b)); a/b -> Prod(a, Inv(b)) (the reference's conventions, graph.hh:2661-2664, 2710-2713)
#define MOREFIT_BINOP(OP, MAKE)                                                                                               \
  template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> operator OP(GraphPtr<kernelT, evalT> A, GraphPtr<kernelT, evalT> B) { return MAKE(std::move(A), std::move(B)); } \
  template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> operator OP(evalT A, GraphPtr<kernelT, evalT> B) { return MAKE(Constant<kernelT, evalT>(A), std::move(B)); } \
  template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> operator OP(GraphPtr<kernelT, evalT> A, evalT B) { return MAKE(std::move(A), Constant<kernelT, evalT>(B)); } \
  template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> operator OP(std::string A, GraphPtr<kernelT, evalT> B) { return MAKE(Variable<kernelT, evalT>(A), std::move(B)); } \
  template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> operator OP(GraphPtr<kernelT, evalT> A, std::string B) { return MAKE(std::move(A), Variable<kernelT, evalT>(B)); }

template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> mf_add(GraphPtr<kernelT, evalT> A, GraphPtr<kernelT, evalT> B) { return Sum<kernelT, evalT>(std::move(A), std::move(B)); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> mf_sub(GraphPtr<kernelT, evalT> A, GraphPtr<kernelT, evalT> B) { return Sum<kernelT, evalT>(std::move(A), Neg<kernelT, evalT>(std::move(B))); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> mf_mul(GraphPtr<kernelT, evalT> A, GraphPtr<kernelT, evalT> B) { return Prod<kernelT, evalT>(std::move(A), std::move(B)); }
template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> mf_div(GraphPtr<kernelT, evalT> A, GraphPtr<kernelT, evalT> B) { return Prod<kernelT, evalT>(std::move(A), Inv<kernelT, evalT>(std::move(B))); }
MOREFIT_BINOP(+, mf_add)
MOREFIT_BINOP(-, mf_sub)
MOREFIT_BINOP(*, mf_mul)
MOREFIT_BINOP(/, mf_div)
#undef MOREFIT_BINOP

template <typename kernelT, typename evalT> GraphPtr<kernelT, evalT> operator-(GraphPtr<kernelT, evalT> A) { return Neg<kernelT, evalT>(std::move(A)); }

// ==================================================================== graph transforms used by the backends

// Is the node free of every name in `observables` (i.e. depends only on parameters/constants)?
template <typename kernelT, typename evalT>
bool parameter_only(const ComputeGraphNode<kernelT, evalT>& n, const std::vector<std::string>& observables) {
  return !n.contains_any(observables);
}

// One pass of the rewrites that let hoisting pull constant offsets out of centred observables:
//  (1) p1*...*pk * (e + q)  ->  p1*...*pk*e + p1*...*pk*q   (single event-dependent Sum factor
//      with a parameter-only part, all other factors parameter-only)
//  (2) exp(e + q)           ->  exp(e) * exp(q)              (q parameter-only)
template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> centring_rewrite(const ComputeGraphNode<kernelT, evalT>& n, const std::vector<std::string>& obs) {
  typedef GraphPtr<kernelT, evalT> Ptr;
  std::vector<Ptr> ch;
  for (const auto& c : n.children_) ch.emplace_back(centring_rewrite(*c, obs));
  if (n.kind() == NodeKind::Prod) {
    int nonparam = -1, count = 0;
    for (size_t i = 0; i < ch.size(); i++)
      if (!parameter_only(*ch[i], obs)) { nonparam = (int)i; count++; }
    if (count == 1 && ch.size() > 1 && ch[nonparam]->kind() == NodeKind::Sum) {
      const auto& s = *ch[nonparam];
      bool has_param_part = false, has_event_part = false;
      for (const auto& t : s.children_) (parameter_only(*t, obs) ? has_param_part : has_event_part) = true;
      if (has_param_part && has_event_part) {
        std::vector<Ptr> terms;
        for (const auto& t : s.children_) {
          std::vector<Ptr> f;
          for (size_t i = 0; i < ch.size(); i++) f.emplace_back(i == (size_t)nonparam ? t->copy() : ch[i]->copy());
          terms.emplace_back(make_node<kernelT, evalT>(NodeKind::Prod, std::move(f)));
        }
        return make_node<kernelT, evalT>(NodeKind::Sum, std::move(terms));
      }
    }
  }
  if (n.kind() == NodeKind::Exp && ch[0]->kind() == NodeKind::Sum) {
    std::vector<Ptr> ev, pa;
    for (const auto& t : ch[0]->children_) (parameter_only(*t, obs) ? pa : ev).emplace_back(t->copy());
    if (!ev.empty() && !pa.empty()) {
      std::vector<Ptr> f;
      f.emplace_back(Exp<kernelT, evalT>(ev.size() == 1 ? std::move(ev[0]) : make_node<kernelT, evalT>(NodeKind::Sum, std::move(ev))));
      f.emplace_back(Exp<kernelT, evalT>(pa.size() == 1 ? std::move(pa[0]) : make_node<kernelT, evalT>(NodeKind::Sum, std::move(pa))));
      return make_node<kernelT, evalT>(NodeKind::Prod, std::move(f));
    }
  }
  return make_node<kernelT, evalT>(n.kind(), std::move(ch), n.value(), n.name(), n.comparator());
}

// x -> x + c for each observable (the kernel then reads centred offsets), simplified and
// rewritten until stable so that parameter-only offsets can be hoisted.
template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> centre_observables(const ComputeGraphNode<kernelT, evalT>& g, const std::vector<std::string>& obs,
                                            const std::vector<double>& centres) {
  GraphPtr<kernelT, evalT> r = g.copy();
  for (size_t i = 0; i < obs.size(); i++) {
    if (centres.at(i) == 0.0) continue;
    auto rep = Sum<kernelT, evalT>(Variable<kernelT, evalT>(obs[i]), Constant<kernelT, evalT>(centres[i]));
    r = r->replace_variable(obs[i], *rep);
  }
  r = r->simplify();
  for (int pass = 0; pass < 4; pass++) {
    auto nxt = centring_rewrite(*r, obs)->simplify();
    bool same = nxt->equals(*r);
    r = std::move(nxt);
    if (same) break;
  }
  return r;
}

// structural hash (equal graphs hash equally)
template <typename kernelT, typename evalT>
uint64_t structural_hash(const ComputeGraphNode<kernelT, evalT>& n) {
  uint64_t h = 1469598103934665603ULL ^ (uint64_t)n.kind() * 0x9e3779b97f4a7c15ULL;
  auto mix = [&h](uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
  };
  if (n.kind() == NodeKind::Constant) {
    uint64_t b;
    double v = n.value();
    std::memcpy(&b, &v, sizeof(b));
    mix(b);
  }
  if (n.kind() == NodeKind::Variable) mix(std::hash<std::string>()(n.name()));
  mix((uint64_t)n.comparator());
  for (const auto& c : n.children_) mix(structural_hash(*c));
  return h;
}

// Pulls factors shared by every term of a Sum out in front:  a*b*c + a*d  ->  a*(b*c + d).
// Applied bottom-up to derivative graphs before hoisting, so event-dependent factors
// (exp(...), 1/p) are multiplied once and the parameter-only remainders of a sum can be
// hoisted into one FP64 term instead of being combined per event in FP32.
template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> factor_common(const ComputeGraphNode<kernelT, evalT>& n) {
  typedef GraphPtr<kernelT, evalT> Ptr;
  std::vector<Ptr> ch;
  for (const auto& c : n.children_) ch.emplace_back(factor_common(*c));
  if (n.kind() != NodeKind::Sum || ch.size() < 2)
    return make_node<kernelT, evalT>(n.kind(), std::move(ch), n.value(), n.name(), n.comparator())->simplify();
  // factor lists of each term
  std::vector<std::vector<Ptr>> fac(ch.size());
  std::vector<std::vector<uint64_t>> hs(ch.size());
  for (size_t t = 0; t < ch.size(); t++) {
    if (ch[t]->kind() == NodeKind::Prod)
      for (auto& f : ch[t]->children_) fac[t].emplace_back(f->copy());
    else
      fac[t].emplace_back(ch[t]->copy());
    for (auto& f : fac[t]) hs[t].push_back(structural_hash(*f));
  }
  std::vector<Ptr> common;
  for (size_t i = 0; i < fac[0].size(); i++) {
    if (fac[0][i]->is_constant()) continue;
    std::vector<int> pos(ch.size(), -1);
    pos[0] = (int)i;
    bool everywhere = true;
    for (size_t t = 1; t < ch.size() && everywhere; t++) {
      for (size_t j = 0; j < fac[t].size(); j++)
        if (fac[t][j] && hs[t][j] == hs[0][i] && fac[t][j]->equals(*fac[0][i])) {
          pos[t] = (int)j;
          break;
        }
      everywhere = pos[t] >= 0;
    }
    if (!everywhere) continue;
    common.emplace_back(fac[0][i]->copy());
    for (size_t t = 0; t < ch.size(); t++) fac[t][pos[t]].reset();
  }
  if (common.empty()) return make_node<kernelT, evalT>(NodeKind::Sum, std::move(ch))->simplify();
  std::vector<Ptr> terms;
  for (size_t t = 0; t < ch.size(); t++) {
    std::vector<Ptr> rest;
    for (auto& f : fac[t])
      if (f) rest.emplace_back(std::move(f));
    if (rest.empty()) terms.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, 1.0));
    else if (rest.size() == 1) terms.emplace_back(std::move(rest[0]));
    else terms.emplace_back(make_node<kernelT, evalT>(NodeKind::Prod, std::move(rest)));
  }
  auto inner = factor_common(*make_node<kernelT, evalT>(NodeKind::Sum, std::move(terms))->simplify());
  common.emplace_back(std::move(inner));
  return make_node<kernelT, evalT>(NodeKind::Prod, std::move(common))->simplify();
}

// ==================================================================== FP32 kernel rewrites
// exp(a) -> exp2(log2(e) a) and log(a) -> ln(2) log2(a): the hardware evaluates exp2 and
// log2, so the conversion constants are otherwise multiplied in per event.  After
// simplify() they sit in products next to parameter-only factors, which hoisting folds into
// one FP64-computed kernel parameter (inside exp2) or into an output scale (split_output_scale).
template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> to_base2(const ComputeGraphNode<kernelT, evalT>& n) {
  std::vector<GraphPtr<kernelT, evalT>> ch;
  for (const auto& c : n.children_) ch.emplace_back(to_base2(*c));
  if (n.kind() == NodeKind::Exp) {
    std::vector<GraphPtr<kernelT, evalT>> p;
    p.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, 1.4426950408889634));
    p.emplace_back(std::move(ch[0]));
    std::vector<GraphPtr<kernelT, evalT>> e;
    e.emplace_back(make_node<kernelT, evalT>(NodeKind::Prod, std::move(p)));
    return make_node<kernelT, evalT>(NodeKind::Exp2, std::move(e));
  }
  if (n.kind() == NodeKind::Log) {
    std::vector<GraphPtr<kernelT, evalT>> p;
    p.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, 0.6931471805599453));
    p.emplace_back(make_node<kernelT, evalT>(NodeKind::Log2, std::move(ch)));
    return make_node<kernelT, evalT>(NodeKind::Prod, std::move(p));
  }
  return make_node<kernelT, evalT>(n.kind(), std::move(ch), n.value(), n.name(), n.comparator());
}

// Splits an output (a per-event term whose sum over events is wanted) into a factor that
// depends only on `params` and constants, and the event-dependent rest: the sum is taken of
// the rest and multiplied by the factor once, in FP64, per segment.  Returns the rest and
// sets `scale` (null when the output has no such factor or no event-dependent part).
template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> split_output_scale(const ComputeGraphNode<kernelT, evalT>& g, const std::vector<std::string>& params,
                                            GraphPtr<kernelT, evalT>& scale) {
  scale.reset();
  const ComputeGraphNode<kernelT, evalT>* n = &g;
  bool neg = false;
  if (n->kind() == NodeKind::Neg) {
    neg = true;
    n = n->children_[0].get();
  }
  if (!n->contains_other_variables(params)) return g.copy();
  std::vector<GraphPtr<kernelT, evalT>> sc, ev;
  if (n->kind() == NodeKind::Prod) {
    for (const auto& c : n->children_) (c->contains_other_variables(params) ? ev : sc).emplace_back(c->copy());
  } else {
    ev.emplace_back(n->copy());
  }
  if (neg) sc.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, -1.0));
  if (sc.empty()) return g.copy();
  scale = sc.size() == 1 ? std::move(sc[0]) : make_node<kernelT, evalT>(NodeKind::Prod, std::move(sc))->simplify();
  return ev.size() == 1 ? std::move(ev[0]) : make_node<kernelT, evalT>(NodeKind::Prod, std::move(ev));
}

// ==================================================================== linear form of an output
// An output (a per-event term whose sum over events is wanted) written as
//   sum_j c_j * m_j
// with every c_j free of the event variables (a function of `params` and constants) and every
// m_j a product of event-dependent factors (an empty product: the constant 1, whose sum is the
// event count).  Products are distributed over event-dependent sums and differences; terms with
// the same monomial (its factors sorted by their emitted text) are merged.  The kernels then sum
// the distinct monomials of all outputs of a kind and form the outputs from those sums once per
// segment in FP64 -- for the mass gradient, (G r)(a (m - mu) + b) and its siblings become
// G r, G r (m - mu), ... : fewer operations per event.
template <typename kernelT, typename evalT>
struct LinearTerm {
  GraphPtr<kernelT, evalT> coef;                   // parameter-only (null: 1)
  std::vector<GraphPtr<kernelT, evalT>> factors;   // event-dependent (empty: 1)
};

template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> linear_mul(GraphPtr<kernelT, evalT> a, GraphPtr<kernelT, evalT> b) {
  if (!a) return b;
  if (!b) return a;
  std::vector<GraphPtr<kernelT, evalT>> v;
  v.emplace_back(std::move(a));
  v.emplace_back(std::move(b));
  return make_node<kernelT, evalT>(NodeKind::Prod, std::move(v));
}

// A shifted event variable, x + c(params): kept as one factor (the centred differences the
// graphs are written in), not split into x and c -- that would trade one add for extra
// monomials (x^2, x, 1 for (x + c)^2) and undo the centring that keeps FP32 sums accurate
template <typename kernelT, typename evalT>
bool linear_shift(const ComputeGraphNode<kernelT, evalT>& n, const std::vector<std::string>& params) {
  if (n.kind() != NodeKind::Sum) return false;
  int ev = 0;
  for (const auto& c : n.children_) {
    if (!c->contains_other_variables(params)) continue;
    if (c->kind() != NodeKind::Variable) return false;
    ev++;
  }
  return ev == 1;
}

// false when the expansion would exceed max_terms terms
template <typename kernelT, typename evalT>
bool linear_expand(const ComputeGraphNode<kernelT, evalT>& n, const std::vector<std::string>& params,
                   std::vector<LinearTerm<kernelT, evalT>>& out, size_t max_terms) {
  typedef GraphPtr<kernelT, evalT> Ptr;
  if (!n.contains_other_variables(params)) {
    out.push_back({n.copy(), {}});
    return out.size() <= max_terms;
  }
  switch (n.kind()) {
    case NodeKind::Sum:
      for (const auto& c : n.children_)
        if (!linear_expand(*c, params, out, max_terms)) return false;
      return true;
    case NodeKind::Neg: {
      std::vector<LinearTerm<kernelT, evalT>> t;
      if (!linear_expand(*n.children_[0], params, t, max_terms)) return false;
      for (auto& x : t) {
        x.coef = linear_mul<kernelT, evalT>(make_node<kernelT, evalT>(NodeKind::Constant, {}, -1.0), std::move(x.coef));
        out.push_back(std::move(x));
      }
      return out.size() <= max_terms;
    }
    case NodeKind::Prod: {
      std::vector<LinearTerm<kernelT, evalT>> acc;
      acc.push_back({nullptr, {}});
      for (const auto& c : n.children_) {
        std::vector<LinearTerm<kernelT, evalT>> ft;
        if (!c->contains_other_variables(params)) {
          ft.push_back({c->copy(), {}});
        } else if ((c->kind() == NodeKind::Sum && !linear_shift(*c, params)) || c->kind() == NodeKind::Neg ||
                   c->kind() == NodeKind::Prod) {
          if (!linear_expand(*c, params, ft, max_terms)) return false;
        } else {
          LinearTerm<kernelT, evalT> t;
          t.factors.push_back(c->copy());
          ft.push_back(std::move(t));
        }
        std::vector<LinearTerm<kernelT, evalT>> nx;
        for (const auto& a : acc)
          for (const auto& b : ft) {
            LinearTerm<kernelT, evalT> t;
            t.coef = linear_mul<kernelT, evalT>(a.coef ? a.coef->copy() : Ptr(), b.coef ? b.coef->copy() : Ptr());
            for (const auto& f : a.factors) t.factors.push_back(f->copy());
            for (const auto& f : b.factors) t.factors.push_back(f->copy());
            nx.push_back(std::move(t));
            if (nx.size() > max_terms) return false;
          }
        acc = std::move(nx);
      }
      for (auto& t : acc) out.push_back(std::move(t));
      return out.size() <= max_terms;
    }
    default: {
      LinearTerm<kernelT, evalT> t;
      t.factors.push_back(n.copy());
      out.push_back(std::move(t));
      return out.size() <= max_terms;
    }
  }
}

// One shift per event variable.  Terms of a kind's outputs can hold the same event variable
// under different shifts, x + c1 and x + c2 (the mass gradient: m - mu from the Gaussian and
// m - <m> from the exponential's normalisation), each a factor of its own monomials.  With
// the shift that occurs in the most terms as x + c*, every other one is written
//   x + c = (x + c*) + (c - c*)
// (the difference is parameter-only), so monomials such as (m - <m>) r become (m - mu) r and
// r, which a kernel accumulates with one multiply-add instead of an add and a multiply-add,
// and the centred difference that keeps FP32 sums accurate is kept.  false: the expansion
// would exceed max_terms terms in one output.
template <typename kernelT, typename evalT>
bool linear_canonical_shifts(std::vector<std::vector<LinearTerm<kernelT, evalT>>>& outputs, const std::vector<std::string>& params,
                             size_t max_terms) {
  typedef GraphPtr<kernelT, evalT> Ptr;
  typedef ComputeGraphNode<kernelT, evalT> Node;
  // the event variable of a shifted factor (x + c(params)), "" for any other factor
  auto shifted_var = [&](const Node& f) -> std::string {
    if (!linear_shift(f, params)) return "";
    for (const auto& c : f.children_)
      if (c->contains_other_variables(params)) return c->name();
    return "";
  };
  // the parameter-only part c of x + c
  auto shift_of = [&](const Node& f) {
    std::vector<Ptr> c;
    for (const auto& ch : f.children_)
      if (!ch->contains_other_variables(params)) c.emplace_back(ch->copy());
    return c.size() == 1 ? std::move(c[0]) : make_node<kernelT, evalT>(NodeKind::Sum, std::move(c));
  };
  std::map<std::string, std::map<std::string, int>> count;  // variable -> shift text -> terms holding it
  std::map<std::string, const Node*> repr;
  for (const auto& terms : outputs)
    for (const auto& t : terms) {
      std::set<std::string> seen;
      for (const auto& f : t.factors) {
        const std::string v = shifted_var(*f);
        if (v.empty()) continue;
        const std::string key = f->get_kernel();
        if (seen.insert(key).second) count[v][key]++;
        repr.emplace(key, f.get());
      }
    }
  std::map<std::string, std::string> canon;  // variable -> text of its canonical shift
  for (const auto& [v, shifts] : count) {
    if (shifts.size() < 2) continue;
    std::string best;
    int most = -1;
    for (const auto& [key, n] : shifts)
      if (n > most) {
        most = n;
        best = key;
      }
    canon[v] = best;
  }
  if (canon.empty()) return true;
  std::map<std::string, Ptr> canon_node;
  for (const auto& [v, key] : canon) canon_node[v] = repr.at(key)->copy();
  size_t nterms = 0;
  std::function<bool(LinearTerm<kernelT, evalT>, std::vector<LinearTerm<kernelT, evalT>>&)> expand =
      [&](LinearTerm<kernelT, evalT> t, std::vector<LinearTerm<kernelT, evalT>>& done) {
        int at = -1;
        std::string v;
        for (size_t i = 0; i < t.factors.size() && at < 0; i++) {
          v = shifted_var(*t.factors[i]);
          if (!v.empty() && canon.count(v) && t.factors[i]->get_kernel() != canon[v]) at = (int)i;
        }
        if (at < 0) {
          done.push_back(std::move(t));
          return ++nterms <= max_terms;
        }
        // (x + c) F = (x + c*) F + (c - c*) F
        std::vector<Ptr> d;
        d.emplace_back(shift_of(*t.factors[at]));
        std::vector<Ptr> neg;
        neg.emplace_back(shift_of(*canon_node[v]));
        d.emplace_back(make_node<kernelT, evalT>(NodeKind::Neg, std::move(neg)));
        Ptr delta = make_node<kernelT, evalT>(NodeKind::Sum, std::move(d))->simplify();
        LinearTerm<kernelT, evalT> a, b;
        a.coef = t.coef ? t.coef->copy() : Ptr();
        b.coef = linear_mul<kernelT, evalT>(t.coef ? t.coef->copy() : Ptr(), std::move(delta));
        for (size_t i = 0; i < t.factors.size(); i++) {
          if ((int)i == at) {
            a.factors.push_back(canon_node[v]->copy());
            continue;
          }
          a.factors.push_back(t.factors[i]->copy());
          b.factors.push_back(t.factors[i]->copy());
        }
        return expand(std::move(a), done) && expand(std::move(b), done);
      };
  for (auto& terms : outputs) {
    std::vector<LinearTerm<kernelT, evalT>> done;
    nterms = 0;
    for (auto& t : terms)
      if (!expand(std::move(t), done)) return false;
    terms = std::move(done);
  }
  return true;
}

// Canonical monomial of a term: factors sorted by their emitted text (key "" for the constant 1)
template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> linear_monomial(std::vector<GraphPtr<kernelT, evalT>>& factors, std::string& key) {
  std::vector<std::pair<std::string, size_t>> order;
  for (size_t i = 0; i < factors.size(); i++) order.push_back({factors[i]->get_kernel(), i});
  std::sort(order.begin(), order.end());
  key.clear();
  if (factors.empty()) return nullptr;
  std::vector<GraphPtr<kernelT, evalT>> f;
  for (auto& o : order) {
    key += o.first + "*";
    f.push_back(std::move(factors[o.second]));
  }
  return f.size() == 1 ? std::move(f[0]) : make_node<kernelT, evalT>(NodeKind::Prod, std::move(f));
}

// ==================================================================== double angles
// cos(2a) -> 2 cos(a)^2 - 1 and sin(2a) -> 2 sin(a) cos(a) where the graph evaluates cos(a) (and,
// for the sine, sin(a)) anyway: the K*mumu density's cos(2 phi) and sin(2 phi) then cost products
// of the cos(phi) and sin(phi) it already has instead of two more library calls per event.
template <typename kernelT, typename evalT>
void collect_trig_args(const ComputeGraphNode<kernelT, evalT>& n, std::set<uint64_t>& cos_args, std::set<uint64_t>& sin_args) {
  if (n.kind() == NodeKind::Cos) cos_args.insert(structural_hash(*n.children_[0]));
  if (n.kind() == NodeKind::Sin) sin_args.insert(structural_hash(*n.children_[0]));
  for (const auto& c : n.children_) collect_trig_args(*c, cos_args, sin_args);
}

template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> double_angle_rewrite(const ComputeGraphNode<kernelT, evalT>& n, const std::set<uint64_t>& cos_args,
                                              const std::set<uint64_t>& sin_args) {
  typedef GraphPtr<kernelT, evalT> Ptr;
  std::vector<Ptr> ch;
  for (const auto& c : n.children_) ch.emplace_back(double_angle_rewrite(*c, cos_args, sin_args));
  if ((n.kind() == NodeKind::Cos || n.kind() == NodeKind::Sin) && ch[0]->kind() == NodeKind::Prod) {
    // the argument as 2 * a
    const auto& f = ch[0]->children_;
    int two = -1;
    for (size_t i = 0; i < f.size(); i++)
      if (f[i]->is_constant() && f[i]->value() == 2.0 && two < 0) two = (int)i;
    if (two >= 0 && f.size() >= 2) {
      std::vector<Ptr> rest;
      for (size_t i = 0; i < f.size(); i++)
        if ((int)i != two) rest.emplace_back(f[i]->copy());
      Ptr a = rest.size() == 1 ? std::move(rest[0]) : make_node<kernelT, evalT>(NodeKind::Prod, std::move(rest));
      const uint64_t h = structural_hash(*a);
      auto one = [&](NodeKind k) {
        std::vector<Ptr> v;
        v.emplace_back(a->copy());
        return make_node<kernelT, evalT>(k, std::move(v));
      };
      if (n.kind() == NodeKind::Cos && cos_args.count(h)) {
        std::vector<Ptr> p;
        p.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, 2.0));
        p.emplace_back(one(NodeKind::Cos));
        p.emplace_back(one(NodeKind::Cos));
        std::vector<Ptr> sum;
        sum.emplace_back(make_node<kernelT, evalT>(NodeKind::Prod, std::move(p)));
        sum.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, -1.0));
        return make_node<kernelT, evalT>(NodeKind::Sum, std::move(sum));
      }
      if (n.kind() == NodeKind::Sin && cos_args.count(h) && sin_args.count(h)) {
        std::vector<Ptr> p;
        p.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, 2.0));
        p.emplace_back(one(NodeKind::Sin));
        p.emplace_back(one(NodeKind::Cos));
        return make_node<kernelT, evalT>(NodeKind::Prod, std::move(p));
      }
    }
  }
  return make_node<kernelT, evalT>(n.kind(), std::move(ch), n.value(), n.name(), n.comparator());
}

template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> double_angle(const ComputeGraphNode<kernelT, evalT>& g) {
  std::set<uint64_t> c, s;
  collect_trig_args(g, c, s);
  return double_angle_rewrite(g, c, s);
}

// ==================================================================== interval bounds
// Bounds of an expression over a box of its variables (observable ranges, parameter limits):
// plain interval arithmetic, so the bounds hold but may be loose where a variable occurs more
// than once.  Unbounded ([-inf, inf]) wherever a variable is outside the box or a node has no
// rule here.
struct Interval {
  double lo = -INFINITY, hi = INFINITY;
  bool bounded() const { return std::isfinite(lo) && std::isfinite(hi); }
};

inline Interval interval_mul(Interval a, Interval b) {
  const double p[4] = {a.lo * b.lo, a.lo * b.hi, a.hi * b.lo, a.hi * b.hi};
  Interval r{p[0], p[0]};
  for (double v : p) {
    if (std::isnan(v)) return Interval();
    r.lo = std::min(r.lo, v);
    r.hi = std::max(r.hi, v);
  }
  return r;
}

// b^e for an integer e (|e| <= 64); an even power is non-negative
inline Interval interval_ipow(Interval b, int e) {
  Interval r{1.0, 1.0};
  for (int k = 0; k < std::abs(e); k++) r = interval_mul(r, b);
  if (e % 2 == 0 && b.lo < 0.0 && b.hi > 0.0) r = Interval{0.0, std::max(std::pow(b.lo, std::abs(e)), std::pow(b.hi, std::abs(e)))};
  if (e >= 0) return r;
  if (!(r.lo > 0.0 || r.hi < 0.0)) return Interval();
  return Interval{1.0 / r.hi, 1.0 / r.lo};
}

template <typename kernelT, typename evalT>
Interval interval_of(const ComputeGraphNode<kernelT, evalT>& n, const std::map<std::string, Interval>& box) {
  auto child = [&](size_t i) { return interval_of(*n.children_[i], box); };
  auto monotone = [](Interval a, double (*f)(double)) { return Interval{f(a.lo), f(a.hi)}; };
  switch (n.kind()) {
    case NodeKind::Constant: return Interval{n.value(), n.value()};
    case NodeKind::Variable: {
      auto it = box.find(n.name());
      return it == box.end() ? Interval() : it->second;
    }
    case NodeKind::Sum: {
      Interval r{0.0, 0.0};
      for (size_t i = 0; i < n.children_.size(); i++) {
        Interval c = child(i);
        r.lo += c.lo;
        r.hi += c.hi;
      }
      if (std::isnan(r.lo) || std::isnan(r.hi)) return Interval();
      return r;
    }
    case NodeKind::Prod: {
      // equal factors bounded as one power: (x - mu) (x - mu) is non-negative
      Interval r{1.0, 1.0};
      std::vector<bool> used(n.children_.size(), false);
      for (size_t i = 0; i < n.children_.size(); i++) {
        if (used[i]) continue;
        int e = 1;
        for (size_t j = i + 1; j < n.children_.size(); j++)
          if (!used[j] && n.children_[j]->equals(*n.children_[i])) {
            used[j] = true;
            e++;
          }
        r = interval_mul(r, e == 1 ? child(i) : interval_ipow(child(i), e));
      }
      return r;
    }
    case NodeKind::Neg: {
      Interval c = child(0);
      return Interval{-c.hi, -c.lo};
    }
    case NodeKind::Inv: {
      Interval c = child(0);
      if (!(c.lo > 0.0 || c.hi < 0.0)) return Interval();
      return Interval{1.0 / c.hi, 1.0 / c.lo};
    }
    case NodeKind::Pow: {
      if (!n.children_[1]->is_constant()) return Interval();
      const double e = n.children_[1]->value();
      Interval b = child(0);
      if (e == std::floor(e) && std::fabs(e) <= 64) return interval_ipow(b, (int)e);
      if (!(b.lo > 0.0)) return Interval();
      Interval r{std::pow(b.lo, e), std::pow(b.hi, e)};
      if (r.lo > r.hi) std::swap(r.lo, r.hi);
      return r;
    }
    case NodeKind::Exp: return monotone(child(0), [](double v) { return std::exp(v); });
    case NodeKind::Exp2: return monotone(child(0), [](double v) { return std::exp2(v); });
    case NodeKind::Log:
    case NodeKind::Log2:
    case NodeKind::Sqrt: {
      Interval c = child(0);
      if (!(c.lo > 0.0)) return Interval();
      if (n.kind() == NodeKind::Sqrt) return monotone(c, [](double v) { return std::sqrt(v); });
      return n.kind() == NodeKind::Log ? monotone(c, [](double v) { return std::log(v); })
                                       : monotone(c, [](double v) { return std::log2(v); });
    }
    case NodeKind::Sin:
    case NodeKind::Cos: return Interval{-1.0, 1.0};
    case NodeKind::Conditional: {
      Interval a = child(1), b = child(2);
      return Interval{std::min(a.lo, b.lo), std::max(a.hi, b.hi)};
    }
    default: return Interval();
  }
}

// ==================================================================== exponential pivot
// A density that is a sum of exponential terms, p = sum_i c_i E(a_i) (E = exp or exp2, c_i free
// of the event variables), enters the per-event monomials as log p and as products
// E(a_i1)...E(a_im) / p^m (its derivatives).  With one term k as pivot and
//   q = sum_i c_i E(a_i - a_k)        (the pivot's own term is c_k)
// these are  log p = a_k + log q  and  E(a_i1)...E(a_im) / p^m = E(a_i1 - a_k)...E(a_im - a_k) / q^m,
// so every event evaluates one exponential fewer (for f Gauss + (1 - f) Exp: exp2, log2 and rcp
// instead of two exp2, log2 and rcp).  The pivot is the term whose exponent bounds (over the
// observable ranges and parameter limits in `box`) keep every E(a_i - a_k) at or below
// 2^max_log2 -- no overflow of it or of the products of two of them in the Hessian monomials,
// where p itself is finite -- and nothing is rewritten when no term
// qualifies, when a monomial has another number of exponentials than powers of 1/p, or when p
// occurs in another form.  Terms whose E(a_i - a_k) underflows are negligible beside c_k, as
// they are in p.  The pivot exponent a_k enters the rewritten monomials as the variable
// `pivot_var` (set to a_k in `pivot`), so that once substituted it is computed once per event,
// not once more inside each a_i - a_k.  Returns whether the monomials were rewritten.
template <typename kernelT, typename evalT>
bool exp_pivot(std::vector<GraphPtr<kernelT, evalT>>& monos, const std::vector<std::string>& params,
               const std::map<std::string, Interval>& box, const std::string& pivot_var, GraphPtr<kernelT, evalT>& pivot,
               double max_log2 = 40.0) {
  typedef GraphPtr<kernelT, evalT> Ptr;
  typedef ComputeGraphNode<kernelT, evalT> Node;
  auto factors_of = [](const Node& m) {
    std::vector<const Node*> f;
    if (m.kind() == NodeKind::Prod)
      for (const auto& c : m.children_) f.push_back(c.get());
    else
      f.push_back(&m);
    return f;
  };
  // the density: the argument of the log factor
  const Node* P = nullptr;
  NodeKind logk = NodeKind::Log;
  for (const auto& m : monos)
    for (const Node* f : factors_of(*m))
      if ((f->kind() == NodeKind::Log || f->kind() == NodeKind::Log2) && f->children_[0]->contains_other_variables(params) && !P) {
        P = f->children_[0].get();
        logk = f->kind();
      }
  if (!P || P->kind() != NodeKind::Sum || P->children_.size() < 2) return false;
  const NodeKind expk = logk == NodeKind::Log2 ? NodeKind::Exp2 : NodeKind::Exp;
  const size_t nt = P->children_.size();
  std::vector<std::vector<const Node*>> coef(nt);
  std::vector<const Node*> ex(nt, nullptr);
  for (size_t i = 0; i < nt; i++) {
    for (const Node* f : factors_of(*P->children_[i])) {
      if (!f->contains_other_variables(params)) coef[i].push_back(f);
      else if (f->kind() == expk && !ex[i]) ex[i] = f;
      else return false;
    }
    if (!ex[i]) return false;
  }
  // pivot: the smallest bound on max_i (a_i - a_k)
  std::vector<Interval> iv(nt);
  for (size_t i = 0; i < nt; i++) iv[i] = interval_of(*ex[i]->children_[0], box);
  const double to_log2 = expk == NodeKind::Exp2 ? 1.0 : 1.4426950408889634;
  int k = -1;
  double best = INFINITY;
  for (size_t j = 0; j < nt; j++) {
    double b = -INFINITY;
    for (size_t i = 0; i < nt; i++)
      if (i != j) b = std::max(b, (iv[i].hi - iv[j].lo) * to_log2);
    if (std::isfinite(b) && b < best) {
      best = b;
      k = (int)j;
    }
  }
  if (k < 0 || !(best <= max_log2)) return false;
  const Node& ak = *ex[k]->children_[0];
  std::vector<Ptr> r(nt);
  std::vector<Ptr> qt;
  for (size_t i = 0; i < nt; i++) {
    std::vector<Ptr> t;
    for (const Node* c : coef[i]) t.emplace_back(c->copy());
    if ((int)i != k) {
      std::vector<Ptr> nk, d;
      nk.emplace_back(make_node<kernelT, evalT>(NodeKind::Variable, {}, 0.0, pivot_var));
      d.emplace_back(ex[i]->children_[0]->copy());
      d.emplace_back(make_node<kernelT, evalT>(NodeKind::Neg, std::move(nk)));
      std::vector<Ptr> e;
      e.emplace_back(make_node<kernelT, evalT>(NodeKind::Sum, std::move(d))->simplify());
      r[i] = make_node<kernelT, evalT>(expk, std::move(e));
      t.emplace_back(r[i]->copy());
    }
    if (t.empty()) t.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, 1.0));
    qt.emplace_back(t.size() == 1 ? std::move(t[0]) : make_node<kernelT, evalT>(NodeKind::Prod, std::move(t)));
  }
  Ptr Q = make_node<kernelT, evalT>(NodeKind::Sum, std::move(qt))->simplify();
  // p or one of its exponentials anywhere but as a rewritable factor: leave everything as it is
  std::function<bool(const Node&)> mentions = [&](const Node& n) {
    if (n.equals(*P)) return true;
    for (size_t i = 0; i < nt; i++)
      if (n.equals(*ex[i])) return true;
    for (const auto& c : n.children_)
      if (mentions(*c)) return true;
    return false;
  };
  std::vector<Ptr> out;
  for (const auto& m : monos) {
    std::vector<Ptr> nf;
    int ninv = 0, nexp = 0;
    bool logp = false;
    auto fs = factors_of(*m);
    for (const Node* f : fs) {
      if (f->kind() == logk && f->children_[0]->equals(*P)) {
        logp = true;
        continue;
      }
      if (f->kind() == NodeKind::Inv && f->children_[0]->equals(*P)) {
        ninv++;
        std::vector<Ptr> v;
        v.emplace_back(Q->copy());
        nf.emplace_back(make_node<kernelT, evalT>(NodeKind::Inv, std::move(v)));
        continue;
      }
      int hit = -1;
      for (size_t i = 0; i < nt && hit < 0; i++)
        if (f->equals(*ex[i])) hit = (int)i;
      if (hit >= 0) {
        nexp++;
        if (hit != k) nf.emplace_back(r[hit]->copy());
        continue;
      }
      if (mentions(*f)) return false;
      nf.emplace_back(f->copy());
    }
    if (logp) {
      if (fs.size() != 1) return false;
      std::vector<Ptr> v;
      v.emplace_back(Q->copy());
      std::vector<Ptr> s;
      s.emplace_back(make_node<kernelT, evalT>(NodeKind::Variable, {}, 0.0, pivot_var));
      s.emplace_back(make_node<kernelT, evalT>(logk, std::move(v)));
      out.emplace_back(make_node<kernelT, evalT>(NodeKind::Sum, std::move(s)));
      continue;
    }
    if (ninv != nexp) return false;
    if (ninv == 0) {
      out.emplace_back(m->copy());
      continue;
    }
    out.emplace_back(nf.size() == 1 ? std::move(nf[0]) : make_node<kernelT, evalT>(NodeKind::Prod, std::move(nf)));
  }
  pivot = ak.copy();  // before the old monomials (which hold p) go
  monos = std::move(out);
  return true;
}

// ==================================================================== polynomials in one event variable
// Coefficients a_0..a_d (parameter-only graphs; null: 0) of n as a polynomial in the variable
// `var`, when n is one: sums, negations, products and non-negative integer powers of `var` and of
// parameter-only subtrees.  false when n depends on any other event variable, on `var` in another
// form, or exceeds degree max_deg.
template <typename kernelT, typename evalT>
bool poly_in(const ComputeGraphNode<kernelT, evalT>& n, const std::string& var, const std::vector<std::string>& params,
             std::vector<GraphPtr<kernelT, evalT>>& a, int max_deg = 4) {
  typedef GraphPtr<kernelT, evalT> Ptr;
  auto add = [](Ptr x, Ptr y) -> Ptr {
    if (!x) return y;
    if (!y) return x;
    std::vector<Ptr> v;
    v.emplace_back(std::move(x));
    v.emplace_back(std::move(y));
    return make_node<kernelT, evalT>(NodeKind::Sum, std::move(v));
  };
  a.clear();
  std::vector<std::string> pv(params);
  if (!n.contains_other_variables(pv)) {
    a.emplace_back(n.copy());
    return true;
  }
  switch (n.kind()) {
    case NodeKind::Variable:
      if (n.name() != var) return false;
      a.emplace_back(nullptr);
      a.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, 1.0));
      return true;
    case NodeKind::Sum: {
      for (const auto& c : n.children_) {
        std::vector<Ptr> b;
        if (!poly_in(*c, var, params, b, max_deg)) return false;
        if (b.size() > a.size()) a.resize(b.size());
        for (size_t j = 0; j < b.size(); j++)
          if (b[j]) a[j] = add(std::move(a[j]), std::move(b[j]));
      }
      return true;
    }
    case NodeKind::Neg: {
      if (!poly_in(*n.children_[0], var, params, a, max_deg)) return false;
      for (auto& c : a)
        if (c) {
          std::vector<Ptr> v;
          v.emplace_back(std::move(c));
          c = make_node<kernelT, evalT>(NodeKind::Neg, std::move(v));
        }
      return true;
    }
    case NodeKind::Pow:
    case NodeKind::Prod: {
      std::vector<const ComputeGraphNode<kernelT, evalT>*> f;
      int reps = 1;
      if (n.kind() == NodeKind::Pow) {
        const auto& e = *n.children_[1];
        if (!e.is_constant() || e.value() != std::floor(e.value()) || e.value() < 0 || e.value() > max_deg) return false;
        reps = (int)e.value();
        f.push_back(n.children_[0].get());
      } else {
        for (const auto& c : n.children_) f.push_back(c.get());
      }
      a.clear();
      a.emplace_back(make_node<kernelT, evalT>(NodeKind::Constant, {}, 1.0));
      for (int r = 0; r < reps; r++)
        for (auto* c : f) {
          std::vector<Ptr> b;
          if (!poly_in(*c, var, params, b, max_deg)) return false;
          if (a.size() + b.size() - 1 > (size_t)max_deg + 1) return false;
          std::vector<Ptr> p(a.size() + b.size() - 1);
          for (size_t i = 0; i < a.size(); i++)
            for (size_t j = 0; j < b.size(); j++) {
              if (!a[i] || !b[j]) continue;
              std::vector<Ptr> v;
              v.emplace_back(a[i]->copy());
              v.emplace_back(b[j]->copy());
              p[i + j] = add(std::move(p[i + j]), make_node<kernelT, evalT>(NodeKind::Prod, std::move(v)));
            }
          a = std::move(p);
        }
      return true;
    }
    default: return false;
  }
}

// Horner form of sum_j a_j x^j in t = x + s: the coefficients b_j of t^j (b_j = sum_{i >= j}
// C(i, j) a_i (-s)^(i - j), parameter-only, simplified) applied as ((b_d t + b_{d-1}) t + ...) t + b_0.
template <typename kernelT, typename evalT>
GraphPtr<kernelT, evalT> horner_in_shift(const std::vector<GraphPtr<kernelT, evalT>>& a, const ComputeGraphNode<kernelT, evalT>& t,
                                         const ComputeGraphNode<kernelT, evalT>* s) {
  typedef GraphPtr<kernelT, evalT> Ptr;
  const int d = (int)a.size() - 1;
  auto C = [](double v) { return make_node<kernelT, evalT>(NodeKind::Constant, {}, v); };
  std::vector<Ptr> b(d + 1);
  for (int j = 0; j <= d; j++) {
    std::vector<Ptr> terms;
    for (int i = j; i <= d; i++) {
      if (!a[i]) continue;
      double binom = 1.0;
      for (int k = 0; k < i - j; k++) binom *= (double)(i - k) / (double)(k + 1);
      std::vector<Ptr> f;
      f.emplace_back(C((i - j) % 2 ? -binom : binom));
      f.emplace_back(a[i]->copy());
      for (int k = 0; k < i - j && s; k++) f.emplace_back(s->copy());
      if (i > j && !s) continue;  // no shift: x = t
      terms.emplace_back(make_node<kernelT, evalT>(NodeKind::Prod, std::move(f)));
    }
    b[j] = terms.empty() ? Ptr() : make_node<kernelT, evalT>(NodeKind::Sum, std::move(terms))->simplify();
  }
  Ptr acc;
  for (int j = d; j >= 0; j--) {
    if (acc) {
      std::vector<Ptr> v;
      v.emplace_back(std::move(acc));
      v.emplace_back(t.copy());
      acc = make_node<kernelT, evalT>(NodeKind::Prod, std::move(v));
    }
    if (b[j] && !(b[j]->is_constant() && b[j]->value() == 0.0)) {
      if (!acc) {
        acc = std::move(b[j]);
      } else {
        std::vector<Ptr> v;
        v.emplace_back(std::move(acc));
        v.emplace_back(std::move(b[j]));
        acc = make_node<kernelT, evalT>(NodeKind::Sum, std::move(v));
      }
    }
  }
  return acc ? std::move(acc) : C(0.0);
}

// ==================================================================== code emission with CSE
// Emits the outputs of several graphs as one straight-line block of SSA statements
//   const T t17 = ...;
// identical subtrees (structural key over already-numbered operands) are computed once.
// `bindings` maps variable names to C expressions (observables / parameter slots).
// Math goes through mf_* wrappers that the kernel prologue defines per precision.
struct KernelEmitter {
  std::string T;          // type of the emitted temporaries: "float", "double", or a vector type
  std::string S;          // the scalar type (literal casts, conditions)
  std::string wrap;       // applied to every temporary's expression (vector code: broadcast of scalars)
  bool fp32 = false;
  std::ostringstream body;
  std::unordered_map<std::string, std::string> memo;
  std::map<std::string, std::string> bindings;
  int next = 0;
  int statements = 0;
  // issue cost of the emitted statements in full-rate VALU operations: a transcendental that the
  // hardware evaluates (FP32 exp2, log2, rcp, sqrt, sin, cos) issues at half rate on gfx950 (8
  // cycles per wave64 against 4, measured); FP64 ones and the rest are library sequences
  double cost = 0.0;
  std::string prefix = "mf_";  // math wrapper prefix ("mfd_" for the FP64 hoisted terms)
  std::string tmp = "t";       // SSA name prefix
  // how often each operand (by structural hash) occurs in the products of the graphs about
  // to be emitted; products are then emitted as chains with their operands in one global order
  // (prod_rank), so operands shared by several products form common prefixes that are computed
  // once.  The order is the one with the fewest distinct chain prefixes over all products: all
  // orders are tried for up to 7 distinct operands, larger sets start from falling frequency and
  // move single operands while that removes prefixes.
  std::unordered_map<uint64_t, int> prod_freq;
  bool search_order = true;  // false: falling frequency, ties by emitted text
  std::unordered_map<uint64_t, int> prod_rank;
  std::vector<std::vector<uint64_t>> prod_lists;  // operand hashes of each distinct product
  std::unordered_map<const void*, uint64_t> hash_memo;

  explicit KernelEmitter(bool is_fp32) : T(is_fp32 ? "float" : "double"), S(T), fp32(is_fp32) {}

  std::string literal(double v) const {
    if (std::isinf(v)) return v > 0 ? "(" + S + ")(1.0/0.0)" : "(" + S + ")(-1.0/0.0)";
    if (std::isnan(v)) return "(" + S + ")(0.0/0.0)";
    char buf[64];
    if (fp32) std::snprintf(buf, sizeof(buf), "%af", (double)(float)v);
    else std::snprintf(buf, sizeof(buf), "%a", v);
    return buf;
  }
  double op_cost(const std::string& expr) const {
    if (!expr.empty() && expr[0] == '-') return 0.0;  // a negation folds into its user (a source modifier)
    if (expr.compare(0, prefix.size(), prefix) != 0) return 1.0;
    const std::string f = expr.substr(prefix.size(), expr.find('(') - prefix.size());
    const bool hw = f == "exp2" || f == "log2" || f == "rcp" || f == "sqrt" || f == "sin" || f == "cos";
    if (fp32) return hw ? 2.0 : (f == "exp" || f == "log") ? 5.0 : 20.0;
    return (f == "rcp" || f == "sqrt") ? 8.0 : 20.0;
  }
  std::string fresh(const std::string& key, const std::string& expr) {
    std::string n = tmp + std::to_string(next++);
    body << "  const " << T << " " << n << " = " << (wrap.empty() ? expr : wrap + "(" + expr + ")") << ";\n";
    statements++;
    cost += op_cost(expr);
    memo[key] = n;
    return n;
  }
  std::string fn(NodeKind k) const {
    switch (k) {
      case NodeKind::Exp: return prefix + "exp";
      case NodeKind::Log: return prefix + "log";
      case NodeKind::Sqrt: return prefix + "sqrt";
      case NodeKind::Sin: return prefix + "sin";
      case NodeKind::Cos: return prefix + "cos";
      case NodeKind::Tan: return prefix + "tan";
      case NodeKind::Asin: return prefix + "asin";
      case NodeKind::Acos: return prefix + "acos";
      case NodeKind::Atan: return prefix + "atan";
      case NodeKind::Erf: return prefix + "erf";
      case NodeKind::Exp2: return prefix + "exp2";
      case NodeKind::Log2: return prefix + "log2";
      case NodeKind::Inv: return prefix + "rcp";
      default: return "";
    }
  }

  template <typename kernelT, typename evalT>
  uint64_t node_hash(const ComputeGraphNode<kernelT, evalT>& n) {
    auto it = hash_memo.find(&n);
    if (it != hash_memo.end()) return it->second;
    return hash_memo[&n] = structural_hash(n);
  }
  // counts each distinct subtree once (the graphs hold many copies of shared subtrees, e.g.
  // the Gaussian factor in every monomial of a linear form)
  std::unordered_set<uint64_t> counted;
  template <typename kernelT, typename evalT>
  void count_products(const ComputeGraphNode<kernelT, evalT>& n) {
    if (n.children_.empty() || !counted.insert(node_hash(n)).second) return;
    if (n.kind() == NodeKind::Prod) {
      // products that contain a factor (a repeated factor counts once per product)
      std::unordered_set<uint64_t> seen;
      std::vector<uint64_t> ops;
      for (const auto& c : n.children_) {
        if (seen.insert(node_hash(*c)).second) prod_freq[node_hash(*c)]++;
        if (!c->is_constant()) ops.push_back(node_hash(*c));
      }
      if (ops.size() > 1) prod_lists.push_back(std::move(ops));
    }
    for (const auto& c : n.children_) count_products(*c);
    prod_rank.clear();
  }
  // multiplications the chains of all counted products take with operands in the order `rank`
  size_t chain_cost(const std::unordered_map<uint64_t, int>& rank) const {
    std::set<std::vector<uint64_t>> prefixes;
    for (auto ops : prod_lists) {
      std::stable_sort(ops.begin(), ops.end(), [&](uint64_t a, uint64_t b) { return rank.at(a) < rank.at(b); });
      for (size_t k = 2; k <= ops.size(); k++) prefixes.insert(std::vector<uint64_t>(ops.begin(), ops.begin() + k));
    }
    return prefixes.size();
  }
  void order_products() {
    std::vector<uint64_t> f;
    for (const auto& [h, n] : prod_freq) f.push_back(h);
    // falling frequency, then hash: the starting order and the tie-break
    std::sort(f.begin(), f.end(), [&](uint64_t a, uint64_t b) { return prod_freq[a] != prod_freq[b] ? prod_freq[a] > prod_freq[b] : a < b; });
    auto ranks = [&](const std::vector<uint64_t>& order) {
      std::unordered_map<uint64_t, int> r;
      for (size_t i = 0; i < order.size(); i++) r[order[i]] = (int)i;
      return r;
    };
    prod_rank = ranks(f);
    // only operands of products with more than two operands take part in chains
    std::set<uint64_t> in_chains;
    for (const auto& ops : prod_lists)
      if (ops.size() > 2) in_chains.insert(ops.begin(), ops.end());
    std::vector<uint64_t> sub;
    for (uint64_t h : f)
      if (in_chains.count(h)) sub.push_back(h);
    if (sub.size() < 2) return;
    auto with_rest = [&](const std::vector<uint64_t>& head) {
      std::vector<uint64_t> o(head);
      for (uint64_t h : f)
        if (!in_chains.count(h)) o.push_back(h);
      return o;
    };
    if (sub.size() > 7) {
      // more operands: moves of one operand to another place while they remove prefixes
      size_t best = chain_cost(prod_rank);
      for (bool improved = true; improved;) {
        improved = false;
        for (size_t i = 0; i < sub.size(); i++)
          for (size_t j = 0; j < sub.size(); j++) {
            if (i == j) continue;
            std::vector<uint64_t> o(sub);
            const uint64_t h = o[i];
            o.erase(o.begin() + i);
            o.insert(o.begin() + j, h);
            const size_t c = chain_cost(ranks(with_rest(o)));
            if (c < best) {
              best = c;
              sub = o;
              improved = true;
            }
          }
      }
      prod_rank = ranks(with_rest(sub));
      return;
    }
    std::vector<int> perm(sub.size());
    for (size_t i = 0; i < perm.size(); i++) perm[i] = (int)i;
    size_t best = chain_cost(prod_rank);
    std::vector<uint64_t> best_order = f;
    while (std::next_permutation(perm.begin(), perm.end())) {
      std::vector<uint64_t> h;
      for (int i : perm) h.push_back(sub[i]);
      const std::vector<uint64_t> o = with_rest(h);
      const size_t c = chain_cost(ranks(o));
      if (c < best) {
        best = c;
        best_order = o;
      }
    }
    prod_rank = ranks(best_order);
  }


  template <typename kernelT, typename evalT>
  std::string emit(const ComputeGraphNode<kernelT, evalT>& n) {
    switch (n.kind()) {
      case NodeKind::Constant: return literal(n.value());
      case NodeKind::Variable: {
        auto it = bindings.find(n.name());
        if (it == bindings.end()) throw std::runtime_error("unbound variable in kernel graph: " + n.name());
        return it->second;
      }
      default: break;
    }
    std::vector<std::string> ops;
    for (const auto& c : n.children_) ops.push_back(emit(*c));
    std::string key = std::to_string((int)n.kind()) + ":" + std::to_string(n.comparator()) + "(";
    {
      std::vector<std::string> k2 = ops;
      if (n.kind() == NodeKind::Sum || n.kind() == NodeKind::Prod) std::sort(k2.begin(), k2.end());
      for (auto& s : k2) key += s + ",";
      key += ")";
    }
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
    if (n.kind() == NodeKind::Prod && !prod_freq.empty() && ops.size() > 2) {
      if (search_order && prod_rank.empty()) order_products();
      std::vector<size_t> idx(ops.size());
      std::vector<int> rank(ops.size());
      for (size_t i = 0; i < ops.size(); i++) {
        idx[i] = i;
        if (search_order) {
          auto r = prod_rank.find(node_hash(*n.children_[i]));
          rank[i] = r == prod_rank.end() ? (int)prod_rank.size() : r->second;
        } else {
          auto f = prod_freq.find(node_hash(*n.children_[i]));
          rank[i] = f == prod_freq.end() ? 0 : -f->second;  // falling frequency
        }
      }
      std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return rank[a] != rank[b] ? rank[a] < rank[b] : ops[a] < ops[b]; });
      std::string acc = ops[idx[0]];
      for (size_t k = 1; k < idx.size(); k++) {
        // each step keyed like a two-factor product node (operands sorted), so chains and
        // binary products share their common prefixes
        const std::string& b = ops[idx[k]];
        const std::string pk = std::to_string((int)NodeKind::Prod) + ":0(" + std::min(acc, b) + "," + std::max(acc, b) + ",)";
        auto pi = memo.find(pk);
        acc = pi != memo.end() ? pi->second : fresh(pk, acc + " * " + b);
      }
      memo[key] = acc;
      return acc;
    }
    std::string e;
    switch (n.kind()) {
      case NodeKind::Sum:
        for (size_t i = 0; i < ops.size(); i++) e += (i ? " + " : "") + ops[i];
        break;
      case NodeKind::Prod:
        for (size_t i = 0; i < ops.size(); i++) e += (i ? " * " : "") + ops[i];
        break;
      case NodeKind::Neg: e = "-" + ops[0]; break;
      case NodeKind::Pow: {
        const auto& ex = n.child(1);
        if (ex.is_constant() && ex.value() == std::floor(ex.value()) && std::fabs(ex.value()) <= 4 && ex.value() != 0) {
          int k = (int)std::fabs(ex.value());
          for (int i = 0; i < k; i++) e += (i ? " * " : "") + ops[0];
          if (ex.value() < 0) e = prefix + "rcp(" + e + ")";
        } else {
          e = prefix + "pow(" + ops[0] + ", " + ops[1] + ")";
        }
        break;
      }
      case NodeKind::Conditional: {
        static const char* cs[] = {"==", "!=", ">", "<", ">=", "<="};
        if (wrap.empty()) e = "(" + ops[0] + " " + cs[n.comparator()] + " (" + S + ")0) ? " + ops[1] + " : " + ops[2];
        else e = "(" + ops[0] + " " + cs[n.comparator()] + " (" + S + ")0) ? " + wrap + "(" + ops[1] + ") : " + wrap + "(" + ops[2] + ")";
        break;
      }
      default: e = fn(n.kind()) + "(" + ops[0] + ")"; break;
    }
    return fresh(key, e);
  }
};

}  // namespace morefit

#endif

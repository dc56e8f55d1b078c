// pivot_check.cc -- CPU check of the expression bounds and the exponential pivot that the code
// generator applies to the per-event monomials (morefit/graph.hh: interval_of, exp_pivot).
//  1. bounds hold: expressions sampled at random points of the box stay inside interval_of's
//     bounds, and a square is bounded below by zero;
//  2. the pivot is exact: for p = c1 exp2(a1) + c2 exp2(a2), the rewritten log p and
//     exp2(a_i)..exp2(a_j) / p^m monomials equal the originals at random points, and the pivot
//     is the term whose exponent bounds qualify;
//  3. the pivot refuses: an unbounded parameter, a monomial with more powers of 1/p than
//     exponentials, and a density term that is not exponential leave the monomials unchanged;
//  4. the host program of a set of expressions (shared subexpressions computed once) returns
//     bitwise the values of evaluating each expression's graph;
//  5. the double-angle rewrite: cos(2x) and sin(2x) become products of cos(x) and sin(x) where
//     those occur, the values stay the same, and cos(2x) alone (no cos(x) in the graph) is kept.
// Prints "ok" or "FAIL <reason>" and exits non-zero on failure.
#include <cmath>
#include <cstdio>
#include <random>

#include "morefit/graph.hh"
#include "morefit/hostprog.hh"

using namespace morefit;

namespace {

typedef GraphPtr<float, double> G;

G V(const char* n) { return make_node<float, double>(NodeKind::Variable, {}, 0.0, n); }
G K(double v) { return make_node<float, double>(NodeKind::Constant, {}, v); }
G U(NodeKind k, G a) {
  std::vector<G> v;
  v.emplace_back(std::move(a));
  return make_node<float, double>(k, std::move(v));
}
G B(NodeKind k, G a, G b) {
  std::vector<G> v;
  v.emplace_back(std::move(a));
  v.emplace_back(std::move(b));
  return make_node<float, double>(k, std::move(v));
}
G P3(G a, G b, G c) {
  std::vector<G> v;
  v.emplace_back(std::move(a));
  v.emplace_back(std::move(b));
  v.emplace_back(std::move(c));
  return make_node<float, double>(NodeKind::Prod, std::move(v));
}

int fails = 0;
void check(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL %s\n", what);
    fails++;
  }
}

const std::vector<std::string> names = {"x", "mu", "s", "l"};
const double lo[4] = {-0.3, 5.0, 0.005, -10.0}, hi[4] = {0.3, 5.6, 0.13, 10.0};

std::map<std::string, Interval> box(bool with_l = true) {
  std::map<std::string, Interval> b;
  for (int i = 0; i < 4; i++)
    if (with_l || names[i] != "l") b[names[i]] = Interval{lo[i], hi[i]};
  return b;
}

// centred Gaussian exponent in base 2: -log2(e) (x - mu + 5.3)^2 / (2 s^2), written as a product
// of the shifted variable with itself
G gauss_exponent() {
  auto d = [] { return B(NodeKind::Sum, V("x"), B(NodeKind::Sum, U(NodeKind::Neg, V("mu")), K(5.3))); };
  return P3(K(-1.4426950408889634), B(NodeKind::Prod, d(), d()), U(NodeKind::Inv, P3(K(2.0), V("s"), V("s"))));
}
G exp_exponent() { return P3(K(1.4426950408889634), V("l"), V("x")); }

// p = c1 exp2(a1) + c2 exp2(a2) (+ extra)
G density(G extra = nullptr) {
  std::vector<G> t;
  t.emplace_back(P3(K(0.3), U(NodeKind::Inv, V("s")), U(NodeKind::Exp2, gauss_exponent())));
  t.emplace_back(B(NodeKind::Prod, B(NodeKind::Sum, K(0.7), U(NodeKind::Neg, K(0.1))), U(NodeKind::Exp2, exp_exponent())));
  if (extra) t.emplace_back(std::move(extra));
  return make_node<float, double>(NodeKind::Sum, std::move(t));
}

G prod(std::vector<G> f) { return make_node<float, double>(NodeKind::Prod, std::move(f)); }

std::vector<G> monomials(G extra = nullptr, bool unbalanced = false) {
  std::vector<G> m;
  m.emplace_back(U(NodeKind::Log2, density(extra ? extra->copy() : nullptr)));
  {
    std::vector<G> f;
    f.emplace_back(U(NodeKind::Exp2, gauss_exponent()));
    f.emplace_back(U(NodeKind::Inv, density(extra ? extra->copy() : nullptr)));
    m.emplace_back(prod(std::move(f)));
  }
  {
    std::vector<G> f;
    f.emplace_back(U(NodeKind::Exp2, exp_exponent()));
    f.emplace_back(U(NodeKind::Inv, density(extra ? extra->copy() : nullptr)));
    f.emplace_back(V("x"));
    m.emplace_back(prod(std::move(f)));
  }
  {
    std::vector<G> f;
    f.emplace_back(U(NodeKind::Exp2, gauss_exponent()));
    f.emplace_back(U(NodeKind::Exp2, exp_exponent()));
    f.emplace_back(U(NodeKind::Inv, density(extra ? extra->copy() : nullptr)));
    f.emplace_back(U(NodeKind::Inv, density(extra ? extra->copy() : nullptr)));
    if (unbalanced) f.emplace_back(U(NodeKind::Inv, density(extra ? extra->copy() : nullptr)));
    m.emplace_back(prod(std::move(f)));
  }
  return m;
}

std::vector<double> sample(std::mt19937_64& rng) {
  std::vector<double> v(4);
  for (int i = 0; i < 4; i++) v[i] = std::uniform_real_distribution<double>(lo[i], hi[i])(rng);
  return v;
}

}  // namespace

int main() {
  std::mt19937_64 rng(7);
  // 1. bounds
  {
    std::vector<G> ex;
    ex.emplace_back(gauss_exponent());
    ex.emplace_back(exp_exponent());
    ex.emplace_back(B(NodeKind::Sum, P3(U(NodeKind::Exp2, exp_exponent()), U(NodeKind::Inv, V("s")), K(-2.0)),
                      B(NodeKind::Pow, B(NodeKind::Sum, V("x"), U(NodeKind::Neg, V("mu"))), K(3.0))));
    ex.emplace_back(U(NodeKind::Log2, B(NodeKind::Sum, K(1.0), B(NodeKind::Prod, V("x"), V("x")))));
    ex.emplace_back(B(NodeKind::Pow, V("s"), K(-2.0)));
    ex.emplace_back(U(NodeKind::Sqrt, B(NodeKind::Prod, V("mu"), U(NodeKind::Cos, V("l")))));  // sqrt of a possibly negative value
    std::vector<Interval> iv;
    for (auto& e : ex) iv.push_back(interval_of(*e, box()));
    check(iv[0].hi == 0.0 && iv[0].lo < -1e4, "square bounded below by zero");
    check(std::fabs(iv[1].hi - 1.4426950408889634 * 3.0) < 1e-12 && std::fabs(iv[1].lo + 1.4426950408889634 * 3.0) < 1e-12,
          "linear exponent bounds");
    check(!iv[5].bounded(), "sqrt of a range across zero is unbounded");
    for (int n = 0; n < 200000; n++) {
      auto v = sample(rng);
      for (size_t k = 0; k < ex.size(); k++) {
        const double y = ex[k]->eval(names, v);
        if (!(y >= iv[k].lo - 1e-9 * std::fabs(iv[k].lo) && y <= iv[k].hi + 1e-9 * std::fabs(iv[k].hi)) && !std::isnan(y)) {
          std::printf("FAIL bound %zu: %g outside [%g, %g]\n", k, y, iv[k].lo, iv[k].hi);
          return 1;
        }
      }
    }
  }
  // 2. the rewrite is exact and picks the exponential term as pivot
  {
    std::vector<std::string> th = {"mu", "s", "l"};
    auto orig = monomials();
    auto rw = monomials();
    G pivot;
    check(exp_pivot(rw, th, box(), "pv", pivot), "pivot applies");
    if (pivot) check(pivot->equals(*exp_exponent()), "pivot is the exponential term (the Gaussian's would overflow)");
    // distinct exponentials the monomials evaluate per event: two before, one after
    std::set<std::string> before, after;
    std::function<void(const ComputeGraphNode<float, double>&, std::set<std::string>&)> collect = [&](const ComputeGraphNode<float, double>& n,
                                                                                                    std::set<std::string>& out) {
      if (n.kind() == NodeKind::Exp2) out.insert(n.get_kernel());
      for (const auto& c : n.children_) collect(*c, out);
    };
    for (auto& m : orig) collect(*m, before);
    for (auto& m : rw) collect(*m, after);
    check(before.size() == 2 && after.size() == 1, "one distinct exponential instead of two");
    std::vector<std::string> nm(names);
    nm.push_back("pv");
    double worst = 0.0;
    for (int n = 0; n < 20000 && pivot; n++) {
      auto v = sample(rng);
      std::vector<double> w(v);
      w.push_back(pivot->eval(names, v));
      for (size_t k = 0; k < orig.size(); k++) {
        const double a = orig[k]->eval(names, v), b = rw[k]->eval(nm, w);
        if (std::isfinite(a) && a != 0.0) worst = std::max(worst, std::fabs(a - b) / std::max(1.0, std::fabs(a)));
      }
    }
    if (!(worst < 1e-12)) std::printf("FAIL rewritten monomials differ by %g\n", worst), fails++;
  }
  // 3. refusals leave the monomials as they are
  {
    std::vector<std::string> th = {"mu", "s", "l"};
    auto same = [](const std::vector<G>& a, const std::vector<G>& b) {
      if (a.size() != b.size()) return false;
      for (size_t i = 0; i < a.size(); i++)
        if (!a[i]->equals(*b[i])) return false;
      return true;
    };
    G pv;
    auto m1 = monomials();
    check(!exp_pivot(m1, th, box(false), "pv", pv) && same(m1, monomials()), "unbounded parameter: no pivot");
    auto m2 = monomials(nullptr, true);
    check(!exp_pivot(m2, th, box(), "pv", pv) && same(m2, monomials(nullptr, true)), "unbalanced monomial: no pivot");
    auto m3 = monomials(B(NodeKind::Prod, K(0.01), V("x")));
    check(!exp_pivot(m3, th, box(), "pv", pv) && same(m3, monomials(B(NodeKind::Prod, K(0.01), V("x")))),
          "non-exponential density term: no pivot");
  }
  // 4. host program == graph evaluation, bitwise
  {
    std::vector<G> ex;
    ex.emplace_back(gauss_exponent());
    ex.emplace_back(U(NodeKind::Exp2, exp_exponent()));
    ex.emplace_back(density());
    ex.emplace_back(U(NodeKind::Erf, P3(V("x"), V("l"), U(NodeKind::Inv, V("s")))));
    ex.emplace_back(make_node<float, double>(NodeKind::Conditional, [] {
      std::vector<G> c;
      c.emplace_back(V("l"));
      c.emplace_back(K(0.6));
      c.emplace_back(U(NodeKind::Inv, V("l")));
      return c;
    }(), 0.0, "", 0));
    ex.emplace_back(B(NodeKind::Pow, V("s"), K(-1.5)));
    std::vector<const ComputeGraphNode<float, double>*> raw;
    for (auto& e : ex) raw.push_back(e.get());
    HostProgram<float, double> prog(raw, names);
    check(prog.outputs() == ex.size(), "host program outputs");
    std::vector<double> out(ex.size());
    for (int n = 0; n < 20000; n++) {
      auto v = sample(rng);
      if (n == 0) v[3] = 0.0;  // the conditional's other branch
      prog.run(v.data(), out.data());
      for (size_t k = 0; k < ex.size(); k++) {
        const double want = ex[k]->eval(names, v);
        if (!(out[k] == want || (std::isnan(out[k]) && std::isnan(want)))) {
          std::printf("FAIL host program %zu: %.17g vs %.17g\n", k, out[k], want);
          return 1;
        }
      }
    }
  }
  // 5. double angles
  {
    auto two_x = [] { return B(NodeKind::Prod, K(2.0), V("x")); };
    // cos(x) sin(x) + cos(2x) + sin(2x) * mu
    std::vector<G> t;
    t.emplace_back(B(NodeKind::Prod, U(NodeKind::Cos, V("x")), U(NodeKind::Sin, V("x"))));
    t.emplace_back(U(NodeKind::Cos, two_x()));
    t.emplace_back(B(NodeKind::Prod, U(NodeKind::Sin, two_x()), V("mu")));
    G g = make_node<float, double>(NodeKind::Sum, std::move(t));
    G r = double_angle(*g);
    std::set<std::string> trig;
    std::function<void(const ComputeGraphNode<float, double>&)> walk = [&](const ComputeGraphNode<float, double>& n) {
      if (n.kind() == NodeKind::Cos || n.kind() == NodeKind::Sin) trig.insert(n.get_kernel());
      for (const auto& c : n.children_) walk(*c);
    };
    walk(*r);
    check(trig.size() == 2, "double angles: only cos(x) and sin(x) left");
    double worst = 0.0;
    for (int n = 0; n < 20000; n++) {
      auto v = sample(rng);
      v[0] *= 10.0;  // x over [-3, 3]
      worst = std::max(worst, std::fabs(g->eval(names, v) - r->eval(names, v)));
    }
    if (!(worst < 1e-14)) std::printf("FAIL double angles differ by %g\n", worst), fails++;
    G lone = U(NodeKind::Cos, two_x());
    check(double_angle(*lone)->equals(*lone), "cos(2x) without cos(x) kept");
    G other = B(NodeKind::Sum, U(NodeKind::Cos, V("x")), U(NodeKind::Cos, B(NodeKind::Prod, K(3.0), V("x"))));
    check(double_angle(*other)->equals(*other), "cos(3x) kept");
  }
  if (fails) return 1;
  std::printf("ok\n");
  return 0;
}

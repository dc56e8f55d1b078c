/**
 * @file pdf.hh
 * Probability densities as symbolic graphs.  Same class names, constructor signatures and
 * virtual interface as the reference pdf.hh:29-136 (prob, norm, logprob, lognorm,
 * prob_normalised, logprob_normalised, definite_integral, get_max, is_extended), so user
 * code written against the reference builds unchanged.
 *
 * Densities implemented (the reference's formulas, written here from the mathematics):
 *   GaussianPDF     exp(-(x-mu)^2/(2 sigma^2)) / (sqrt(2 pi) sigma),
 *                   norm 1/2 [erf((b-mu)/(sqrt2 sigma)) - erf((a-mu)/(sqrt2 sigma))]    (pdf.hh:147-154)
 *   ExponentialPDF  exp(alpha x), norm (exp(alpha b) - exp(alpha a))/alpha, b-a at alpha=0 (pdf.hh:431-440)
 *   SumPDF          sum_i f_i p_i + (1 - sum f) p_last, or extended sum_i N_i p_i / sum N (pdf.hh:526-554)
 *   ProdPDF         product of normalised children over distinct observables          (pdf.hh:654-714)
 *   CrystalballPDF  Gaussian core exp(-t^2/2), t = (x-mu)/sigma, for t > -alpha; power-law tail
 *                   A (B - t)^-n below, A = (n/alpha)^n exp(-alpha^2/2), B = n/alpha - alpha;
 *                   norm = tail integral (n != 1: A sigma^n [(C-m)^(1-n) - (C-a)^(1-n)]/(n-1),
 *                   n = 1: A sigma log((C-a)/(C-m)), C = mu + sigma B, m = mu - alpha sigma) +
 *                   core integral sigma sqrt(pi/2) [erf(alpha/sqrt2) - erf((mu-b)/(sqrt2 sigma))]
 *                                                                                     (pdf.hh:195-419)
 *   PolynomialPDF   1 + c_1 x + ... + c_k x^k (Horner), norm P(b) - P(a) with
 *                   P(x) = x + c_1 x^2/2 + ... + c_k x^(k+1)/(k+1)                     (pdf.hh:716-892)
 */
#ifndef MOREFIT_PDF_HH
#define MOREFIT_PDF_HH

#include <cassert>
#include <cmath>
#include <functional>
#include <set>
#include <vector>

#include "dimensionvector.hh"
#include "graph.hh"
#include "parametervector.hh"

namespace morefit {

template <typename kernelT = double, typename evalT = double>
class PDF {
 public:
  typedef GraphPtr<kernelT, evalT> Ptr;
  std::vector<dimension<evalT>*> dimensions_;
  std::vector<parameter<evalT>*> parameters_;
  std::vector<PDF<kernelT, evalT>*> children_;

  virtual ~PDF() = default;
  unsigned int nparameters() { return (unsigned int)parameters_.size(); }
  unsigned int ndimensions() { return (unsigned int)dimensions_.size(); }
  const std::vector<dimension<evalT>*>& dimensions() const { return dimensions_; }
  const std::vector<parameter<evalT>*>& parameters() const { return parameters_; }
  evalT from(unsigned int idx) const { return dimensions_.at(idx)->get_min(); }
  evalT to(unsigned int idx) const { return dimensions_.at(idx)->get_max(); }
  virtual evalT get_max() const { return -1.0; }
  // unnormalised density
  virtual Ptr prob() const = 0;
  virtual Ptr logprob() const { return Log<kernelT, evalT>(prob()); }
  // integral of prob() over the observable ranges
  virtual Ptr norm() const = 0;
  // integral of prob()/norm() over [from, to] with the boundaries as the variables
  // dimension::get_from_name()/get_to_name()
  virtual Ptr definite_integral() const { return Constant<kernelT, evalT>(0.0); }
  virtual Ptr lognorm() const { return Log<kernelT, evalT>(norm()); }
  virtual Ptr prob_normalised() const { return prob() / norm(); }
  virtual Ptr logprob_normalised() const { return logprob() - lognorm(); }
  virtual std::vector<Ptr> logprob_diffs() const {
    std::vector<Ptr> r;
    for (auto* p : parameters_) r.emplace_back(logprob()->diff(p->get_name()));
    return r;
  }
  virtual std::vector<Ptr> lognorm_diffs() const {
    std::vector<Ptr> r;
    for (auto* p : parameters_) r.emplace_back(lognorm()->diff(p->get_name()));
    return r;
  }
  virtual std::vector<Ptr> logprob_normalised_diffs() const {
    std::vector<Ptr> r;
    for (auto* p : parameters_) r.emplace_back(logprob_normalised()->diff(p->get_name()));
    return r;
  }
  std::string get_kernel() const { return logprob_normalised()->get_kernel(); }
  virtual bool is_extended() const { return false; }
  // every parameter the density depends on, children first (used by generators/fitters)
  std::vector<parameter<evalT>*> all_parameters() const {
    std::vector<parameter<evalT>*> r;
    for (auto* c : children_)
      for (auto* p : c->all_parameters())
        if (std::find(r.begin(), r.end(), p) == r.end()) r.push_back(p);
    for (auto* p : parameters_)
      if (std::find(r.begin(), r.end(), p) == r.end()) r.push_back(p);
    return r;
  }
};

template <typename kernelT = double, typename evalT = double>
class GaussianPDF : public PDF<kernelT, evalT> {
 public:
  typedef GraphPtr<kernelT, evalT> Ptr;
  GaussianPDF(dimension<evalT>* x, parameter<evalT>* mu, parameter<evalT>* sigma) {
    this->dimensions_ = {x};
    this->parameters_ = {mu, sigma};
  }
  dimension<evalT>* x() const { return this->dimensions_.at(0); }
  parameter<evalT>* mu() const { return this->parameters_.at(0); }
  parameter<evalT>* sigma() const { return this->parameters_.at(1); }
  evalT from() const { return x()->get_min(); }
  evalT to() const { return x()->get_max(); }
  Ptr X() const { return Variable<kernelT, evalT>(x()->get_name()); }
  Ptr MU() const { return Variable<kernelT, evalT>(mu()->get_name()); }
  Ptr SG() const { return Variable<kernelT, evalT>(sigma()->get_name()); }
  // erf((bound - mu) / (sqrt(2) sigma))
  Ptr erf_at(Ptr bound) const { return Erf<kernelT, evalT>((std::move(bound) - MU()) / (Sqrt<kernelT, evalT>(2.0) * SG())); }
  Ptr prob() const override {
    Ptr d = X() - MU();
    Ptr d2 = X() - MU();
    return (1.0 / std::sqrt(2.0 * M_PI)) / SG() * Exp<kernelT, evalT>(-(std::move(d) * std::move(d2) / (2.0 * SG() * SG())));
  }
  Ptr norm() const override {
    return 0.5 * (erf_at(Constant<kernelT, evalT>(to())) - erf_at(Constant<kernelT, evalT>(from())));
  }
  Ptr definite_integral() const override {
    return 0.5 * (erf_at(Variable<kernelT, evalT>(x()->get_to_name())) - erf_at(Variable<kernelT, evalT>(x()->get_from_name()))) / norm();
  }
  evalT get_max() const override {
    evalT s = sigma()->get_value(), m = mu()->get_value();
    evalT n = 0.5 * (std::erf((to() - m) / (std::sqrt(2.0) * s)) - std::erf((from() - m) / (std::sqrt(2.0) * s)));
    return 1.0 / std::sqrt(2.0 * M_PI) / s / n;
  }
};

template <typename kernelT = double, typename evalT = double>
class ExponentialPDF : public PDF<kernelT, evalT> {
 public:
  typedef GraphPtr<kernelT, evalT> Ptr;
  ExponentialPDF(dimension<evalT>* x, parameter<evalT>* alpha) {
    this->dimensions_ = {x};
    this->parameters_ = {alpha};
  }
  dimension<evalT>* x() const { return this->dimensions_.at(0); }
  parameter<evalT>* alpha() const { return this->parameters_.at(0); }
  evalT from() const { return x()->get_min(); }
  evalT to() const { return x()->get_max(); }
  Ptr A() const { return Variable<kernelT, evalT>(alpha()->get_name()); }
  Ptr prob() const override { return Exp<kernelT, evalT>(A() * Variable<kernelT, evalT>(x()->get_name())); }
  Ptr norm() const override {
    return ConditionalEqual<kernelT, evalT>(A(), Constant<kernelT, evalT>(to() - from()),
                                            (Exp<kernelT, evalT>(A() * to()) - Exp<kernelT, evalT>(A() * from())) / A());
  }
  Ptr definite_integral() const override {
    // the branches are built first: argument evaluation order is unspecified
    Ptr width = Variable<kernelT, evalT>(x()->get_to_name()) - Variable<kernelT, evalT>(x()->get_from_name());
    Ptr expdiff = Exp<kernelT, evalT>(A() * Variable<kernelT, evalT>(x()->get_to_name())) -
                  Exp<kernelT, evalT>(A() * Variable<kernelT, evalT>(x()->get_from_name()));
    return ConditionalEqual<kernelT, evalT>(A(), std::move(width) / norm(), std::move(expdiff) / (A() * norm()));
  }
  evalT get_max() const override {
    evalT a = alpha()->get_value();
    if (a == 0.0) return 1.0 / (to() - from());
    evalT n = std::exp(a * to()) - std::exp(a * from());
    return a * std::exp(a * (a > 0.0 ? to() : from())) / n;
  }
};

template <typename kernelT, typename evalT = double>
class SumPDF : public PDF<kernelT, evalT> {
  bool extended_ = false;

 public:
  typedef GraphPtr<kernelT, evalT> Ptr;
  SumPDF(const std::vector<PDF<kernelT, evalT>*>& pdfs, const std::vector<parameter<evalT>*>& fractions) {
    this->children_ = pdfs;
    assert(pdfs.size() >= 2);
    assert(pdfs.size() == fractions.size() + 1 || pdfs.size() == fractions.size());
    extended_ = pdfs.size() == fractions.size();
    this->dimensions_ = pdfs.at(0)->dimensions();
    for (auto* c : pdfs) {  // every child must live on the same observables
      assert(c->dimensions().size() == this->dimensions_.size());
      for (auto* d : this->dimensions_) {
        bool found = false;
        for (auto* e : c->dimensions()) found |= e->get_name() == d->get_name();
        assert(found);
        (void)found;
      }
    }
    this->parameters_ = fractions;
  }
  SumPDF(PDF<kernelT, evalT>* a, PDF<kernelT, evalT>* b, parameter<evalT>* fraction)
      : SumPDF(std::vector<PDF<kernelT, evalT>*>{a, b}, std::vector<parameter<evalT>*>{fraction}) {}
  SumPDF(PDF<kernelT, evalT>* a, PDF<kernelT, evalT>* b, parameter<evalT>* na, parameter<evalT>* nb)
      : SumPDF(std::vector<PDF<kernelT, evalT>*>{a, b}, std::vector<parameter<evalT>*>{na, nb}) {}
  const std::vector<parameter<evalT>*>& fractions() const { return this->parameters_; }
  Ptr F(size_t i) const { return Variable<kernelT, evalT>(fractions().at(i)->get_name()); }

  template <typename Term>
  Ptr combine(Term term) const {
    const auto& ch = this->children_;
    std::vector<Ptr> terms;
    if (!extended_) {
      std::vector<Ptr> rest;
      rest.emplace_back(Constant<kernelT, evalT>(1.0));
      for (size_t i = 0; i + 1 < ch.size(); i++) {
        terms.emplace_back(F(i) * term(ch[i]));
        rest.emplace_back(-F(i));
      }
      terms.emplace_back(Ptr(std::make_unique<SumNode<kernelT, evalT>>(std::move(rest))) * term(ch.back()));
      return std::make_unique<SumNode<kernelT, evalT>>(std::move(terms));
    }
    std::vector<Ptr> yields;
    for (size_t i = 0; i < ch.size(); i++) {
      terms.emplace_back(F(i) * term(ch[i]));
      yields.emplace_back(F(i));
    }
    return Ptr(std::make_unique<SumNode<kernelT, evalT>>(std::move(terms))) / Ptr(std::make_unique<SumNode<kernelT, evalT>>(std::move(yields)));
  }
  Ptr prob() const override { return combine([](PDF<kernelT, evalT>* c) { return c->prob_normalised(); }); }
  Ptr norm() const override { return Constant<kernelT, evalT>(1.0); }
  Ptr definite_integral() const override { return combine([](PDF<kernelT, evalT>* c) { return c->definite_integral(); }); }
  evalT get_max() const override {
    const auto& ch = this->children_;
    evalT r = 0.0;
    if (!extended_) {
      evalT last = 1.0;
      for (size_t i = 0; i + 1 < ch.size(); i++) {
        r += fractions()[i]->get_value() * ch[i]->get_max();
        last -= fractions()[i]->get_value();
      }
      return r + last * ch.back()->get_max();
    }
    evalT tot = 0.0;
    for (auto* f : fractions()) tot += f->get_value();
    for (size_t i = 0; i < ch.size(); i++) r += fractions()[i]->get_value() / tot * ch[i]->get_max();
    return r;
  }
  bool is_extended() const override { return extended_; }
};

template <typename kernelT, typename evalT = double>
class ProdPDF : public PDF<kernelT, evalT> {
 public:
  typedef GraphPtr<kernelT, evalT> Ptr;
  explicit ProdPDF(const std::vector<PDF<kernelT, evalT>*>& pdfs) {
    this->children_ = pdfs;
    assert(pdfs.size() >= 2);
    std::set<std::string> names;
    for (auto* c : pdfs)
      for (auto* d : c->dimensions()) {
        this->dimensions_.push_back(d);
        names.insert(d->get_name());
      }
    if (names.size() != this->dimensions_.size()) {
      std::cout << "All dimensions of a product PDF need to be unique" << std::endl;
      assert(0);
    }
  }
  ProdPDF(PDF<kernelT, evalT>* a, PDF<kernelT, evalT>* b) : ProdPDF(std::vector<PDF<kernelT, evalT>*>{a, b}) {}
  Ptr prob() const override {
    std::vector<Ptr> f;
    for (auto* c : this->children_) f.emplace_back(c->prob_normalised());
    return std::make_unique<ProdNode<kernelT, evalT>>(std::move(f));
  }
  Ptr norm() const override { return Constant<kernelT, evalT>(1.0); }
  Ptr definite_integral() const override {
    std::vector<Ptr> f;
    for (auto* c : this->children_) f.emplace_back(c->definite_integral());
    return std::make_unique<ProdNode<kernelT, evalT>>(std::move(f));
  }
  evalT get_max() const override {
    evalT r = 1.0;
    for (auto* c : this->children_) r *= c->get_max();
    return r;
  }
};

template <typename kernelT = double, typename evalT = double>
class CrystalballPDF : public PDF<kernelT, evalT> {
 public:
  typedef GraphPtr<kernelT, evalT> Ptr;
  // the tail joins the core at mu - alpha sigma, which must lie inside the range (alpha > 0)
  CrystalballPDF(dimension<evalT>* x, parameter<evalT>* mu, parameter<evalT>* sigma, parameter<evalT>* alpha, parameter<evalT>* n) {
    assert(x->get_min() < mu->get_value() - alpha->get_value() * sigma->get_value() &&
           x->get_max() > mu->get_value() - alpha->get_value() * sigma->get_value());
    assert(alpha->get_value() > 0.0);
    this->dimensions_ = {x};
    this->parameters_ = {mu, sigma, alpha, n};
  }
  dimension<evalT>* x() const { return this->dimensions_.at(0); }
  parameter<evalT>* mu() const { return this->parameters_.at(0); }
  parameter<evalT>* sigma() const { return this->parameters_.at(1); }
  parameter<evalT>* alpha() const { return this->parameters_.at(2); }
  parameter<evalT>* n() const { return this->parameters_.at(3); }
  evalT from() const { return x()->get_min(); }
  evalT to() const { return x()->get_max(); }
  Ptr V(const std::string& nm) const { return Variable<kernelT, evalT>(nm); }
  Ptr X() const { return V(x()->get_name()); }
  Ptr MU() const { return V(mu()->get_name()); }
  Ptr SG() const { return V(sigma()->get_name()); }
  Ptr AL() const { return V(alpha()->get_name()); }
  Ptr N() const { return V(n()->get_name()); }
  Ptr amp() const { return Pow<kernelT, evalT>(N() / AL(), N()) * Exp<kernelT, evalT>(-0.5 * AL() * AL()); }  // A
  Ptr B() const { return N() / AL() - AL(); }
  Ptr C() const { return SG() * B() + MU(); }
  Ptr join() const { return MU() - AL() * SG(); }  // mu - alpha sigma
  // integral of the power-law tail over [lo, hi] (both at or below the join)
  Ptr tail_integral(Ptr lo, Ptr hi) const {
    Ptr lo2 = lo->copy(), hi2 = hi->copy();
    return ConditionalUnequal<kernelT, evalT>(
        N() - 1.0,
        amp() * Pow<kernelT, evalT>(SG(), N()) *
            (Pow<kernelT, evalT>(C() - std::move(hi), 1.0 - N()) - Pow<kernelT, evalT>(C() - std::move(lo), 1.0 - N())) / (N() - 1.0),
        amp() * SG() * (Log<kernelT, evalT>(C() - std::move(lo2)) - Log<kernelT, evalT>(C() - std::move(hi2))));
  }
  // integral of the Gaussian core over [lo, hi] (both at or above the join)
  Ptr core_integral(Ptr lo, Ptr hi) const {
    const double r2 = std::sqrt(2.0);
    return std::sqrt(M_PI / 2.0) * SG() *
           (Erf<kernelT, evalT>((MU() - std::move(lo)) / (SG() * r2)) - Erf<kernelT, evalT>((MU() - std::move(hi)) / (SG() * r2)));
  }
  Ptr prob() const override {
    Ptr t = (X() - MU()) / SG();
    Ptr t2 = (X() - MU()) / SG();
    return ConditionalLarger<kernelT, evalT>(X() - MU() + AL() * SG(), Exp<kernelT, evalT>(-0.5 * std::move(t) * std::move(t2)),
                                             amp() / Pow<kernelT, evalT>(B() - (X() - MU()) / SG(), N()));
  }
  Ptr norm() const override {
    return tail_integral(Constant<kernelT, evalT>(from()), join()) + core_integral(join(), Constant<kernelT, evalT>(to()));
  }
  // [from, to] below the join: tail only; above: core only; across: both parts
  Ptr definite_integral() const override {
    Ptr lo = V(x()->get_from_name()), hi = V(x()->get_to_name());
    return ConditionalSmaller<kernelT, evalT>(
               hi->copy() - join(), tail_integral(lo->copy(), hi->copy()),
               ConditionalLarger<kernelT, evalT>(lo->copy() - join(), core_integral(lo->copy(), hi->copy()),
                                                 tail_integral(lo->copy(), join()) + core_integral(join(), hi->copy()))) /
           norm();
  }
  // the unnormalised density peaks at 1 (x = mu)
  evalT get_max() const override {
    const evalT nn = n()->get_value(), s = sigma()->get_value(), a = alpha()->get_value(), m = mu()->get_value();
    const evalT A = std::pow(nn / a, nn) * std::exp(-0.5 * a * a), Bv = nn / a - a, Cv = s * Bv + m, mid = m - a * s;
    evalT integral = nn != 1.0 ? A * std::pow(s, nn) * (std::pow(Cv - mid, 1.0 - nn) - std::pow(Cv - from(), 1.0 - nn)) / (nn - 1.0)
                               : A * s * (std::log(Cv - from()) - std::log(Cv - mid));
    integral += std::sqrt(M_PI / 2.0) * s * (std::erf((m - mid) / (s * std::sqrt(2.0))) - std::erf((m - to()) / (s * std::sqrt(2.0))));
    return 1.0 / integral;
  }
};

template <typename kernelT = double, typename evalT = double>
class PolynomialPDF : public PDF<kernelT, evalT> {
 public:
  typedef GraphPtr<kernelT, evalT> Ptr;
  PolynomialPDF(dimension<evalT>* x, std::vector<parameter<evalT>*> coefficients) {
    assert(!coefficients.empty());
    this->dimensions_ = {x};
    this->parameters_ = coefficients;
  }
  dimension<evalT>* x() const { return this->dimensions_.at(0); }
  const std::vector<parameter<evalT>*>& coefficients() const { return this->parameters_; }
  evalT from() const { return x()->get_min(); }
  evalT to() const { return x()->get_max(); }
  Ptr coef(size_t i) const { return Variable<kernelT, evalT>(coefficients().at(i)->get_name()); }
  // 1 + c_1 x + ... + c_k x^k in Horner form: ((c_k x + c_{k-1}) x + ...) x + 1
  Ptr prob() const override {
    const size_t k = coefficients().size();
    Ptr p = coef(k - 1);
    for (size_t i = k - 1; i-- > 0;) p = std::move(p) * Variable<kernelT, evalT>(x()->get_name()) + coef(i);
    return std::move(p) * Variable<kernelT, evalT>(x()->get_name()) + 1.0;
  }
  // the antiderivative x + c_1 x^2/2 + ... + c_k x^(k+1)/(k+1) at the point `at`
  Ptr antiderivative(const std::function<Ptr()>& at) const {
    const size_t k = coefficients().size();
    Ptr p = (1.0 / (double)(k + 1)) * coef(k - 1);
    for (size_t i = k - 1; i-- > 0;) p = std::move(p) * at() + (1.0 / (double)(i + 2)) * coef(i);
    return (std::move(p) * at() + 1.0) * at();
  }
  Ptr norm() const override {
    const evalT a = from(), b = to();
    return antiderivative([b] { return Constant<kernelT, evalT>(b); }) - antiderivative([a] { return Constant<kernelT, evalT>(a); });
  }
  Ptr definite_integral() const override {
    const std::string lo = x()->get_from_name(), hi = x()->get_to_name();
    return (antiderivative([hi] { return Variable<kernelT, evalT>(hi); }) - antiderivative([lo] { return Variable<kernelT, evalT>(lo); })) / norm();
  }
  evalT value(evalT xv) const {
    const size_t k = coefficients().size();
    evalT r = coefficients()[k - 1]->get_value();
    for (size_t i = k - 1; i-- > 0;) r = r * xv + coefficients()[i]->get_value();
    return r * xv + 1.0;
  }
  evalT integral_value(evalT xv) const {
    const size_t k = coefficients().size();
    evalT r = coefficients()[k - 1]->get_value() / (evalT)(k + 1);
    for (size_t i = k - 1; i-- > 0;) r = r * xv + coefficients()[i]->get_value() / (evalT)(i + 2);
    return (r * xv + 1.0) * xv;
  }
  // maximum of the normalised density: the largest value on a 4096-point grid over the range,
  // refined by golden-section search around it, with the reference's 1.001 margin
  evalT get_max() const override {
    const evalT a = from(), b = to();
    const int G = 4096;
    int best = 0;
    evalT bv = value(a);
    for (int i = 1; i <= G; i++) {
      const evalT v = value(a + (b - a) * i / G);
      if (v > bv) {
        bv = v;
        best = i;
      }
    }
    evalT lo = a + (b - a) * std::max(0, best - 1) / G, hi = a + (b - a) * std::min(G, best + 1) / G;
    const evalT g = 0.5 * (std::sqrt(5.0) - 1.0);
    for (int it = 0; it < 60; it++) {
      const evalT c = hi - g * (hi - lo), d = lo + g * (hi - lo);
      if (value(c) > value(d)) hi = d;
      else lo = c;
    }
    bv = std::max(bv, value(0.5 * (lo + hi)));
    return 1.001 * bv / (integral_value(b) - integral_value(a));
  }
};

}  // namespace morefit
#endif

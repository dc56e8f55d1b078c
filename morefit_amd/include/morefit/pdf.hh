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
 */
#ifndef MOREFIT_PDF_HH
#define MOREFIT_PDF_HH

#include <cassert>
#include <cmath>
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
    Ptr hi = Variable<kernelT, evalT>(x()->get_to_name()), lo = Variable<kernelT, evalT>(x()->get_from_name());
    return ConditionalEqual<kernelT, evalT>(
        A(), (hi->copy() - lo->copy()) / norm(),
        (Exp<kernelT, evalT>(A() * std::move(hi)) - Exp<kernelT, evalT>(A() * std::move(lo))) / (A() * norm()));
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

}  // namespace morefit
#endif

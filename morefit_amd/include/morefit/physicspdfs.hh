/**
 * @file physicspdfs.hh
 * KstarmumuAngularPDF: the P-wave B0 -> K*0 l+ l- angular density in (cos theta_l,
 * cos theta_K, phi) with FL, S3, S4, S5, AFB, S7, S8, S9 (reference physicspdfs.hh:28-214):
 *   9/(32 pi) [ 3/4 (1-FL) sK^2 + FL cK^2 + 1/4 (1-FL) sK^2 cos2l - FL cK^2 cos2l
 *             + S3 sK^2 sl^2 cos2phi + S4 sin2K sin2l cosphi + S5 sin2K sl cosphi
 *             + 4/3 AFB sK^2 cl + S7 sin2K sl sinphi + S8 sin2K sin2l sinphi + S9 sK^2 sl^2 sin2phi ]
 * normalised to one over the full box (norm() = 1, physicspdfs.hh:78-81).
 */
#ifndef MOREFIT_PHYSICSPDFS_HH
#define MOREFIT_PHYSICSPDFS_HH

#include <cmath>
#include <functional>

#include "pdf.hh"

namespace morefit {

template <typename kernelT = double, typename evalT = double>
class KstarmumuAngularPDF : public PDF<kernelT, evalT> {
 public:
  typedef GraphPtr<kernelT, evalT> Ptr;
  KstarmumuAngularPDF(dimension<evalT>* ctl, dimension<evalT>* ctk, dimension<evalT>* phi, parameter<evalT>* Fl,
                      parameter<evalT>* S3, parameter<evalT>* S4, parameter<evalT>* S5, parameter<evalT>* Afb,
                      parameter<evalT>* S7, parameter<evalT>* S8, parameter<evalT>* S9) {
    this->dimensions_ = {ctl, ctk, phi};
    this->parameters_ = {Fl, S3, S4, S5, Afb, S7, S8, S9};
  }
  dimension<evalT>* ctl() const { return this->dimensions_.at(0); }
  dimension<evalT>* ctk() const { return this->dimensions_.at(1); }
  dimension<evalT>* phi() const { return this->dimensions_.at(2); }
  parameter<evalT>* Fl() const { return this->parameters_.at(0); }
  parameter<evalT>* S3() const { return this->parameters_.at(1); }
  parameter<evalT>* S4() const { return this->parameters_.at(2); }
  parameter<evalT>* S5() const { return this->parameters_.at(3); }
  parameter<evalT>* Afb() const { return this->parameters_.at(4); }
  parameter<evalT>* S7() const { return this->parameters_.at(5); }
  parameter<evalT>* S8() const { return this->parameters_.at(6); }
  parameter<evalT>* S9() const { return this->parameters_.at(7); }
  Ptr V(const std::string& n) const { return Variable<kernelT, evalT>(n); }
  Ptr P(int i) const { return V(this->parameters_.at(i)->get_name()); }

  Ptr prob() const override {
    const std::string l = ctl()->get_name(), k = ctk()->get_name(), f = phi()->get_name();
    auto cl = [&] { return V(l); };
    auto ck = [&] { return V(k); };
    auto cl2 = [&] { return V(l) * V(l); };
    auto ck2 = [&] { return V(k) * V(k); };
    auto sl2 = [&] { return 1.0 - cl2(); };
    auto sk2 = [&] { return 1.0 - ck2(); };
    auto sl = [&] { return Sqrt<kernelT, evalT>(sl2()); };
    auto sk = [&] { return Sqrt<kernelT, evalT>(sk2()); };
    auto c2l = [&] { return 2.0 * cl2() - 1.0; };
    auto s2l = [&] { return 2.0 * sl() * cl(); };
    auto s2k = [&] { return 2.0 * sk() * ck(); };
    auto cphi = [&] { return Cos<kernelT, evalT>(V(f)); };
    auto sphi = [&] { return Sin<kernelT, evalT>(V(f)); };
    auto c2phi = [&] { return Cos<kernelT, evalT>(2.0 * V(f)); };
    auto s2phi = [&] { return Sin<kernelT, evalT>(2.0 * V(f)); };
    const double c = 9.0 / 32.0 / M_PI;
    std::vector<Ptr> t;
    t.emplace_back(c * 0.75 * sk2() * (1.0 - P(0)));
    t.emplace_back(c * ck2() * P(0));
    t.emplace_back(c * 0.25 * sk2() * c2l() * (1.0 - P(0)));
    t.emplace_back(c * ck2() * c2l() * (-P(0)));
    t.emplace_back(c * sk2() * sl2() * c2phi() * P(1));
    t.emplace_back(c * s2k() * s2l() * cphi() * P(2));
    t.emplace_back(c * s2k() * sl() * cphi() * P(3));
    t.emplace_back(c * (4.0 / 3.0) * sk2() * cl() * P(4));
    t.emplace_back(c * s2k() * sl() * sphi() * P(5));
    t.emplace_back(c * s2k() * s2l() * sphi() * P(6));
    t.emplace_back(c * sk2() * sl2() * s2phi() * P(7));
    return std::make_unique<SumNode<kernelT, evalT>>(std::move(t));
  }
  Ptr norm() const override { return Constant<kernelT, evalT>(1.0); }
  // Integral over the box [from, to] of each observable (the variables get_from_name() /
  // get_to_name()): every term of the density is a product of one function of each observable,
  // so the box integral is the sum of products of one-dimensional antiderivative differences
  // (reference physicspdfs.hh:82-164).  With u = cos theta_K, v = cos theta_l, w = phi:
  //   sK^2 -> u - u^3/3, cK^2 -> u^3/3, sin2K -> -2/3 (1-u^2)^(3/2);
  //   1 -> v, cos2l -> 2v^3/3 - v, sl^2 -> v - v^3/3, sin2l -> -2/3 (1-v^2)^(3/2),
  //   sl -> (v sqrt(1-v^2) + asin v)/2, cl -> v^2/2;
  //   1 -> w, cos2w -> sin(2w)/2, cos w -> sin w, sin w -> -cos w, sin2w -> -cos(2w)/2.
  Ptr definite_integral() const override {
    auto diff = [this](dimension<evalT>* d, const std::function<Ptr(Ptr)>& F) {
      return F(V(d->get_to_name())) - F(V(d->get_from_name()));
    };
    auto cube = [](Ptr a) { Ptr b = a->copy(), c = a->copy(); return std::move(a) * std::move(b) * std::move(c); };
    auto sq = [](Ptr a) { Ptr b = a->copy(); return std::move(a) * std::move(b); };
    auto k_s2 = [&] { return diff(ctk(), [&](Ptr u) { Ptr u3 = cube(u->copy()); return std::move(u) - std::move(u3) / 3.0; }); };
    auto k_c2 = [&] { return diff(ctk(), [&](Ptr u) { return cube(std::move(u)) / 3.0; }); };
    auto three_halves = [&](Ptr a) {  // (1 - a^2)^(3/2)
      Ptr s2 = 1.0 - sq(std::move(a));
      Ptr s = Sqrt<kernelT, evalT>(s2->copy());
      return std::move(s2) * std::move(s);
    };
    auto k_s2k = [&] { return diff(ctk(), [&](Ptr u) { return (-2.0 / 3.0) * three_halves(std::move(u)); }); };
    auto l_1 = [&] { return diff(ctl(), [](Ptr v) { return v; }); };
    auto l_c2 = [&] { return diff(ctl(), [&](Ptr v) { Ptr v3 = cube(v->copy()); return (2.0 / 3.0) * std::move(v3) - std::move(v); }); };
    auto l_s2 = [&] { return diff(ctl(), [&](Ptr v) { Ptr v3 = cube(v->copy()); return std::move(v) - std::move(v3) / 3.0; }); };
    auto l_s2l = [&] { return diff(ctl(), [&](Ptr v) { return (-2.0 / 3.0) * three_halves(std::move(v)); }); };
    auto l_s = [&] {
      return diff(ctl(), [&](Ptr v) {
        Ptr root = Sqrt<kernelT, evalT>(1.0 - sq(v->copy()));
        Ptr as = Asin<kernelT, evalT>(v->copy());
        return 0.5 * (std::move(v) * std::move(root) + std::move(as));
      });
    };
    auto l_c = [&] { return diff(ctl(), [&](Ptr v) { return 0.5 * sq(std::move(v)); }); };
    auto f_1 = [&] { return diff(phi(), [](Ptr w) { return w; }); };
    auto f_c2 = [&] { return diff(phi(), [](Ptr w) { return 0.5 * Sin<kernelT, evalT>(2.0 * std::move(w)); }); };
    auto f_c = [&] { return diff(phi(), [](Ptr w) { return Sin<kernelT, evalT>(std::move(w)); }); };
    auto f_s = [&] { return diff(phi(), [](Ptr w) { return -Cos<kernelT, evalT>(std::move(w)); }); };
    auto f_s2 = [&] { return diff(phi(), [](Ptr w) { return -0.5 * Cos<kernelT, evalT>(2.0 * std::move(w)); }); };
    const double c = 9.0 / 32.0 / M_PI;
    std::vector<Ptr> t;
    t.emplace_back(c * 0.75 * k_s2() * l_1() * f_1() * (1.0 - P(0)));
    t.emplace_back(c * k_c2() * l_1() * f_1() * P(0));
    t.emplace_back(c * 0.25 * k_s2() * l_c2() * f_1() * (1.0 - P(0)));
    t.emplace_back(c * k_c2() * l_c2() * f_1() * (-P(0)));
    t.emplace_back(c * k_s2() * l_s2() * f_c2() * P(1));
    t.emplace_back(c * k_s2k() * l_s2l() * f_c() * P(2));
    t.emplace_back(c * k_s2k() * l_s() * f_c() * P(3));
    t.emplace_back(c * (4.0 / 3.0) * k_s2() * l_c() * f_1() * P(4));
    t.emplace_back(c * k_s2k() * l_s() * f_s() * P(5));
    t.emplace_back(c * k_s2k() * l_s2l() * f_s() * P(6));
    t.emplace_back(c * k_s2() * l_s2() * f_s2() * P(7));
    return std::make_unique<SumNode<kernelT, evalT>>(std::move(t));
  }
  evalT get_max() const override {
    auto a = [this](int i) { return std::fabs(this->parameters_.at(i)->get_value()); };
    return 9.0 / 32.0 / M_PI * (4.0 + a(1) + a(2) + a(3) + 4.0 / 3.0 * a(4) + a(5) + a(6) + a(7));
  }
};

}  // namespace morefit
#endif

// diff_rules_check.cc -- CPU check of the symbolic derivative rules of morefit/graph.hh (the
// rules of the reference graph.hh diff(): Pow 1101-1150, the trigonometric and inverse
// trigonometric nodes, Erf 2067-2079, Conditional 2291-2299; the reference's asin / acos / tan
// rules are wrong and this build's are the textbook ones).  For every node kind, with an inner
// argument a(t) that depends on the variable (so the chain rule is exercised):
//   1. the first derivative (as built, and simplified) equals the closed form d/dt f(a(t)),
//      written out here from the calculus, to 1e-12 relative;
//   2. it equals a central difference of the value, (f(t+h) - f(t-h)) / 2h (Richardson
//      extrapolated), to 1e-8 relative;
//   3. the second derivative (diff twice, simplified) equals a central difference of the
//      closed-form first derivative to 1e-7 relative;
// and the two-variable power a(t)^b(u) gives the closed-form mixed derivative.
// Prints "ok" or "FAIL <case> <check>" lines and exits non-zero on failure.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "morefit/graph.hh"

using namespace morefit;

namespace {

typedef double K;
typedef GraphPtr<K, double> G;

G T() { return Variable<K, double>("t"); }
G U() { return Variable<K, double>("u"); }
G C(double v) { return Constant<K, double>(v); }

int fails = 0;
void check(bool ok, const std::string& name, const char* what, double got, double want) {
  if (!ok) {
    std::printf("FAIL %s %s got %.17g want %.17g\n", name.c_str(), what, got, want);
    fails++;
  }
}
double rel(double a, double b) { return std::fabs(a - b) / std::max(1.0, std::fabs(b)); }

double at(const G& g, double t, double u = 0.0) { return g->eval({"t", "u"}, {t, u}); }

struct Case {
  std::string name;
  std::function<G()> build;                   // f(a(t)) as a graph in t
  std::function<double(double)> f, df;        // closed forms of f and df/dt
  std::vector<double> points;
};

// central difference with one Richardson step: error O(h^4)
double central(const std::function<double(double)>& f, double t, double h) {
  const double d1 = (f(t + h) - f(t - h)) / (2 * h), d2 = (f(t + 2 * h) - f(t - 2 * h)) / (4 * h);
  return (4 * d1 - d2) / 3;
}

void run(const Case& c) {
  G f = c.build();
  G d = f->diff("t");
  G ds = d->simplify();
  G dd = ds->diff("t")->simplify();
  for (double t : c.points) {
    check(rel(at(f, t), c.f(t)) < 1e-13, c.name, "value", at(f, t), c.f(t));
    const double want = c.df(t);
    check(rel(at(d, t), want) < 1e-12, c.name, "d/dt (as built)", at(d, t), want);
    check(rel(at(ds, t), want) < 1e-12, c.name, "d/dt (simplified)", at(ds, t), want);
    const double fd = central(c.f, t, 1e-3 * std::max(1.0, std::fabs(t)));
    check(rel(at(ds, t), fd) < 1e-8, c.name, "d/dt vs central difference", at(ds, t), fd);
    const double fdd = central(c.df, t, 1e-3 * std::max(1.0, std::fabs(t)));
    check(rel(at(dd, t), fdd) < 1e-7, c.name, "d2/dt2 vs central difference of d/dt", at(dd, t), fdd);
  }
}

}  // namespace

int main() {
  const std::vector<double> pts = {-0.9, -0.35, 0.0, 0.27, 0.8};
  const double ln2 = std::log(2.0), sqpi = std::sqrt(M_PI);
  std::vector<Case> cases;
  // a(t) = 0.3 t + 0.1 t^2 in (-1, 1) on the points: asin, acos
  auto a1 = [] { return C(0.3) * T() + C(0.1) * T() * T(); };
  auto a1v = [](double t) { return 0.3 * t + 0.1 * t * t; };
  auto a1d = [](double t) { return 0.3 + 0.2 * t; };
  cases.push_back({"asin", [&] { return Asin<K, double>(a1()); }, [&](double t) { return std::asin(a1v(t)); },
                   [&](double t) { return a1d(t) / std::sqrt(1 - a1v(t) * a1v(t)); }, pts});
  cases.push_back({"acos", [&] { return Acos<K, double>(a1()); }, [&](double t) { return std::acos(a1v(t)); },
                   [&](double t) { return -a1d(t) / std::sqrt(1 - a1v(t) * a1v(t)); }, pts});
  // b(t) = 0.5 t + 0.2: tan, atan, sin, cos, erf, exp
  auto a2 = [] { return C(0.5) * T() + C(0.2); };
  auto a2v = [](double t) { return 0.5 * t + 0.2; };
  cases.push_back({"tan", [&] { return Tan<K, double>(a2()); }, [&](double t) { return std::tan(a2v(t)); },
                   [&](double t) { return 0.5 / (std::cos(a2v(t)) * std::cos(a2v(t))); }, pts});
  cases.push_back({"atan", [&] { return Atan<K, double>(C(2.0) * T() - C(0.3) * T() * T()); },
                   [](double t) { return std::atan(2 * t - 0.3 * t * t); },
                   [](double t) { double v = 2 * t - 0.3 * t * t; return (2 - 0.6 * t) / (1 + v * v); }, pts});
  cases.push_back({"sin", [&] { return Sin<K, double>(a2()); }, [&](double t) { return std::sin(a2v(t)); },
                   [&](double t) { return 0.5 * std::cos(a2v(t)); }, pts});
  cases.push_back({"cos", [&] { return Cos<K, double>(a2()); }, [&](double t) { return std::cos(a2v(t)); },
                   [&](double t) { return -0.5 * std::sin(a2v(t)); }, pts});
  cases.push_back({"erf", [&] { return Erf<K, double>(a2()); }, [&](double t) { return std::erf(a2v(t)); },
                   [&](double t) { return 0.5 * 2 / sqpi * std::exp(-a2v(t) * a2v(t)); }, pts});
  cases.push_back({"exp", [&] { return Exp<K, double>(a2() * T()); }, [](double t) { return std::exp((0.5 * t + 0.2) * t); },
                   [](double t) { return (t + 0.2) * std::exp((0.5 * t + 0.2) * t); }, pts});
  // c(t) = 1 + t^2 > 0: log, sqrt, inverse, powers with a constant exponent
  auto a3 = [] { return C(1.0) + T() * T(); };
  auto a3v = [](double t) { return 1 + t * t; };
  cases.push_back({"log", [&] { return Log<K, double>(a3()); }, [&](double t) { return std::log(a3v(t)); },
                   [&](double t) { return 2 * t / a3v(t); }, pts});
  cases.push_back({"sqrt", [&] { return Sqrt<K, double>(a3()); }, [&](double t) { return std::sqrt(a3v(t)); },
                   [&](double t) { return t / std::sqrt(a3v(t)); }, pts});
  cases.push_back({"inv", [&] { return Inv<K, double>(a3()); }, [&](double t) { return 1 / a3v(t); },
                   [&](double t) { return -2 * t / (a3v(t) * a3v(t)); }, pts});
  cases.push_back({"pow base", [&] { return Pow<K, double>(a3(), C(2.5)); }, [&](double t) { return std::pow(a3v(t), 2.5); },
                   [&](double t) { return 2.5 * std::pow(a3v(t), 1.5) * 2 * t; }, pts});
  cases.push_back({"pow negative exponent", [&] { return Pow<K, double>(a3(), C(-1.5)); },
                   [&](double t) { return std::pow(a3v(t), -1.5); }, [&](double t) { return -1.5 * std::pow(a3v(t), -2.5) * 2 * t; },
                   pts});
  // exponent depending on t, base constant: 2^(0.5 t + 0.2)
  cases.push_back({"pow exponent", [&] { return Pow<K, double>(C(2.0), a2()); }, [&](double t) { return std::pow(2.0, a2v(t)); },
                   [&](double t) { return std::pow(2.0, a2v(t)) * ln2 * 0.5; }, pts});
  // base and exponent depending on t: (1 + t^2)^(0.5 t + 0.2)
  cases.push_back({"pow both", [&] { return Pow<K, double>(a3(), a2()); }, [&](double t) { return std::pow(a3v(t), a2v(t)); },
                   [&](double t) { return std::pow(a3v(t), a2v(t)) * (0.5 * std::log(a3v(t)) + a2v(t) * 2 * t / a3v(t)); }, pts});
  // the exponential normalisation's conditional away from its branch point
  cases.push_back({"conditional", [&] { return ConditionalEqual<K, double>(T(), C(0.6), (Exp<K, double>(T() * C(5.6)) - Exp<K, double>(T() * C(5.0))) / T()); },
                   [](double t) { return (std::exp(5.6 * t) - std::exp(5.0 * t)) / t; },
                   [](double t) { return (5.6 * std::exp(5.6 * t) - 5.0 * std::exp(5.0 * t)) / t - (std::exp(5.6 * t) - std::exp(5.0 * t)) / (t * t); },
                   {-0.9, -0.35, 0.27, 0.8}});
  for (auto& c : cases) run(c);

  // the conditional at its branch point: the rule differentiates the branch taken (a constant
  // here), the reference's behaviour at lambda = 0 exactly
  {
    G f = ConditionalEqual<K, double>(T(), C(0.6), (Exp<K, double>(T() * C(5.6)) - Exp<K, double>(T() * C(5.0))) / T());
    G d = f->diff("t")->simplify();
    check(at(f, 0.0) == 0.6, "conditional", "value at the branch point", at(f, 0.0), 0.6);
    check(at(d, 0.0) == 0.0, "conditional", "d/dt at the branch point", at(d, 0.0), 0.0);
  }
  // mixed second derivative of a(t)^b(u): d2/dt du [(1+t^2)^(0.5u+0.2)]
  {
    G f = Pow<K, double>(C(1.0) + T() * T(), C(0.5) * U() + C(0.2));
    G dtu = f->diff("t")->simplify()->diff("u")->simplify();
    G dut = f->diff("u")->simplify()->diff("t")->simplify();
    for (double t : pts)
      for (double u : {-0.7, 0.3, 1.1}) {
        const double A = 1 + t * t, e = 0.5 * u + 0.2;
        // d/dt = e A^(e-1) 2t ;  d/du of that = 0.5 A^(e-1) 2t + e A^(e-1) ln A 0.5 2t
        const double want = std::pow(A, e - 1) * 2 * t * 0.5 * (1 + e * std::log(A));
        check(rel(at(dtu, t, u), want) < 1e-12, "pow mixed", "d2/dt du", at(dtu, t, u), want);
        check(rel(at(dut, t, u), want) < 1e-12, "pow mixed", "d2/du dt", at(dut, t, u), want);
      }
  }
  if (fails == 0) std::printf("ok\n");
  return fails == 0 ? 0 : 1;
}

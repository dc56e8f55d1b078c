// refgraph.cc -- golden-vector driver around the REFERENCE's own graph.hh.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from the
// reference headers where they lie (nothing is copied).  pdf.hh cannot be compiled here
// (it includes Eigen, which is absent), so this driver assembles the densities with the
// reference's graph builders exactly as pdf.hh:147-154, 431-440, 498-554, 207-260 (Crystal
// Ball), 725-755 (polynomial) and physicspdfs.hh:39-77 assemble them, then uses the reference's own
// substitute/simplify/diff/eval (graph.hh) the way fitter.hh:336, 405-407, 434-440 and
// 1071-1072 use them.  Output: JSON golden vectors on stdout.
#include <cassert>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>
#include "graph.hh"

using namespace morefit;
typedef double K;
typedef double E;
typedef std::unique_ptr<ComputeGraphNode<K, E>> Ptr;

static Ptr V(const std::string& n) { return Variable<K, E>(n); }
static Ptr C(double c) { return Constant<K, E>(c); }

// GaussianPDF prob/norm (pdf.hh:147-154) on [a,b]
static Ptr gauss_prob_normalised(double a, double b) {
  Ptr prob = 1.0 / sqrt(2.0 * M_PI) / V("sigma") *
             Exp<K, E>(-((V("m") - V("mu")) * (V("m") - V("mu")) / (2.0 * V("sigma") * V("sigma"))));
  Ptr norm = C(0.5) * (Erf<K, E>((C(b) - V("mu")) / (Sqrt<K, E>(2.0) * V("sigma"))) -
                       Erf<K, E>((C(a) - V("mu")) / (Sqrt<K, E>(2.0) * V("sigma"))));
  return std::make_unique<ProdNode<K, E>>(std::move(prob), std::make_unique<InvNode<K, E>>(std::move(norm)));
}
// ExponentialPDF prob/norm (pdf.hh:431-440)
static Ptr expo_prob_normalised(double a, double b) {
  Ptr prob = Exp<K, E>(V("lambda") * V("m"));
  Ptr norm = ConditionalEqual(V("lambda"), C(b - a), (Exp<K, E>(V("lambda") * C(b)) - Exp<K, E>(V("lambda") * C(a))) / V("lambda"));
  return std::make_unique<ProdNode<K, E>>(std::move(prob), std::make_unique<InvNode<K, E>>(std::move(norm)));
}
// SumPDF with one fraction (pdf.hh:526-554), then PDF::logprob_normalised (pdf.hh:103-106)
static Ptr mass_logprob(double a, double b) {
  std::vector<Ptr> args;
  args.emplace_back(std::make_unique<ProdNode<K, E>>(V("f"), gauss_prob_normalised(a, b)));
  std::vector<Ptr> frac;
  frac.emplace_back(C(1.0));
  frac.emplace_back(std::make_unique<NegNode<K, E>>(V("f")));
  args.emplace_back(std::make_unique<ProdNode<K, E>>(std::make_unique<SumNode<K, E>>(std::move(frac)), expo_prob_normalised(a, b)));
  Ptr prob = std::make_unique<SumNode<K, E>>(std::move(args));
  Ptr logprob = std::make_unique<LogNode<K, E>>(std::move(prob));
  Ptr lognorm = std::make_unique<LogNode<K, E>>(C(1.0));
  return std::make_unique<SumNode<K, E>>(std::move(logprob), std::make_unique<NegNode<K, E>>(std::move(lognorm)));
}
// KstarmumuAngularPDF::prob (physicspdfs.hh:39-77), norm = 1
static Ptr kstar_logprob() {
  Ptr c = C(9.0 / 32.0 / M_PI);
  Ptr costhetal = V("ctl"), costhetak = V("ctk");
  Ptr costhetal2 = V("ctl") * V("ctl");
  Ptr costhetak2 = V("ctk") * V("ctk");
  Ptr cos2thetal = 2.0 * costhetal2->copy() - 1.0;
  Ptr sinthetal2 = 1.0 - costhetal2->copy();
  Ptr sinthetak2 = 1.0 - costhetak2->copy();
  Ptr sinthetal = Sqrt<K, E>(sinthetal2->copy());
  Ptr sinthetak = Sqrt<K, E>(sinthetak2->copy());
  Ptr sin2thetal = 2.0 * sinthetal->copy() * V("ctl");
  Ptr sin2thetak = 2.0 * sinthetak->copy() * V("ctk");
  Ptr cosphi = Cos<K, E>(V("phi"));
  Ptr cos2phi = Cos<K, E>(2.0 * V("phi"));
  Ptr sinphi = Sin<K, E>(V("phi"));
  Ptr sin2phi = Sin<K, E>(2.0 * V("phi"));
  Ptr prob = c->copy() * sinthetak2->copy() * 3.0 / 4.0 * (1.0 - V("FL")) +
             c->copy() * costhetak2->copy() * V("FL") +
             c->copy() * sinthetak2->copy() * cos2thetal->copy() * 1.0 / 4.0 * (1.0 - V("FL")) +
             c->copy() * costhetak2->copy() * cos2thetal->copy() * (-V("FL")) +
             c->copy() * sinthetak2->copy() * sinthetal2->copy() * cos2phi->copy() * V("S3") +
             c->copy() * sin2thetak->copy() * sin2thetal->copy() * cosphi->copy() * V("S4") +
             c->copy() * sin2thetak->copy() * sinthetal->copy() * cosphi->copy() * V("S5") +
             c->copy() * sinthetak2->copy() * costhetal->copy() * 4.0 / 3.0 * V("AFB") +
             c->copy() * sin2thetak->copy() * sinthetal->copy() * sinphi->copy() * V("S7") +
             c->copy() * sin2thetak->copy() * sin2thetal->copy() * sinphi->copy() * V("S8") +
             c->copy() * sinthetak2->copy() * sinthetal2->copy() * sin2phi->copy() * V("S9");
  Ptr logprob = std::make_unique<LogNode<K, E>>(std::move(prob));
  Ptr lognorm = std::make_unique<LogNode<K, E>>(C(1.0));
  return std::make_unique<SumNode<K, E>>(std::move(logprob), std::make_unique<NegNode<K, E>>(std::move(lognorm)));
}

// CrystalballPDF prob/norm (pdf.hh:207-260) on [a,b], then PDF::logprob_normalised (pdf.hh:103-106)
static Ptr cb_logprob(double a, double b) {
  Ptr prob = ConditionalLarger(V("m") - V("mu") + V("alpha") * V("sigma"),
                               Exp<K, E>(-0.5 * (V("m") - V("mu")) * (V("m") - V("mu")) / (V("sigma") * V("sigma"))),
                               Pow<K, E>(V("n") / V("alpha"), V("n")) * Exp<K, E>(-0.5 * V("alpha") * V("alpha")) /
                                   Pow<K, E>(V("n") / V("alpha") - V("alpha") - (V("m") - V("mu")) / V("sigma"), V("n")));
  Ptr A = Pow<K, E>(V("n") / V("alpha"), V("n")) * Exp<K, E>(-0.5 * V("alpha") * V("alpha"));
  Ptr B = V("n") / V("alpha") - V("alpha");
  Ptr Cc = V("sigma") * B->copy() + V("mu");
  Ptr mid = V("mu") - V("alpha") * V("sigma");
  Ptr gaussint = Sqrt<K, E>(M_PI / 2.0) * V("sigma") *
                 (Erf<K, E>((V("mu") - mid->copy()) / (V("sigma") * sqrt(2.0))) - Erf<K, E>((V("mu") - C(b)) / (V("sigma") * sqrt(2.0))));
  Ptr powint = ConditionalUnequal(V("n") - 1.0,
                                  A->copy() * Pow<K, E>(V("sigma"), V("n")) *
                                      (Pow<K, E>(Cc->copy() - mid->copy(), 1.0 - V("n")) - Pow<K, E>(Cc->copy() - C(a), 1.0 - V("n"))) /
                                      (V("n") - 1.0),
                                  A->copy() * V("sigma") * (Log<K, E>(Cc->copy() - C(a)) - Log<K, E>(Cc->copy() - mid->copy())));
  Ptr norm = powint->copy() + gaussint->copy();
  Ptr logprob = std::make_unique<LogNode<K, E>>(std::move(prob));
  Ptr lognorm = std::make_unique<LogNode<K, E>>(std::move(norm));
  return std::make_unique<SumNode<K, E>>(std::move(logprob), std::make_unique<NegNode<K, E>>(std::move(lognorm)));
}
// PolynomialPDF prob/norm (pdf.hh:725-755) with coefficients c1..c3 on [a,b] (Horner forms)
static Ptr poly_logprob(double a, double b) {
  const char* c[3] = {"c1", "c2", "c3"};
  const int k = 3;
  Ptr prob = V(c[k - 1]);
  for (int i = 1; i < k; i++) prob = std::move(prob) * V("x") + V(c[k - i - 1]);
  prob = std::move(prob) * V("x") + 1.0;
  auto anti = [&](double x) {  // x (1 + c1 x / 2 + c2 x^2 / 3 + c3 x^3 / 4)
    Ptr q = C(1.0 / (k + 1)) * V(c[k - 1]);
    for (int i = 1; i < k; i++) q = std::move(q) * C(x) + C(1.0 / (k + 1 - i)) * V(c[k - i - 1]);
    q = std::move(q) * C(x) + 1.0;
    return std::move(q) * C(x);
  };
  Ptr norm = anti(b) - anti(a);
  Ptr logprob = std::make_unique<LogNode<K, E>>(std::move(prob));
  Ptr lognorm = std::make_unique<LogNode<K, E>>(std::move(norm));
  return std::make_unique<SumNode<K, E>>(std::move(logprob), std::make_unique<NegNode<K, E>>(std::move(lognorm)));
}

static void emit(const char* model, Ptr lp, const std::vector<std::string>& obs, const std::vector<std::string>& pars,
                 const std::vector<std::vector<double>>& points, bool last) {
  // fitter.hh:336 (simplify), 405-407 (first derivatives), 434-440 (second derivatives, i>=j)
  Ptr f = lp->simplify();
  std::vector<Ptr> d1, d2;
  for (size_t i = 0; i < pars.size(); i++) d1.emplace_back(f->diff(pars[i])->simplify());
  for (size_t i = 0; i < pars.size(); i++)
    for (size_t j = 0; j <= i; j++) d2.emplace_back(d1[i]->diff(pars[j])->simplify());
  std::vector<std::string> names(obs);
  names.insert(names.end(), pars.begin(), pars.end());
  printf("  \"%s\": {\"observables\": [", model);
  for (size_t i = 0; i < obs.size(); i++) printf("%s\"%s\"", i ? ", " : "", obs[i].c_str());
  printf("], \"parameters\": [");
  for (size_t i = 0; i < pars.size(); i++) printf("%s\"%s\"", i ? ", " : "", pars[i].c_str());
  printf("], \"points\": [\n");
  for (size_t p = 0; p < points.size(); p++) {
    const auto& v = points[p];
    printf("    {\"x\": [");
    for (size_t i = 0; i < obs.size(); i++) printf("%s%.17g", i ? ", " : "", v[i]);
    printf("], \"theta\": [");
    for (size_t i = 0; i < pars.size(); i++) printf("%s%.17g", i ? ", " : "", v[obs.size() + i]);
    printf("], \"logp\": %.17g, \"grad\": [", f->eval(names, v));
    for (size_t i = 0; i < d1.size(); i++) printf("%s%.17g", i ? ", " : "", d1[i]->eval(names, v));
    printf("], \"hess\": [");
    for (size_t i = 0; i < d2.size(); i++) printf("%s%.17g", i ? ", " : "", d2[i]->eval(names, v));
    printf("]}%s\n", p + 1 < points.size() ? "," : "");
  }
  printf("  ]}%s\n", last ? "" : ",");
}

int main() {
  // deterministic point sets (a fixed LCG; no dependence on any library RNG)
  uint64_t s = 0x9e3779b97f4a7c15ULL;
  auto u01 = [&s]() { s = s * 6364136223846793005ULL + 1442695040888963407ULL; return (double)(s >> 11) * 0x1.0p-53; };
  const double a = 5.0, b = 5.6;
  std::vector<std::vector<double>> mass_pts;
  const double thetas[4][4] = {{5.28, 0.02, 0.3, -1.5}, {5.34, 0.022, 0.33, -1.65}, {5.25, 0.05, 0.7, 0.8}, {5.02, 0.1, 0.1, -3.0}};
  for (int t = 0; t < 4; t++)
    for (int k = 0; k < 8; k++) {
      double m = a + (b - a) * u01();
      mass_pts.push_back({m, thetas[t][0], thetas[t][1], thetas[t][2], thetas[t][3]});
    }
  std::vector<std::vector<double>> kstar_pts;
  const double kth[3][8] = {{0.7, 0.0, -0.1, -0.2, 0.1, 0.0, 0.0, 0.0}, {0.6, 0.05, -0.05, -0.3, 0.2, 0.02, -0.03, 0.01}, {0.3, -0.1, 0.1, 0.1, -0.15, 0.1, 0.05, -0.05}};
  for (int t = 0; t < 3; t++)
    for (int k = 0; k < 8; k++) {
      std::vector<double> v = {-1.0 + 2.0 * u01(), -1.0 + 2.0 * u01(), -M_PI + 2.0 * M_PI * u01()};
      for (int i = 0; i < 8; i++) v.push_back(kth[t][i]);
      kstar_pts.push_back(v);
    }
  // lambda = 0: the ConditionalEqual branch of the exponential's normalisation (pdf.hh:437-439);
  // drawn last so that the points above keep their values
  for (int k = 0; k < 8; k++) {
    double m = a + (b - a) * u01();
    mass_pts.push_back({m, 5.3, 0.03, 0.5, 0.0});
  }
  // Crystal Ball on the mass window (both sides of the join; n = 1 takes the tail integral's
  // logarithmic branch) and the third-order polynomial on [-1, 1]; drawn after the points above
  std::vector<std::vector<double>> cb_pts, poly_pts;
  const double cbth[4][4] = {{5.28, 0.02, 1.5, 3.0}, {5.30, 0.05, 0.5, 1.2}, {5.25, 0.015, 3.0, 10.0}, {5.40, 0.08, 1.0, 1.0}};
  for (int t = 0; t < 4; t++)
    for (int k = 0; k < 8; k++) {
      double m = a + (b - a) * u01();
      cb_pts.push_back({m, cbth[t][0], cbth[t][1], cbth[t][2], cbth[t][3]});
    }
  const double pth[4][3] = {{0.3, -0.2, 0.1}, {0.0, 0.0, 0.0}, {-0.9, 0.5, 0.2}, {0.2, 0.9, -0.3}};
  for (int t = 0; t < 4; t++)
    for (int k = 0; k < 8; k++) {
      double x = -1.0 + 2.0 * u01();
      poly_pts.push_back({x, pth[t][0], pth[t][1], pth[t][2]});
    }
  printf("{\n  \"generator\": \"oracle/ref/refgraph.cc over the reference graph.hh\",\n  \"mass_range\": [%.17g, %.17g],\n", a, b);
  emit("mass", mass_logprob(a, b), {"m"}, {"mu", "sigma", "f", "lambda"}, mass_pts, false);
  emit("kstar", kstar_logprob(), {"ctl", "ctk", "phi"}, {"FL", "S3", "S4", "S5", "AFB", "S7", "S8", "S9"}, kstar_pts, false);
  emit("cb", cb_logprob(a, b), {"m"}, {"mu", "sigma", "alpha", "n"}, cb_pts, false);
  emit("poly", poly_logprob(-1.0, 1.0), {"x"}, {"c1", "c2", "c3"}, poly_pts, true);
  printf("}\n");
  return 0;
}

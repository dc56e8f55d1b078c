/**
 * @file minimizer.hh
 * Variable-metric minimiser with MIGRAD's control flow and HESSE-style errors, written from
 * the published MINUIT algorithm description (F. James, "MINUIT Function Minimization and
 * Error Analysis", CERN D506; F. James & M. Winkler, Minuit2 user's guide):
 *
 *  - parameters with two-sided limits live in an internal space through
 *    ext = lo + (hi - lo)/2 (sin(int) + 1); one-sided limits use the square-root maps;
 *  - seed: analytic Hessian (strategy 2 with an analytic Hessian, the setting the reference
 *    fitter uses, fitter.hh:762-767, 806-813) transformed to the internal space and
 *    inverted, forced positive-definite by a diagonal shift when needed;
 *  - iteration: Newton step -V g taken whole on sufficient decrease (one value+gradient
 *    call), else a parabolic line search on the value; where the value's rounding noise
 *    (FP32 sums) exceeds the predicted decrease, the directional derivative decides instead
 *    (secant search for its zero); BFGS update of V;
 *  - convergence: EDM = g^T V g / 2 < 0.002 * tolerance * up (tolerance 0.01, up 1 as in the
 *    reference's Minuit2 defaults), confirmed with a fresh analytic Hessian (strategy 2);
 *  - polish (on by default, not in MINUIT): from the converged point, Newton steps with that
 *    exact Hessian down to EDM 1e-10, so the result is the minimum itself and not any point
 *    inside MIGRAD's EDM tolerance (which allows ~6e-3 sigma);
 *  - HESSE: covariance = 2 up H^-1 in the internal space, mapped back through the Jacobian;
 *  - objectives without a gradient or a Hessian (the reference fitter's analytic_gradient /
 *    analytic_hessian = false, fitter.hh:796-814: Minuit2 then differentiates numerically):
 *    central differences in the internal space, steps a fraction of each parameter's error
 *    estimate (refreshed from every second difference), the Hessian from differences of the
 *    analytic gradient where there is one, else of the value.
 *
 * The function interface is synchronous (value, value+gradient, Hessian).  The batched toy
 * engine runs one instance per toy in a fibre and answers the calls from shared launches.
 */
#ifndef MOREFIT_MINIMIZER_HH
#define MOREFIT_MINIMIZER_HH

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

namespace morefit {

// objective: all calls take the full external parameter vector (fixed ones included)
struct MinimizerFcn {
  virtual ~MinimizerFcn() = default;
  virtual double value(const double* x) = 0;
  virtual double value_grad(const double* x, double* grad) = 0;  // returns the value as well
  virtual bool hessian(const double* x, double* h) = 0;             // n x n, row-major, all params
  // rounding noise of value() (e.g. FP32 per-event sums); value differences below it carry
  // no information and the search then relies on the gradient
  virtual double value_noise() const { return 0.0; }
  // false: value_grad() is not available and the minimiser differentiates value() numerically
  // (the reference's analytic_gradient = false, a Minuit2 FCN without a gradient)
  virtual bool has_gradient() const { return true; }
  // false: hessian() is not available; the metric seed, the strategy-2 confirmation and HESSE
  // use second differences (of the analytic gradient where there is one, else of value())
  virtual bool has_hessian() const { return true; }
};

// Rounding floor, in -2 lnL units, of a likelihood summed in FP32 over n events.  The kernel
// parameters (floating values and hoisted terms) are rounded to FP32 once per call, so all
// events see the same relative error of up to 2^-24 in each: as theta moves, the sum jumps
// coherently by up to ~n 2^-24 per parameter-dependent term, which dominates the
// random-walk rounding of the per-event terms.  Four such terms, times 2 for -2 lnL.
inline double fp32_sum_noise(double nevents) { return 8.0 * std::ldexp(1.0, -24) * std::max(1.0, nevents); }

struct MinimizerParam {
  std::string name;
  double value = 0.0, step = 0.1, lo = 0.0, hi = 0.0;
  bool has_lo = false, has_hi = false, fixed = false;
};

// small dense linear algebra (n <= ~50)
namespace la {
// in-place inverse of a symmetric positive-definite matrix by Cholesky; false if not PD
inline bool spd_inverse(std::vector<double>& a, int n) {
  std::vector<double> L(n * n, 0.0);
  for (int j = 0; j < n; j++) {
    double d = a[j * n + j];
    for (int k = 0; k < j; k++) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 0.0) || !std::isfinite(d)) return false;
    L[j * n + j] = std::sqrt(d);
    for (int i = j + 1; i < n; i++) {
      double s = a[i * n + j];
      for (int k = 0; k < j; k++) s -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = s / L[j * n + j];
    }
  }
  // inverse of L
  std::vector<double> Li(n * n, 0.0);
  for (int i = 0; i < n; i++) {
    Li[i * n + i] = 1.0 / L[i * n + i];
    for (int j = 0; j < i; j++) {
      double s = 0.0;
      for (int k = j; k < i; k++) s -= L[i * n + k] * Li[k * n + j];
      Li[i * n + j] = s / L[i * n + i];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= i; j++) {
      double s = 0.0;
      for (int k = i; k < n; k++) s += Li[k * n + i] * Li[k * n + j];
      a[i * n + j] = a[j * n + i] = s;
    }
  return true;
}
// symmetric Jacobi eigenvalues (for the positive-definite check/shift)
inline double min_eigenvalue(std::vector<double> a, int n) {
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) off += a[i * n + j] * a[i * n + j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = a[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
      }
  }
  double m = a[0];
  for (int i = 1; i < n; i++) m = std::min(m, a[i * n + i]);
  return m;
}
}  // namespace la

class Migrad {
 public:
  // settings (reference fitter_options / Minuit2 defaults)
  int strategy = 2;
  int max_calls = 4000;
  int max_iterations = 4000;
  double tolerance = 0.01;
  double up = 1.0;
  int print_level = 0;         // >= 3: one line per iteration on stderr
  int trace_id = -1;           // printed with the trace lines (fit id in a batch)
  // after convergence: Newton steps with a fresh exact Hessian down to this EDM
  bool polish = true;
  double polish_edm = 1e-10;
  int polish_steps = 4;
  // numerical derivatives, for objectives without a gradient or a Hessian: central differences
  // in the internal space with steps of these fractions of each parameter's current error
  // estimate (sqrt(2 up / d2f/du2), first from the parameter's step size), raised where the
  // value's rounding noise would otherwise dominate the difference
  double grad_step_frac = 0.01;
  double hess_step_frac = 0.05;

  // results
  std::vector<double> x;      // external values (all parameters)
  std::vector<double> errors; // external errors (0 for fixed)
  std::vector<double> cov;    // covariance of the floating parameters (nf x nf, row-major)
  double fval = NAN, edm = NAN;
  int ncalls = 0, ngrad = 0, nhess = 0, iterations = 0;
  int status = -1;            // 0 ok, 1 call limit, 2 line search failure, 3 not positive-definite, 4 non-finite
  bool valid = false, cov_accurate = false;

  explicit Migrad(std::vector<MinimizerParam> params) : p_(std::move(params)) {
    for (size_t i = 0; i < p_.size(); i++)
      if (!p_[i].fixed) free_.push_back((int)i);
    x.resize(p_.size());
    for (size_t i = 0; i < p_.size(); i++) x[i] = p_[i].value;
    errors.assign(p_.size(), 0.0);
  }
  int nfree() const { return (int)free_.size(); }

  // ---------------------------------------------------------------- transforms
  double int2ext(int k, double u) const {
    const auto& q = p_[free_[k]];
    if (q.has_lo && q.has_hi) return q.lo + 0.5 * (q.hi - q.lo) * (std::sin(u) + 1.0);
    if (q.has_lo) return q.lo - 1.0 + std::sqrt(u * u + 1.0);
    if (q.has_hi) return q.hi + 1.0 - std::sqrt(u * u + 1.0);
    return u;
  }
  double ext2int(int k, double v) const {
    const auto& q = p_[free_[k]];
    if (q.has_lo && q.has_hi) {
      double y = 2.0 * (v - q.lo) / (q.hi - q.lo) - 1.0;
      y = std::max(-1.0, std::min(1.0, y));
      return std::asin(y);
    }
    if (q.has_lo) { double y = v - q.lo + 1.0; return std::sqrt(std::max(0.0, y * y - 1.0)); }
    if (q.has_hi) { double y = q.hi - v + 1.0; return std::sqrt(std::max(0.0, y * y - 1.0)); }
    return v;
  }
  double dext(int k, double u) const {  // d ext / d int
    const auto& q = p_[free_[k]];
    if (q.has_lo && q.has_hi) return 0.5 * (q.hi - q.lo) * std::cos(u);
    if (q.has_lo) return u / std::sqrt(u * u + 1.0);
    if (q.has_hi) return -u / std::sqrt(u * u + 1.0);
    return 1.0;
  }
  double d2ext(int k, double u) const {  // d2 ext / d int2
    const auto& q = p_[free_[k]];
    if (q.has_lo && q.has_hi) return -0.5 * (q.hi - q.lo) * std::sin(u);
    double r = std::pow(u * u + 1.0, -1.5);
    if (q.has_lo) return r;
    if (q.has_hi) return -r;
    return 0.0;
  }

  // ---------------------------------------------------------------- minimisation
  int minimize(MinimizerFcn& fcn) {
    const int n = nfree();
    status = -1;
    valid = false;
    std::vector<double> u(n), g(n), V(n * n);
    std::vector<double> d(n), u1(n), g1(n), s(n), y(n), Vy(n);
    for (int k = 0; k < n; k++) u[k] = ext2int(k, p_[free_[k]].value);
    init_sigma(u);
    // value and internal gradient (analytic, or central differences of the value)
    auto eval_fg = [&](const std::vector<double>& uu, std::vector<double>& gg) {
      if (fcn.has_gradient()) return grad_at(fcn, uu, gg);
      double f = value_at(fcn, uu);
      if (std::isfinite(f)) numeric_gradient(fcn, uu, f, gg);
      return f;
    };
    auto eval_f = [&](const std::vector<double>& uu) { return value_at(fcn, uu); };
    // internal Hessian at uu (value ff, internal gradient gg there) -> V = H^-1 (shifted to
    // positive-definite); false on failure
    auto seed_metric = [&](const std::vector<double>& uu, const std::vector<double>& gg, double ff) {
      std::vector<double> H;
      if (!internal_hessian(fcn, uu, gg, ff, H)) return false;
      last_hint_ = H;
      std::vector<double> Hi = H;
      if (!la::spd_inverse(Hi, n)) {
        double lmin = la::min_eigenvalue(H, n);
        double dmax = 0.0;
        for (int i = 0; i < n; i++) dmax = std::max(dmax, std::fabs(H[i * n + i]));
        double shift = -lmin + 1e-3 * std::max(dmax, 1e-10);
        Hi = H;
        for (int i = 0; i < n; i++) Hi[i * n + i] += shift;
        if (!la::spd_inverse(Hi, n)) return false;
        pd_forced_ = true;
      }
      V = Hi;
      return true;
    };
    auto compute_edm = [&](const std::vector<double>& gg) {
      double e = 0.0;
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) e += gg[i] * V[i * n + j] * gg[j];
      return 0.5 * e;
    };
    const double edm_goal = 0.002 * tolerance * up;

    double f = eval_fg(u, g);
    if (!std::isfinite(f)) return finish(4, f, NAN, u, n);
    if (!seed_metric(u, g, f)) {  // diagonal fallback from the step sizes
      std::fill(V.begin(), V.end(), 0.0);
      for (int k = 0; k < n; k++) {
        double s = p_[free_[k]].step / std::max(1e-300, std::fabs(dext(k, u[k])));
        V[k * n + k] = s * s;
      }
    }
    double e = compute_edm(g);
    bool confirmed = false;
    // Newton steps with the exact metric until EDM < polish_edm (at most polish_steps); a
    // step is kept only when it lowers EDM.  Takes a converged point to the minimum itself,
    // so fits in different precisions land within rounding of each other.
    auto polish_at_minimum = [&]() {
      for (int k = 0; k < polish_steps && e > polish_edm; k++) {
        for (int i = 0; i < n; i++) {
          double t = 0.0;
          for (int j = 0; j < n; j++) t -= V[i * n + j] * g[j];
          u1[i] = u[i] + t;
        }
        double f1 = eval_fg(u1, g1);
        if (!std::isfinite(f1)) break;
        double e1 = compute_edm(g1);
        if (print_level >= 3) std::fprintf(stderr, "[migrad %d]   polish edm %.3g -> %.3g f1-f %.3g\n", trace_id, e, e1, f1 - f);
        if (!(e1 < e)) break;
        u = u1;
        g = g1;
        f = f1;
        e = e1;
      }
    };
    for (iterations = 0; iterations < max_iterations; iterations++) {
      if (e < edm_goal) {
        if (strategy >= 2 && !confirmed) {  // confirm with a fresh Hessian
          confirmed = true;
          if (seed_metric(u, g, f)) {
            e = compute_edm(g);
            if (e < edm_goal) {
              if (polish) polish_at_minimum();
              break;
            }
          } else {
            break;
          }
        } else {
          if (polish && seed_metric(u, g, f)) {
            e = compute_edm(g);
            polish_at_minimum();
          }
          break;
        }
      }
      if (ncalls >= max_calls) return finish(1, f, e, u, n);
      // Newton direction
      double slope = 0.0;
      for (int i = 0; i < n; i++) {
        double t = 0.0;
        for (int j = 0; j < n; j++) t -= V[i * n + j] * g[j];
        d[i] = t;
        slope += g[i] * t;
      }
      if (!(slope < 0.0)) {  // not a descent direction: restart from the Hessian
        if (!seed_metric(u, g, f)) return finish(3, f, e, u, n);
        continue;
      }
      // Full Newton step first, with value and gradient in one call, accepted on sufficient
      // decrease (Armijo, widened by the value's rounding noise).  Where the predicted
      // decrease is itself inside the noise (FP32 sums near the minimum) the value says
      // nothing, and the step is judged by the directional derivative instead: accepted when
      // it moved toward the line minimum, else a secant search for the derivative's zero.
      for (int i = 0; i < n; i++) u1[i] = u[i] + d[i];
      double f1 = eval_fg(u1, g1);
      double slope1 = 0.0;
      for (int i = 0; i < n; i++) slope1 += g1[i] * d[i];
      const double noise = fcn.value_noise();
      const bool noisy = noise > 0.0 && -slope <= 16.0 * noise;
      bool accept = std::isfinite(f1) && f1 <= f + 1e-4 * slope + noise &&
                    (!noisy || std::fabs(slope1) < std::fabs(slope));
      double best_a = accept ? 1.0 : 0.0;
      if (print_level >= 3)
        std::fprintf(stderr, "[migrad %d] it %d f %.10g edm %.3g slope %.3g f1-f %.3g slope1 %.3g noise %.3g newton %s\n", trace_id,
                     iterations, f, e, slope, f1 - f, slope1, noise, accept ? "accepted" : "rejected");
      if (!accept && noisy) {
        best_a = derivative_search(d, slope, f, noise, 1.0, f1, slope1, u, u1, g1, eval_fg, f1);
        if (print_level >= 3) std::fprintf(stderr, "[migrad %d]   derivative search alpha %.3g\n", trace_id, best_a);
      } else if (!accept) {
        // parabolic search on the value along d
        double alpha = 1.0, best_f = f;
        double fa = f1;
        for (int ls = 0; ls < 12; ls++) {
          if (ls > 0) {
            for (int i = 0; i < n; i++) u1[i] = u[i] + alpha * d[i];
            fa = eval_f(u1);
          }
          if (std::isfinite(fa) && fa < best_f) { best_f = fa; best_a = alpha; }
          double c2 = std::isfinite(fa) ? (fa - f - slope * alpha) / (alpha * alpha) : INFINITY;
          double next;
          if (std::isfinite(c2) && c2 > 0.0) next = -slope / (2.0 * c2);
          else next = best_a > 0.0 ? 2.0 * alpha : 0.25 * alpha;
          next = std::max(0.1 * alpha, std::min(next, 4.0 * std::max(alpha, best_a)));
          if (best_a > 0.0 && std::fabs(next - best_a) < 0.15 * best_a) break;  // parabola agrees: done
          if (best_a > 0.0 && ls >= 2) break;
          if (std::fabs(next - alpha) < 1e-12 * alpha) break;
          alpha = next;
        }
        if (print_level >= 3) std::fprintf(stderr, "[migrad %d]   line search best alpha %.3g\n", trace_id, best_a);
        if (best_a > 0.0) {
          for (int i = 0; i < n; i++) u1[i] = u[i] + best_a * d[i];
          f1 = eval_fg(u1, g1);
          if (!std::isfinite(f1)) return finish(4, f, e, u, n);
        }
      }
      if (!accept) {
        if (best_a == 0.0) {
          // no acceptable point along the Newton direction.  Inside the value's noise floor
          // the gradient decides: Newton steps that lower EDM; otherwise refresh the metric once.
          if (noisy && seed_metric(u, g, f)) {
            e = compute_edm(g);
            polish_at_minimum();
            return finish(e < edm_goal ? 0 : 2, f, e, u, n);
          }
          if (!refreshed_) {
            refreshed_ = true;
            if (seed_metric(u, g, f)) { e = compute_edm(g); continue; }
          }
          return finish(e < edm_goal ? 0 : 2, f, e, u, n);
        }
      }
      refreshed_ = false;
      for (int i = 0; i < n; i++) s[i] = best_a * d[i];
      // BFGS update of the inverse metric
      double sy = 0.0;
      for (int i = 0; i < n; i++) { y[i] = g1[i] - g[i]; sy += s[i] * y[i]; }
      if (sy > 0.0) {
        double yVy = 0.0;
        for (int i = 0; i < n; i++) {
          double t = 0.0;
          for (int j = 0; j < n; j++) t += V[i * n + j] * y[j];
          Vy[i] = t;
          yVy += y[i] * t;
        }
        double rho = 1.0 / sy;
        for (int i = 0; i < n; i++)
          for (int j = 0; j < n; j++)
            V[i * n + j] += (1.0 + rho * yVy) * rho * s[i] * s[j] - rho * (Vy[i] * s[j] + s[i] * Vy[j]);
      }
      u = u1;
      g = g1;
      f = f1;
      e = compute_edm(g);
      confirmed = false;
    }
    if (iterations >= max_iterations) return finish(1, f, e, u, n);
    return finish(0, f, e, u, n);
  }

  // errors and covariance from the Hessian at the current point (analytic, or second
  // differences when the objective has none)
  bool hesse(MinimizerFcn& fcn) {
    const int n = nfree();
    std::vector<double> u(n), gi(n), H;
    for (int k = 0; k < n; k++) u[k] = ext2int(k, x[free_[k]]);
    if ((int)sig_.size() != n) init_sigma(u);
    double f;
    if (fcn.has_gradient()) {
      f = grad_at(fcn, u, gi);
    } else {
      f = value_at(fcn, u);
      if (std::isfinite(f)) numeric_gradient(fcn, u, f, gi);
    }
    if (!std::isfinite(f) || !internal_hessian(fcn, u, gi, f, H)) return false;
    std::vector<double> C = H;
    cov_accurate = la::spd_inverse(C, n);
    if (!cov_accurate) return false;
    cov.assign(n * n, 0.0);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) cov[i * n + j] = 2.0 * up * C[i * n + j] * dext(i, u[i]) * dext(j, u[j]);
    for (int k = 0; k < n; k++) errors[free_[k]] = std::sqrt(cov[k * n + k]);
    double e = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) e += gi[i] * C[i * n + j] * gi[j];
    edm = 0.5 * e;
    fval = f;
    return true;
  }

 private:
  std::vector<MinimizerParam> p_;
  std::vector<int> free_;
  std::vector<double> last_hint_;
  bool pd_forced_ = false;
  bool refreshed_ = false;
  std::vector<double> sig_;  // internal error estimate per floating parameter (numerical steps)
  std::vector<double> xw_;   // external parameter vector handed to the objective

  bool bounded(int k) const { return p_[free_[k]].has_lo || p_[free_[k]].has_hi; }
  // the transformed (sin / sqrt) internal coordinates have no meaningful scale above ~1
  void set_sigma(int k, double v) {
    if (!(std::isfinite(v) && v > 0.0)) return;
    sig_[k] = bounded(k) ? std::min(v, 1.0) : v;
  }
  void init_sigma(const std::vector<double>& uu) {
    const int n = nfree();
    sig_.assign(n, 1.0);
    for (int k = 0; k < n; k++) set_sigma(k, p_[free_[k]].step / std::max(1e-300, std::fabs(dext(k, uu[k]))));
  }
  const double* ext_of(const std::vector<double>& uu) {
    xw_ = x;
    for (int k = 0; k < nfree(); k++) xw_[free_[k]] = int2ext(k, uu[k]);
    return xw_.data();
  }
  double value_at(MinimizerFcn& fcn, const std::vector<double>& uu) {
    ncalls++;
    return fcn.value(ext_of(uu));
  }
  // value and analytic gradient, chained into the internal space
  double grad_at(MinimizerFcn& fcn, const std::vector<double>& uu, std::vector<double>& gg) {
    std::vector<double> ge(p_.size());
    const double f = fcn.value_grad(ext_of(uu), ge.data());
    ncalls++;
    ngrad++;
    for (int k = 0; k < nfree(); k++) gg[k] = ge[free_[k]] * dext(k, uu[k]);
    return f;
  }
  // Rounding floor of value differences: the objective's own noise, at least 8 ulp of |f|.
  // With f ~ f0 + up (du/sigma)^2 a central difference over h = a sigma has a gradient error of
  // noise / (2 a up) sigma and a curvature error of 2 noise / (a^2 up) relative: the step
  // fractions are raised until those stay below 1e-4 sigma and 1e-3 (at most a whole sigma).
  double noise_floor(const MinimizerFcn& fcn, double f) const {
    return std::max(fcn.value_noise(), 8.0 * std::numeric_limits<double>::epsilon() * (std::fabs(f) + up));
  }
  double grad_frac(double noise) const { return std::min(1.0, std::max(grad_step_frac, 5000.0 * noise / up)); }
  double hess_frac(double noise) const { return std::min(1.0, std::max(hess_step_frac, std::sqrt(2000.0 * noise / up))); }
  void note_curvature(int k, double g2) {
    if (std::isfinite(g2) && g2 > 0.0) set_sigma(k, std::sqrt(2.0 * up / g2));
  }
  // internal gradient of value() at uu (value f0 there) by central differences, 2 n calls;
  // the second differences refresh the step scales
  void numeric_gradient(MinimizerFcn& fcn, const std::vector<double>& uu, double f0, std::vector<double>& gg) {
    const int n = nfree();
    std::vector<double> w(uu);
    const double a = grad_frac(noise_floor(fcn, f0));
    for (int k = 0; k < n; k++) {
      const double h = a * sig_[k];
      w[k] = uu[k] + h;
      const double fp = value_at(fcn, w);
      w[k] = uu[k] - h;
      const double fm = value_at(fcn, w);
      w[k] = uu[k];
      gg[k] = (fp - fm) / (2.0 * h);
      note_curvature(k, (fp + fm - 2.0 * f0) / (h * h));
    }
    ngrad++;
  }
  // Internal Hessian at uu (value f0, internal gradient gg there).  Analytic: the external
  // Hessian through the Jacobian plus the transform's curvature times the gradient.  Without
  // one: central differences of the analytic gradient (2 n calls), or of the value (diagonal
  // from f(u +- h_k), off-diagonal from the four corners f(u +- h_j +- h_k)), symmetrised.
  bool internal_hessian(MinimizerFcn& fcn, const std::vector<double>& uu, const std::vector<double>& gg, double f0,
                        std::vector<double>& H) {
    const int n = nfree();
    const int N = (int)p_.size();
    H.assign((size_t)n * n, 0.0);
    if (fcn.has_hessian()) {
      std::vector<double> he((size_t)N * N);
      if (!fcn.hessian(ext_of(uu), he.data())) return false;
      nhess++;
      ncalls++;
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
          double v = he[free_[i] * N + free_[j]] * dext(i, uu[i]) * dext(j, uu[j]);
          if (i == j && dext(i, uu[i]) != 0.0) v += (gg[i] / dext(i, uu[i])) * d2ext(i, uu[i]);
          H[i * n + j] = v;
        }
    } else {
      const double a = hess_frac(noise_floor(fcn, f0));
      std::vector<double> h(n), w(uu);
      for (int k = 0; k < n; k++) h[k] = a * sig_[k];
      if (fcn.has_gradient()) {
        std::vector<double> gp(n), gm(n);
        for (int k = 0; k < n; k++) {
          w[k] = uu[k] + h[k];
          const double fp = grad_at(fcn, w, gp);
          w[k] = uu[k] - h[k];
          const double fm = grad_at(fcn, w, gm);
          w[k] = uu[k];
          if (!std::isfinite(fp) || !std::isfinite(fm)) return false;
          for (int j = 0; j < n; j++) H[j * n + k] = (gp[j] - gm[j]) / (2.0 * h[k]);
        }
      } else {
        std::vector<double> fp(n), fm(n);
        for (int k = 0; k < n; k++) {
          w[k] = uu[k] + h[k];
          fp[k] = value_at(fcn, w);
          w[k] = uu[k] - h[k];
          fm[k] = value_at(fcn, w);
          w[k] = uu[k];
          H[k * n + k] = (fp[k] + fm[k] - 2.0 * f0) / (h[k] * h[k]);
        }
        for (int k = 0; k < n; k++)
          for (int j = 0; j < k; j++) {
            double c[4];
            const int sk[4] = {1, 1, -1, -1}, sj[4] = {1, -1, 1, -1};
            for (int q = 0; q < 4; q++) {
              w[k] = uu[k] + sk[q] * h[k];
              w[j] = uu[j] + sj[q] * h[j];
              c[q] = value_at(fcn, w);
            }
            w[k] = uu[k];
            w[j] = uu[j];
            H[k * n + j] = H[j * n + k] = (c[0] - c[1] - c[2] + c[3]) / (4.0 * h[k] * h[j]);
          }
      }
      for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++) H[i * n + j] = H[j * n + i] = 0.5 * (H[i * n + j] + H[j * n + i]);
      nhess++;
    }
    for (int i = 0; i < n * n; i++)
      if (!std::isfinite(H[i])) return false;
    for (int k = 0; k < n; k++) note_curvature(k, H[k * n + k]);
    return true;
  }

  // Zero of the directional derivative phi'(a) = g(u + a d) . d by safeguarded secant steps,
  // starting from phi'(0) = slope0 < 0 and the probe (a1, f_a1, s_a1).  Accepts a point whose
  // |phi'| is at most half of |slope0| and whose value is within the noise of f0; on success
  // u1/g1/f_out hold that point and the step length is returned, else 0.
  template <class EvalFG>
  double derivative_search(const std::vector<double>& d, double slope0, double f0, double noise, double a1, double f_a1,
                           double s_a1, const std::vector<double>& u, std::vector<double>& u1, std::vector<double>& g1,
                           EvalFG& eval_fg, double& f_out) {
    const int n = (int)d.size();
    double a_lo = 0.0, s_lo = slope0;        // phi' < 0 here
    double a_hi = NAN, s_hi = NAN;           // phi' > 0 (or not finite) here, once bracketed
    double a_prev = 0.0, s_prev = slope0;    // previous descending point, for extrapolation
    double a = a1, fa = f_a1, sa = s_a1;
    for (int it = 0; it < 10; it++) {
      if (std::isfinite(fa) && std::fabs(sa) <= 0.5 * std::fabs(slope0) && fa <= f0 + noise) {
        f_out = fa;
        return a;
      }
      if (!std::isfinite(fa) || sa > 0.0) {
        a_hi = a;
        s_hi = std::isfinite(fa) ? sa : NAN;
      } else {
        a_prev = a_lo;
        s_prev = s_lo;
        a_lo = a;
        s_lo = sa;
      }
      double next;
      if (std::isfinite(a_hi)) {
        const double w = a_hi - a_lo;
        next = std::isfinite(s_hi) ? a_lo - s_lo * w / (s_hi - s_lo) : a_lo + 0.5 * w;
        next = std::min(std::max(next, a_lo + 0.1 * w), a_hi - 0.1 * w);
      } else {
        const double c = (s_lo - s_prev) / (a_lo - a_prev);  // curvature along d
        next = c > 0.0 ? a_lo - s_lo / c : 2.0 * a_lo;
        next = std::min(std::max(next, 1.5 * a_lo), 4.0 * a_lo);
      }
      a = next;
      for (int i = 0; i < n; i++) u1[i] = u[i] + a * d[i];
      fa = eval_fg(u1, g1);
      sa = 0.0;
      for (int i = 0; i < n; i++) sa += g1[i] * d[i];
    }
    return 0.0;
  }

  int finish(int st, double f, double e, const std::vector<double>& u, int n) {
    if (print_level >= 3)
      std::fprintf(stderr, "[migrad %d] finish status %d f %.10g edm %.3g calls %d iterations %d\n", trace_id, st, f, e, ncalls,
                   iterations);
    for (int k = 0; k < n; k++) x[free_[k]] = int2ext(k, u[k]);
    fval = f;
    edm = e;
    status = st;
    valid = st == 0;
    // approximate errors from the last internal Hessian (hesse() replaces them)
    if (!last_hint_.empty()) {
      std::vector<double> C = last_hint_;
      if (la::spd_inverse(C, n)) {
        cov.assign(n * n, 0.0);
        for (int i = 0; i < n; i++)
          for (int j = 0; j < n; j++) cov[i * n + j] = 2.0 * up * C[i * n + j] * dext(i, u[i]) * dext(j, u[j]);
        for (int k = 0; k < n; k++) errors[free_[k]] = std::sqrt(cov[k * n + k]);
      }
    }
    return st;
  }
};

}  // namespace morefit
#endif

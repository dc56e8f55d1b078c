/*
 * oracle.c -- FP64 CPU restatement of MoreFit's likelihood hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP backend and the CPU
 * baseline that bench.py times.  Never linked into the product.
 */
#include "oracle.h"
#include "jet.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define SQRT2PI 2.5066282746310002
#define SQRT2 1.4142135623730951
#define SQRTPI 1.7724538509055160
#define KSTAR_C (9.0 / 32.0 / M_PI)

int orc_npar(int model) {
  switch (model) {
    case ORC_MASS: return 4;
    case ORC_KSTAR: return 8;
    case ORC_CB: return 4;
    case ORC_POLY: return 3;
  }
  return -1;
}
int orc_nobs(int model) { return model == ORC_KSTAR ? 3 : (model >= 0 && model <= ORC_POLY ? 1 : -1); }

/* ------------------------------------------------------------------------------------
 * Mass model (pdf.hh:147-154, 431-440, 526-558):
 *   p(m) = f * G(m)/N_G + (1-f) * E(m)/N_E
 *   G = exp(-(m-mu)^2 / (2 sigma^2)) / (sqrt(2 pi) sigma)
 *   N_G = 1/2 [erf((b-mu)/(sqrt2 sigma)) - erf((a-mu)/(sqrt2 sigma))]
 *   E = exp(lambda m),  N_E = (exp(lambda b) - exp(lambda a)) / lambda  (b-a at lambda=0)
 * ---------------------------------------------------------------------------------- */
static jet mass_jet(double x, const double* th, const double* r) {
  const int n = 4;
  jet mu = jet_var(n, th[0], 0), sg = jet_var(n, th[1], 1);
  jet f = jet_var(n, th[2], 2), lam = jet_var(n, th[3], 3);
  double a = r[0], b = r[1];
  /* Gaussian */
  jet t = jet_div(jet_sub(jet_const(n, x), mu), sg);
  jet G = jet_div(jet_exp(jet_scale(jet_mul(t, t), -0.5)), jet_scale(sg, SQRT2PI));
  jet s2 = jet_scale(sg, SQRT2);
  jet NG = jet_scale(jet_sub(jet_erf(jet_div(jet_sub(jet_const(n, b), mu), s2)),
                             jet_erf(jet_div(jet_sub(jet_const(n, a), mu), s2))), 0.5);
  /* Exponential */
  jet E = jet_exp(jet_scale(lam, x));
  jet NE;
  if (th[3] == 0.0)
    NE = jet_const(n, b - a); /* ConditionalEqual branch, pdf.hh:437-439 */
  else
    NE = jet_div(jet_sub(jet_exp(jet_scale(lam, b)), jet_exp(jet_scale(lam, a))), lam);
  jet one_m_f = jet_addc(jet_scale(f, -1.0), 1.0);
  jet p = jet_add(jet_mul(f, jet_div(G, NG)), jet_mul(one_m_f, jet_div(E, NE)));
  return p;
}

/* parameter-only terms, computed once per evaluation (what the fitter hoists) */
typedef struct {
  double mu, sigma, f, lam, inv_s, A, B, dlnNG_dmu, dlnNG_ds, dlnNE_dl, Gnorm, Enorm;
} mass_hoist;

static void mass_prepare(const double* th, const double* r, mass_hoist* h) {
  double mu = th[0], s = th[1], f = th[2], lam = th[3], a = r[0], b = r[1];
  double za = (a - mu) / (SQRT2 * s), zb = (b - mu) / (SQRT2 * s);
  double NG = 0.5 * (erf(zb) - erf(za));
  double ea2 = exp(-za * za), eb2 = exp(-zb * zb);
  double dNG_dmu = -(eb2 - ea2) / (SQRT2PI * s);
  double dNG_ds = -(zb * eb2 - za * ea2) / (SQRTPI * s);
  double NE, dNE_dl;
  if (lam == 0.0) {
    /* the reference's ConditionalEqual branch (pdf.hh:437-439) is the constant b - a, and its
     * derivative rule (graph.hh:2291-2299) differentiates the branch taken: at lambda = 0
     * exactly the normalisation contributes nothing to d/dlambda (pinned by the lambda = 0
     * points of tests/golden/ref_graph.json; the limit lambda -> 0 would give (b^2 - a^2)/2) */
    NE = b - a;
    dNE_dl = 0.0;
  } else {
    double eb = exp(lam * b), ea = exp(lam * a);
    NE = (eb - ea) / lam;
    dNE_dl = (b * eb - a * ea) / lam - NE / lam;
  }
  h->mu = mu; h->sigma = s; h->f = f; h->lam = lam;
  h->inv_s = 1.0 / s;
  h->Gnorm = 1.0 / (SQRT2PI * s * NG);
  h->Enorm = 1.0 / NE;
  h->A = f * h->Gnorm;
  h->B = (1.0 - f) * h->Enorm;
  h->dlnNG_dmu = dNG_dmu / NG;
  h->dlnNG_ds = dNG_ds / NG;
  h->dlnNE_dl = dNE_dl / NE;
}

static inline double mass_event(double x, const mass_hoist* h, double* grad) {
  double t = (x - h->mu) * h->inv_s;
  double g = exp(-0.5 * t * t);
  double e = exp(h->lam * x);
  double Gn = g * h->Gnorm, En = e * h->Enorm;
  double p = h->f * Gn + (1.0 - h->f) * En;
  if (grad) {
    double ip = 1.0 / p;
    grad[0] = h->f * Gn * (t * h->inv_s - h->dlnNG_dmu) * ip;
    grad[1] = h->f * Gn * ((t * t - 1.0) * h->inv_s - h->dlnNG_ds) * ip;
    grad[2] = (Gn - En) * ip;
    grad[3] = (1.0 - h->f) * En * (x - h->dlnNE_dl) * ip;
  }
  return log(p);
}

static inline double mass_p(double x, const mass_hoist* h) {
  double t = (x - h->mu) * h->inv_s;
  return h->f * (exp(-0.5 * t * t) * h->Gnorm) + (1.0 - h->f) * (exp(h->lam * x) * h->Enorm);
}

/* ------------------------------------------------------------------------------------
 * Angular model (physicspdfs.hh:39-81), linear in the parameters:
 *   p = c [ (1-FL) k_t + FL k_l + S3 k3 + S4 k4 + S5 k5 + AFB kA + S7 k7 + S8 k8 + S9 k9 ]
 * ---------------------------------------------------------------------------------- */
static void kstar_basis(const double* x, double* k /* 9: k_t, k_l, k3, k4, k5, kA, k7, k8, k9 */) {
  double cl = x[0], ck = x[1], ph = x[2];
  double cl2 = cl * cl, ck2 = ck * ck;
  double sl2 = 1.0 - cl2, sk2 = 1.0 - ck2;
  double sl = sqrt(sl2), sk = sqrt(sk2);
  double c2l = 2.0 * cl2 - 1.0;
  double s2l = 2.0 * sl * cl, s2k = 2.0 * sk * ck;
  double cph = cos(ph), sph = sin(ph), c2ph = cos(2.0 * ph), s2ph = sin(2.0 * ph);
  k[0] = 0.75 * sk2 + 0.25 * sk2 * c2l;
  k[1] = ck2 - ck2 * c2l;
  k[2] = sk2 * sl2 * c2ph;
  k[3] = s2k * s2l * cph;
  k[4] = s2k * sl * cph;
  k[5] = 4.0 / 3.0 * sk2 * cl;
  k[6] = s2k * sl * sph;
  k[7] = s2k * s2l * sph;
  k[8] = sk2 * sl2 * s2ph;
}

static inline double kstar_event(const double* x, const double* th, double* grad) {
  double k[9];
  kstar_basis(x, k);
  double q = (1.0 - th[0]) * k[0] + th[0] * k[1];
  for (int i = 1; i < 8; i++) q += th[i] * k[i + 1];
  if (grad) {
    double iq = 1.0 / q; /* the constant c cancels in dlog p */
    grad[0] = (k[1] - k[0]) * iq;
    for (int i = 1; i < 8; i++) grad[i] = k[i + 1] * iq;
  }
  return log(KSTAR_C * q);
}

static inline double kstar_p(const double* x, const double* th) {
  double k[9];
  kstar_basis(x, k);
  double q = (1.0 - th[0]) * k[0] + th[0] * k[1];
  for (int i = 1; i < 8; i++) q += th[i] * k[i + 1];
  return KSTAR_C * q;
}

static jet kstar_jet(const double* x, const double* th) {
  const int n = 8;
  double k[9];
  kstar_basis(x, k);
  jet q = jet_add(jet_scale(jet_addc(jet_scale(jet_var(n, th[0], 0), -1.0), 1.0), k[0]),
                  jet_scale(jet_var(n, th[0], 0), k[1]));
  for (int i = 1; i < 8; i++) q = jet_add(q, jet_scale(jet_var(n, th[i], i), k[i + 1]));
  return jet_scale(q, KSTAR_C);
}

/* ------------------------------------------------------------------------------------
 * Crystal Ball (pdf.hh:195-419) on [a, b], theta = [mu, sigma, alpha, n]:
 *   t = (x - mu)/sigma; core exp(-t^2/2) for x - mu + alpha sigma > 0, else tail
 *   A (B - t)^-n, A = (n/alpha)^n exp(-alpha^2/2), B = n/alpha - alpha;
 *   norm = tail on [a, m] + core on [m, b], m = mu - alpha sigma, C = mu + sigma B:
 *   tail = A sigma^n [(C-m)^(1-n) - (C-a)^(1-n)]/(n-1)   (n != 1; A sigma log((C-a)/(C-m)) at n = 1)
 *   core = sigma sqrt(pi/2) [erf((mu-m)/(sqrt2 sigma)) - erf((mu-b)/(sqrt2 sigma))]
 * The branches are chosen by value and differentiated as taken (the reference's Conditional
 * derivative rule).
 * ---------------------------------------------------------------------------------- */
static jet jet_pow(jet u, jet v) { return jet_exp(jet_mul(v, jet_log(u))); }

static jet cb_jet(double x, const double* th, const double* r) {
  const int n = 4;
  jet mu = jet_var(n, th[0], 0), sg = jet_var(n, th[1], 1), al = jet_var(n, th[2], 2), nn = jet_var(n, th[3], 3);
  double a = r[0], b = r[1];
  jet A = jet_mul(jet_pow(jet_div(nn, al), nn), jet_exp(jet_scale(jet_mul(al, al), -0.5)));
  jet B = jet_sub(jet_div(nn, al), al);
  jet C = jet_add(mu, jet_mul(sg, B));
  jet m = jet_sub(mu, jet_mul(al, sg));
  jet t = jet_div(jet_addc(jet_scale(mu, -1.0), x), sg);
  jet p;
  if (x - th[0] + th[2] * th[1] > 0.0) p = jet_exp(jet_scale(jet_mul(t, t), -0.5));
  else p = jet_div(A, jet_pow(jet_sub(B, t), nn));
  jet tail;
  if (th[3] - 1.0 != 0.0)
    tail = jet_div(jet_mul(jet_mul(A, jet_pow(sg, nn)),
                           jet_sub(jet_pow(jet_sub(C, m), jet_addc(jet_scale(nn, -1.0), 1.0)),
                                   jet_pow(jet_addc(C, -a), jet_addc(jet_scale(nn, -1.0), 1.0)))),
                   jet_addc(nn, -1.0));
  else
    tail = jet_mul(jet_mul(A, sg), jet_sub(jet_log(jet_addc(C, -a)), jet_log(jet_sub(C, m))));
  jet s2 = jet_scale(sg, SQRT2);
  jet core = jet_mul(jet_scale(sg, sqrt(M_PI / 2.0)),
                     jet_sub(jet_erf(jet_div(jet_sub(mu, m), s2)), jet_erf(jet_div(jet_addc(mu, -b), s2))));
  return jet_div(p, jet_add(tail, core));
}

/* ------------------------------------------------------------------------------------
 * Polynomial (pdf.hh:716-892) on [a, b], theta = [c1, c2, c3]: p = 1 + c1 x + c2 x^2 + c3 x^3
 * over its integral P(b) - P(a), P(x) = x + c1 x^2/2 + c2 x^3/3 + c3 x^4/4.
 * ---------------------------------------------------------------------------------- */
static jet poly_jet(double x, const double* th, const double* r) {
  const int n = 3;
  jet c1 = jet_var(n, th[0], 0), c2 = jet_var(n, th[1], 1), c3 = jet_var(n, th[2], 2);
  double a = r[0], b = r[1];
  jet p = jet_addc(jet_add(jet_add(jet_scale(c1, x), jet_scale(c2, x * x)), jet_scale(c3, x * x * x)), 1.0);
  jet P = jet_addc(jet_add(jet_add(jet_scale(c1, (b * b - a * a) / 2.0), jet_scale(c2, (b * b * b - a * a * a) / 3.0)),
                           jet_scale(c3, (b * b * b * b - a * a * a * a) / 4.0)),
                   b - a);
  return jet_div(p, P);
}

/* ------------------------------------------------------------------------------------ */
static jet model_jet(int model, const double* x, const double* theta, const double* range) {
  switch (model) {
    case ORC_MASS: return mass_jet(x[0], theta, range);
    case ORC_CB: return cb_jet(x[0], theta, range);
    case ORC_POLY: return poly_jet(x[0], theta, range);
  }
  return kstar_jet(x, theta);
}

void orc_event_jet(int model, const double* x, const double* theta, const double* range,
                   double* logp, double* grad, double* hess_full) {
  jet p = model_jet(model, x, theta, range);
  jet lp = jet_log(p);
  int n = lp.n;
  if (logp) *logp = lp.v;
  if (grad) for (int i = 0; i < n; i++) grad[i] = lp.g[i];
  if (hess_full)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) hess_full[i * n + j] = lp.h[i][j];
}

double orc_event_grad(int model, const double* x, const double* theta, const double* range, double* grad) {
  if (model == ORC_CB || model == ORC_POLY) {
    double lp;
    orc_event_jet(model, x, theta, range, &lp, grad, NULL);
    return lp;
  }
  if (model == ORC_MASS) {
    mass_hoist h;
    mass_prepare(theta, range, &h);
    return mass_event(x[0], &h, grad);
  }
  return kstar_event(x, theta, grad);
}

double orc_prob(int model, const double* x, const double* theta, const double* range) {
  if (model == ORC_CB || model == ORC_POLY) return model_jet(model, x, theta, range).v;
  if (model == ORC_MASS) {
    mass_hoist h;
    mass_prepare(theta, range, &h);
    return mass_p(x[0], &h);
  }
  return kstar_p(x, theta);
}

double orc_get_max(int model, const double* th, const double* r) {
  if (model == ORC_CB) { /* the unnormalised density peaks at 1 (x = mu) */
    double x = th[0];
    return model_jet(model, &x, th, r).v;
  }
  if (model == ORC_POLY) { /* largest value on a fine grid, with a 0.1 % margin */
    double best = 0.0;
    for (int i = 0; i <= 100000; i++) {
      double x = r[0] + (r[1] - r[0]) * i / 100000.0, v = model_jet(model, &x, th, r).v;
      if (v > best) best = v;
    }
    return 1.001 * best;
  }
  if (model == ORC_MASS) {
    double mu = th[0], s = th[1], f = th[2], lam = th[3], a = r[0], b = r[1];
    double NG = 0.5 * (erf((b - mu) / (SQRT2 * s)) - erf((a - mu) / (SQRT2 * s)));
    double gmax = 1.0 / SQRT2PI / s / NG;
    double emax;
    double nrm = exp(lam * b) - exp(lam * a);
    if (lam == 0.0) emax = 1.0 / (b - a);
    else if (lam > 0.0) emax = lam * exp(lam * b) / nrm;
    else emax = lam * exp(lam * a) / nrm;
    return f * gmax + (1.0 - f) * emax;
  }
  return KSTAR_C * (4.0 + fabs(th[1]) + fabs(th[2]) + fabs(th[3]) + 4.0 / 3.0 * fabs(th[4]) +
                    fabs(th[5]) + fabs(th[6]) + fabs(th[7]));
}

/* ------------------------------------------------------------------------------------
 * Sums
 * ---------------------------------------------------------------------------------- */
#define ORC_MAXOUT (1 + 8 + 36)

typedef struct {
  int model, mode, nout, P, absval;
  const double* data;
  long stride, from, to;
  const double* theta;
  const double* range;
  double sum[ORC_MAXOUT];
} slice_job;

static int n_outputs(int model, int mode) {
  int P = orc_npar(model);
  return mode == 0 ? 1 : (mode == 1 ? 1 + P : 1 + P + P * (P + 1) / 2);
}

/* one event's outputs in the reference layout */
static void event_outputs(int model, int mode, const double* x, const double* theta,
                          const double* range, const mass_hoist* mh, double* o) {
  int P = orc_npar(model);
  if (mode <= 1 && (model == ORC_MASS || model == ORC_KSTAR)) {
    if (model == ORC_MASS) o[0] = mass_event(x[0], mh, mode ? o + 1 : NULL);
    else o[0] = kstar_event(x, theta, mode ? o + 1 : NULL);
    return;
  }
  double H[64];
  orc_event_jet(model, x, theta, range, o, o + 1, H);
  int k = 1 + P;
  for (int i = 0; i < P; i++)
    for (int j = 0; j <= i; j++) o[k++] = H[i * P + j];
}

static void* slice_run(void* arg) {
  slice_job* J = (slice_job*)arg;
  double sum[ORC_MAXOUT], c[ORC_MAXOUT], o[ORC_MAXOUT], x[3];
  mass_hoist mh;
  if (J->model == ORC_MASS) mass_prepare(J->theta, J->range, &mh);
  int nobs = orc_nobs(J->model);
  for (int k = 0; k < J->nout; k++) sum[k] = c[k] = 0.0;
  for (long i = J->from; i < J->to; i++) {
    for (int d = 0; d < nobs; d++) x[d] = J->data[d * J->stride + i];
    event_outputs(J->model, J->mode, x, J->theta, J->range, &mh, o);
    if (J->absval)
      for (int k = 0; k < J->nout; k++) o[k] = fabs(o[k]);
    for (int k = 0; k < J->nout; k++) { /* in-loop Kahan, compute_llvm.hh:694-700 */
      double y = o[k] - c[k];
      volatile double t = sum[k] + y;
      c[k] = (t - sum[k]) - y;
      sum[k] = t;
    }
  }
  for (int k = 0; k < J->nout; k++) J->sum[k] = sum[k];
  return NULL;
}

static int nll_precise(int model, const double* data, long n, long stride, const double* theta,
                       const double* range, int mode, double* out) {
  int nout = n_outputs(model, mode), nobs = orc_nobs(model);
  long double sum[ORC_MAXOUT], c[ORC_MAXOUT];
  double o[ORC_MAXOUT], x[3];
  mass_hoist mh;
  if (model == ORC_MASS) mass_prepare(theta, range, &mh);
  for (int k = 0; k < nout; k++) sum[k] = c[k] = 0.0L;
  for (long i = 0; i < n; i++) {
    for (int d = 0; d < nobs; d++) x[d] = data[d * stride + i];
    event_outputs(model, mode, x, theta, range, &mh, o);
    for (int k = 0; k < nout; k++) {
      long double y = (long double)o[k] - c[k];
      volatile long double t = sum[k] + y;
      c[k] = (t - sum[k]) - y;
      sum[k] = t;
    }
  }
  for (int k = 0; k < nout; k++) out[k] = (double)sum[k];
  return 0;
}

int orc_nll(int model, const double* data, long n, long stride, const double* theta,
            const double* range, int mode, int nthreads, double* out) {
  int absval = mode >= 10;
  if (absval) {
    mode -= 10;
    if (nthreads == 0) nthreads = 1;
  }
  if (model < ORC_MASS || model > ORC_POLY) return -1;
  if (mode < 0 || mode > 2 || n < 0) return -1;
  if (nthreads == 0) return nll_precise(model, data, n, stride, theta, range, mode, out);
  if (nthreads > 256) nthreads = 256;
  if (n < nthreads) nthreads = n > 0 ? (int)n : 1;
  int nout = n_outputs(model, mode);
  slice_job* jobs = (slice_job*)calloc((size_t)nthreads, sizeof(slice_job));
  pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
  long size = n / nthreads; /* compute_llvm.hh:800-808: the last thread takes the remainder */
  for (int t = 0; t < nthreads; t++) {
    slice_job* J = &jobs[t];
    J->model = model; J->mode = mode; J->nout = nout; J->P = orc_npar(model); J->absval = absval;
    J->data = data; J->stride = stride; J->theta = theta; J->range = range;
    J->from = t * size;
    J->to = (t == nthreads - 1) ? n : t * size + size;
    if (nthreads == 1) slice_run(J);
    else pthread_create(&th[t], NULL, slice_run, J);
  }
  if (nthreads > 1)
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  for (int k = 0; k < nout; k++) { /* host Kahan over the partials, compute_llvm.hh:812-820 */
    double s = 0.0, c = 0.0;
    for (int t = 0; t < nthreads; t++) {
      double y = jobs[t].sum[k] - c;
      volatile double tt = s + y;
      c = (tt - s) - y;
      s = tt;
    }
    out[k] = s;
  }
  free(jobs);
  free(th);
  return 0;
}

/* ------------------------------------------------------------------------------------
 * Bin integrals (the plotter's integral_kernel, plotter.hh:138-216, over
 * PDF::definite_integral, pdf.hh:75-82 and physicspdfs.hh:82-164): the normalised density
 * integrated over [lo, hi] in observable `var`, the other observables over their full ranges.
 * Restated by quadrature of orc_prob, not by the antiderivatives the product uses:
 * composite Gauss-Legendre (4 panels of 20 nodes, split at the Crystal Ball join), and
 * 16 x 16 nodes over the angular model's other observables.
 * ---------------------------------------------------------------------------------- */
#define GL_N 20
#define GL_M 16

static void gauss_legendre(int n, double* x, double* w) { /* Newton on P_n, textbook */
  for (int i = 0; i < n; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p1 = 1.0, p0 = 0.0;
      for (int j = 1; j <= n; j++) {
        double pm = p0;
        p0 = p1;
        p1 = ((2.0 * j - 1.0) * z * p0 - (j - 1.0) * pm) / j;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

static double marginal(int model, const double* th, const double* r, int var, double v,
                       const double* mx, const double* mw) {
  if (orc_nobs(model) == 1) return orc_prob(model, &v, th, r);
  int o1 = (var + 1) % 3, o2 = (var + 2) % 3;
  double h1 = 0.5 * (r[2 * o1 + 1] - r[2 * o1]), c1 = 0.5 * (r[2 * o1 + 1] + r[2 * o1]);
  double h2 = 0.5 * (r[2 * o2 + 1] - r[2 * o2]), c2 = 0.5 * (r[2 * o2 + 1] + r[2 * o2]);
  double pt[3], s = 0.0;
  pt[var] = v;
  for (int i = 0; i < GL_M; i++)
    for (int j = 0; j < GL_M; j++) {
      pt[o1] = c1 + h1 * mx[i];
      pt[o2] = c2 + h2 * mx[j];
      s += mw[i] * mw[j] * orc_prob(model, pt, th, r);
    }
  return h1 * h2 * s;
}

double orc_bin_integral(int model, const double* th, const double* r, int var, double lo, double hi) {
  double gx[GL_N], gw[GL_N], mx[GL_M], mw[GL_M];
  gauss_legendre(GL_N, gx, gw);
  gauss_legendre(GL_M, mx, mw);
  double cut[3];
  int nc = 0;
  cut[nc++] = lo;
  if (model == ORC_CB) {
    double join = th[0] - th[2] * th[1];
    if (join > lo && join < hi) cut[nc++] = join;
  }
  cut[nc++] = hi;
  double total = 0.0;
  for (int c = 0; c + 1 < nc; c++) {
    const int panels = 4;
    double w = (cut[c + 1] - cut[c]) / panels;
    for (int p = 0; p < panels; p++) {
      double a = cut[c] + p * w, h = 0.5 * w, m = a + h, s = 0.0;
      for (int i = 0; i < GL_N; i++) s += gw[i] * marginal(model, th, r, var, m + h * gx[i], mx, mw);
      total += h * s;
    }
  }
  return total;
}

/* ------------------------------------------------------------------------------------
 * PRNG (random.hh:160-185, 333-465, 42-83)
 * ---------------------------------------------------------------------------------- */
uint64_t orc_splitmix64_next(uint64_t* state) {
  uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

void orc_xoshiro128pp_seed(uint32_t* s, uint64_t seed) {
  uint64_t st = seed;
  for (int i = 0; i < 4; i++) s[i] = (uint32_t)(orc_splitmix64_next(&st) >> 32);
}

static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

uint32_t orc_xoshiro128pp_next(uint32_t* s) {
  uint32_t r = rotl32(s[0] + s[3], 7) + s[0];
  uint32_t t = s[1] << 9;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl32(s[3], 11);
  return r;
}

void orc_xoshiro128pp_jump(uint32_t* s) {
  static const uint32_t J[4] = {0x8764000b, 0xf542d2d3, 0x6fa035c3, 0x77f2db5b};
  uint32_t a = 0, b = 0, c = 0, d = 0;
  for (int i = 0; i < 4; i++)
    for (int bit = 0; bit < 32; bit++) {
      if (J[i] & (1u << bit)) { a ^= s[0]; b ^= s[1]; c ^= s[2]; d ^= s[3]; }
      orc_xoshiro128pp_next(s);
    }
  s[0] = a; s[1] = b; s[2] = c; s[3] = d;
}

double orc_xoshiro128pp_random(uint32_t* s) {
  return (double)(float)(orc_xoshiro128pp_next(s) * 2.3283064365386963e-10);
}

unsigned orc_poisson(uint32_t* s, double mean) {
  if (mean <= 0.0) return 0;
  if (mean < 25.0) { /* Knuth */
    double p = 1.0, L = exp(-mean);
    int k;
    for (k = 0; p > L; k++) p *= orc_xoshiro128pp_random(s);
    return (unsigned)(k - 1);
  }
  if (mean < 1.e9) { /* rejection from a Lorentzian envelope */
    double sq = sqrt(2.0 * mean), lm = log(mean), g = mean * lm - lgamma(mean + 1.0);
    double r, m, y;
    do {
      do {
        y = tan(M_PI * orc_xoshiro128pp_random(s));
        m = sq * y + mean;
      } while (m < 0.0);
      m = floor(m);
      r = 0.9 * (1.0 + y * y) * exp(m * lm - lgamma(m + 1.0) - g);
    } while (orc_xoshiro128pp_random(s) > r);
    return (unsigned)m;
  }
  return (unsigned)mean; /* not reached by any configuration here */
}

long orc_generate(int model, const double* theta, const double* range, double maxprob,
                  uint32_t* states, int nstreams, const long* offsets, double* out, long stride) {
  int nobs = orc_nobs(model);
  long proposals = 0;
  double x[3];
  mass_hoist mh;
  if (model == ORC_MASS) mass_prepare(theta, range, &mh);
  for (int k = 0; k < nstreams; k++) {
    uint32_t* st = states + 4 * k;
    for (long i = offsets[k]; i < offsets[k + 1]; i++) {
      for (;;) {
        proposals++;
        for (int d = 0; d < nobs; d++) {
          double lo = range[2 * d], hi = range[2 * d + 1];
          x[d] = lo + (hi - lo) * (orc_xoshiro128pp_next(st) * 2.3283064365386963e-10);
        }
        double p = model == ORC_MASS ? mass_p(x[0], &mh) : model == ORC_KSTAR ? kstar_p(x, theta) : orc_prob(model, x, theta, range);
        double u = orc_xoshiro128pp_next(st) * 2.3283064365386963e-10;
        if (maxprob * u < p) break;
      }
      for (int d = 0; d < nobs; d++) out[d * stride + i] = x[d];
    }
  }
  return proposals;
}

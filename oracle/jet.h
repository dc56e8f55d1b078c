/*
 * jet.h -- second-order forward-mode differentiation for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links this file; it is used by
 * oracle.c, which only tests/, bench.py's CPU-baseline leg and the smoke check load.
 *
 * A jet carries f, grad f and the (full, symmetric) Hessian of f with respect to up to
 * JET_MAXP independent parameters.  Every operation applies the exact chain rule
 *     h(u):  grad = h'(u) grad u,   H = h'(u) H_u + h''(u) grad u grad u^T
 * so value, gradient and Hessian of a closed-form log-density come out to rounding
 * accuracy.  This is deliberately a different method from the symbolic differentiation
 * the product uses (graph.hh diff rules, reference graph.hh:321-2299), so that the two
 * cross-check each other.
 */
#ifndef MOREFIT_ORACLE_JET_H
#define MOREFIT_ORACLE_JET_H

#include <math.h>
#include <string.h>

#define JET_MAXP 8

typedef struct {
  int n;                         /* number of active parameters */
  double v;                      /* value */
  double g[JET_MAXP];            /* gradient */
  double h[JET_MAXP][JET_MAXP];  /* Hessian */
} jet;

static inline jet jet_const(int n, double c) {
  jet r; memset(&r, 0, sizeof(r)); r.n = n; r.v = c; return r;
}
static inline jet jet_var(int n, double v, int idx) {
  jet r = jet_const(n, v); r.g[idx] = 1.0; return r;
}
/* generic unary rule with f = h(u), d1 = h'(u), d2 = h''(u) */
static inline jet jet_unary(const jet* u, double f, double d1, double d2) {
  jet r; r.n = u->n; r.v = f;
  for (int i = 0; i < u->n; i++) r.g[i] = d1 * u->g[i];
  for (int i = 0; i < u->n; i++)
    for (int j = 0; j < u->n; j++)
      r.h[i][j] = d1 * u->h[i][j] + d2 * u->g[i] * u->g[j];
  return r;
}
static inline jet jet_add(jet a, jet b) {
  jet r; r.n = a.n; r.v = a.v + b.v;
  for (int i = 0; i < a.n; i++) r.g[i] = a.g[i] + b.g[i];
  for (int i = 0; i < a.n; i++) for (int j = 0; j < a.n; j++) r.h[i][j] = a.h[i][j] + b.h[i][j];
  return r;
}
static inline jet jet_sub(jet a, jet b) {
  jet r; r.n = a.n; r.v = a.v - b.v;
  for (int i = 0; i < a.n; i++) r.g[i] = a.g[i] - b.g[i];
  for (int i = 0; i < a.n; i++) for (int j = 0; j < a.n; j++) r.h[i][j] = a.h[i][j] - b.h[i][j];
  return r;
}
static inline jet jet_scale(jet a, double c) {
  jet r; r.n = a.n; r.v = c * a.v;
  for (int i = 0; i < a.n; i++) r.g[i] = c * a.g[i];
  for (int i = 0; i < a.n; i++) for (int j = 0; j < a.n; j++) r.h[i][j] = c * a.h[i][j];
  return r;
}
static inline jet jet_addc(jet a, double c) { a.v += c; return a; }
static inline jet jet_mul(jet a, jet b) {
  jet r; r.n = a.n; r.v = a.v * b.v;
  for (int i = 0; i < a.n; i++) r.g[i] = a.v * b.g[i] + b.v * a.g[i];
  for (int i = 0; i < a.n; i++)
    for (int j = 0; j < a.n; j++)
      r.h[i][j] = a.v * b.h[i][j] + b.v * a.h[i][j] + a.g[i] * b.g[j] + b.g[i] * a.g[j];
  return r;
}
static inline jet jet_inv(jet u) {
  double f = 1.0 / u.v;
  return jet_unary(&u, f, -f * f, 2.0 * f * f * f);
}
static inline jet jet_div(jet a, jet b) { return jet_mul(a, jet_inv(b)); }
static inline jet jet_exp(jet u) { double e = exp(u.v); return jet_unary(&u, e, e, e); }
static inline jet jet_log(jet u) { return jet_unary(&u, log(u.v), 1.0 / u.v, -1.0 / (u.v * u.v)); }
static inline jet jet_sqrt(jet u) {
  double s = sqrt(u.v);
  return jet_unary(&u, s, 0.5 / s, -0.25 / (s * u.v));
}
static inline jet jet_erf(jet u) {
  /* d/dx erf = 2/sqrt(pi) exp(-x^2);  d2/dx2 = -2x * that */
  double d1 = 1.1283791670955126 * exp(-u.v * u.v);
  return jet_unary(&u, erf(u.v), d1, -2.0 * u.v * d1);
}

#endif

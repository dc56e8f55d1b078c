/*
 * oracle.h -- FP64 CPU restatement of MoreFit's likelihood hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load liboracle.so.  The
 * product (libmorefit_hip.so and the morefit C++ headers) never links or calls it.
 *
 * What it restates (reference files, relative to the reference checkout):
 *   mass model      pdf.hh:147-154 (Gaussian prob/norm), pdf.hh:431-440 (Exponential),
 *                   pdf.hh:526-558 (SumPDF with one fraction), pdf.hh:103-106 (log p - log norm)
 *   angular model   physicspdfs.hh:39-81 (P-wave B->K*ll in FL,S3,S4,S5,AFB,S7,S8,S9)
 *   per-output sums compute_llvm.hh:785-850 (contiguous thread slices, in-loop Kahan per
 *                   output, host Kahan over the thread partials), utils.hh:14-41
 *   output layout   fitter.hh:1094-1106 (element 0 = sum log p, 1..P = sum dlog p/dtheta_k)
 *                   fitter.hh:1290-1301 (Hessian packed i>=j, row by row)
 *   PRNG            random.hh:160-185 (Splitmix64), random.hh:333-465 (Xoshiro128++ with
 *                   jump), random.hh:42-83 (Poisson), compute_opencl.hh:540-575 (accept-reject)
 *   Crystal Ball    pdf.hh:195-419; polynomial pdf.hh:716-892 (value, jets, maxima)
 *   maxima          pdf.hh:185-189, 467-476, 621-631; physicspdfs.hh:209-213
 *
 * Parity pins (tests/test_oracle.py): golden vectors written by the reference's own
 * graph.hh / random.hh (oracle/ref/, tests/golden/), closed forms, central differences,
 * numerical normalisation, and the jet (forward-mode) vs closed-form cross-check.
 */
#ifndef MOREFIT_ORACLE_H
#define MOREFIT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MASS = 0, ORC_KSTAR = 1, ORC_CB = 2, ORC_POLY = 3 };

/* number of parameters / observables of a model */
int orc_npar(int model);
int orc_nobs(int model);

/* range: [lo_0, hi_0, lo_1, hi_1, ...] of the observables
 * mass theta  = [mu, sigma, f, lambda]           (order of benchmarking/morefit_massfit.cc:62)
 * kstar theta = [FL, S3, S4, S5, AFB, S7, S8, S9] (order of physicspdfs.hh:33-34)
 * cb theta    = [mu, sigma, alpha, n]             (CrystalballPDF, pdf.hh:195-419)
 * poly theta  = [c1, c2, c3]                      (PolynomialPDF of degree 3, pdf.hh:716-892)
 * The Crystal Ball and polynomial densities are evaluated by jets on every path.         */

/* value, gradient and full Hessian of log p(x; theta) for one event, by jets */
void orc_event_jet(int model, const double* x, const double* theta, const double* range,
                   double* logp, double* grad, double* hess_full);

/* closed-form value and gradient of log p for one event (the fast CPU path) */
double orc_event_grad(int model, const double* x, const double* theta, const double* range, double* grad);

/* normalised density p(x; theta) */
double orc_prob(int model, const double* x, const double* theta, const double* range);
/* PDF::get_max() of the reference (upper bound used by accept-reject) */
double orc_get_max(int model, const double* theta, const double* range);

/* integral of the normalised density over [lo, hi] in observable var, the others over their
 * full ranges (plotter bins), by Gauss-Legendre quadrature of orc_prob */
double orc_bin_integral(int model, const double* theta, const double* range, int var, double lo, double hi);

/* Per-output sums over events.  data is SoA: observable d of event i at data[d*stride + i].
 * mode 0: out[0] = sum log p
 * mode 1: + out[1..P] = sum dlog p/dtheta
 * mode 2: + out[1+P ..] = sum d2 log p/dtheta_i dtheta_j, i>=j packed row by row
 * mode + 10: the same outputs summed as absolute values (error scales for tolerances)
 * nthreads >= 1: compute_llvm.hh semantics (contiguous slices, Kahan per output, Kahan of
 *                partials); nthreads == 0: long-double Kahan sum in one pass (the precise sum).
 * Returns 0 on success. */
int orc_nll(int model, const double* data, long n, long stride, const double* theta,
            const double* range, int mode, int nthreads, double* out);

/* ---------------- PRNG ---------------- */
uint64_t orc_splitmix64_next(uint64_t* state);
void orc_xoshiro128pp_seed(uint32_t* s, uint64_t seed);   /* Xoshiro128pp::setSeed(uint64_t) */
uint32_t orc_xoshiro128pp_next(uint32_t* s);               /* randomInt32 */
void orc_xoshiro128pp_jump(uint32_t* s);                    /* jump(): 2^64 steps */
double orc_xoshiro128pp_random(uint32_t* s);                /* host random(): float-rounded */
unsigned orc_poisson(uint32_t* s, double mean);              /* RandomGenerator::poisson */

/* Accept-reject generation (compute_opencl.hh:540-575 semantics, FP64 uniforms
 * u = r * 2^-32).  Stream k owns events [offsets[k], offsets[k+1]) and the 4-word
 * xoshiro128++ state states[4k..4k+3] (advanced in place).  out is SoA with stride
 * `stride`.  Returns the number of proposals drawn. */
long orc_generate(int model, const double* theta, const double* range, double maxprob,
                  uint32_t* states, int nstreams, const long* offsets, double* out, long stride);

#ifdef __cplusplus
}
#endif

#endif

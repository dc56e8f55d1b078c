/*
 * morefit_capi.h -- host-level C ABI of libmorefit_capi.so: the configuration densities of
 * MoreFit's likelihood path compiled for the MI355X backend, resident datasets, fused
 * -2lnL/gradient/Hessian evaluation, single fits, batched toy fits and the fit service.
 *
 * This is the binding surface a non-C++ caller (the Python tests and bench.py via ctypes)
 * uses; C++ callers use the morefit headers directly (morefit_amd/include/morefit.hh).
 * Reference interfaces each entry point replaces:
 *   mfh_model_create     PDF construction + fitter::make_kernels (benchmarking/morefit_massfit.cc:48-62,
 *                        benchmarking/morefit_kstarmumu.cc:41-55, fitter.hh:321-658)
 *   mfh_batch_*          EventVector upload + ComputeBlock CopyTo.. / SubmitKernel / CopyFromKahanBuffer
 *                        (fitter.hh:704-757, 1054-1326)
 *   mfh_batch_generate   generator::generate on the accelerator (generator.hh:71-175)
 *   mfh_fit_plain        fitter::fit with the drop-in HIPBlock (FP64 or FP32), one launch per call
 *   mfh_fit_cpu          fitter::fit with the host-compiled CPU backend (reference LLVM semantics)
 *   mfh_fit_weighted     fitter::fit of weighted events with correct_weighted_fit (fitter.hh:894-947, 1327-1432)
 *   mfh_fit_options      fitter::fit with analytic_gradient / analytic_hessian / kahan_on_accelerator /
 *                        postrun_hesse switched (fitter.hh:38-106, 796-814, 1110-1124)
 *   mfh_plain_*          fitter::fit per dataset with kernels reused (refit, toystudy.hh:76-77), device or CPU
 *   mfh_batch_fit        toystudy::toy's fit loop (toystudy.hh:55-111) as one batched study
 *   mfh_serve_run        no counterpart: continuous-batching fit service (serve.hh)
 * Status: 0 on success, else nonzero with mfh_last_error().
 */
#ifndef MOREFIT_CAPI_H
#define MOREFIT_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

#define MFH_MASS 0   /* f Gauss(mu, sigma) + (1-f) Exp(lambda) on [lo, hi]; theta = mu, sigma, f, lambda */
#define MFH_KSTAR 1  /* P-wave B->K*ll angular density; theta = FL, S3, S4, S5, AFB, S7, S8, S9 */
#define MFH_CB 2     /* Crystal Ball (pdf.hh:195-419) on [lo, hi] (5.0, 5.6); theta = mu, sigma, alpha, n */
#define MFH_POLY 3   /* 1 + c1 x + c2 x^2 + c3 x^3 (pdf.hh:716-892) on [lo, hi] (-1, 1); theta = c1, c2, c3 */

typedef struct mfh_context mfh_context;
typedef struct mfh_model mfh_model;
typedef struct mfh_batch mfh_batch;

typedef struct mfh_options {
  int device;
  int block_size;          /* evaluation work-group size (multiple of 64) */
  int events_per_thread;   /* chunk = block_size * events_per_thread */
  int grid;                /* 0 = one work-group per chunk, else persistent grid */
  int fast_math;           /* FP32 hardware exp/log/rcp/sqrt/sin/cos */
  int accumulate;          /* FP32 per-lane sums: 0 FP64 per event, 1 Kahan FP32, 2 FP32 per unrolled group then FP64 */
  int centre;              /* observables as offsets from the range centre (FP32 stored so, FP64 computed per event) */
  int unroll;              /* events per thread whose loads are issued together */
  int balance;             /* large segments: chunk count a multiple of the resident work-groups */
  int base2;               /* exp/log as exp2/log2: 0 never, 1 FP32 kernels, 2 FP32 and FP64 */
  int fused_final;         /* 1: segment totals in the work-group finishing a segment; 0: final kernel */
  int event_columns;       /* 1: event-only terms precomputed into data columns (optimize_dimensions) */
  double event_cost;       /* cost above which an event-only term becomes a column (graph cost units) */
  int linear;              /* outputs as linear forms over shared monomials: 0 never, 1 when shorter, 2 always */
  int exp_pivot;           /* linear forms: a sum of exponential terms evaluated around one of them */
  int host_params;         /* one-segment calls: hoisted terms on the host, passed by value (no prep kernel) */
  int prefetch;            /* evaluation kernels load the next event batch while computing the current one */
} mfh_options;

typedef struct mfh_fit_out {
  double fval;
  double edm;
  int status;
  int ncalls;
  int iterations;
  int hesse_ok;
  int thread_moved;  /* batched fits: the fit's fibre resumed on another OS thread (never expected) */
  int nonfinite_evals;         /* evaluations whose sums came back non-finite (failure detection) */
  long long nonfinite_events;  /* events with non-finite per-event terms at such a point (-1: unknown) */
  long long first_nonfinite;   /* the smallest such event index (-1: none or unknown) */
  long long out_of_range;      /* events with an observable outside its range or not a number (plain
                                  fits: of the bound data; batched fits: found by the diagnostic pass,
                                  which runs only after a non-finite evaluation) */
} mfh_fit_out;

typedef struct mfh_batch_stats {
  int ticks;
  int launches_lh, launches_grad, launches_hess;
  double wall_ms, gpu_wait_ms, host_ms;
  int host_threads;        /* host threads that ran the fits' MIGRAD code */
  int pad;
  double setup_ms;         /* creating the fits (fibres and their stacks) */
  double step_ms;          /* running the fits' host code between launches */
} mfh_batch_stats;

const char* mfh_last_error(void);
void mfh_default_options(mfh_options* o);

int mfh_open(const mfh_options* opts, mfh_context** out);
int mfh_close(mfh_context* ctx);
int mfh_device_name(mfh_context* ctx, char* name, int len, int* compute_units);
int mfh_copy_ceiling(mfh_context* ctx, double bytes, int iters, double* gbps);

/* model: `range` = [lo, hi] per observable (NULL: the configuration ranges) */
int mfh_model_create(mfh_context* ctx, int kind, int fp32, const double* range, int with_hessian, mfh_model** out);
int mfh_model_destroy(mfh_model* m);
/* info[0..7] = ntheta, nhoist, nobs, nout_lh, nout_grad, nout_hess, statements, graph_nodes;
 * info[8..10] = unroll of the value, gradient and Hessian kernels (fitted to the register
 * budget, 0 when not compiled), info[11..13] = their VGPR counts, info[14..16] = 1 where the
 * kernel reduces multi-chunk segments itself (0: a final kernel follows it), info[17] = data
 * columns per event (observables, then precomputed event terms);
 * ms[0..1] = codegen ms, hipRTC compile + load ms (small when served from the cache) */
int mfh_model_info(mfh_model* m, int* info, double* ms);
int mfh_model_source(mfh_model* m, char* buf, int len);
/* generated source without a device (warms the on-disk code-object cache) */
int mfh_model_precompile(int kind, int fp32, int with_hessian, const mfh_options* opts, double* compile_ms);
/* without a device: compile a configuration module and report, for the value, gradient and
 * Hessian kernels, info[0..2] = unroll, info[3..5] = VGPRs, info[6..8] = AGPRs, info[9..11] =
 * scratch bytes per lane, info[12..14] = 1 where the kernel reduces multi-chunk segments
 * itself.  fit = 1 fits each kernel's unroll to the register budget (as a model does), fit = 0
 * compiles every kernel at opts->unroll as given. */
int mfh_model_kernels(int kind, int fp32, int with_hessian, const mfh_options* opts, int fit, int* info);
/* without a device: the per-event outputs of the value (which = 1), gradient (2) or Hessian (3)
 * kernel of a configuration module exactly as generated (observables centred for fp32, base-2
 * transcendentals, hoisted terms, output scales applied), evaluated on the host in FP64 at
 * theta for events data (SoA nobs x n): out[i * nout + k].  event_cost >= 0: event-only terms
 * costing more become columns computed from the observables first (event_columns); linear as
 * mfh_options.linear (the accumulated terms combined per event as the kernels combine their
 * sums).  Checks the code generator's rewrites against the density's definition. */
int mfh_model_host_outputs(int kind, int fp32, int which, const double* theta, const double* data, long long n, double event_cost,
                           int linear, double* out);
/* launch-shape autotuner: for each candidate (block, events_per_thread, unroll) triple in
 * cand[3 * ncand] (NULL: the built-in shapes of mfh_autotune_candidates for the sizes' total) times `eval_kind` evaluations (per-chunk kernel + final pass, mean of `iters`
 * back-to-back launches) of the configuration density on a batch of nseg segments of the
 * given sizes generated on the device at its generating values; sets the context's launch
 * options to the fastest (used by models created afterwards), returns it in best[3] and the
 * per-candidate us in us_per_launch[ncand] (may be NULL).  No counterpart in the reference. */
int mfh_autotune(mfh_context* ctx, int kind, int fp32, const long long* sizes, int nseg, int eval_kind, const int* cand, int ncand,
                 int iters, int* best, double* us_per_launch);
/* the built-in candidate shapes for a workload of total_events events: block 256/512 x 8..64
 * events per thread x unroll 4/8 from 4e6 events, block 128..512 x 2..16 events per thread
 * below; up to `cap` triples into cand, their number in *count */
int mfh_autotune_candidates(long long total_events, int* cand, int cap, int* count);
/* without a device: xoshiro128++ states of streams first .. first+count-1 of `seed` by the
 * jump-ahead matrices the device seeding uses (out: 4 words per stream) */
int mfh_stream_states(unsigned long long seed, long long first, int count, unsigned int* out);

int mfh_batch_create(mfh_model* m, const long long* sizes, int nseg, mfh_batch** out);
int mfh_batch_destroy(mfh_batch* b);
int mfh_batch_upload(mfh_batch* b, int seg, const double* cols, long long col_stride);
int mfh_batch_download(mfh_batch* b, int seg, double* cols);
int mfh_batch_generate(mfh_batch* b, const double* theta_gen, const unsigned long long* seeds, int events_per_stream,
                       unsigned long long* proposals);
/* kind 1: value, 2: value+gradient, 3: value+gradient+Hessian (packed i>=j).  theta: nseg x ntheta.
 * out: nseg x nout(kind) raw sums of log p and its derivatives (not scaled by -2). */
int mfh_batch_eval(mfh_batch* b, int kind, const double* theta, int use_graph, double* out);
/* mean device time per launch of the evaluation kernels (HIP events on the library's stream
 * around `iters` back-to-back launches): the per-chunk kernel and its final pass, or with
 * main_only = 1 the per-chunk kernel alone (totals left unreduced) */
int mfh_batch_time_kernel(mfh_batch* b, int kind, int iters, int main_only, double* ms_per_launch);
/* mean wall time of the full per-call sequence (upload, prep, kernel, totals, sync) */
int mfh_batch_time_eval(mfh_batch* b, int kind, int iters, int use_graph, double* us_per_eval);
/* mean device time per call of the captured call graph (prep, kernel, final pass) replayed
 * `iters` times back to back with no host synchronisation in between (HIP events) */
int mfh_batch_time_graph(mfh_batch* b, int kind, int iters, double* us_per_eval);
long long mfh_batch_bytes(mfh_batch* b);
/* fits of the listed segments (NULL = all) from start values (nfit x ntheta) */
int mfh_batch_fit(mfh_batch* b, const int* segs, int nfit, const double* start, int use_graph, double* values, double* errors,
                  mfh_fit_out* outs, mfh_batch_stats* stats);
/* failure detection (the reference's NaN check of the host sum, fitter.hh:1197-1204, as a
 * device pass): for each listed segment at theta (nseg x ntheta), out[3 i ..] = events whose
 * per-event log-density terms are not all finite, the smallest such event index (-1: none),
 * events with an observable outside its range or not a number */
int mfh_batch_check(mfh_batch* b, const int* segs, int nseg, const double* theta, long long* out);

/* Fit service with continuous batching over the given models (request class = index into
 * `models`, all of one precision): requests arrive at arrival_s (seconds after the start),
 * are admitted within `budget_bytes` of HBM and `slots` per class, get their dataset
 * generated on the device from `seeds` at the class's generation values theta_gen (per class,
 * ntheta values) and are fitted from `start` (per class).  Per request: timings[3] =
 * arrival, admission, completion (s); values/errors (max ntheta per request, NaN-padded);
 * status.  stats[0..7] = ticks, launches, peak live fits, rejected, wall s, peak bytes (event
 * data of the admitted requests, all columns), generation ms, device bytes of the event arena
 * that all classes share (the budget: one allocation bounds every class's data together).
 * request_start (nreq rows of max ntheta values, may be NULL) replaces the class's start
 * values per request. */
int mfh_serve_run(mfh_model* const* models, int nmodels, const double* theta_gen, const double* start, const int* cls,
                  const long long* n, const unsigned long long* seeds, const double* arrival_s, int nreq, double budget_bytes,
                  int slots, double* timings, double* values, double* errors, int* status, double* stats,
                  const double* request_start);

/* one fit through the drop-in HIPBlock (reference fitter flow); data SoA nobs x n.  The
 * context's event_columns option turns on the fitter's optimize_dimensions (a store kernel
 * precomputes event-only terms into extra columns, fitter.hh:339-375). */
int mfh_fit_plain(mfh_context* ctx, int kind, int fp32, const double* data, long long n, const double* start, double* values,
                  double* errors, mfh_fit_out* out, double* ms);
/* one fit through the host-compiled CPU backend (no device needed); fp32 = 1 runs the FP32
 * kernels (FP32-rounded parameters, centred FP32 observables) as the device's FP32 path does;
 * optimize_dimensions as the fitter option of that name */
int mfh_fit_cpu(int kind, int fp32, const double* data, long long n, const double* start, int nthreads, int optimize_dimensions,
                double* values, double* errors, mfh_fit_out* out, double* ms);
/* weighted fit (log p of event i multiplied by weights[i]) with the reference's covariance
 * correction -2 H^-1 F H^-1 (Fisher matrix F from analytic kernels): through the drop-in
 * HIPBlock when ctx is set, else through the CPU backend with nthreads threads */
int mfh_fit_weighted(mfh_context* ctx, int kind, int fp32, const double* data, const double* weights, long long n, const double* start,
                     int nthreads, double* values, double* errors, mfh_fit_out* out, double* ms);
/* -2lnL, gradient (-2 scale) and Hessian through the CPU backend at theta (no device needed) */
/* fitter::fit with the reference's fitter_options switches (fitter.hh:38-106; its mass benchmark
 * runs analytic_gradient = analytic_hessian = false, benchmarking/morefit_massfit.cc:84-95):
 * flags is an OR of MFH_FIT_* bits; ctx NULL fits on the host CPU backend with nthreads. */
#define MFH_FIT_ANALYTIC_GRADIENT 1
#define MFH_FIT_ANALYTIC_HESSIAN 2
#define MFH_FIT_KAHAN_ON_ACCELERATOR 4
#define MFH_FIT_POSTRUN_HESSE 8
#define MFH_FIT_ALL 15
#define MFH_FIT_DEFAULT MFH_FIT_ALL
int mfh_fit_options(mfh_context* ctx, int kind, int fp32, const double* data, long long n, const double* start, int nthreads, int flags,
                    double* values, double* errors, mfh_fit_out* out, double* ms);
/* A plain fitter session (fitter::fit one dataset at a time, kernels built at the first fit and
 * reused by the next ones with refit, as the reference's toy study does, toystudy.hh:76-77):
 * ctx given: the device's drop-in HIPBlock path; ctx NULL: the host CPU backend with nthreads
 * threads (cpu_flags: host compiler flags of its kernels, NULL for the default).  flags as for
 * mfh_fit_options.  mfh_plain_eval binds a dataset and times `iters` -2lnL + gradient
 * evaluations at theta (seconds), returning the last value and gradient.  A session is not
 * thread-safe; sessions are independent of each other. */
typedef struct mfh_plain mfh_plain;
int mfh_plain_create(mfh_context* ctx, int kind, int fp32, int nthreads, int flags, const char* cpu_flags, mfh_plain** out);
int mfh_plain_fit(mfh_plain* p, const double* data, long long n, const double* start, double* values, double* errors, mfh_fit_out* out,
                  double* ms);
int mfh_plain_eval(mfh_plain* p, const double* data, long long n, const double* theta, int iters, double* value, double* grad,
                   double* seconds);
int mfh_plain_destroy(mfh_plain* p);
int mfh_eval_cpu(int kind, const double* data, long long n, const double* theta, int nthreads, double* value, double* grad, double* hess);

/* Projection of a density onto observable `var` (plotter<...>::plot, reference
 * plotter.hh:218-640, integral kernel plotter.hh:138-216): the integral of the density at theta
 * over nbins * multiplier fine bins of [range_min, range_max] (equal values: the observable's
 * range), run by the integral kernel on the device (ctx) or on the CPU backend (ctx NULL,
 * nthreads threads); the histogram of data (SoA, n events, weights NULL or per event),
 * residuals and pulls.  Outputs, each may be NULL: pdf_integrals [nbins * multiplier],
 * counts, errors, pulls [nbins].  filename (NULL: no file): a matplotlib script (".py",
 * plot_type MFH_PLOT_MATPLOTLIB) or a ROOT macro (".C", MFH_PLOT_ROOT); flags MFH_PLOT_PULLS
 * adds the pull panel.  replots > 0 repeats the projection that many times with the kernel of
 * the first (plot with replot); ms: duration of the last call. */
#define MFH_PLOT_ROOT 0
#define MFH_PLOT_MATPLOTLIB 1
#define MFH_PLOT_PULLS 1
int mfh_plot(mfh_context* ctx, int kind, int fp32, int nthreads, const double* theta, const double* data, const double* weights,
             long long n, int var, int nbins, int multiplier, double range_min, double range_max, int plot_type, int flags,
             const char* filename, int replots, double* pdf_integrals, double* counts, double* errors, double* pulls, double* ms);

#ifdef __cplusplus
}
#endif

#endif

// capi.cc -- libmorefit_capi.so: the host C ABI of include/morefit_capi.h over the morefit
// C++ headers (graph, densities, fitter, HIP and CPU backends, batched fits).
#include <dlfcn.h>
#include <sys/stat.h>
#include <fstream>
#include <sstream>
#include <chrono>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "morefit.hh"
#include "morefit/autotune.hh"
#include "morefit/plotter.hh"
#include "morefit/batch.hh"
#include "morefit/serve.hh"
#include "morefit_capi.h"

using namespace morefit;

namespace {
thread_local std::string g_err;

int fail(const std::string& what) {
  g_err = what;
  return 1;
}

#define MFH_TRY try {
#define MFH_CATCH                                  \
  }                                                \
  catch (const std::exception& e) {                \
    return fail(e.what());                         \
  }                                                \
  catch (...) {                                    \
    return fail("unknown C++ exception");          \
  }                                                \
  return 0;

// the two configuration densities with the parameter limits and steps of the reference
// benchmarks (benchmarking/morefit_massfit.cc:48-57, benchmarking/morefit_kstarmumu.cc:41-52),
// and the Crystal Ball and third-order polynomial densities
template <typename K>
struct ModelDef {
  typedef K kernel_type;
  int kind;
  std::vector<std::unique_ptr<dimension<double>>> dims;
  std::vector<std::unique_ptr<parameter<double>>> pars;
  std::vector<std::unique_ptr<PDF<K, double>>> pdfs;
  PDF<K, double>* pdf = nullptr;
  std::vector<parameter<double>*> params;
  std::vector<dimension<double>*> dimv;

  ModelDef(int k, const double* range) : kind(k) {
    if (k == MFH_MASS) {
      double lo = range ? range[0] : 5.0, hi = range ? range[1] : 5.6;
      dims.emplace_back(new dimension<double>("m", lo, hi));
      pars.emplace_back(new parameter<double>("mu", 5.28, lo, hi, 0.01));
      pars.emplace_back(new parameter<double>("sigma", 0.02, 0.005, 0.13, 0.001));
      pars.emplace_back(new parameter<double>("f", 0.3, 0.0, 1.0, 0.01));
      pars.emplace_back(new parameter<double>("lambda", -1.5, -10.0, 10.0, 0.01));
      auto* g = new GaussianPDF<K, double>(dims[0].get(), pars[0].get(), pars[1].get());
      auto* e = new ExponentialPDF<K, double>(dims[0].get(), pars[3].get());
      pdfs.emplace_back(g);
      pdfs.emplace_back(e);
      pdfs.emplace_back(new SumPDF<K, double>(g, e, pars[2].get()));
    } else if (k == MFH_KSTAR) {
      const double dl[6] = {-1.0, 1.0, -1.0, 1.0, -M_PI, M_PI};
      const double* r = range ? range : dl;
      dims.emplace_back(new dimension<double>("ctl", r[0], r[1]));
      dims.emplace_back(new dimension<double>("ctk", r[2], r[3]));
      dims.emplace_back(new dimension<double>("phi", r[4], r[5]));
      const char* names[8] = {"FL", "S3", "S4", "S5", "AFB", "S7", "S8", "S9"};
      const double truth[8] = {0.7, 0.0, -0.1, -0.2, 0.1, 0.0, 0.0, 0.0};
      for (int i = 0; i < 8; i++) pars.emplace_back(new parameter<double>(names[i], truth[i], i == 0 ? 0.0 : -1.0, 1.0, 0.01));
      pdfs.emplace_back(new KstarmumuAngularPDF<K, double>(dims[0].get(), dims[1].get(), dims[2].get(), pars[0].get(), pars[1].get(),
                                                           pars[2].get(), pars[3].get(), pars[4].get(), pars[5].get(), pars[6].get(),
                                                           pars[7].get()));
    } else if (k == MFH_CB) {
      // Crystal Ball signal on the mass range: its tail joins the core at mu - alpha sigma
      double lo = range ? range[0] : 5.0, hi = range ? range[1] : 5.6;
      dims.emplace_back(new dimension<double>("m", lo, hi));
      pars.emplace_back(new parameter<double>("mu", 5.28, lo, hi, 0.01));
      pars.emplace_back(new parameter<double>("sigma", 0.02, 0.005, 0.13, 0.001));
      pars.emplace_back(new parameter<double>("alpha", 1.5, 0.2, 5.0, 0.01));
      pars.emplace_back(new parameter<double>("n", 3.0, 1.05, 20.0, 0.1));
      pdfs.emplace_back(new CrystalballPDF<K, double>(dims[0].get(), pars[0].get(), pars[1].get(), pars[2].get(), pars[3].get()));
    } else if (k == MFH_POLY) {
      // third-order polynomial 1 + c1 x + c2 x^2 + c3 x^3 on [-1, 1]
      double lo = range ? range[0] : -1.0, hi = range ? range[1] : 1.0;
      dims.emplace_back(new dimension<double>("x", lo, hi));
      const char* names[3] = {"c1", "c2", "c3"};
      const double truth[3] = {0.3, -0.2, 0.1};
      for (int i = 0; i < 3; i++) pars.emplace_back(new parameter<double>(names[i], truth[i], -1.0, 1.0, 0.01));
      pdfs.emplace_back(new PolynomialPDF<K, double>(dims[0].get(), {pars[0].get(), pars[1].get(), pars[2].get()}));
    } else {
      throw std::invalid_argument("unknown model kind " + std::to_string(k));
    }
    pdf = pdfs.back().get();
    for (auto& p : pars) params.push_back(p.get());
    for (auto& d : dims) dimv.push_back(d.get());
  }
};

void apply_options(compute_options& co, const mfh_options* o) {
  mfh_options d;
  mfh_default_options(&d);
  if (!o) o = &d;
  co.hip_device = o->device;
  co.hip_block_size = o->block_size > 0 ? o->block_size : 256;
  co.hip_events_per_thread = o->events_per_thread > 0 ? o->events_per_thread : 64;
  co.hip_grid = o->grid;
  co.hip_fast_math = o->fast_math != 0;
  co.hip_accumulate = o->accumulate;
  co.hip_centre_observables = o->centre != 0;
  co.hip_unroll = o->unroll > 0 ? o->unroll : 8;
  co.hip_balance_chunks = o->balance != 0;
  co.hip_base2 = o->base2;
  co.hip_fused_final = o->fused_final != 0;
  co.hip_event_columns = o->event_columns != 0;
  co.hip_event_cost_threshold = o->event_cost;
  co.hip_linear_outputs = o->linear;
  co.hip_exp_pivot = o->exp_pivot != 0;
  co.hip_host_params = o->host_params != 0;
  co.hip_prefetch = o->prefetch != 0;
}
}  // namespace

struct mfh_context {
  compute_options co;
  std::unique_ptr<HIPBackend> be;
};

struct mfh_model {
  mfh_context* ctx = nullptr;
  int kind = 0;
  bool fp32 = false;
  std::unique_ptr<ModelDef<float>> df;
  std::unique_ptr<ModelDef<double>> dd;
  std::unique_ptr<HipModel<float>> mf;
  std::unique_ptr<HipModel<double>> md;
  template <typename F>
  auto with(F f) {
    return fp32 ? f(*mf) : f(*md);
  }
};

struct mfh_batch {
  mfh_model* model = nullptr;
  std::unique_ptr<HipBatch<float>> bf;
  std::unique_ptr<HipBatch<double>> bd;
  template <typename F>
  auto with(F f) {
    return bf ? f(*bf) : f(*bd);
  }
};

extern "C" {

const char* mfh_last_error(void) { return g_err.c_str(); }

void mfh_default_options(mfh_options* o) {
  o->device = 0;
  o->block_size = 256;
  o->events_per_thread = 32;
  o->grid = 0;
  o->fast_math = 1;
  o->accumulate = 2;
  o->centre = 1;
  o->unroll = 8;
  o->balance = 0;
  o->base2 = 1;
  o->fused_final = 0;
  o->event_columns = 0;
  o->event_cost = 10.0;
  o->linear = 1;
  o->exp_pivot = 1;
  o->host_params = 1;
  o->prefetch = 1;
}

int mfh_open(const mfh_options* opts, mfh_context** out) {
  MFH_TRY
  auto* c = new mfh_context();
  apply_options(c->co, opts);
  try {
    c->be.reset(new HIPBackend(&c->co));
  } catch (...) {
    delete c;
    throw;
  }
  *out = c;
  MFH_CATCH
}

int mfh_close(mfh_context* ctx) {
  delete ctx;
  return 0;
}

int mfh_device_name(mfh_context* ctx, char* name, int len, int* cus) {
  MFH_TRY
  auto p = ctx->be->props();
  std::snprintf(name, len, "%s (%s)", p.name, p.arch);
  if (cus) *cus = p.compute_units;
  MFH_CATCH
}

int mfh_copy_ceiling(mfh_context* ctx, double bytes, int iters, double* gbps) {
  MFH_TRY
  hip_check(mf_bench_copy(ctx->be->device(), (size_t)bytes, iters, gbps), "mf_bench_copy");
  MFH_CATCH
}

int mfh_model_create(mfh_context* ctx, int kind, int fp32, const double* range, int with_hessian, mfh_model** out) {
  MFH_TRY
  std::unique_ptr<mfh_model> m(new mfh_model());
  m->ctx = ctx;
  m->kind = kind;
  m->fp32 = fp32 != 0;
  if (m->fp32) {
    m->df.reset(new ModelDef<float>(kind, range));
    m->mf.reset(new HipModel<float>(ctx->be.get(), m->df->pdf, m->df->params, with_hessian != 0));
  } else {
    m->dd.reset(new ModelDef<double>(kind, range));
    m->md.reset(new HipModel<double>(ctx->be.get(), m->dd->pdf, m->dd->params, with_hessian != 0));
  }
  *out = m.release();
  MFH_CATCH
}

int mfh_model_destroy(mfh_model* m) {
  delete m;
  return 0;
}

int mfh_model_info(mfh_model* m, int* info, double* ms) {
  MFH_TRY
  m->with([&](auto& hm) {
    info[0] = hm.ntheta;
    info[1] = hm.nhoist;
    info[2] = hm.nobs;
    info[3] = hm.nout[EVAL_LH];
    info[4] = hm.nout[EVAL_GRAD];
    info[5] = hm.nout[EVAL_HESS];
    info[6] = (int)hm.statements;
    info[7] = (int)hm.graph_nodes;
    const char* names[3] = {"mf_eval_lh", "mf_eval_grad", "mf_eval_hess"};
    for (int k = 0; k < 3; k++) {
      info[8 + k] = hm.kernel_unroll(names[k]);
      info[11 + k] = hm.kernel_vgprs(names[k]);
      info[14 + k] = hm.kernel_fused(names[k]) ? 1 : 0;
    }
    info[17] = hm.ncol;
    if (ms) {
      ms[0] = hm.codegen_ms;
      ms[1] = hm.compile_ms;
    }
    return 0;
  });
  MFH_CATCH
}

int mfh_model_source(mfh_model* m, char* buf, int len) {
  MFH_TRY
  m->with([&](auto& hm) {
    std::snprintf(buf, len, "%s", hm.source.c_str());
    return 0;
  });
  MFH_CATCH
}

int mfh_model_precompile(int kind, int fp32, int with_hessian, const mfh_options* opts, double* compile_ms) {
  MFH_TRY
  compute_options co;
  apply_options(co, opts);
  std::string src;
  if (fp32) {
    ModelDef<float> d(kind, nullptr);
    src = HipModel<float>::make_source(d.pdf, d.params, co, with_hessian != 0).device_source();
  } else {
    ModelDef<double> d(kind, nullptr);
    src = HipModel<double>::make_source(d.pdf, d.params, co, with_hessian != 0).device_source();
  }
  auto t0 = std::chrono::steady_clock::now();
  size_t bytes = 0;
  hip_check(mf_module_compile_only(src.c_str(), co.hip_compile_options.c_str(), &bytes), "mf_module_compile_only");
  if (compile_ms) *compile_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  MFH_CATCH
}

int mfh_model_kernels(int kind, int fp32, int with_hessian, const mfh_options* opts, int fit, int* info) {
  MFH_TRY
  compute_options co;
  apply_options(co, opts);
  auto report = [&](auto& d) {
    typedef typename std::remove_reference<decltype(d)>::type::kernel_type K;
    ModuleSource ms;
    std::vector<mf_kernel_resources> res;
    if (fit) {
      ms = HipModel<K>::make_source(d.pdf, d.params, co, with_hessian != 0, nullptr, &res);
    } else {
      const bool f32 = std::is_same<K, float>::value;
      LikelihoodGraphs<K> g(d.pdf, d.params, co.hip_centre_observables, with_hessian != 0,
                            co.hip_base2 >= 2 || (f32 && co.hip_base2 == 1), 0.0,
                            co.hip_event_columns ? co.hip_event_cost_threshold : -1.0, co.hip_linear_outputs, co.hip_exp_pivot, !f32);
      ms = g.module(co, with_hessian != 0);
      mf_kernel_resources r[64];
      int n = 0;
      hip_check(mf_module_resources(ms.device_source().c_str(), co.hip_compile_options.c_str(), r, 64, &n), "mf_module_resources");
      res.assign(r, r + std::min(n, 64));
    }
    const char* names[3] = {"mf_eval_lh", "mf_eval_grad", "mf_eval_hess"};
    for (int k = 0; k < 15; k++) info[k] = 0;
    for (int k = 0; k < 3; k++) {
      for (auto& e : ms.evals)
        if (e.name == names[k]) {
          info[k] = e.unroll;
          info[12 + k] = e.fused ? 1 : 0;
        }
      for (auto& r : res)
        if (std::string(r.name) == names[k]) {
          info[3 + k] = r.vgpr_count;
          info[6 + k] = r.agpr_count;
          info[9 + k] = r.private_segment_bytes;
        }
    }
  };
  if (fp32) {
    ModelDef<float> d(kind, nullptr);
    report(d);
  } else {
    ModelDef<double> d(kind, nullptr);
    report(d);
  }
  MFH_CATCH
}

int mfh_model_host_outputs(int kind, int fp32, int which, const double* theta, const double* data, long long n, double event_cost,
                           int linear, double* out) {
  MFH_TRY
  if (which < 1 || which > 3) throw std::invalid_argument("which must be 1, 2 or 3");
  auto run = [&](auto& d) {
    typedef typename std::remove_reference<decltype(d)>::type::kernel_type K;
    LikelihoodGraphs<K> g(d.pdf, d.params, true, true, fp32 != 0, 0.0, event_cost, linear, true, fp32 == 0);
    std::vector<std::string> names(g.theta);
    std::vector<double> vals(theta, theta + g.theta.size());
    const size_t nt = names.size();
    std::vector<double> hv;
    for (auto& e : g.hoist_exprs) hv.push_back(e->eval(std::vector<std::string>(names.begin(), names.begin() + nt),
                                                       std::vector<double>(vals.begin(), vals.begin() + nt)));
    names.insert(names.end(), g.hoist_names.begin(), g.hoist_names.end());
    vals.insert(vals.end(), hv.begin(), hv.end());
    const size_t off = names.size();
    names.insert(names.end(), g.obs.begin(), g.obs.end());
    const size_t eoff = names.size();
    names.insert(names.end(), g.ev_names.begin(), g.ev_names.end());
    vals.resize(names.size());
    const std::vector<std::string> obs_names(g.obs);
    std::vector<double> xv(g.obs.size());
    const auto& kf = which == 1 ? g.lh : (which == 2 ? g.grad : g.hess);
    const size_t na = kf.acc.size(), no = kf.form.size(), nm = g.mom_exprs.size();
    std::vector<double> av(na), mv(nm);
    for (long long i = 0; i < n; i++) {
      for (size_t dd = 0; dd < g.obs.size(); dd++) xv[dd] = vals[off + dd] = data[dd * n + i] - g.centres[dd];
      for (size_t e = 0; e < g.ev_exprs.size(); e++) vals[eoff + e] = g.ev_exprs[e]->eval(obs_names, xv);
      for (size_t j = 0; j < na; j++) av[j] = (double)kf.acc[j]->eval(names, vals);
      for (size_t j = 0; j < nm; j++) mv[j] = (double)g.mom_exprs[j]->eval(names, vals);
      // the kernels' FP64 combination of the accumulated terms and data moments, for one event
      // (n = 1)
      for (size_t k = 0; k < no; k++) {
        double v = 0.0;
        for (const auto& t : kf.form[k])
          v += t.mult * (t.hoist >= 0 ? hv[t.hoist] : 1.0) * (t.mom >= 0 ? mv[t.mom] : t.acc >= 0 ? av[t.acc] : 1.0);
        out[i * no + k] = v;
      }
    }
  };
  if (fp32) {
    ModelDef<float> d(kind, nullptr);
    run(d);
  } else {
    ModelDef<double> d(kind, nullptr);
    run(d);
  }
  MFH_CATCH
}

}  // extern "C"

namespace {
// Autotuner results persist in <kernel cache>/autotune.txt (MOREFIT_KCACHE, else kcache/ next
// to the library directory), one line per workload key, so later processes on the same
// machine reuse the same launch shape instead of timing again.
std::string autotune_path() {
  const char* env = getenv("MOREFIT_KCACHE");
  if (env && *env) return std::string(env) + "/autotune.txt";
  Dl_info info;
  std::string dir = ".";
  if (dladdr((void*)&mfh_last_error, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    size_t k = p.rfind('/');
    if (k != std::string::npos) dir = p.substr(0, k);
  }
  return dir + "/../kcache/autotune.txt";
}
uint64_t fnv64(const std::string& s) {
  uint64_t h = 1469598103934665603ULL;
  for (unsigned char ch : s) {
    h ^= ch;
    h *= 1099511628211ULL;
  }
  return h;
}
bool autotune_lookup(const std::string& key, std::vector<double>& us, int* best) {
  std::ifstream f(autotune_path());
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string k;
    int n = 0;
    if (!(is >> k >> best[0] >> best[1] >> best[2] >> n) || k != key || n != (int)us.size()) continue;
    for (int i = 0; i < n; i++) is >> us[i];
    if (is) return true;
  }
  return false;
}
void autotune_store(const std::string& key, const std::vector<double>& us, const int* best) {
  std::string path = autotune_path();
  mkdir(path.substr(0, path.rfind('/')).c_str(), 0755);
  std::ofstream f(path, std::ios::app);
  if (!f) return;
  f << key << " " << best[0] << " " << best[1] << " " << best[2] << " " << us.size();
  for (double u : us) f << " " << u;
  f << "\n";
}
}  // namespace

extern "C" {

int mfh_autotune(mfh_context* ctx, int kind, int fp32, const long long* sizes, int nseg, int eval_kind, const int* cand, int ncand,
                 int iters, int* best, double* us_per_launch) {
  MFH_TRY
  if (!ctx || !sizes || nseg <= 0 || eval_kind < 1 || eval_kind > 3 || iters <= 0 || !best)
    throw std::invalid_argument("mfh_autotune: bad argument");
  std::vector<LaunchShape> cs;
  if (cand && ncand > 0)
    for (int i = 0; i < ncand; i++) cs.push_back({cand[3 * i], cand[3 * i + 1], cand[3 * i + 2]});
  else {
    long long tot = 0;
    for (int i = 0; i < nseg; i++) tot += sizes[i];
    cs = launch_candidates(tot);
  }
  for (auto& s : cs)
    if (s.block < 64 || s.block > 1024 || s.block % 64 || s.events_per_thread < 1 || s.unroll < 1)
      throw std::invalid_argument("mfh_autotune: bad candidate launch shape");
  std::vector<long long> sz(sizes, sizes + nseg);
  LaunchShape win;
  std::vector<double> us(cs.size(), NAN);
  auto run = [&](auto& d) {
    typedef typename std::remove_reference<decltype(d)>::type::kernel_type K;
    // key: device, workload, candidates and the generated module (so a code change re-tunes)
    std::ostringstream k;
    k << ctx->be->props().name << "|" << ctx->be->props().compute_units << "|" << kind << "|" << fp32 << "|" << eval_kind << "|"
      << iters << "|" << nseg;
    long long tot = 0, mx = 0;
    for (auto v : sz) {
      tot += v;
      mx = std::max(mx, v);
    }
    k << "|" << tot << "|" << mx;
    for (auto& s : cs) k << "|" << s.block << "," << s.events_per_thread << "," << s.unroll;
    k << "|" << HipModel<K>::make_source(d.pdf, d.params, ctx->co, eval_kind == EVAL_HESS).device_source();
    char key[32];
    std::snprintf(key, sizeof(key), "%016llx", (unsigned long long)fnv64(k.str()));
    int b3[3];
    if (autotune_lookup(key, us, b3)) {
      win = {b3[0], b3[1], b3[2]};
      ctx->co.hip_block_size = win.block;
      ctx->co.hip_events_per_thread = win.events_per_thread;
      ctx->co.hip_unroll = win.unroll;
      return;
    }
    std::vector<double> tg;
    for (auto* p : d.params)
      if (!p->is_constant()) tg.push_back(p->get_value());
    us = autotune_launch_shape(ctx->be.get(), d.pdf, d.params, sz, eval_kind, tg, cs, iters, &win);
    b3[0] = win.block;
    b3[1] = win.events_per_thread;
    b3[2] = win.unroll;
    autotune_store(key, us, b3);
  };
  if (fp32) {
    ModelDef<float> d(kind, nullptr);
    run(d);
  } else {
    ModelDef<double> d(kind, nullptr);
    run(d);
  }
  best[0] = win.block;
  best[1] = win.events_per_thread;
  best[2] = win.unroll;
  if (us_per_launch)
    for (size_t i = 0; i < us.size(); i++) us_per_launch[i] = us[i];
  MFH_CATCH
}

int mfh_autotune_candidates(long long total_events, int* cand, int cap, int* count) {
  MFH_TRY
  if (!count || (cap > 0 && !cand)) throw std::invalid_argument("mfh_autotune_candidates: null argument");
  auto cs = launch_candidates(total_events);
  *count = (int)cs.size();
  for (int i = 0; i < cap && i < (int)cs.size(); i++) {
    cand[3 * i] = cs[i].block;
    cand[3 * i + 1] = cs[i].events_per_thread;
    cand[3 * i + 2] = cs[i].unroll;
  }
  MFH_CATCH
}

int mfh_stream_states(unsigned long long seed, long long first, int count, unsigned int* out) {
  MFH_TRY
  if (first < 0 || count < 0 || (count > 0 && !out)) throw std::invalid_argument("bad stream range");
  static const std::vector<uint32_t> powers = xoshiro_jump_powers(40);
  Xoshiro128pp r;
  r.setSeed((uint64_t)seed);
  uint32_t base[4];
  for (int w = 0; w < 4; w++) base[w] = r.getSeed(w);
  for (int i = 0; i < count; i++) xoshiro_stream_state(powers, base, (unsigned long long)(first + i), out + 4 * (size_t)i);
  MFH_CATCH
}

int mfh_batch_create(mfh_model* m, const long long* sizes, int nseg, mfh_batch** out) {
  MFH_TRY
  std::vector<long long> s(sizes, sizes + nseg);
  std::unique_ptr<mfh_batch> b(new mfh_batch());
  b->model = m;
  if (m->fp32) b->bf.reset(new HipBatch<float>(m->mf.get(), s));
  else b->bd.reset(new HipBatch<double>(m->md.get(), s));
  *out = b.release();
  MFH_CATCH
}

int mfh_batch_destroy(mfh_batch* b) {
  delete b;
  return 0;
}

int mfh_batch_upload(mfh_batch* b, int seg, const double* cols, long long col_stride) {
  MFH_TRY
  b->with([&](auto& bb) {
    if (seg < 0 || seg >= bb.S) throw std::out_of_range("segment index");
    bb.upload(seg, cols, col_stride);
    return 0;
  });
  MFH_CATCH
}

int mfh_batch_download(mfh_batch* b, int seg, double* cols) {
  MFH_TRY
  b->with([&](auto& bb) {
    if (seg < 0 || seg >= bb.S) throw std::out_of_range("segment index");
    bb.download(seg, cols);
    return 0;
  });
  MFH_CATCH
}

int mfh_batch_generate(mfh_batch* b, const double* theta_gen, const unsigned long long* seeds, int events_per_stream,
                       unsigned long long* proposals) {
  MFH_TRY
  b->with([&](auto& bb) {
    std::vector<double> th(theta_gen, theta_gen + (size_t)bb.S * bb.m->ntheta);
    std::vector<uint64_t> sd(seeds, seeds + bb.S);
    auto p = bb.generate(th, sd, events_per_stream > 0 ? events_per_stream : 64);
    if (proposals) *proposals = p;
    return 0;
  });
  MFH_CATCH
}

int mfh_batch_eval(mfh_batch* b, int kind, const double* theta, int use_graph, double* out) {
  MFH_TRY
  b->with([&](auto& bb) {
    if (kind < 1 || kind > 3 || bb.m->nout[kind] == 0) throw std::invalid_argument("bad evaluation kind");
    for (int s = 0; s < bb.S; s++) std::memcpy(bb.theta(s), theta + (size_t)s * bb.m->ntheta, bb.m->ntheta * sizeof(double));
    bb.run(kind, true, use_graph != 0);
    for (int s = 0; s < bb.S; s++) std::memcpy(out + (size_t)s * bb.m->nout[kind], bb.result(s, kind), bb.m->nout[kind] * sizeof(double));
    return 0;
  });
  MFH_CATCH
}

int mfh_batch_time_kernel(mfh_batch* b, int kind, int iters, int main_only, double* ms_per_launch) {
  MFH_TRY
  b->with([&](auto& bb) {
    if (kind < 1 || kind > 3 || bb.m->nout[kind] == 0) throw std::invalid_argument("bad evaluation kind");
    bb.set_modes(kind, true);
    bb.prep(kind);
    bb.launch_kernel(kind, main_only != 0);  // warm-up
    hip_check(mf_sync(bb.dev), "sync");
    hip_check(mf_timer_begin(bb.dev), "timer");
    for (int i = 0; i < iters; i++) bb.launch_kernel(kind, main_only != 0);
    float ms = 0;
    hip_check(mf_timer_end(bb.dev, &ms), "timer");
    *ms_per_launch = (double)ms / iters;
    return 0;
  });
  MFH_CATCH
}

int mfh_batch_time_eval(mfh_batch* b, int kind, int iters, int use_graph, double* us_per_eval) {
  MFH_TRY
  b->with([&](auto& bb) {
    bb.run(kind, true, use_graph != 0);  // warm-up (and graph capture)
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < iters; i++) bb.run(kind, true, use_graph != 0);
    *us_per_eval = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / iters;
    return 0;
  });
  MFH_CATCH
}

int mfh_batch_time_graph(mfh_batch* b, int kind, int iters, double* us_per_eval) {
  MFH_TRY
  b->with([&](auto& bb) {
    bb.run(kind, true, true);  // capture
    hip_check(mf_timer_begin(bb.dev), "timer");
    for (int i = 0; i < iters; i++) hip_check(mf_graph_launch(bb.graphs[kind]), "mf_graph_launch");
    float ms = 0.0f;
    hip_check(mf_timer_end(bb.dev, &ms), "timer");
    *us_per_eval = 1e3 * (double)ms / iters;
    return 0;
  });
  MFH_CATCH
}

long long mfh_batch_bytes(mfh_batch* b) {
  return b->with([&](auto& bb) { return (long long)bb.data_bytes(); });
}

int mfh_batch_check(mfh_batch* b, const int* segs, int nseg, const double* theta, long long* out) {
  MFH_TRY
  if (!b || nseg < 0 || (nseg > 0 && (!segs || !theta || !out))) throw std::invalid_argument("mfh_batch_check: bad arguments");
  b->with([&](auto& bb) {
    std::vector<int> list(segs, segs + nseg);
    std::vector<double> th(theta, theta + (size_t)nseg * bb.m->ntheta);
    auto d = bb.diagnose(list, th);
    for (int i = 0; i < nseg; i++) {
      out[3 * i] = d[i].nonfinite;
      out[3 * i + 1] = d[i].first;
      out[3 * i + 2] = d[i].outside;
    }
  });
  MFH_CATCH
}

int mfh_batch_fit(mfh_batch* b, const int* segs, int nfit, const double* start, int use_graph, double* values, double* errors,
                  mfh_fit_out* outs, mfh_batch_stats* stats) {
  MFH_TRY
  b->with([&](auto& bb) {
    std::vector<int> list;
    if (nfit < 0 || (nfit > 0 && !start)) throw std::invalid_argument("mfh_batch_fit: bad fit count or no start values");
    if (segs) list.assign(segs, segs + nfit);
    else
      for (int s = 0; s < bb.S; s++) list.push_back(s);
    const int P = bb.m->ntheta;
    std::vector<double> st(start, start + (size_t)list.size() * P);
    typedef typename std::remove_reference<decltype(bb)>::type BT;
    typedef typename std::remove_pointer<decltype(bb.m)>::type MT;
    (void)sizeof(MT);
    BatchFitter<typename std::conditional<std::is_same<BT, HipBatch<float>>::value, float, double>::type> bf(bb);
    bf.use_graph = use_graph != 0;
    auto res = bf.fit(list, st);
    for (size_t i = 0; i < res.size(); i++) {
      for (int k = 0; k < P; k++) {
        values[i * P + k] = res[i].values.size() > (size_t)k ? res[i].values[k] : NAN;
        errors[i * P + k] = res[i].errors.size() > (size_t)k ? res[i].errors[k] : NAN;
      }
      if (outs) {
        outs[i].fval = res[i].fval;
        outs[i].edm = res[i].edm;
        outs[i].status = res[i].status;
        outs[i].ncalls = res[i].ncalls;
        outs[i].iterations = res[i].iterations;
        outs[i].hesse_ok = res[i].hesse_ok;
        outs[i].thread_moved = res[i].thread_moved;
        outs[i].nonfinite_evals = res[i].nonfinite_evals;
        outs[i].nonfinite_events = res[i].nonfinite_evals > 0 ? res[i].nonfinite_events : 0;
        outs[i].first_nonfinite = res[i].first_nonfinite_event;
        outs[i].out_of_range = res[i].out_of_range_events;
      }
    }
    if (stats) {
      stats->ticks = bf.stats.ticks;
      stats->launches_lh = bf.stats.launches[1];
      stats->launches_grad = bf.stats.launches[2];
      stats->launches_hess = bf.stats.launches[3];
      stats->wall_ms = bf.stats.wall_ms;
      stats->gpu_wait_ms = bf.stats.gpu_wait_ms;
      stats->host_ms = bf.stats.host_ms;
      stats->host_threads = bf.stats.host_threads;
      stats->setup_ms = bf.stats.setup_ms;
      stats->step_ms = bf.stats.step_ms;
    }
    return 0;
  });
  MFH_CATCH
}

int mfh_serve_run(mfh_model* const* models, int nmodels, const double* theta_gen, const double* start, const int* cls,
                  const long long* n, const unsigned long long* seeds, const double* arrival_s, int nreq, double budget_bytes,
                  int slots, double* timings, double* values, double* errors, int* status, double* stats, const double* request_start) {
  MFH_TRY
  if (nmodels <= 0) throw std::invalid_argument("no models");
  const bool fp32 = models[0]->fp32;
  int maxp = 0;
  for (int i = 0; i < nmodels; i++) {
    if (models[i]->fp32 != fp32) throw std::invalid_argument("all service models must share one precision");
    maxp = std::max(maxp, models[i]->with([](auto& hm) { return hm.ntheta; }));
  }
  std::vector<ServeRequest> reqs(nreq);
  for (int i = 0; i < nreq; i++) {
    reqs[i].id = i;
    reqs[i].cls = cls[i];
    reqs[i].n = n[i];
    reqs[i].seed = seeds[i];
    reqs[i].arrival_s = arrival_s[i];
    if (cls[i] < 0 || cls[i] >= nmodels) throw std::out_of_range("request class");
    if (request_start) {
      const int p = models[cls[i]]->with([](auto& hm) { return hm.ntheta; });
      reqs[i].start.assign(request_start + (size_t)i * maxp, request_start + (size_t)i * maxp + p);
    }
  }
  std::stable_sort(reqs.begin(), reqs.end(), [](const ServeRequest& a, const ServeRequest& b) { return a.arrival_s < b.arrival_s; });
  auto serve = [&](auto tag) {
    typedef decltype(tag) K;
    FitServer<K> srv((size_t)budget_bytes, slots);
    for (int c = 0; c < nmodels; c++) {
      HipModel<K>* hm;
      if constexpr (std::is_same<K, float>::value) hm = models[c]->mf.get();
      else hm = models[c]->md.get();
      std::vector<double> tg(theta_gen + (size_t)c * maxp, theta_gen + (size_t)c * maxp + hm->ntheta);
      std::vector<double> st(start + (size_t)c * maxp, start + (size_t)c * maxp + hm->ntheta);
      srv.add_class(hm, tg, st);
    }
    auto res = srv.run(reqs);
    for (auto& r : res) {
      const long long i = r.id;
      timings[3 * i + 0] = r.arrival_s;
      timings[3 * i + 1] = r.admit_s;
      timings[3 * i + 2] = r.done_s;
      for (int k = 0; k < maxp; k++) {
        values[i * maxp + k] = k < (int)r.fit.values.size() ? r.fit.values[k] : NAN;
        errors[i * maxp + k] = k < (int)r.fit.errors.size() ? r.fit.errors[k] : NAN;
      }
      status[i] = r.fit.status;
    }
    stats[0] = srv.stats.ticks;
    stats[1] = srv.stats.launches;
    stats[2] = srv.stats.peak_live;
    stats[3] = srv.stats.rejected;
    stats[4] = srv.stats.wall_s;
    stats[5] = srv.stats.peak_bytes;
    stats[6] = srv.stats.generate_ms;
    stats[7] = srv.stats.arena_bytes;
    return 0;
  };
  if (fp32) serve(float());
  else serve(double());
  MFH_CATCH
}

}  // extern "C"

namespace {
template <typename F>
void fill_failure(mfh_fit_out& out, const F& f) {
  out.thread_moved = 0;
  out.nonfinite_evals = f.nonfinite_evals_;
  out.nonfinite_events = f.nonfinite_evals_ > 0 ? f.nonfinite_events_ : 0;
  out.first_nonfinite = f.first_nonfinite_event_;
  out.out_of_range = f.out_of_range_events_;
}

template <typename K, typename Backend, typename Block>
void run_fit(Backend* be, int kind, const double* data, long long n, const double* start, double* values, double* errors, mfh_fit_out* out,
             double* ms, const double* weights = nullptr, bool optimize_dimensions = false, int flags = MFH_FIT_DEFAULT) {
  ModelDef<K> def(kind, nullptr);
  // weights ride along as an extra column named "w" that the fitter multiplies log p by
  dimension<double> wdim("w", -1e300, 1e300);
  std::vector<dimension<double>*> dims = def.dimv;
  if (weights) dims.push_back(&wdim);
  EventVector<double, double> ev(dims, (unsigned long)n);
  for (size_t d = 0; d < def.dimv.size(); d++)
    for (long long i = 0; i < n; i++) ev(i, (unsigned)d) = data[d * n + i];
  if (weights) {
    for (long long i = 0; i < n; i++) ev(i, (unsigned)def.dimv.size()) = weights[i];
    ev.set_event_weight("w");
  }
  for (size_t k = 0; k < def.params.size(); k++) {
    def.params[k]->set_value(start[k]);
    def.params[k]->set_start_value(start[k]);
  }
  fitter_options fo;
  fo.analytic_fisher = weights != nullptr;
  fo.optimize_dimensions = optimize_dimensions;
  fo.analytic_gradient = (flags & MFH_FIT_ANALYTIC_GRADIENT) != 0;
  fo.analytic_hessian = (flags & MFH_FIT_ANALYTIC_HESSIAN) != 0;
  fo.kahan_on_accelerator = (flags & MFH_FIT_KAHAN_ON_ACCELERATOR) != 0;
  fo.postrun_hesse = (flags & MFH_FIT_POSTRUN_HESSE) != 0;
  fitter<K, double, Backend, Block> f(&fo, be);
  f.fit(def.pdf, def.params, &ev);
  for (size_t k = 0; k < def.params.size(); k++) {
    values[k] = def.params[k]->get_value();
    errors[k] = def.params[k]->get_error();
  }
  if (out) {
    out->fval = f.fval_;
    out->edm = f.edm_;
    out->status = f.status_;
    out->ncalls = f.ncalls_;
    out->iterations = 0;
    out->hesse_ok = !f.cov_.empty();
    fill_failure(*out, f);
  }
  if (ms) *ms = f.fit_ms_;
}
}  // namespace

extern "C" {

int mfh_fit_plain(mfh_context* ctx, int kind, int fp32, const double* data, long long n, const double* start, double* values,
                  double* errors, mfh_fit_out* out, double* ms) {
  MFH_TRY
  // the context's event_columns option selects the fitter's optimize_dimensions
  const bool od = ctx->be->opts_->hip_event_columns;
  if (fp32) run_fit<float, HIPBackend, HIPBlock<float, double>>(ctx->be.get(), kind, data, n, start, values, errors, out, ms, nullptr, od);
  else run_fit<double, HIPBackend, HIPBlock<double, double>>(ctx->be.get(), kind, data, n, start, values, errors, out, ms, nullptr, od);
  MFH_CATCH
}

int mfh_fit_cpu(int kind, int fp32, const double* data, long long n, const double* start, int nthreads, int optimize_dimensions,
                double* values, double* errors, mfh_fit_out* out, double* ms) {
  MFH_TRY
  compute_options co;
  co.llvm_nthreads = nthreads > 0 ? nthreads : 1;
  CPUBackend be(&co);
  const bool od = optimize_dimensions != 0;
  if (fp32) run_fit<float, CPUBackend, CPUBlock<float, double>>(&be, kind, data, n, start, values, errors, out, ms, nullptr, od);
  else run_fit<double, CPUBackend, CPUBlock<double, double>>(&be, kind, data, n, start, values, errors, out, ms, nullptr, od);
  MFH_CATCH
}

int mfh_fit_weighted(mfh_context* ctx, int kind, int fp32, const double* data, const double* weights, long long n, const double* start,
                     int nthreads, double* values, double* errors, mfh_fit_out* out, double* ms) {
  MFH_TRY
  if (!weights) throw std::invalid_argument("mfh_fit_weighted: no weights");
  if (ctx) {
    if (fp32) run_fit<float, HIPBackend, HIPBlock<float, double>>(ctx->be.get(), kind, data, n, start, values, errors, out, ms, weights);
    else run_fit<double, HIPBackend, HIPBlock<double, double>>(ctx->be.get(), kind, data, n, start, values, errors, out, ms, weights);
  } else {
    compute_options co;
    co.llvm_nthreads = nthreads > 0 ? nthreads : 1;
    CPUBackend be(&co);
    if (fp32) run_fit<float, CPUBackend, CPUBlock<float, double>>(&be, kind, data, n, start, values, errors, out, ms, weights);
    else run_fit<double, CPUBackend, CPUBlock<double, double>>(&be, kind, data, n, start, values, errors, out, ms, weights);
  }
  MFH_CATCH
}

int mfh_fit_options(mfh_context* ctx, int kind, int fp32, const double* data, long long n, const double* start, int nthreads, int flags,
                    double* values, double* errors, mfh_fit_out* out, double* ms) {
  MFH_TRY
  if (flags & ~MFH_FIT_ALL) throw std::invalid_argument("mfh_fit_options: unknown flag bits");
  if (ctx) {
    if (fp32) run_fit<float, HIPBackend, HIPBlock<float, double>>(ctx->be.get(), kind, data, n, start, values, errors, out, ms, nullptr, false, flags);
    else run_fit<double, HIPBackend, HIPBlock<double, double>>(ctx->be.get(), kind, data, n, start, values, errors, out, ms, nullptr, false, flags);
  } else {
    compute_options co;
    co.llvm_nthreads = nthreads > 0 ? nthreads : 1;
    CPUBackend be(&co);
    if (fp32) run_fit<float, CPUBackend, CPUBlock<float, double>>(&be, kind, data, n, start, values, errors, out, ms, nullptr, false, flags);
    else run_fit<double, CPUBackend, CPUBlock<double, double>>(&be, kind, data, n, start, values, errors, out, ms, nullptr, false, flags);
  }
  MFH_CATCH
}

}  // extern "C"

// A plain fitter session: one density, one backend, kernels built at the first fit and reused
// by later ones (fitter::fit with refit, as the reference's toy study reuses them,
// toystudy.hh:76-77).
struct mfh_plain {
  virtual ~mfh_plain() = default;
  virtual void fit(const double* data, long long n, const double* start, double* values, double* errors, mfh_fit_out* out, double* ms) = 0;
  virtual void eval(const double* data, long long n, const double* theta, int iters, double* value, double* grad, double* seconds) = 0;
};

namespace {
template <typename K, typename Backend, typename Block>
struct PlainSession : mfh_plain {
  compute_options co;
  std::unique_ptr<Backend> own;
  Backend* be;
  ModelDef<K> def;
  fitter_options fo;
  fitter<K, double, Backend, Block> f;
  bool built = false;
  // shared: the context's backend; null: a backend of this session's own (the CPU backend with
  // nthreads threads and, when given, the host compiler flags cpu_flags)
  PlainSession(Backend* shared, int kind, int flags, int nthreads, const char* cpu_flags)
      : co(cpu_options(nthreads, cpu_flags)),
        own(shared ? nullptr : new Backend(&co)),
        be(shared ? shared : own.get()),
        def(kind, nullptr),
        fo(options(flags)),
        f(&fo, be) {}
  static compute_options cpu_options(int nthreads, const char* cpu_flags) {
    compute_options c;
    c.llvm_nthreads = nthreads > 0 ? nthreads : 1;
    if (cpu_flags && *cpu_flags) c.cpu_flags = cpu_flags;
    return c;
  }
  static fitter_options options(int flags) {
    fitter_options o;
    o.analytic_gradient = (flags & MFH_FIT_ANALYTIC_GRADIENT) != 0;
    o.analytic_hessian = (flags & MFH_FIT_ANALYTIC_HESSIAN) != 0;
    o.kahan_on_accelerator = (flags & MFH_FIT_KAHAN_ON_ACCELERATOR) != 0;
    o.postrun_hesse = (flags & MFH_FIT_POSTRUN_HESSE) != 0;
    return o;
  }
  EventVector<double, double> events(const double* data, long long n) {
    if (!data || n <= 0) throw std::invalid_argument("plain session: no events");
    EventVector<double, double> ev(def.dimv, (unsigned long)n);
    for (size_t d = 0; d < def.dimv.size(); d++)
      for (long long i = 0; i < n; i++) ev(i, (unsigned)d) = data[d * n + i];
    return ev;
  }
  void fit(const double* data, long long n, const double* start, double* values, double* errors, mfh_fit_out* out, double* ms) override {
    auto ev = events(data, n);
    for (size_t k = 0; k < def.params.size(); k++) {
      def.params[k]->set_value(start[k]);
      def.params[k]->set_start_value(start[k]);
    }
    f.fit(def.pdf, def.params, &ev, built);
    built = true;
    for (size_t k = 0; k < def.params.size(); k++) {
      values[k] = def.params[k]->get_value();
      errors[k] = def.params[k]->get_error();
    }
    if (out) {
      out->fval = f.fval_;
      out->edm = f.edm_;
      out->status = f.status_;
      out->ncalls = f.ncalls_;
      out->iterations = 0;
      out->hesse_ok = !f.cov_.empty();
      fill_failure(*out, f);
    }
    if (ms) *ms = f.fit_ms_;
  }
  void eval(const double* data, long long n, const double* theta, int iters, double* value, double* grad, double* seconds) override {
    if (!fo.analytic_gradient) throw std::invalid_argument("plain session: evaluation timing needs the analytic gradient");
    auto ev = events(data, n);
    f.prepare(def.pdf, def.params, &ev, built);
    built = true;
    for (size_t k = 0; k < def.params.size(); k++) def.params[k]->set_value(theta[k]);
    const size_t N = def.params.size();
    std::vector<double> g(N);
    double v = f.likelihood_and_gradient((int)N, g.data());
    auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it < iters; it++) v = f.likelihood_and_gradient((int)N, g.data());
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (value) *value = v;
    if (grad) std::memcpy(grad, g.data(), N * sizeof(double));
  }
};
}  // namespace

extern "C" {

int mfh_plain_create(mfh_context* ctx, int kind, int fp32, int nthreads, int flags, const char* cpu_flags, mfh_plain** out) {
  MFH_TRY
  if (!out) throw std::invalid_argument("mfh_plain_create: null output");
  if (flags & ~MFH_FIT_ALL) throw std::invalid_argument("mfh_plain_create: unknown flag bits");
  if (ctx) {
    if (fp32) *out = new PlainSession<float, HIPBackend, HIPBlock<float, double>>(ctx->be.get(), kind, flags, 0, nullptr);
    else *out = new PlainSession<double, HIPBackend, HIPBlock<double, double>>(ctx->be.get(), kind, flags, 0, nullptr);
  } else {
    if (fp32) *out = new PlainSession<float, CPUBackend, CPUBlock<float, double>>(nullptr, kind, flags, nthreads, cpu_flags);
    else *out = new PlainSession<double, CPUBackend, CPUBlock<double, double>>(nullptr, kind, flags, nthreads, cpu_flags);
  }
  MFH_CATCH
}

int mfh_plain_fit(mfh_plain* p, const double* data, long long n, const double* start, double* values, double* errors, mfh_fit_out* out,
                  double* ms) {
  MFH_TRY
  if (!p || !start || !values || !errors) throw std::invalid_argument("mfh_plain_fit: null argument");
  p->fit(data, n, start, values, errors, out, ms);
  MFH_CATCH
}

int mfh_plain_eval(mfh_plain* p, const double* data, long long n, const double* theta, int iters, double* value, double* grad,
                   double* seconds) {
  MFH_TRY
  if (!p || !theta || iters < 0) throw std::invalid_argument("mfh_plain_eval: bad argument");
  p->eval(data, n, theta, iters, value, grad, seconds);
  MFH_CATCH
}

int mfh_plain_destroy(mfh_plain* p) {
  delete p;
  return 0;
}

int mfh_eval_cpu(int kind, const double* data, long long n, const double* theta, int nthreads, double* value, double* grad, double* hess) {
  MFH_TRY
  compute_options co;
  co.llvm_nthreads = nthreads > 0 ? nthreads : 1;
  CPUBackend be(&co);
  ModelDef<double> def(kind, nullptr);
  EventVector<double, double> ev(def.dimv, (unsigned long)n);
  for (size_t d = 0; d < def.dimv.size(); d++)
    for (long long i = 0; i < n; i++) ev(i, (unsigned)d) = data[d * n + i];
  fitter_options fo;
  fitter<double, double, CPUBackend, CPUBlock<double, double>> f(&fo, &be);
  f.prepare(def.pdf, def.params, &ev, false);
  for (size_t k = 0; k < def.params.size(); k++) def.params[k]->set_value(theta[k]);
  const size_t N = def.params.size();
  std::vector<double> g(N), h(N * N);
  double v = f.likelihood_and_gradient((int)N, g.data());
  f.hessian((int)N, h.data());
  if (value) *value = v;
  if (grad) std::memcpy(grad, g.data(), N * sizeof(double));
  if (hess) std::memcpy(hess, h.data(), N * N * sizeof(double));
  MFH_CATCH
}

}  // extern "C"

namespace {
template <typename K, typename Backend, typename Block>
void run_plot(Backend* be, int kind, const double* theta, const double* data, const double* weights, long long n, int var, int nbins,
              int multiplier, double range_min, double range_max, int plot_type, int flags, const char* filename, int replots,
              double* pdf_integrals, double* counts, double* errors, double* pulls, double* ms) {
  ModelDef<K> def(kind, nullptr);
  if (var < 0 || var >= (int)def.dimv.size()) throw std::invalid_argument("mfh_plot: observable index out of range");
  if (nbins <= 0 || multiplier <= 0) throw std::invalid_argument("mfh_plot: nbins and multiplier must be positive");
  if (n < 0 || n > (long long)std::numeric_limits<int>::max()) throw std::invalid_argument("mfh_plot: event count out of range");
  dimension<double> wdim("w", -1e300, 1e300);
  std::vector<dimension<double>*> dims = def.dimv;
  if (weights) dims.push_back(&wdim);
  EventVector<K, double> ev(dims, (unsigned long)n);
  for (size_t d = 0; d < def.dimv.size(); d++)
    for (long long i = 0; i < n; i++) ev(i, (unsigned)d) = (K)data[d * n + i];
  if (weights) {
    for (long long i = 0; i < n; i++) ev(i, (unsigned)def.dimv.size()) = (K)weights[i];
    ev.set_event_weight("w");
  }
  for (size_t k = 0; k < def.params.size(); k++) def.params[k]->set_value(theta[k]);
  plotter_options po;
  po.plotter = plot_type == MFH_PLOT_MATPLOTLIB ? plotter_options::MatPlotLib : plotter_options::Root;
  po.pdf_bin_multiplier = multiplier;
  po.plot_pulls = (flags & MFH_PLOT_PULLS) != 0;
  po.print_level = 0;
  plotter<K, double, Backend, Block> pl(&po, be);
  const std::string var_name = def.dimv[(size_t)var]->get_name();
  pl.plot(def.pdf, def.params, &ev, var_name, filename ? filename : "", "", (unsigned)nbins, range_min, range_max, false);
  for (int r = 0; r < replots; r++) pl.plot(def.pdf, def.params, &ev, var_name, "", "", (unsigned)nbins, range_min, range_max, true);
  const auto& P = pl.last;
  if (pdf_integrals) std::copy(P.pdf_integrals.begin(), P.pdf_integrals.end(), pdf_integrals);
  if (counts) std::copy(P.counts.begin(), P.counts.end(), counts);
  if (errors) std::copy(P.errors.begin(), P.errors.end(), errors);
  if (pulls) std::copy(P.pulls.begin(), P.pulls.end(), pulls);
  if (ms) *ms = pl.last_ms;
}
}  // namespace

extern "C" {

int mfh_plot(mfh_context* ctx, int kind, int fp32, int nthreads, const double* theta, const double* data, const double* weights,
             long long n, int var, int nbins, int multiplier, double range_min, double range_max, int plot_type, int flags,
             const char* filename, int replots, double* pdf_integrals, double* counts, double* errors, double* pulls, double* ms) {
  MFH_TRY
  if (!theta || (n > 0 && !data)) throw std::invalid_argument("mfh_plot: null theta or data");
  if (ctx) {
    if (fp32)
      run_plot<float, HIPBackend, HIPBlock<float, double>>(ctx->be.get(), kind, theta, data, weights, n, var, nbins, multiplier, range_min,
                                                            range_max, plot_type, flags, filename, replots, pdf_integrals, counts, errors,
                                                            pulls, ms);
    else
      run_plot<double, HIPBackend, HIPBlock<double, double>>(ctx->be.get(), kind, theta, data, weights, n, var, nbins, multiplier,
                                                              range_min, range_max, plot_type, flags, filename, replots, pdf_integrals,
                                                              counts, errors, pulls, ms);
  } else {
    compute_options co;
    co.llvm_nthreads = nthreads > 0 ? nthreads : 1;
    CPUBackend be(&co);
    if (fp32)
      run_plot<float, CPUBackend, CPUBlock<float, double>>(&be, kind, theta, data, weights, n, var, nbins, multiplier, range_min, range_max,
                                                            plot_type, flags, filename, replots, pdf_integrals, counts, errors, pulls, ms);
    else
      run_plot<double, CPUBackend, CPUBlock<double, double>>(&be, kind, theta, data, weights, n, var, nbins, multiplier, range_min,
                                                              range_max, plot_type, flags, filename, replots, pdf_integrals, counts, errors,
                                                              pulls, ms);
  }
  MFH_CATCH
}

}  // extern "C"

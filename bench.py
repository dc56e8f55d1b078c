"""Benchmark of the MoreFit likelihood hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): the mass-fit density
f*Gauss(mu, sigma) + (1-f)*Exp(lambda) on m in [5.0, 5.6], N = 10,000,000 events generated
on the device by seeded accept-reject, one dataset. A step is one -2lnL + gradient
evaluation over all N events at the fit's start values: FP64 hoisted terms computed on the
host and passed with theta in the kernel arguments, the per-chunk kernel and the
deterministic final reduction launched directly, completion polled, totals read from mapped
host memory. Before the W warm-up steps the same evaluation runs untimed for SETTLE_S
seconds, so that the K timed steps see the device's sustained clock. `value` is evaluations
per second of the FP32-compensated path; the FP64 path, accuracy against the FP64 oracle,
the kernel's HBM roofline fraction, the angular configuration, the batched toy study (4096
toys, Poisson(5000) events) and the reference's toy protocol (100 toys x N = 1e3..1e6)
against the plain per-fit FP64 path (kernels reused from fit to fit, and built anew per
fit), the fit service, and the CPU baselines (the host CPU backend with the reference LLVM
backend's semantics, on 1 thread and all threads; a config-1 fit; CPU toys) are reported
beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-toys] [--no-cpu]
       python bench.py --profile   (counter workload: the headline FP32 gradient and value
                                    kernels at N = 10^7 and the copy kernel over 40 MB, 20
                                    launches each, no JSON; for rocprofv3 --pmc passes)
       python bench.py --profile others   (the same for the other gradient kernels: FP64 mass at
                                    N = 10^7, angular FP32 and FP64 at N = 10^6, the 4096-toy
                                    FP32 launch; each in its autotuned shape, 20 launches)
Progress goes to stderr; stdout carries exactly one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_EVENTS = 10_000_000
SETTLE_S = 0.25  # untimed evaluations before the warm-up steps


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def shape(model, tuned):
    """launch shape of a model's evaluation kernels: the autotuned (block, events per thread,
    requested unroll) of its context, the unrolls fitted per kernel and their VGPRs"""
    import morefit_amd as M
    o = M.options()
    return {"block": tuned[0], "events_per_thread": tuned[1], "requested_unroll": tuned[2], "accumulate": o.accumulate,
            "fast_math": bool(o.fast_math), "unroll": {"lh": model.unroll[M.EVAL_LH], "grad": model.unroll[M.EVAL_GRAD],
                                                       "hess": model.unroll[M.EVAL_HESS]},
            "vgprs": {"lh": model.vgprs[M.EVAL_LH], "grad": model.vgprs[M.EVAL_GRAD], "hess": model.vgprs[M.EVAL_HESS]},
            "segment_reduction": "in-kernel" if model.fused[M.EVAL_GRAD] else "final kernel"}


def tuned_context(M, kind, fp32, sizes, label, candidates=None, **options):
    """a context whose launch shape the autotuner picked for this workload (among `candidates`,
    default: the built-in set for its size)"""
    ctx = M.Context(device=0, **options)
    t0 = time.perf_counter()
    r = ctx.autotune(kind, fp32, sizes, eval_kind=M.EVAL_GRAD, candidates=candidates)
    log(f"autotune {label}: block {r['best'][0]}, {r['best'][1]} events/thread, unroll {r['best'][2]} "
        f"({min(u for _, u in r['table']):.2f} us; {len(r['table'])} shapes in {time.perf_counter() - t0:.1f} s)")
    return ctx, r


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "mf_eval_grad_fp32_mass_1e7_traffic.txt")


def measured_traffic(current_shape):
    """HBM bytes per launch of the FP32 mf_eval_grad at N = 10^7 from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the copy kernel's reading) and the launch
    shape it was taken at (the kernel reads each event once at every shape, so the figure
    carries over; the shape is reported beside it); (None, None) without a summary"""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None, None
    kv = dict(l.split(":", 1) for l in lines if l.startswith(("traffic_bytes_per_launch:", "launch_shape:")))
    if "traffic_bytes_per_launch" not in kv:
        return None, None
    return float(kv["traffic_bytes_per_launch"]), kv.get("launch_shape", "").strip()


# VALU issue rate of MI355X: 256 CUs x 4 SIMDs, one wave64 VALU instruction per 4 cycles at 2.4 GHz;
# an FP32 transcendental (v_exp/v_log/v_rcp/v_sqrt/v_sin/v_cos_f32) occupies its SIMD for 8 cycles,
# two issue slots (half rate on gfx950: morefit_amd/csrc/calib/issue_rate.hip,
# profiles/valu_issue_rate_calibration.txt)
VALU_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4
TRANS_EXTRA_SLOTS = 1.0


def issue_floor(kernel_ms):
    """the FP32 mf_eval_grad's instruction-issue floor at N = 10^7 from the committed counter
    summary (VALU wave instructions per launch, transcendentals counted 2x) and the measured
    kernel's fraction of it; None without the summary"""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    kv = dict(l.split(":", 1) for l in lines if l.startswith(("valu_wave_instructions_per_launch:",
                                                             "valu_trans_wave_instructions_per_launch:")))
    if len(kv) < 2:
        return None
    valu, trans = float(kv["valu_wave_instructions_per_launch"]), float(kv["valu_trans_wave_instructions_per_launch"])
    floor_ms = (valu + TRANS_EXTRA_SLOTS * trans) / VALU_WAVE_INSTR_PER_S * 1e3
    return {"floor_ms": floor_ms, "kernel_ms": kernel_ms, "frac": floor_ms / kernel_ms,
            "valu_wave_instructions": valu, "transcendental_wave_instructions": trans,
            "basis": "counter summary profiles/mf_eval_grad_fp32_mass_1e7_traffic.txt (SQ_INSTS_VALU, SQ_INSTS_VALU_TRANS_F32); "
                     "6.144e11 VALU wave instructions/s, transcendentals at half rate (calib/issue_rate.hip)"}


OTHERS_FILE = os.path.join(ROOT, "profiles", "mf_eval_grad_fp64_mass_angular_toys_counters.txt")


def others_counters(key):
    """a per-workload figure of the secondary gradient kernels from their committed counter
    summary (lines `key[workload]: value`), by workload label ({} without it)"""
    try:
        lines = open(OTHERS_FILE).read().splitlines()
    except OSError:
        return {}
    out = {}
    for l in lines:
        if l.startswith(key + "["):
            k, v = l.split("]:", 1)
            out[k[len(key) + 1:]] = float(v)
    return out


def others_traffic():
    """HBM bytes per launch of the secondary gradient kernels (FETCH_SIZE + WRITE_SIZE, scaled)"""
    return others_counters("traffic_bytes_per_launch")


def roofline_row(label, kernel_ms, algorithmic_bytes, traffic, floors=None):
    """the kernel against the HBM roofline; with its VALU issue floor from the counter summary
    (VALU wave instructions at the measured issue rates) the fraction of that floor beside it"""
    gbs = algorithmic_bytes / (kernel_ms * 1e-3) / 1e9
    row = {"workload": label, "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
           "traffic": traffic.get(label), "kernel_ms": kernel_ms, "algorithmic_bytes": algorithmic_bytes}
    floor = (floors or {}).get(label)
    if floor:
        row["valu_issue_floor_ms"] = floor
        row["valu_issue_floor_frac"] = floor / kernel_ms
    return row


def plain_context(M, O):
    """a default-shape context for the plain per-fit FP64 path, its modules compiled by one
    untimed fit per density (the plain path builds its kernels at the first fit), and one
    plain fitter session per density that reuses its kernels from fit to fit (refit, as the
    reference's toy study does, toystudy.hh:76-77), warmed by one fit"""
    cp = M.Context(device=0)
    cp.sessions = {}
    for kind in (M.MASS, M.KSTAR):
        d = O.generate(kind, M.TRUTH[kind], 2000, seed=1, nstreams=4)
        cp.fit_plain(kind, d, M.TRUTH[kind], fp32=False)
        cp.sessions[kind] = M.PlainFitter(kind, fp32=False, ctx=cp)
        cp.sessions[kind].fit(d, M.TRUTH[kind])
    return cp


def close_plain(cp):
    for sess in cp.sessions.values():
        sess.close()
    cp.close()


def serve_workload(ctx, cplain, M, O, nreq, rate, seed=11, plain_subset=12):
    """Fit requests: 70 % mass fits, 30 % angular fits, event counts log-uniform in [1e3, 2e6],
    data generated on the device per request (seeded), FP32 kernels, evaluations of concurrent
    fits in shared launches, admission bounded by a 64 GB HBM budget.  First all requests at
    once (the sustainable throughput), then Poisson arrivals at 50 %, 80 % and 95 % of it and at
    `rate` per second, each with p50/p99 latency from arrival to result.  Compared with
    submitting the same kinds of fits one at a time through the plain FP64 per-fit path (a
    subset of the requests, data generated by the oracle's generator)."""
    rs = np.random.default_rng(seed)
    cls = (rs.random(nreq) < 0.3).astype(int)
    n = np.exp(rs.uniform(np.log(1e3), np.log(2e6), nreq)).astype(np.int64)
    seeds = rs.integers(1, 2**62, nreq)
    gaps = rs.exponential(1.0, nreq)
    mass, kst = ctx.model(M.MASS, fp32=True), ctx.model(M.KSTAR, fp32=True)
    truth = [M.MASS_TRUTH, M.KSTAR_TRUTH]

    def run(arrival):
        r = M.serve([mass, kst], truth, truth, cls, n, seeds, arrival, budget_bytes=64 << 30, slots=512)
        lat = (r["done"] - r["arrival"]) * 1e3
        span = float(r["done"].max() - r["arrival"].min())
        return {"fits_per_s": nreq / span, "p50_latency_ms": float(np.percentile(lat, 50)),
                "p99_latency_ms": float(np.percentile(lat, 99)), "converged": int(np.sum(r["status"] == 0)),
                "rejected": r["rejected"], "ticks": r["ticks"], "launches": r["launches"], "peak_live": r["peak_live"],
                "generate_ms": r["generate_ms"], "span_s": span}

    sat = run(np.zeros(nreq))
    sustained = sat["fits_per_s"]
    rows = []
    for label, lam in [("50%", 0.5 * sustained), ("80%", 0.8 * sustained), ("95%", 0.95 * sustained), ("fixed", rate)]:
        row = run(np.cumsum(gaps / lam))
        row.update({"arrival_rate_per_s": lam, "load": label})
        rows.append(row)
    t_plain = t_scratch = 0.0
    for i in range(plain_subset):
        d = O.generate(int(cls[i]), truth[cls[i]], int(n[i]), seed=int(seeds[i]) % (2**63), nstreams=64)
        t0 = time.perf_counter()
        cplain.sessions[int(cls[i])].fit(d, truth[cls[i]])
        t1 = time.perf_counter()
        cplain.fit_plain(int(cls[i]), d, truth[cls[i]], fp32=False)
        t_plain += t1 - t0
        t_scratch += time.perf_counter() - t1
    # the subset's mix of sizes and classes differs from the whole stream's: scale by events
    ev_sub = float(np.sum(n[:plain_subset] * np.where(cls[:plain_subset] == 1, 3, 1)))
    ev_all = float(np.sum(n * np.where(cls == 1, 3, 1)))
    plain_rate = nreq / (t_plain * ev_all / ev_sub)
    scratch_rate = nreq / (t_scratch * ev_all / ev_sub)
    mass.close()
    kst.close()
    fixed = rows[-1]
    return {"requests": nreq, "mass_fraction": float(np.mean(cls == 0)), "events_log_uniform": [1000, 2000000],
            "hbm_budget_bytes": 64 << 30, "saturation": sat, "sustained_fits_per_s": sustained, "rates": rows,
            "p50_latency_ms": fixed["p50_latency_ms"], "p99_latency_ms": fixed["p99_latency_ms"],
            "fits_per_s": fixed["fits_per_s"], "arrival_rate_per_s": rate,
            "plain_fp64_fits_per_s": plain_rate, "plain_fp64_from_scratch_fits_per_s": scratch_rate,
            "plain_fp64_sample": f"{plain_subset} requests fitted one at a time through the plain FP64 path with its kernels "
                                 "reused from fit to fit (and, beside it, built anew per fit), time scaled to the whole "
                                 "stream by its event count (angular events weighted x3)",
            "speedup_vs_plain": sustained / plain_rate, "speedup_vs_plain_from_scratch": sustained / scratch_rate}


def reference_protocol(M, cplain, repeats):
    """The reference's toy-study benchmark (benchmarking/morefit_massfit.cc:64-107): 100 toys of
    the mass density at N = 1e3, 1e4, 1e5 and 1e6 events, repeated.  Batched FP32 path: the
    toys generated on the device (one seed per toy) and fitted in one batched study, from the
    generating values as the reference does; plain FP64 path: the same datasets fitted one at a
    time with kernels reused (a subset at 1e6)."""
    rows = []
    ctx = M.Context(device=0)
    m = ctx.model(M.MASS, fp32=True)
    for n in (1000, 10000, 100000, 1000000):
        ntoys = 100
        b = m.batch([n] * ntoys)
        walls = []
        for r in range(repeats):
            t0 = time.perf_counter()
            b.generate(M.MASS_TRUTH, list(range(1000 * r + 1, 1000 * r + ntoys + 1)))
            v, e, o, st = b.fit(M.MASS_TRUTH)
            walls.append(time.perf_counter() - t0)
        conv = sum(1 for x in o if x.status == 0)
        nplain = ntoys if n < 1000000 else 10
        tp = 0.0
        for i in range(nplain):
            d = b.download(i)
            t1 = time.perf_counter()
            cplain.sessions[M.MASS].fit(d, M.MASS_TRUTH)
            tp += time.perf_counter() - t1
        rows.append({"events": n, "toys": ntoys, "repeats": repeats, "batched_fp32_study_ms_mean": 1e3 * float(np.mean(walls)),
                     "batched_fp32_study_ms_rms": 1e3 * float(np.std(walls, ddof=1)) if repeats > 1 else 0.0,
                     "batched_fp32_toy_fits_per_s": ntoys / float(np.mean(walls)), "converged": conv,
                     "plain_fp64_toy_fits_per_s": nplain / tp, "plain_fp64_sample_toys": nplain})
        log(f"reference protocol N={n}: batched {rows[-1]['batched_fp32_toy_fits_per_s']:.0f} toys/s "
            f"(study {rows[-1]['batched_fp32_study_ms_mean']:.1f} ms incl. generation), plain FP64 {rows[-1]['plain_fp64_toy_fits_per_s']:.1f} toys/s")
        b.close()
    m.close()
    ctx.close()
    return rows


def cpu_baselines(M, O, data, start, seconds):
    """The host CPU backend (generated C++ per density with the semantics of the reference's
    LLVM backend: FP64, 4 events per AVX2 vector with libmvec, 4 interleaved Kahan lanes,
    contiguous slices per thread), on this machine's cores: -2lnL+gradient evaluations/s at
    N = 1e7 (a 2e6-event sample, scaled) and of the angular density at N = 1e6, built for the
    host (-march=native) and for the reference's default target (x86-64-v3), on 1 thread and
    on all threads; a config-1 fit (N = 1e5) with the analytic and with the reference
    benchmark's numerical derivatives; toys of Poisson(5000) events fitted one after another
    (kernels reused), on 1 thread and on all threads (one toy per thread)."""
    import threading
    nproc = os.cpu_count() or 1
    # the one-GPU share of the machine's host cores that work may use (16), or fewer if the
    # process is limited to fewer
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    native = "-O3 -march=native -fno-math-errno"
    v3 = "-O3 -march=x86-64-v3 -fno-math-errno"
    sample = 2_000_000
    d = data[:, :sample]
    out = {"threads_all": cores, "nproc": nproc}

    def eval_rate(kind, dd, theta, threads, flags, scale_to):
        p = M.PlainFitter(kind, fp32=False, nthreads=threads, cpu_flags=flags)
        p.time_eval(dd, theta, 1)
        iters, el = 2, 0.0
        while True:
            _, _, el = p.time_eval(dd, theta, iters)
            if el > seconds / 8 or iters >= 4096:
                break
            iters *= 2
        p.close()
        return iters / el * dd.shape[1] / scale_to, iters

    for label, threads, flags in (("native_1thread", 1, native), ("native_all", cores, native), ("v3_1thread", 1, v3), ("v3_all", cores, v3)):
        r, it = eval_rate(M.MASS, d, start, threads, flags, N_EVENTS)
        out["mass_1e7_" + label] = {"evals_per_s": r, "threads": threads, "flags": flags, "sample": f"{it} evaluations over {sample} events, scaled to 1e7"}
    dk = O.generate(M.KSTAR, M.KSTAR_TRUTH, 200000, seed=3, nstreams=16)
    for label, threads in (("native_1thread", 1), ("native_all", cores)):
        r, it = eval_rate(M.KSTAR, dk, M.KSTAR_START, threads, native, 1_000_000)
        out["angular_1e6_" + label] = {"evals_per_s": r, "threads": threads, "flags": native, "sample": f"{it} evaluations over 200000 events, scaled to 1e6"}
    # config 1: a full MIGRAD + HESSE fit at N = 1e5
    d1 = O.generate(M.MASS, M.MASS_TRUTH, 100000, seed=229387429, nstreams=16)
    for label, flags in (("analytic", M.FIT_DEFAULT), ("numerical", M.FIT_KAHAN_ON_ACCELERATOR | M.FIT_POSTRUN_HESSE)):
        for threads in (1, cores):
            M.fit_options(M.MASS, d1, M.MASS_START, flags=flags, nthreads=threads)  # kernels compiled and cached
            t0 = time.perf_counter()
            v, e, o, ms = M.fit_options(M.MASS, d1, M.MASS_START, flags=flags, nthreads=threads)
            out[f"fit_1e5_{label}_{threads}threads"] = {"wall_ms": 1e3 * (time.perf_counter() - t0), "fit_ms": ms,
                                                         "status": o.status, "calls": o.ncalls}
    # config 4 on the CPU: toys of Poisson(5000) events, one fit after another, kernels reused
    rs = np.random.default_rng(4)
    toys = [O.generate(M.MASS, M.MASS_TRUTH, int(k), seed=100 + i, nstreams=4) for i, k in enumerate(rs.poisson(5000, 64))]
    starts = M.toy_starts(M.MASS, [t.shape[1] for t in toys], seed=4)

    def toy_worker(idx, res, barrier=None):
        # a session per thread, its kernels built by one untimed fit; then the timed fits
        p = M.PlainFitter(M.MASS, fp32=False, nthreads=1)
        p.fit(toys[0], starts[0])
        if barrier is not None:
            barrier.wait()
        t0 = time.perf_counter()
        for i in idx:
            p.fit(toys[i], starts[i])
        res.append((t0, time.perf_counter()))
        p.close()

    res1 = []
    toy_worker(range(1, 33), res1)
    out["toys_1thread"] = {"toy_fits_per_s": 32 / (res1[0][1] - res1[0][0]), "toys": 32, "events_mean": 5000}
    resn, ths = [], []
    per = 16
    bar = threading.Barrier(cores)
    for w in range(cores):
        ths.append(threading.Thread(target=toy_worker, args=([1 + (w * per + j) % 63 for j in range(per)], resn, bar)))
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    span = max(e for _, e in resn) - min(b for b, _ in resn)
    out["toys_all_threads"] = {"toy_fits_per_s": cores * per / span, "threads": cores, "toys": cores * per, "events_mean": 5000,
                               "note": "one fitter session per thread, toys fitted concurrently (the ctypes calls release the GIL)"}
    return out


def call_latency_and_extras(M, O, ctx, data, start):
    """Per-call latency of the FP32 mass evaluation at smaller N (the same launch shape; the
    fixed cost of a call shows below 10^6 events), the plotter (integral kernel over 100 x 10
    fine bins of the mass density on the device, histogram and pulls of 10^6 events; bin
    integrals checked against the oracle's quadrature) and the failure-detection pass at
    N = 10^7."""
    out = {"call_latency": []}
    m = ctx.model(M.MASS, fp32=True)
    for n in (1000, 10_000, 100_000, 1_000_000):
        b = m.batch([n])
        b.upload(0, data[:, :n])
        for _ in range(20):
            b.eval(M.EVAL_GRAD, start)
        iters = 300
        t0 = time.perf_counter()
        for _ in range(iters):
            b.eval(M.EVAL_GRAD, start)
        py = (time.perf_counter() - t0) / iters * 1e6
        out["call_latency"].append({"events": n, "python_us": py, "cpp_loop_us": b.time_eval(M.EVAL_GRAD, iters=iters),
                                    "kernels_of_a_call_us": b.time_kernel(M.EVAL_GRAD, iters=iters) * 1e3})
        b.close()
    b = m.batch([data.shape[1]])
    b.upload(0, data)
    b.check(start)
    t0 = time.perf_counter()
    for _ in range(20):
        r = b.check(start)
    out["diagnostic_check_us"] = (time.perf_counter() - t0) / 20 * 1e6
    out["diagnostic_check_result"] = r[0].tolist()
    b.close()
    m.close()
    sub = data[:, :1_000_000]
    # untimed: the first plot of a process builds its integral kernel (hipRTC, ~130 ms uncached)
    M.plot(M.MASS, M.MASS_TRUTH, sub[:, :1000], nbins=100, multiplier=10, ctx=ctx, fp32=False)
    t0 = time.perf_counter()
    p = M.plot(M.MASS, M.MASS_TRUTH, sub, nbins=100, multiplier=10, ctx=ctx, fp32=False)
    plot_ms = (time.perf_counter() - t0) * 1e3
    rp = M.plot(M.MASS, M.MASS_TRUTH, sub, nbins=100, multiplier=10, ctx=ctx, fp32=False, replots=10)
    e = np.linspace(5.0, 5.6, 1001)
    ref = O.bin_integrals(O.MASS, M.MASS_TRUTH, 0, e[:-1], e[1:])
    ok = p["counts"] > 0
    out["plot"] = {"events": sub.shape[1], "bins": 100, "fine_bins": 1000, "plot_ms": plot_ms, "replot_ms": rp["ms"],
                   "fp64_max_abs_bin_err_vs_oracle": float(np.max(np.abs(p["pdf_integrals"] - ref))),
                   "pull_mean": float(np.mean(p["pulls"][ok])), "pull_rms": float(np.sqrt(np.mean(p["pulls"][ok] ** 2)))}
    return out


def angular_workload(M, O):
    """The angular configuration: K*mumu P-wave density in (cos theta_l, cos theta_K, phi), 8 free
    parameters, N = 10^6 events generated on the device (FP64 generator), FP32 kernels in the
    autotuned shape and FP64 kernels: -2lnL+gradient per call and per kernel, accuracy against
    the oracle (FP32 on the same stored observables), fits of both from a 1 sigma smeared start."""
    n = 1_000_000
    ctx, tune = tuned_context(M, M.KSTAR, True, [n], "angular N=1e6 FP32")
    ctx64 = M.Context(device=0)  # FP64 at the default launch shape
    m64 = ctx64.model(M.KSTAR, fp32=False)
    b64 = m64.batch([n])
    b64.generate(M.KSTAR_TRUTH, [20240607])
    data = b64.download(0)
    m32 = ctx.model(M.KSTAR, fp32=True)
    b32 = m32.batch([n])
    b32.upload(0, data)
    theta = np.array(M.KSTAR_START)
    out = {"events": n, "launch_shape": shape(m32, tune["best"])}
    for label, b in (("fp32", b32), ("fp64", b64)):
        us = b.time_eval(M.EVAL_GRAD, iters=100)
        k = b.time_kernel(M.EVAL_GRAD, iters=100, main_only=not m32.fused[M.EVAL_GRAD])
        out[label] = {"evals_per_s": 1e6 / us, "us_per_eval": us, "kernel_us": k * 1e3,
                      "GBps": n * 3 * (4 if label == "fp32" else 8) / (k * 1e-3) / 1e9}
    ref = O.nll_sums(M.KSTAR, data, theta, mode=1, nthreads=0)
    stored = data.astype(np.float32).astype(np.float64)  # the angular ranges are centred on 0
    ref32 = O.nll_sums(M.KSTAR, stored, theta, mode=1, nthreads=0)
    absref = O.nll_sums(M.KSTAR, data, theta, mode=11, nthreads=8)
    r64 = b64.eval(M.EVAL_GRAD, theta)[0]
    r32 = b32.eval(M.EVAL_GRAD, theta)[0]
    out["fp64"]["rel_err_m2lnL"] = float(abs(r64[0] - ref[0]) / abs(ref[0]))
    out["fp64"]["max_grad_err_scaled"] = float(np.max(np.abs(r64[1:] - ref[1:]) / absref[1:]))
    out["fp32"]["rel_err_m2lnL_same_stored_data"] = float(abs(r32[0] - ref32[0]) / abs(ref32[0]))
    out["fp32"]["max_grad_err_scaled"] = float(np.max(np.abs(r32[1:] - ref32[1:]) / absref[1:]))
    start = M.toy_starts(M.KSTAR, [n], seed=3)[0]
    v64, e64, o64, s64 = b64.fit(start)
    v32, e32, o32, s32 = b32.fit(start)
    out["fit"] = {"status": [o32[0].status, o64[0].status], "fp32_vs_fp64_max_dtheta_over_sigma":
                  float(np.max(np.abs(v32[0] - v64[0]) / e64[0])), "fit_ms": [s32.wall_ms, s64.wall_ms],
                  "fp64_values": v64[0].tolist(), "fp64_errors": e64[0].tolist()}
    log(f"angular N=1e6: fp32 {out['fp32']['evals_per_s']:.0f} evals/s (kernel {out['fp32']['kernel_us']:.1f} us), "
        f"fp64 {out['fp64']['evals_per_s']:.0f} evals/s (kernel {out['fp64']['kernel_us']:.1f} us); "
        f"fits fp32/fp64 dtheta/sigma {out['fit']['fp32_vs_fp64_max_dtheta_over_sigma']:.2e}")
    # the reference's K*mumu benchmark setting (optimize_dimensions): event-only terms of the
    # density precomputed once into data columns, read at every call instead of recomputed
    cxe, tunee = tuned_context(M, M.KSTAR, True, [n], "angular N=1e6 FP32, event columns", event_columns=1)
    cxe64 = M.Context(device=0, event_columns=1)
    ev = {}
    for label, cx, fp32 in (("fp32", cxe, True), ("fp64", cxe64, False)):
        me = cx.model(M.KSTAR, fp32=fp32)
        be = me.batch([n])
        be.upload(0, data)
        us = be.time_eval(M.EVAL_GRAD, iters=100)
        k = be.time_kernel(M.EVAL_GRAD, iters=100, main_only=not me.fused[M.EVAL_GRAD])
        r = be.eval(M.EVAL_GRAD, theta)[0]
        rr, st = (ref32, "same_stored_data") if fp32 else (ref, "oracle")
        ve, ee, oe, se = be.fit(start)
        ev[label] = {"columns": me.ncol, "evals_per_s": 1e6 / us, "us_per_eval": us, "kernel_us": k * 1e3,
                     "GBps_columns_stored": n * me.ncol * (4 if fp32 else 8) / (k * 1e-3) / 1e9,
                     "rel_err_m2lnL_" + st: float(abs(r[0] - rr[0]) / abs(rr[0])),
                     "max_grad_err_scaled": float(np.max(np.abs(r[1:] - rr[1:]) / absref[1:])),
                     "fit_status": oe[0].status, "fit_ms": se.wall_ms,
                     "fit_vs_fp64_plain_max_dtheta_over_sigma": float(np.max(np.abs(ve[0] - v64[0]) / e64[0]))}
        if fp32:
            ev[label]["launch_shape"] = shape(me, tunee["best"])
        be.close()
        me.close()
    out["event_columns"] = ev
    log(f"angular N=1e6 with {ev['fp32']['columns']} data columns: fp32 {ev['fp32']['evals_per_s']:.0f} evals/s "
        f"(kernel {ev['fp32']['kernel_us']:.1f} us), fp64 {ev['fp64']['evals_per_s']:.0f} evals/s (kernel {ev['fp64']['kernel_us']:.1f} us)")
    for x in (b32, b64):
        x.close()
    m32.close()
    m64.close()
    for c in (ctx, ctx64, cxe, cxe64):
        c.close()
    return out


def profile_workload():
    """Headline kernels only, for hardware-counter passes: FP32 mf_eval_grad and mf_eval_lh at
    N = 10^7 (generated on the device) in the benched launch shape (block 256, 32 events per
    thread, unroll 4), then the copy kernel over 40 MB (the counters' scale)."""
    import morefit_amd as M
    # the launch shape the unprofiled bench picks (a profiler's per-dispatch cost skews tuning)
    ctx, tune = tuned_context(M, M.MASS, True, [N_EVENTS], "mass N=1e7 FP32", candidates=[(256, 32, 4)])
    m32 = ctx.model(M.MASS, fp32=True)
    s = shape(m32, tune["best"])
    log("profile launch_shape: block=%d ept=%d acc=%d unroll=%d reduction=%s" % (
        s["block"], s["events_per_thread"], s["accumulate"], s["unroll"]["grad"], s["segment_reduction"].replace(" ", "-")))
    b32 = m32.batch([N_EVENTS])
    b32.generate(M.MASS_TRUTH, [229387429])
    log("profile: grad", b32.time_kernel(M.EVAL_GRAD, iters=20) * 1e3, "us")
    log("profile: lh", b32.time_kernel(M.EVAL_LH, iters=20) * 1e3, "us")
    log("profile: copy 40 MB", ctx.copy_ceiling(40_000_000, iters=20), "GB/s")


def profile_others():
    """Gradient kernels below the headline, for hardware-counter passes: FP64 mass at N = 10^7,
    angular FP32 and FP64 at N = 10^6, and the 4096-toy FP32 batch (Poisson(5000) events per
    toy), each in a fixed launch shape (the one the autotuner picked for it without a profiler
    attached: tuning under counter collection, which serialises the dispatches, picks
    others), 20 back-to-back launches of the per-chunk kernel, labelled on stderr by grid size"""
    import morefit_amd as M
    jobs = [("mass FP64 N=1e7", M.MASS, False, [N_EVENTS], (256, 32, 4)),
            ("angular FP32 N=1e6", M.KSTAR, True, [1_000_000], (256, 16, 8)),
            ("angular FP64 N=1e6", M.KSTAR, False, [1_000_000], (256, 16, 8)),
            ("4096 toys FP32", M.MASS, True, list(np.random.default_rng(229387429).poisson(5000, 4096)), (256, 64, 8))]
    for label, kind, fp32, sizes, fixed in jobs:
        ctx, tune = tuned_context(M, kind, fp32, sizes, label, candidates=[fixed])
        m = ctx.model(kind, fp32=fp32)
        b = m.batch(sizes)
        b.generate(M.TRUTH[kind], list(range(1, len(sizes) + 1)))
        s = shape(m, tune["best"])
        ms = b.time_kernel(M.EVAL_GRAD, iters=20, main_only=True)
        chunk = s["block"] * s["events_per_thread"]
        nch = int(sum((int(n) + chunk - 1) // chunk for n in sizes))
        log(f"profile {label}: block={s['block']} ept={s['events_per_thread']} unroll={s['unroll']['grad']} vgprs={s['vgprs']['grad']} "
            f"grid={nch * s['block']} ({nch} chunks), {ms * 1e3:.2f} us per launch, {sum(sizes)} events")
        b.close()
        m.close()
        ctx.close()


def single_thread_blas():
    """numpy's BLAS on one thread in this process: its pool threads keep spinning for a while
    after each call and take the cores that the batched fits' host threads run MIGRAD on (a
    500 x 500 product before a 4096-toy study made the study's host steps 3-4x slower); the
    bench's own BLAS work is small"""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        return None
    return threadpool_limits(limits=1, user_api="blas")


def main():
    single_thread_blas()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-toys", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--toys", type=int, default=4096)
    ap.add_argument("--plain-toys", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--protocol-repeats", type=int, default=3)
    ap.add_argument("--profile", nargs="?", const="headline", default=None, choices=["headline", "others"])
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-angular", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=1000.0)
    # a fixed launch shape "block,events_per_thread,unroll" for the N = 1e7 FP32 / FP64 workloads
    # instead of the autotuner's pick (profiled runs: a profiler's per-dispatch cost skews tuning)
    ap.add_argument("--shape32", default=None)
    ap.add_argument("--shape64", default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
              "(the likelihood path is measured on one MI355X's share; multi-GPU sharding is a design note only).",
              file=sys.stderr)
        sys.exit(2)
    if args.profile == "headline":
        profile_workload()
        return
    if args.profile == "others":
        profile_others()
        return

    import morefit_amd as M
    from oracle import oracle as O

    n = N_EVENTS
    start = M.MASS_START.copy()
    # launch shapes picked by the autotuner for each workload
    fixed = lambda v: [tuple(int(x) for x in v.split(","))] if v else None
    ctx, tune32 = tuned_context(M, M.MASS, True, [n], "mass N=1e7 FP32", candidates=fixed(args.shape32))
    ctx64, tune64 = tuned_context(M, M.MASS, False, [n], "mass N=1e7 FP64", candidates=fixed(args.shape64))
    name, cus = ctx.device_name()
    log("device", name, cus, "CUs")

    # ---------------------------------------------------------------- data (on device)
    m64 = ctx64.model(M.MASS, fp32=False)
    b64 = m64.batch([n])
    t0 = time.perf_counter()
    b64.generate(M.MASS_TRUTH, [229387429])
    data = b64.download(0)
    log(f"generated {n} events on device in {time.perf_counter() - t0:.2f} s")
    m32 = ctx.model(M.MASS, fp32=True)
    b32 = m32.batch([n])
    b32.upload(0, data)

    # ---------------------------------------------------------------- timed steps (FP32)
    # the device at its sustained clock first: evaluations for SETTLE_S seconds, untimed, then the
    # W warm-up steps (after idle the first few hundred calls run up to 15 % slower: 24.5 against
    # 20.8 us per call, kernel 9.5 against 9.0 us)
    t_settle = time.perf_counter() + SETTLE_S
    while time.perf_counter() < t_settle:
        b32.eval(M.EVAL_GRAD, start)
    for _ in range(args.warmup):
        b32.eval(M.EVAL_GRAD, start)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r32 = b32.eval(M.EVAL_GRAD, start)
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    value = args.steps / dt
    log(f"fp32 -2lnL+grad: {value:.1f} evals/s ({ms_per_step:.4f} ms/eval)")

    # device time (HIP events on the library's stream, around back-to-back launches) of the
    # dominant kernel, mf_eval_grad -- the per-chunk kernel alone when a final kernel reduces the
    # segments, the whole kernel when it reduces them itself -- and of a call's kernels together
    fused32 = m32.fused[M.EVAL_GRAD]
    # each figure the median of 5 timing loops (a single loop now and then catches interference)
    med = lambda f: float(np.median([f() for _ in range(5)]))
    k32_pair_ms = med(lambda: b32.time_kernel(M.EVAL_GRAD, iters=max(50, args.steps)))
    k32_ms = k32_pair_ms if fused32 else med(lambda: b32.time_kernel(M.EVAL_GRAD, iters=max(50, args.steps), main_only=True))
    k32_lh_ms = med(lambda: b32.time_kernel(M.EVAL_LH, iters=max(50, args.steps), main_only=not fused32))
    bytes32 = n * 4
    achieved = bytes32 / (k32_ms * 1e-3) / 1e9
    log(f"fp32 grad kernel {k32_ms * 1e3:.2f} us -> {achieved:.0f} GB/s (with the segment reduction {k32_pair_ms * 1e3:.2f} us); "
        f"lh kernel {k32_lh_ms * 1e3:.2f} us")
    traffic, traffic_shape = measured_traffic(shape(m32, tune32["best"]))
    # where a call's time goes: C++ loop of full calls (launch + sync), device time of the call
    # graph replayed back to back, and the same through the direct-launch path
    step = {"python_loop_us": ms_per_step * 1e3,
            "cpp_loop_direct_launch_us": b32.time_eval(M.EVAL_GRAD, iters=max(50, args.steps)),
            "cpp_loop_graph_replay_us": b32.time_eval(M.EVAL_GRAD, iters=max(50, args.steps), use_graph=True),
            "device_graph_us": b32.time_graph(M.EVAL_GRAD, iters=max(50, args.steps)),
            "kernel_us": k32_ms * 1e3, "kernels_of_a_call_us": k32_pair_ms * 1e3}
    log("call breakdown (us): " + ", ".join(f"{k} {v:.1f}" for k, v in step.items()))

    # ---------------------------------------------------------------- FP64 path
    for _ in range(max(5, args.warmup // 2)):
        b64.eval(M.EVAL_GRAD, start)
    t0 = time.perf_counter()
    steps64 = max(20, args.steps // 2)
    for _ in range(steps64):
        r64 = b64.eval(M.EVAL_GRAD, start)
    v64 = steps64 / (time.perf_counter() - t0)
    k64_ms = b64.time_kernel(M.EVAL_GRAD, iters=max(50, args.steps), main_only=not m64.fused[M.EVAL_GRAD])
    log(f"fp64 -2lnL+grad: {v64:.1f} evals/s, kernel {k64_ms * 1e3:.2f} us")

    # ---------------------------------------------------------------- accuracy vs oracle
    t0 = time.perf_counter()
    ref = O.nll_sums(O.MASS, data, start, mode=1, nthreads=0)
    stored = (data - 5.3).astype(np.float32).astype(np.float64) + 5.3
    ref_st = O.nll_sums(O.MASS, stored, start, mode=1, nthreads=0)
    absref = O.nll_sums(O.MASS, data, start, mode=11, nthreads=8)
    log(f"oracle sums in {time.perf_counter() - t0:.1f} s")
    acc = {
        "fp64_rel_err_m2lnL": float(abs(r64[0, 0] - ref[0]) / abs(ref[0])),
        "fp64_max_grad_err_scaled": float(np.max(np.abs(r64[0, 1:] - ref[1:]) / absref[1:])),
        "fp32_rel_err_m2lnL": float(abs(r32[0, 0] - ref[0]) / abs(ref[0])),
        "fp32_rel_err_m2lnL_same_stored_data": float(abs(r32[0, 0] - ref_st[0]) / abs(ref_st[0])),
        "fp32_max_grad_err_scaled": float(np.max(np.abs(r32[0, 1:] - ref_st[1:]) / absref[1:])),
    }
    # fitted parameters: FP32-compensated vs FP64 on the same data, in units of sigma, both
    # started 1 sigma (N(0,1)-smeared) from the generating values
    t0 = time.perf_counter()
    near = M.toy_starts(M.MASS, [n], seed=5)[0]
    v32f, e32f, o32, st32 = b32.fit(near)
    v64f, e64f, o64, st64 = b64.fit(near)
    acc["fit_fp32_vs_fp64_max_dtheta_over_sigma"] = float(np.max(np.abs(v32f[0] - v64f[0]) / e64f[0]))
    acc["fit_status"] = [o32[0].status, o64[0].status]
    acc["fit_fp32"] = v32f[0].tolist()
    acc["fit_fp64"] = v64f[0].tolist()
    acc["fit_sigma_fp64"] = e64f[0].tolist()
    acc["fit_fp32_ms"] = st32.wall_ms
    acc["fit_fp64_ms"] = st64.wall_ms
    log(f"fits fp32/fp64 in {time.perf_counter() - t0:.1f} s: dtheta/sigma {acc['fit_fp32_vs_fp64_max_dtheta_over_sigma']:.2e}")
    # the same from the configuration's start values, thousands of sigma away at this N
    f32, _, fo32, _ = b32.fit(start)
    f64, _, fo64, _ = b64.fit(start)
    acc["far_start"] = {"start": list(map(float, start)), "status": [fo32[0].status, fo64[0].status],
                        "fp32_vs_fp64_max_dtheta_over_sigma": float(np.max(np.abs(f32[0] - f64[0]) / e64f[0])),
                        "fp64_vs_near_start_max_dtheta_over_sigma": float(np.max(np.abs(f64[0] - v64f[0]) / e64f[0]))}

    # ---------------------------------------------------------------- smaller fits, plotter, diagnostics
    extras = call_latency_and_extras(M, O, ctx, data, start)
    log("call latency (us): " + ", ".join(f"N={r['events']:.0e} {r['python_us']:.1f}" for r in extras["call_latency"]) +
        f"; plot of 1e6 events {extras['plot']['plot_ms']:.2f} ms (re-plot {extras['plot']['replot_ms']:.2f} ms); "
        f"diagnostic pass {extras['diagnostic_check_us']:.1f} us")

    # ---------------------------------------------------------------- HBM ceilings
    copy_gbs = ctx.copy_ceiling(1 << 31, iters=5)
    stream = None
    if not args.no_stream:
        # the same kernels in steady state: 2 GB of events (beyond the 256 MB last-level cache, so
        # every event comes from HBM), launch shape tuned for that size; the per-chunk kernel alone
        # (the per-event kernel's rate) and with the final pass
        ns = 500_000_000
        cts, tunes = tuned_context(M, M.MASS, True, [ns], "mass N=5e8 FP32 (streaming)")
        ms = cts.model(M.MASS, fp32=True)
        bs = ms.batch([ns])
        bs.generate(M.MASS_TRUTH, [7])
        ks = med(lambda: bs.time_kernel(M.EVAL_GRAD, iters=20, main_only=True))
        ksf = med(lambda: bs.time_kernel(M.EVAL_GRAD, iters=20))
        kls = med(lambda: bs.time_kernel(M.EVAL_LH, iters=20, main_only=True))
        stream = {"events": ns, "bytes": ns * 4, "launch_shape": shape(ms, tunes["best"]), "grad_kernel_ms": ks,
                  "grad_GBps": ns * 4 / (ks * 1e-3) / 1e9, "grad_with_final_ms": ksf, "lh_kernel_ms": kls,
                  "lh_GBps": ns * 4 / (kls * 1e-3) / 1e9}
        stream["grad_frac"] = stream["grad_GBps"] / HBM_PEAK_GBS
        stream["lh_frac"] = stream["lh_GBps"] / HBM_PEAK_GBS
        log(f"streaming 2 GB: grad {stream['grad_GBps']:.0f} GB/s ({stream['grad_frac']:.2f} of peak; with the final pass "
            f"{ns * 4 / (ksf * 1e-3) / 1e9:.0f} GB/s), lh {stream['lh_GBps']:.0f} GB/s, copy ceiling {copy_gbs:.0f} GB/s")
        bs.close()
        ms.close()
        cts.close()

    # ---------------------------------------------------------------- angular configuration
    angular = None if args.no_angular else angular_workload(M, O)

    # ---------------------------------------------------------------- plain per-fit FP64 path
    cplain = plain_context(M, O)

    # ---------------------------------------------------------------- toy study
    toys = None
    if not args.no_toys:
        rs = np.random.default_rng(229387429)
        sizes = rs.poisson(5000, args.toys)
        ctxt, tunet = tuned_context(M, M.MASS, True, sizes, f"{args.toys} toys FP32")
        mt = ctxt.model(M.MASS, fp32=True)
        bt = mt.batch(sizes)
        t0 = time.perf_counter()
        bt.generate(M.MASS_TRUTH, list(range(1, args.toys + 1)))
        tgen = time.perf_counter() - t0
        starts = M.toy_starts(M.MASS, sizes, seed=229387429)
        # untimed warm-up study: the process's fibre stacks and host threads are created once
        bt.fit(starts)
        t0 = time.perf_counter()
        tv, te, to, tst = bt.fit(starts)
        tfit = time.perf_counter() - t0
        ok = np.array([o.status == 0 for o in to])
        pulls = (tv[ok] - M.MASS_TRUTH) / te[ok]
        batched_rate = args.toys / tfit
        toy_kernel_ms = bt.time_kernel(M.EVAL_GRAD, iters=50, main_only=not mt.fused[M.EVAL_GRAD])
        # plain per-fit FP64 GPU path on a subset (same toys, one fit at a time, one launch per
        # call): with its kernels reused from toy to toy (refit, the reference toy study's way)
        # and, beside it, built anew for every fit
        tplain = tscratch = 0.0
        plain_dev = []
        for i in range(args.plain_toys):
            d = bt.download(i)
            t1 = time.perf_counter()
            pv, pe, po, _ = cplain.sessions[M.MASS].fit(d, starts[i])
            t2 = time.perf_counter()
            cplain.fit_plain(M.MASS, d, starts[i], fp32=False)
            tplain += t2 - t1
            tscratch += time.perf_counter() - t2
            if po.status == 0 and to[i].status == 0:
                plain_dev.append(float(np.max(np.abs(tv[i] - pv) / pe)))
        plain_rate = args.plain_toys / tplain
        scratch_rate = args.plain_toys / tscratch
        toys = {"toys": args.toys, "mean_events": float(sizes.mean()), "generate_s": tgen, "fit_s": tfit,
                "toy_fits_per_s": batched_rate, "study_fits_per_s_incl_generation": args.toys / (tgen + tfit),
                "plain_fp64_toy_fits_per_s": plain_rate, "plain_fp64_from_scratch_toy_fits_per_s": scratch_rate,
                "speedup_vs_plain": batched_rate / plain_rate, "speedup_vs_plain_from_scratch": batched_rate / scratch_rate,
                "batched_fp32_vs_plain_fp64_max_dtheta_over_sigma": max(plain_dev) if plain_dev else None,
                "converged": int(ok.sum()),
                "pull_mean": pulls.mean(axis=0).round(4).tolist(), "pull_width": pulls.std(axis=0).round(4).tolist(),
                "ticks": tst.ticks, "launches": [tst.launches_lh, tst.launches_grad, tst.launches_hess],
                "host_ms": tst.host_ms, "gpu_wait_ms": tst.gpu_wait_ms, "host_threads": tst.host_threads,
                "setup_ms": tst.setup_ms, "step_ms": tst.step_ms,
                "launch_shape": shape(mt, tunet["best"]), "grad_kernel_ms": toy_kernel_ms,
                "autotune_table_us": [[list(c), round(u, 2)] for c, u in tunet["table"]]}
        log(f"toys: {batched_rate:.0f} fits/s batched vs {plain_rate:.1f} plain with kernel reuse ({batched_rate / plain_rate:.1f}x), "
            f"{scratch_rate:.1f} plain from scratch; FP32 batched vs FP64 plain max {max(plain_dev) if plain_dev else float('nan'):.2e} sigma")
        toys["reference_protocol"] = reference_protocol(M, cplain, args.protocol_repeats)
        bt.close()
        mt.close()
        ctxt.close()

    # ---------------------------------------------------------------- fit service
    serving = None
    if not args.no_serve:
        serving = serve_workload(ctx, cplain, M, O, args.serve_requests, args.serve_rate)
        log(f"serving: sustained {serving['sustained_fits_per_s']:.0f} fits/s vs {serving['plain_fp64_fits_per_s']:.1f} "
            f"one at a time; " + "; ".join(f"{r['load']} ({r['arrival_rate_per_s']:.0f}/s): p50 {r['p50_latency_ms']:.1f} ms "
                                           f"p99 {r['p99_latency_ms']:.1f} ms" for r in serving["rates"]))

    # ---------------------------------------------------------------- CPU baselines
    cpu = cpu1 = cpu_port = cpu_more = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu_more = cpu_baselines(M, O, data, start, args.cpu_seconds)
        # the faster of the two target flags on each thread count (the reference's LLVM backend
        # targets x86-64-v3 by default, compute.hh llvm_cpu)
        pick = lambda a, b: a if a["evals_per_s"] >= b["evals_per_s"] else b
        nat = pick(cpu_more["mass_1e7_native_all"], cpu_more["mass_1e7_v3_all"])
        nat1 = pick(cpu_more["mass_1e7_native_1thread"], cpu_more["mass_1e7_v3_1thread"])
        what = ("-2lnL+grad of the mass density through the host CPU backend (generated C++, FP64, AVX2 4-event vectors with "
                "libmvec, 4 interleaved Kahan lanes, contiguous slices per thread: the reference LLVM backend's semantics), "
                f"{nat['flags']}; ")
        cpu = {"value": nat["evals_per_s"], "unit": "evals/s", "cores": nat["threads"], "kind": "port",
               "sample": what + nat["sample"] + f"; {cpu_more['nproc']} host CPUs seen, {nat['threads']} threads used"}
        cpu1 = {"value": nat1["evals_per_s"], "unit": "evals/s", "cores": 1, "kind": "port",
                "sample": what.replace(nat["flags"], nat1["flags"]) + nat1["sample"]}
        # the scalar oracle loop, as before, for comparison
        sample = 2_000_000
        d = data[:, :sample]
        t1 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t1 < 2.0:
            O.nll_sums(O.MASS, d, start, mode=1, nthreads=1)
            reps += 1
        cpu_port = {"value": 1.0 / ((time.perf_counter() - t1) / reps * (n / sample)), "unit": "evals/s", "cores": 1,
                    "kind": "port", "sample": f"{reps} evaluations of the scalar oracle loop over {sample} events, scaled to N={n}"}
        log(f"cpu baseline ({time.perf_counter() - t0:.0f} s): {cpu['value']:.2f} evals/s ({cpu['cores']} threads), {cpu1['value']:.2f} "
            f"(1 thread), x86-64-v3 1 thread {cpu_more['mass_1e7_v3_1thread']['evals_per_s']:.2f}; scalar oracle {cpu_port['value']:.2f}; "
            f"config-1 fit {cpu_more['fit_1e5_analytic_1threads']['wall_ms']:.0f} ms (analytic, 1 thread), "
            f"{cpu_more['fit_1e5_numerical_1threads']['wall_ms']:.0f} ms (numerical); toys {cpu_more['toys_1thread']['toy_fits_per_s']:.1f}/s "
            f"(1 thread), {cpu_more['toys_all_threads']['toy_fits_per_s']:.1f}/s ({cpu_more['toys_all_threads']['threads']} threads)")

    # the other gradient kernels against the same roofline (traffic from their counter summary)
    ot, fl = others_traffic(), others_counters("valu_issue_floor_ms")
    other_rows = [roofline_row("mass FP64 N=1e7", k64_ms, n * 8, ot, fl)]
    if angular:
        other_rows.append(roofline_row("angular FP32 N=1e6", angular["fp32"]["kernel_us"] * 1e-3, 12_000_000, ot, fl))
        other_rows.append(roofline_row("angular FP64 N=1e6", angular["fp64"]["kernel_us"] * 1e-3, 24_000_000, ot, fl))
    if toys:
        other_rows.append(roofline_row("4096 toys FP32", toys["grad_kernel_ms"], int(round(toys["mean_events"] * toys["toys"])) * 4, ot,
                                       fl))
        other_rows[-1]["launch_shape_bench"] = toys["launch_shape"]

    out = {
        "metric": "-2lnL+grad evals/s; toy fits/s vs plain GPU path; p50/p99 fit latency",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded accept-reject sample of the density, generated on the device",
        "config": {"workload": "mass fit f*Gauss+(1-f)*Exp on [5.0,5.6], N=1e7 events, one -2lnL+gradient evaluation per step",
                   "events": n, "precision": "FP32 per-event arithmetic, FP32 sums of 8-event groups added in FP64 per lane, FP64 totals",
                   "launch_shape": shape(m32, tune32["best"]),
                   "autotune_table_us": [[list(c), round(u, 2)] for c, u in tune32["table"]]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_launch_shape": traffic_shape,
                     "kernel": "mf_eval_grad (fp32, " + ("event loop and segment reduction)" if fused32 else "per-chunk kernel; mf_eval_grad_final reduces the segment)"),
                     "kernel_ms": k32_ms, "algorithmic_bytes": bytes32, "kernels_of_a_call_ms": k32_pair_ms,
                     "issue_floor": issue_floor(k32_ms)},
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "cpu_baseline_scalar_oracle": cpu_port,
        "cpu_baselines": cpu_more,
        "fp64": {"evals_per_s": v64, "kernel_ms": k64_ms, "GBps": n * 8 / (k64_ms * 1e-3) / 1e9,
                 "launch_shape": shape(m64, tune64["best"])},
        "lh_kernel_ms_fp32": k32_lh_ms,
        "roofline_other_kernels": other_rows,
        "call_breakdown": step,
        "accuracy": acc,
        "copy_ceiling_GBps": copy_gbs,
        "streaming": stream,
        "toys": toys,
        "serving": serving,
        "call_latency": extras["call_latency"],
        "plotter": extras["plot"],
        "diagnostic_check_us": extras["diagnostic_check_us"],
        "angular": angular,
        "device": name,
    }
    for x in (b32, b64):
        x.close()
    m32.close()
    m64.close()
    ctx.close()
    ctx64.close()
    close_plain(cplain)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""GPU parity of the MI355X likelihood path against the FP64 oracle, through the C ABI.

Bars (written here as the tolerances):
  FP64 kernels: -2lnL within 1e-12 relative of the oracle's long-double sum; gradient and
    Hessian sums within 1e-12 of the sum of absolute per-event values (the rounding scale).
  FP32 kernels (observables stored as centred FP32, parameter-only terms FP64, compensated
    per-lane sums, FP64 totals): every sum within 2e-6 of the sum of absolute per-event
    values, checked against the oracle evaluated on the same stored FP32 observables.
  Fits: FP64 GPU fit == CPU-backend fit of the same data to 1e-4 sigma; FP32 fits within
    1e-3 sigma of the FP64 fit.
  Segments: a segment's result is bitwise the same alone and in a mixed launch.
  Generator (FP64): the event stream equals the oracle's accept-reject stream exactly.
"""
import sys

import numpy as np
import pytest

import morefit_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def _hess_full(P, packed):
    H = np.zeros((P, P))
    k = 0
    for i in range(P):
        for j in range(i + 1):
            H[i, j] = H[j, i] = packed[k]
            k += 1
    return H


def _stored(model, data, fp32):
    """the observables exactly as the FP32 kernel stores them (centred, rounded), back in FP64"""
    if not fp32:
        return data
    rng = O.RANGES[model].reshape(-1, 2)
    c = rng.mean(axis=1)[:, None]
    return (data - c).astype(np.float32).astype(np.float64) + c


@pytest.mark.parametrize("model", [M.MASS, M.KSTAR])
@pytest.mark.parametrize("fp32", [False, True])
def test_eval_matches_oracle(ctx, model, fp32):
    n = 100000
    data = O.generate(model, M.TRUTH[model], n, seed=1234, nstreams=16)
    theta = np.array(M.START[model])
    say("eval parity", "model", model, "fp32", fp32, "n", n)
    mdl = ctx.model(model, fp32=fp32)
    b = mdl.batch([n])
    b.upload(0, data)
    stored = _stored(model, data, fp32)
    ref = O.nll_sums(model, stored, theta, mode=2, nthreads=0)
    absref = O.nll_sums(model, stored, theta, mode=12, nthreads=4)
    P = O.NPAR[model]
    tol = 2e-6 if fp32 else 1e-12
    for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
        say("  kind", kind)
        out = b.eval(kind, theta)[0]
        k = len(out)
        if fp32:
            assert np.all(np.abs(out - ref[:k]) <= tol * absref[:k]), (kind, out - ref[:k], absref[:k])
        else:
            assert abs(out[0] - ref[0]) <= 1e-12 * abs(ref[0])
            assert np.all(np.abs(out - ref[:k]) <= tol * absref[:k]), (kind, out - ref[:k])
    b.close()
    mdl.close()


@pytest.mark.parametrize("model", [M.MASS, M.KSTAR])
@pytest.mark.parametrize("fp32", [False, True])
def test_prefetch_off_matches_prefetch_on(model, fp32):
    """the evaluation kernels without the prefetch of the next event batch (one register set,
    prefetch=0) at a size with many chunks and a ragged tail: within the oracle bars, and the
    value kernel bitwise the same as with the prefetch (the same events in the same order per
    lane; only when the loads are issued changes)"""
    n = 1_000_003
    data = O.generate(model, M.TRUTH[model], n, seed=31, nstreams=64)
    theta = np.array(M.START[model])
    say("prefetch off", "model", model, "fp32", fp32, "n", n)
    stored = _stored(model, data, fp32)
    ref = O.nll_sums(model, stored, theta, mode=2, nthreads=8)
    absref = O.nll_sums(model, stored, theta, mode=12, nthreads=8)
    tol = 2e-6 if fp32 else 1e-12
    outs = []
    for pf in (0, 1):
        cx = M.Context(prefetch=pf)
        mdl = cx.model(model, fp32=fp32, with_hessian=False)
        b = mdl.batch([n])
        b.upload(0, data)
        r = [b.eval(kind, theta)[0] for kind in (M.EVAL_LH, M.EVAL_GRAD)]
        for out in r:
            k = len(out)
            assert np.all(np.abs(out - ref[:k]) <= tol * absref[:k]), (pf, out - ref[:k], absref[:k])
        outs.append((r, mdl.unroll[M.EVAL_LH]))
        b.close()
        mdl.close()
        cx.close()
    if outs[0][1] == outs[1][1]:
        assert np.array_equal(outs[0][0][0], outs[1][0][0])


@pytest.mark.parametrize("fp32", [False, True])
def test_data_moments_follow_new_data(ctx, fp32):
    """the mass density's pivot exponent (lambda m) is summed from per-segment data moments
    (sum m, computed once per upload or generation): new data in a slot -- uploaded, or
    generated on the device -- must give the sums of the new data, not the old ones"""
    n = 200_001
    mdl = ctx.model(M.MASS, fp32=fp32)
    assert "#define MF_NMOM 1" in mdl.source()
    b = mdl.batch([n, 5000])
    theta = np.array(M.START[M.MASS])
    tol = 2e-6 if fp32 else 1e-12

    def check(data, seg):
        stored = _stored(M.MASS, data, fp32)
        ref = O.nll_sums(M.MASS, stored, theta, mode=2, nthreads=8)
        absref = O.nll_sums(M.MASS, stored, theta, mode=12, nthreads=8)
        for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
            out = b.eval(kind, theta)[seg]
            k = len(out)
            assert np.all(np.abs(out - ref[:k]) <= tol * absref[:k]), (kind, seg, out - ref[:k])

    small = O.generate(M.MASS, M.MASS_TRUTH, 5000, seed=8, nstreams=4)
    b.upload(1, small)
    for seed in (3, 4):
        # a shifted sample the second time, so its sum of m differs clearly from the first
        data = O.generate(M.MASS, M.MASS_TRUTH if seed == 3 else np.array([5.1, 0.03, 0.6, 2.0]), n, seed=seed, nstreams=16)
        say("moments", "fp32", fp32, "upload seed", seed)
        b.upload(0, data)
        check(data, 0)
        check(small, 1)
    # generated on the device into both slots, read back: the sums are those of the downloaded data
    b.generate(np.array([5.3, 0.05, 0.5, -3.0]), [99, 100])
    check(b.download(0), 0)
    check(b.download(1), 1)
    b.close()
    mdl.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_mass_at_lambda_zero_matches_oracle(ctx, fp32):
    """lambda = 0 takes the ConditionalEqual branch of the exponential's normalisation
    (reference pdf.hh:437-439): b - a, whose derivative rule contributes nothing at lambda = 0
    exactly (pinned by the lambda = 0 points of tests/golden/ref_graph.json).  Through both
    places the hoisted terms are computed: the host program of a one-segment call (passed by
    value) and the device prep kernel of a two-segment batch."""
    n = 60000
    theta = np.array([5.3, 0.03, 0.5, 0.0])
    data = O.generate(M.MASS, theta, n, seed=99, nstreams=8)
    stored = _stored(M.MASS, data, fp32)
    ref = O.nll_sums(M.MASS, stored, theta, mode=2, nthreads=0)
    absref = O.nll_sums(M.MASS, stored, theta, mode=12, nthreads=4)
    tol = 2e-6 if fp32 else 1e-12
    mdl = ctx.model(M.MASS, fp32=fp32)
    one = mdl.batch([n])
    one.upload(0, data)
    two = mdl.batch([n, 1000])
    two.upload(0, data)
    two.upload(1, data[:, :1000])
    for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
        say("lambda = 0", "fp32", fp32, "kind", kind)
        for b in (one, two):
            out = b.eval(kind, theta)[0]
            k = len(out)
            assert np.all(np.isfinite(out))
            assert np.all(np.abs(out - ref[:k]) <= tol * absref[:k]), (kind, out - ref[:k], absref[:k])
    one.close()
    two.close()
    mdl.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_event_columns_match_oracle(fp32):
    """Angular density with event-only terms precomputed into data columns (event_columns=1):
    after an upload, every sum within the bars of test_eval_matches_oracle; after generation
    on the device, the same sums as a batch without columns holding the same stored events
    (within the FP32 bar for both precisions: the columns are computed in FP64 and rounded once
    in place of per-call FP32 arithmetic)."""
    model, n = M.KSTAR, 200_000
    data = O.generate(model, M.TRUTH[model], n, seed=4242, nstreams=16)
    theta = np.array(M.START[model])
    say("event columns fp32", fp32)
    cx = M.Context(event_columns=1)
    mdl = cx.model(model, fp32=fp32)
    assert mdl.ncol > mdl.nobs
    b = mdl.batch([n])
    b.upload(0, data)
    stored = _stored(model, data, fp32)
    ref = O.nll_sums(model, stored, theta, mode=2, nthreads=8)
    absref = O.nll_sums(model, stored, theta, mode=12, nthreads=8)
    tol = 2e-6 if fp32 else 1e-12
    for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
        out = b.eval(kind, theta)[0]
        k = len(out)
        assert np.all(np.abs(out - ref[:k]) <= tol * absref[:k]), (kind, (out - ref[:k]) / absref[:k])
    # generated on the device: columns filled by the precompute kernel after generation
    g = mdl.batch([n])
    g.generate(M.TRUTH[model], [77])
    ev = g.download(0)
    plain = ctx_plain = M.Context()
    pm = plain.model(model, fp32=fp32)
    pb = pm.batch([n])
    pb.upload(0, ev)
    gabs = O.nll_sums(model, _stored(model, ev, fp32), theta, mode=12, nthreads=8)
    for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
        a = g.eval(kind, theta)[0]
        r = pb.eval(kind, theta)[0]
        assert np.all(np.abs(a - r) <= 2e-6 * gabs[:len(a)]), (kind, (a - r) / gabs[:len(a)])
    for h in (pb, pm, ctx_plain, g, b, mdl, cx):
        h.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_fit_plain_optimize_dimensions(fp32):
    """The drop-in fitter flow with optimize_dimensions (HIPBlock store kernel filling the
    expanded data buffer, fitter.hh:339-375): the fit equals the one without it (FP64: 1e-9
    sigma, FP32: 1e-4 sigma)"""
    data = O.generate(M.KSTAR, M.TRUTH[M.KSTAR], 50000, seed=6, nstreams=8)
    say("fit_plain optimize_dimensions fp32", fp32)
    cp, ce = M.Context(), M.Context(event_columns=1)
    a = cp.fit_plain(M.KSTAR, data, M.START[M.KSTAR], fp32=fp32)
    b = ce.fit_plain(M.KSTAR, data, M.START[M.KSTAR], fp32=fp32)
    assert a[2].status == 0 and b[2].status == 0
    assert np.max(np.abs(a[0] - b[0]) / a[1]) < (1e-4 if fp32 else 1e-9)
    cp.close()
    ce.close()


@pytest.mark.parametrize("model", [M.MASS, M.KSTAR])
def test_fp64_base2_matches_oracle(model):
    """FP64 kernels with exp/log rewritten as exp2/log2 (option base2=2) keep the FP64 bar, 1e-12."""
    n = 100000
    data = O.generate(model, M.TRUTH[model], n, seed=4321, nstreams=16)
    theta = np.array(M.START[model])
    say("fp64 base2 parity", "model", model, "n", n)
    cx = M.Context(base2=2)
    mdl = cx.model(model, fp32=False)
    b = mdl.batch([n])
    b.upload(0, data)
    ref = O.nll_sums(model, data, theta, mode=2, nthreads=0)
    absref = O.nll_sums(model, data, theta, mode=12, nthreads=4)
    for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
        out = b.eval(kind, theta)[0]
        k = len(out)
        assert abs(out[0] - ref[0]) <= 1e-12 * abs(ref[0])
        assert np.all(np.abs(out - ref[:k]) <= 1e-12 * absref[:k]), (kind, out - ref[:k])
    b.close()
    mdl.close()
    cx.close()


@pytest.mark.parametrize("model", [M.MASS, M.KSTAR])
def test_fp32_library_math_matches_oracle(model):
    """FP32 kernels with the device library's exp, log, sqrt, sin and cos (option fast_math=0) in place of
    the hardware instructions keep the FP32 bar, 2e-6 of the sum of absolute terms."""
    n = 100000
    data = O.generate(model, M.TRUTH[model], n, seed=2468, nstreams=16)
    theta = np.array(M.START[model])
    say("fp32 library math parity", "model", model, "n", n)
    cx = M.Context(fast_math=0)
    mdl = cx.model(model, fp32=True)
    b = mdl.batch([n])
    b.upload(0, data)
    stored = _stored(model, data, True)
    ref = O.nll_sums(model, stored, theta, mode=2, nthreads=0)
    absref = O.nll_sums(model, stored, theta, mode=12, nthreads=4)
    for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
        out = b.eval(kind, theta)[0]
        k = len(out)
        assert np.all(np.abs(out - ref[:k]) <= 2e-6 * absref[:k]), (kind, out - ref[:k], absref[:k])
    b.close()
    mdl.close()
    cx.close()


@pytest.mark.parametrize("pivot", [1, 0])
@pytest.mark.parametrize("fp32", [False, True])
def test_exponential_pivot_matches_oracle(fp32, pivot):
    """the mass density evaluated around its exponential term (exp_pivot=1: exp, log and rcp per
    event) and as written (exp_pivot=0: two exp, log and rcp), at the start point and at points
    on the parameter limits: FP64 within 1e-12, FP32 within 2e-6 of the sums of absolute values.
    At the limit points the FP64 Hessian is held to 1e-11: there its lambda-lambda entry is
    ill-conditioned, and the oracle and the kernels' graphs are each ~7e-13 off the 50-digit value
    (test_cpu_pipeline.py::test_hessian_conditioning_at_parameter_limits).  FP32 is held to 1e-5
    there: with sigma = 0.005 and mu at the range edge, the FP32 shift 5.3 - mu (absolute
    rounding ~1.5e-8) resolves the distances m - mu ~ sigma to ~3e-6"""
    n = 1_000_003
    data = O.generate(M.MASS, M.TRUTH[M.MASS], n, seed=91, nstreams=64)
    say("exp pivot", pivot, "fp32", fp32, "n", n)
    cx = M.Context(exp_pivot=pivot)
    mdl = cx.model(M.MASS, fp32=fp32)
    b = mdl.batch([n])
    b.upload(0, data)
    stored = _stored(M.MASS, data, fp32)
    thetas = [np.array(M.START[M.MASS]), np.array([5.0, 0.005, 0.02, -10.0]), np.array([5.6, 0.13, 0.98, 10.0]),
              np.array([5.3, 0.02, 0.999, -9.9])]
    for t, theta in enumerate(thetas):
        ref = O.nll_sums(M.MASS, stored, theta, mode=2, nthreads=8)
        absref = O.nll_sums(M.MASS, stored, theta, mode=12, nthreads=8)
        tol = np.full(len(ref), 2e-6 if fp32 else 1e-12)
        if t > 0:
            if fp32:
                tol[:] = 1e-5
            else:
                tol[5:] = 1e-11
        for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
            out = b.eval(kind, theta)[0]
            k = len(out)
            assert np.all(np.abs(out - ref[:k]) <= tol[:k] * absref[:k]), (theta, kind, (out - ref[:k]) / absref[:k])
    b.close()
    mdl.close()
    cx.close()


@pytest.mark.parametrize("model", [M.MASS, M.KSTAR])
@pytest.mark.parametrize("accumulate", [0, 1, 2])
def test_fp32_accumulation_modes(model, accumulate):
    """FP32 kernels with each per-lane accumulation (0: FP64 per event, 1: Kahan FP32, 2: FP32
    sums of each unrolled group added in FP64), at a size with many chunks per segment and
    a ragged tail: every sum within 2e-6 of the sum of absolute per-event values"""
    n = 1_000_003
    data = O.generate(model, M.TRUTH[model], n, seed=77, nstreams=64)
    theta = np.array(M.START[model])
    say("fp32 accumulation", accumulate, "model", model, "n", n)
    cx = M.Context(accumulate=accumulate)
    mdl = cx.model(model, fp32=True)
    b = mdl.batch([n])
    b.upload(0, data)
    stored = _stored(model, data, True)
    ref = O.nll_sums(model, stored, theta, mode=2, nthreads=8)
    absref = O.nll_sums(model, stored, theta, mode=12, nthreads=8)
    for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
        out = b.eval(kind, theta)[0]
        k = len(out)
        assert np.all(np.abs(out - ref[:k]) <= 2e-6 * absref[:k]), (kind, (out - ref[:k]) / absref[:k])
    b.close()
    mdl.close()
    cx.close()


def test_graph_replays_match_eager(ctx):
    """hipGraph replays (several, with eager work between them) == eager launches: bitwise when
    both prepare the parameters on the device (host_params=0); with the default one-segment path
    (hoisted terms computed on the host and passed by value) within 1e-12 per output"""
    n = 300000
    data = O.generate(M.MASS, M.MASS_TRUTH, n, seed=5, nstreams=16)
    thetas = [M.MASS_START, M.MASS_TRUTH, np.array([5.3, 0.025, 0.25, -1.2])]
    cdev = M.Context(host_params=0)
    for cx, exact in ((cdev, True), (ctx, False)):
        mdl = cx.model(M.MASS, fp32=True)
        b = mdl.batch([n])
        b.upload(0, data)
        say("graph replay, host parameters", not exact)
        for t in thetas:
            eager = b.eval(M.EVAL_GRAD, t, use_graph=False)
            for _ in range(3):
                g = b.eval(M.EVAL_GRAD, t, use_graph=True)
                if exact:
                    assert np.array_equal(g, eager)
                else:
                    assert np.all(np.abs(g - eager) <= 1e-12 * np.abs(eager) + 1e-14 * np.max(np.abs(eager))), g - eager
                b.eval(M.EVAL_LH, t, use_graph=False)
        b.close()
        mdl.close()
    cdev.close()


@pytest.mark.parametrize("fp32", [True, False])
def test_polled_completion_returns_complete_totals(ctx, fp32):
    """A one-segment multi-chunk batch: the final kernel publishes a completion flag that the host
    polls instead of waiting on the stream.  200 calls alternating two parameter points, all
    kinds: each result bitwise equal to the first polled result at that point and within 1e-12
    of the stream-synchronised result there (a flag published before the totals would show the
    other point's totals).  Every 64th polled call also checks the stream without waiting
    (mf_query), so the 600 calls run that path nine times."""
    n = 1_000_003
    data = O.generate(M.MASS, M.MASS_TRUTH, n, seed=11, nstreams=16)
    mdl = ctx.model(M.MASS, fp32=fp32)
    b = mdl.batch([n])
    b.upload(0, data)
    thetas = [np.array(M.MASS_START), np.array([5.27, 0.021, 0.31, -1.4])]
    say("polled completion fp32", fp32)
    for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
        want = [b.eval(kind, t, use_graph=True) for t in thetas]
        assert not np.array_equal(want[0], want[1])
        first = [None, None]
        for i in range(200):
            got = b.eval(kind, thetas[i % 2])
            w = want[i % 2]
            # stream-synchronised graph result (parameters prepared on the device): equal to the
            # polled call's within the host/device hoisted-term rounding; the other point's
            # totals differ by far more
            assert np.all(np.abs(got - w) <= 1e-12 * np.abs(w) + 1e-14 * np.max(np.abs(w))), (kind, i, got - w)
            if first[i % 2] is None:
                first[i % 2] = got
            assert np.array_equal(got, first[i % 2]), (kind, i)
    b.close()
    mdl.close()


@pytest.mark.parametrize("fused", [1, 0])
@pytest.mark.parametrize("fp32", [False, True])
def test_segment_reduction_in_kernel_and_final_pass(fp32, fused):
    """Multi-chunk segments reduced by the work-group that completes them (fused=1: per-segment
    counters that the kernel sets back to zero) or by the separate final kernel (fused=0):
    three evaluations in a row at different parameters, each kind, a batch of multi-chunk,
    exactly-one-chunk, tiny and empty segments -- every total within the oracle bars (a counter
    left non-zero would leave the previous call's totals in place and fail the later rounds)."""
    sizes = [1_000_003, 0, 77, 8192, 8193, 300_001]
    datas = [O.generate(M.MASS, M.MASS_TRUTH, n, seed=500 + i, nstreams=8) if n else None for i, n in enumerate(sizes)]
    say("segment reduction fused", fused, "fp32", fp32, "sizes", sizes)
    cx = M.Context(fused_final=fused)
    mdl = cx.model(M.MASS, fp32=fp32)
    b = mdl.batch(sizes)
    for i, d in enumerate(datas):
        if d is not None:
            b.upload(i, d)
    tol = 2e-6 if fp32 else 1e-12
    for rep in range(3):
        theta = np.array(M.MASS_START) * (1.0 + 1e-3 * rep)
        for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
            out = b.eval(kind, theta)
            for i, d in enumerate(datas):
                if d is None:
                    assert np.all(out[i] == 0)
                    continue
                st = _stored(M.MASS, d, fp32)
                k = out.shape[1]
                ref = O.nll_sums(M.MASS, st, theta, mode=2, nthreads=8)[:k]
                absref = O.nll_sums(M.MASS, st, theta, mode=12, nthreads=8)[:k]
                assert np.all(np.abs(out[i] - ref) <= tol * absref), (rep, kind, i, out[i] - ref)
    b.close()
    mdl.close()
    cx.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_segments_alone_equal_mixed(ctx, fp32):
    """ragged segments (0, 1, 63, 4096, 4097, 100000 events): each segment's sums are bitwise
    identical evaluated alone and together with the others when both prepare the parameters on
    the device (host_params=0), within 1e-12 of the sums of absolute values with the default
    one-segment path (hoisted terms computed on the host, passed by value: an ulp of difference
    in a hoisted normalisation term reaches ~1e-13 in a one-event Hessian), and match the
    oracle."""
    sizes = [0, 1, 63, 4096, 4097, 100000]
    rs = np.random.default_rng(9)
    datas = [O.generate(M.MASS, M.MASS_TRUTH, s, seed=100 + i, nstreams=2) if s else np.zeros((1, 0)) for i, s in enumerate(sizes)]
    thetas = np.array([M.MASS_START + rs.normal(0, 1e-3, 4) * [1, 0.1, 1, 1] for _ in sizes])
    say("segments", sizes, "fp32", fp32)
    mdl = ctx.model(M.MASS, fp32=fp32)
    cdev = M.Context(host_params=0)
    mdev = cdev.model(M.MASS, fp32=fp32)
    mixed = mdl.batch(sizes)
    for i, d in enumerate(datas):
        if sizes[i]:
            mixed.upload(i, d)
    both = mixed.eval(M.EVAL_HESS, thetas)
    for i, d in enumerate(datas):
        alone_dev = mdev.batch([sizes[i]])
        alone = mdl.batch([sizes[i]])
        if sizes[i]:
            alone_dev.upload(0, d)
            alone.upload(0, d)
        for _ in range(2):
            r = alone_dev.eval(M.EVAL_HESS, thetas[i:i + 1])
            assert np.array_equal(r[0], both[i]), (i, r[0] - both[i])
        alone_dev.close()
        if sizes[i] == 0:
            assert np.all(r[0] == 0)
            absref = np.zeros_like(r[0])
        else:
            ref = O.nll_sums(M.MASS, _stored(M.MASS, d, fp32), thetas[i], mode=2, nthreads=0)
            absref = O.nll_sums(M.MASS, _stored(M.MASS, d, fp32), thetas[i], mode=12, nthreads=1)
            tol = 2e-6 if fp32 else 1e-12
            assert np.all(np.abs(r[0] - ref) <= tol * absref + 1e-300)
        # the default one-segment path: theta and host-computed hoisted terms by value (no prep
        # kernel), the same sums up to the hoisted terms' rounding, bitwise the same call to call
        r1 = alone.eval(M.EVAL_HESS, thetas[i:i + 1])
        for _ in range(2):
            r2 = alone.eval(M.EVAL_HESS, thetas[i:i + 1])
            assert np.all(np.abs(r2[0] - both[i]) <= 1e-12 * absref), (i, r2[0] - both[i])
            assert np.array_equal(r2, r1)
        alone.close()
    mixed.close()
    mdev.close()
    cdev.close()
    mdl.close()


@pytest.mark.parametrize("model", [M.MASS, M.KSTAR])
def test_generator_fp64_matches_oracle_stream(ctx, model):
    """on-device accept-reject in FP64 reproduces the oracle's xoshiro128++ streams"""
    n, eps = 20000, 64
    seed = 229387429
    say("generator", model)
    mdl = ctx.model(model, fp32=False)
    b = mdl.batch([n])
    props = b.generate(M.TRUTH[model], [seed], events_per_stream=eps)
    got = b.download(0)
    ref = O.generate(model, M.TRUTH[model], n, seed=seed, events_per_stream=eps)
    same = np.mean(np.all(got == ref, axis=0))
    assert same > 0.999, same
    assert props >= n
    b.close()
    mdl.close()


def test_generator_fp32_statistics(ctx):
    """FP32 on-device generation: right count, inside the range, mean/width as the oracle sample"""
    n = 200000
    say("generator fp32")
    mdl = ctx.model(M.MASS, fp32=True)
    b = mdl.batch([n, 5000, 1])
    b.generate(M.MASS_TRUTH, [1, 2, 3])
    g = b.download(0)[0]
    assert g.min() >= 5.0 and g.max() <= 5.6
    ref = O.generate(M.MASS, M.MASS_TRUTH, n, seed=77, nstreams=8)[0]
    assert abs(g.mean() - ref.mean()) < 5 * ref.std() / np.sqrt(n) * 1.5
    assert abs(g.std() - ref.std()) < 0.01 * ref.std()
    b.close()
    mdl.close()


def test_fit_fp64_gpu_equals_cpu_backend(ctx):
    n = 100000
    data = O.generate(M.MASS, M.MASS_TRUTH, n, seed=229387429, nstreams=16)
    say("plain fit fp64 gpu")
    v, e, out, ms = ctx.fit_plain(M.MASS, data, M.MASS_START, fp32=False)
    assert out.status == 0
    vc, ec, outc, msc = M.fit_cpu(M.MASS, data, M.MASS_START, nthreads=4)
    assert np.max(np.abs(v - vc) / ec) < 1e-4
    np.testing.assert_allclose(e, ec, rtol=1e-6)
    say("plain fit fp32 gpu")
    v32, e32, out32, _ = ctx.fit_plain(M.MASS, data, M.MASS_START, fp32=True)
    assert out32.status == 0
    assert np.max(np.abs(v32 - vc) / ec) < 1e-3


@pytest.mark.parametrize("fp32", [False, True])
def test_batched_fits_match_single_fits(ctx, fp32):
    """each toy fitted inside one batched study == the same toy fitted alone (CPU backend,
    FP64) within 1e-3 sigma (FP32) / 1e-4 sigma (FP64)"""
    rs = np.random.default_rng(4)
    sizes = list(rs.poisson(5000, 24))
    datas = [O.generate(M.MASS, M.MASS_TRUTH, int(s), seed=1000 + i, nstreams=4) for i, s in enumerate(sizes)]
    say("batched fits", len(sizes), "fp32", fp32)
    mdl = ctx.model(M.MASS, fp32=fp32)
    b = mdl.batch(sizes)
    for i, d in enumerate(datas):
        b.upload(i, d)
    starts = M.toy_starts(M.MASS, sizes, seed=4)
    vals, errs, outs, stats = b.fit(starts)
    lim = 1e-3 if fp32 else 1e-4
    for i, d in enumerate(datas):
        assert outs[i].status == 0, i
        vc, ec, oc, _ = M.fit_cpu(M.MASS, d, starts[i], nthreads=1)
        assert oc.status == 0 and np.all(ec > 0), i
        assert np.max(np.abs(vals[i] - vc) / ec) < lim, (i, (vals[i] - vc) / ec)
        np.testing.assert_allclose(errs[i], ec, rtol=1e-3 if fp32 else 1e-6)
        # each fit's fibre stayed on the host thread it started on, also once the live count
        # dropped below the worker count
        assert outs[i].thread_moved == 0, i
    assert stats.ticks > 0
    b.close()
    mdl.close()


def test_batch_fit_rejects_bad_segment_lists(ctx):
    """an index outside the batch or listed twice is refused before any fit runs (two fits
    of one segment would share its parameter slot)"""
    mdl = ctx.model(M.MASS, fp32=True)
    b = mdl.batch([1000, 1000])
    say("bad segment lists")
    for segs in ([0, 2], [-1], [1, 1]):
        with pytest.raises(RuntimeError, match="outside|twice"):
            b.fit(M.MASS_START, segs=segs)
    b.close()
    mdl.close()


def test_kstar_fit_gpu(ctx):
    n = 200000
    data = O.generate(M.KSTAR, M.KSTAR_TRUTH, n, seed=7, nstreams=16)
    say("kstar plain fit")
    v, e, out, ms = ctx.fit_plain(M.KSTAR, data, M.KSTAR_START, fp32=False)
    assert out.status == 0
    vc, ec, oc, _ = M.fit_cpu(M.KSTAR, data, M.KSTAR_START, nthreads=4)
    assert np.max(np.abs(v - vc) / ec) < 1e-4
    mdl = ctx.model(M.KSTAR, fp32=True)
    b = mdl.batch([n])
    b.upload(0, data)
    vals, errs, outs, stats = b.fit(M.KSTAR_START)
    assert outs[0].status == 0
    assert np.max(np.abs(vals[0] - vc) / ec) < 1e-3
    b.close()
    mdl.close()


def test_full_size_properties(ctx):
    """N = 1e7 (the headline configuration): FP32 and FP64 agree within the FP32 bar,
    repeated evaluations are bitwise identical, the value sum is extensive (two halves)."""
    n = 10_000_000
    say("full size generate+eval", n)
    m64 = ctx.model(M.MASS, fp32=False)
    b64 = m64.batch([n])
    b64.generate(M.MASS_TRUTH, [42])
    data = b64.download(0)
    r1 = b64.eval(M.EVAL_GRAD, M.MASS_START)
    r2 = b64.eval(M.EVAL_GRAD, M.MASS_START)
    assert np.array_equal(r1, r2)
    halves = m64.batch([n // 2, n - n // 2])
    halves.upload(0, data[:, : n // 2])
    halves.upload(1, data[:, n // 2:])
    rh = halves.eval(M.EVAL_GRAD, M.MASS_START)
    assert abs(rh[:, 0].sum() - r1[0, 0]) <= 1e-12 * abs(r1[0, 0])
    m32 = ctx.model(M.MASS, fp32=True)
    b32 = m32.batch([n])
    b32.upload(0, data)
    r32 = b32.eval(M.EVAL_GRAD, M.MASS_START)
    absref = O.nll_sums(M.MASS, data[:, :200000], M.MASS_START, mode=11, nthreads=8) * (n / 200000)
    assert np.all(np.abs(r32[0] - r1[0]) <= 4e-6 * absref), (r32[0] - r1[0]) / absref
    for x in (b64, halves, b32):
        x.close()
    m64.close()
    m32.close()


def test_autotuner_picks_a_candidate_and_results_hold():
    """the autotuner times every candidate, leaves the context at the fastest, and a model
    built afterwards (that launch shape) still matches the oracle"""
    n = 300_000
    cands = [(256, 8, 4), (256, 32, 8), (512, 16, 8)]
    cx = M.Context()
    say("autotune", cands)
    r = cx.autotune(M.MASS, True, [n, 7000, 1], eval_kind=M.EVAL_GRAD, candidates=cands, iters=5)
    us = [u for _, u in r["table"]]
    assert all(np.isfinite(u) and u > 0 for u in us)
    assert r["best"] == r["table"][int(np.argmin(us))][0]
    data = O.generate(M.MASS, M.MASS_TRUTH, n, seed=5, nstreams=16)
    theta = np.array(M.MASS_START)
    mdl = cx.model(M.MASS, fp32=True)
    b = mdl.batch([n])
    b.upload(0, data)
    out = b.eval(M.EVAL_GRAD, theta)[0]
    stored = _stored(M.MASS, data, True)
    ref = O.nll_sums(M.MASS, stored, theta, mode=1, nthreads=4)
    absref = O.nll_sums(M.MASS, stored, theta, mode=11, nthreads=4)
    assert np.all(np.abs(out - ref) <= 2e-6 * absref)
    b.close()
    mdl.close()
    cx.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_weighted_fit_gpu_equals_cpu_backend(ctx, fp32):
    """weighted fit with the covariance correction through the drop-in HIP block equals the CPU
    backend's (FP64: 1e-4 sigma, errors 1e-6; FP32: 1e-3 sigma, errors 1e-3)"""
    n = 50_000
    data = O.generate(M.MASS, M.MASS_TRUTH, n, seed=91, nstreams=16)
    w = np.random.default_rng(6).uniform(0.4, 1.6, n)
    say("weighted fit", "fp32", fp32)
    v, e, out, _ = M.fit_weighted(M.MASS, data, w, M.MASS_START, ctx=ctx, fp32=fp32)
    assert out.status == 0
    vc, ec, oc, _ = M.fit_weighted(M.MASS, data, w, M.MASS_START, nthreads=4)
    assert oc.status == 0
    assert np.max(np.abs(v - vc) / ec) < (1e-3 if fp32 else 1e-4)
    np.testing.assert_allclose(e, ec, rtol=1e-3 if fp32 else 1e-6)

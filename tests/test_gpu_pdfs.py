"""GPU parity of the Crystal Ball and polynomial densities (reference pdf.hh:195-419 and
716-892) through the C ABI, at the bars of the mass and angular densities:
  FP64 kernels: -2lnL within 1e-12 relative of the oracle's long-double sum, gradient and
    Hessian sums within 1e-12 of the sum of absolute per-event values;
  FP32 kernels: every sum within 2e-6 of the sum of absolute per-event values, against the
    oracle on the same stored (centred, rounded) FP32 observables;
  generator (FP64): the oracle's accept-reject streams;
  fits: the FP64 device fit equals the CPU-backend fit to 1e-4 sigma, batched fits equal
    single fits.
"""
import sys

import numpy as np
import pytest

import morefit_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MODELS = [M.CB, M.POLY]


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def _stored(model, data, fp32):
    if not fp32:
        return data
    c = O.RANGES[model].reshape(-1, 2).mean(axis=1)[:, None]
    return (data - c).astype(np.float32).astype(np.float64) + c


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("fp32", [False, True])
def test_eval_matches_oracle(ctx, model, fp32):
    n = 60000
    data = O.generate(model, M.TRUTH[model], n, seed=4321, nstreams=16)
    theta = np.array(M.START[model])
    say("pdf eval parity", "model", model, "fp32", fp32, "n", n)
    mdl = ctx.model(model, fp32=fp32)
    b = mdl.batch([n])
    b.upload(0, data)
    stored = _stored(model, data, fp32)
    ref = O.nll_sums(model, stored, theta, mode=2, nthreads=0)
    absref = O.nll_sums(model, stored, theta, mode=12, nthreads=4)
    tol = 2e-6 if fp32 else 1e-12
    for kind in (M.EVAL_LH, M.EVAL_GRAD, M.EVAL_HESS):
        out = b.eval(kind, theta)[0]
        k = len(out)
        if not fp32:
            assert abs(out[0] - ref[0]) <= 1e-12 * abs(ref[0])
        assert np.all(np.abs(out - ref[:k]) <= tol * absref[:k]), (kind, out - ref[:k], absref[:k])
    b.close()
    mdl.close()


@pytest.mark.parametrize("model", MODELS)
def test_generator_fp64_matches_oracle_stream(ctx, model):
    n, eps, seed = 20000, 64, 229387429
    say("pdf generator", model)
    mdl = ctx.model(model, fp32=False)
    b = mdl.batch([n])
    props = b.generate(M.TRUTH[model], [seed], events_per_stream=eps)
    got = b.download(0)
    ref = O.generate(model, M.TRUTH[model], n, seed=seed, events_per_stream=eps)
    assert np.mean(np.all(got == ref, axis=0)) > 0.999
    assert props >= n
    b.close()
    mdl.close()


@pytest.mark.parametrize("model", MODELS)
def test_fit_gpu_equals_cpu_backend(ctx, model):
    n = 100000
    data = O.generate(model, M.TRUTH[model], n, seed=229387429, nstreams=16)
    say("pdf plain fit", model)
    v, e, out, _ = ctx.fit_plain(model, data, M.START[model], fp32=False)
    assert out.status == 0
    vc, ec, outc, _ = M.fit_cpu(model, data, M.START[model], nthreads=4)
    assert outc.status == 0
    assert np.max(np.abs(v - vc) / ec) < 1e-4
    np.testing.assert_allclose(e, ec, rtol=1e-5)
    v32, e32, out32, _ = ctx.fit_plain(model, data, M.START[model], fp32=True)
    assert out32.status == 0
    assert np.max(np.abs(v32 - vc) / ec) < 1e-3


@pytest.mark.parametrize("model", MODELS)
def test_batched_fits_match_single_fits(ctx, model):
    sizes = [30000, 5000, 12345]
    say("pdf batched fits", model, sizes)
    mdl = ctx.model(model, fp32=False)
    b = mdl.batch(sizes)
    data = [O.generate(model, M.TRUTH[model], n, seed=11 + i, nstreams=4) for i, n in enumerate(sizes)]
    for i, d in enumerate(data):
        b.upload(i, d)
    start = M.toy_starts(model, sizes, seed=3)
    vals, errs, outs, stats = b.fit(start)
    for i, d in enumerate(data):
        vc, ec, outc, _ = M.fit_cpu(model, d, start[i], nthreads=4)
        assert outs[i].status == 0 and outc.status == 0, i
        assert np.max(np.abs(vals[i] - vc) / ec) < 1e-4, (i, (vals[i] - vc) / ec)
        np.testing.assert_allclose(errs[i], ec, rtol=1e-5)
    b.close()
    mdl.close()

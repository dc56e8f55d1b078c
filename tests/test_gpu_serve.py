"""The fit service: requests of both classes arriving while others are on the GPU are
admitted under an HBM budget, their evaluations share launches, and every result is the
same alone and in the mix (bitwise: a segment's sums depend only on its own chunks).
The configuration mix (70 % mass / 30 % angular, 1e3..2e6 events log-uniform, Poisson
arrivals, smeared start values) is checked against FP64 fits of the same data."""
import sys

import numpy as np
import pytest

import morefit_amd as M

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def test_service_alone_equals_mixed(ctx):
    rs = np.random.default_rng(17)
    nreq = 24
    cls = (rs.random(nreq) < 0.3).astype(int)  # 30 % angular
    n = np.exp(rs.uniform(np.log(1e3), np.log(2e5), nreq)).astype(np.int64)
    seeds = rs.integers(1, 2**62, nreq)
    arrival = np.cumsum(rs.exponential(0.002, nreq))
    say("service models")
    mass = ctx.model(M.MASS, fp32=True)
    kst = ctx.model(M.KSTAR, fp32=True)
    models = [mass, kst]
    tg = [M.MASS_TRUTH, M.KSTAR_TRUTH]
    st = [M.MASS_TRUTH, M.KSTAR_TRUTH]  # start at the truth: requests as small as 1e3 events
    say("service mixed", nreq, "requests")
    mixed = M.serve(models, tg, st, cls, n, seeds, arrival, budget_bytes=256 << 20, slots=64)
    assert np.all(mixed["status"] == 0), mixed["status"]
    assert mixed["launches"] < sum(1 for _ in range(nreq)) * 40  # launches are shared
    assert mixed["peak_live"] > 1
    assert np.all(mixed["done"] >= mixed["admit"]) and np.all(mixed["admit"] >= mixed["arrival"] - 1e-9)
    for i in range(0, nreq, 5):
        say("service alone", i)
        alone = M.serve(models, tg, st, [cls[i]], [n[i]], [seeds[i]], [0.0], budget_bytes=256 << 20, slots=4)
        P = M.NTHETA[cls[i]]
        assert np.array_equal(alone["values"][0, :P], mixed["values"][i, :P]), (i, alone["values"][0, :P] - mixed["values"][i, :P])
        assert np.array_equal(alone["errors"][0, :P], mixed["errors"][i, :P])
    mass.close()
    kst.close()


def test_service_budget_limits_concurrency(ctx):
    """a budget that holds about two datasets: fits queue, all finish, oversize ones are rejected"""
    mass = ctx.model(M.MASS, fp32=True)
    n = [100000] * 6 + [10_000_000]
    budget = 2.2 * 100000 * 4
    say("service budget")
    r = M.serve([mass], [M.MASS_TRUTH], [M.MASS_TRUTH], [0] * len(n), n, list(range(1, len(n) + 1)), [0.0] * len(n),
                budget_bytes=budget, slots=16)
    assert r["peak_bytes"] <= budget
    assert r["rejected"] == 1 and r["status"][-1] == -2
    assert np.all(r["status"][:-1] == 0)
    assert r["peak_live"] <= 2
    mass.close()


# the generator stream lengths the fit service uses (FitServer::small_request_events,
# events_per_stream_small / _large): a request's data depend only on its size and seed
SERVE_SMALL_REQUEST_EVENTS, SERVE_EPS_SMALL, SERVE_EPS_LARGE = 65536, 16, 64


def _regenerate(model, kind, sizes, seeds):
    """the datasets the service generated for these requests, downloaded as FP64"""
    out = [None] * len(sizes)
    for large in (False, True):
        idx = [i for i, n in enumerate(sizes) if (n > SERVE_SMALL_REQUEST_EVENTS) == large]
        if not idx:
            continue
        b = model.batch([int(sizes[i]) for i in idx])
        b.generate(np.tile(M.TRUTH[kind], (len(idx), 1)), [int(seeds[i]) for i in idx],
                   events_per_stream=SERVE_EPS_LARGE if large else SERVE_EPS_SMALL)
        for j, i in enumerate(idx):
            out[i] = b.download(j)
        b.close()
    return out


def test_service_configuration_mix_matches_fp64_fits(ctx):
    """about 40 requests of the configuration mix, two of them above 1e6 events, each started
    from its own smeared point: every served FP32 fit converges and lands within 1e-3 sigma of
    the FP64 fit of the same data from the same start, errors within 1e-3 relative (reference
    behaviour per request: fitter.hh:660-1002)"""
    rs = np.random.default_rng(2025)
    nreq = 40
    cls = (rs.random(nreq) < 0.3).astype(int)
    n = np.exp(rs.uniform(np.log(1e3), np.log(2e6), nreq)).astype(np.int64)
    n[rs.choice(np.flatnonzero(cls == 0), 1)] = 2_000_000
    n[rs.choice(np.flatnonzero(cls == 1), 1)] = 1_500_000
    assert np.sum(n > 1_000_000) >= 2
    seeds = rs.integers(1, 2**62, nreq)
    arrival = np.cumsum(rs.exponential(1.0 / 2000.0, nreq))
    starts = [M.toy_starts(int(c), [int(k)], seed=int(sd))[0] for c, k, sd in zip(cls, n, seeds)]
    say("config-5 mix models")
    mass = ctx.model(M.MASS, fp32=True)
    kst = ctx.model(M.KSTAR, fp32=True)
    models = [mass, kst]
    tg = [M.MASS_TRUTH, M.KSTAR_TRUTH]
    say("config-5 mix served", nreq, "requests, largest", int(n.max()))
    r = M.serve(models, tg, tg, cls, n, seeds, arrival, budget_bytes=1 << 30, slots=64, request_start=starts)
    assert np.all(r["status"] == 0), r["status"]
    assert r["peak_live"] > 1
    for kind, mdl in ((M.MASS, mass), (M.KSTAR, kst)):
        idx = np.flatnonzero(cls == kind)
        say("config-5 FP64 refits, class", kind, len(idx), "requests")
        datas = _regenerate(mdl, kind, [int(n[i]) for i in idx], [int(seeds[i]) for i in idx])
        m64 = ctx.model(kind, fp32=False)
        b64 = m64.batch([int(n[i]) for i in idx])
        for j, d in enumerate(datas):
            assert d.shape[1] == n[idx[j]]
            b64.upload(j, d)
        v64, e64, o64, _ = b64.fit(np.array([starts[i] for i in idx]))
        P = M.NTHETA[kind]
        for j, i in enumerate(idx):
            assert o64[j].status == 0, (i, o64[j].status)
            pull = (r["values"][i, :P] - v64[j]) / e64[j]
            assert np.max(np.abs(pull)) < 1e-3, (i, int(n[i]), pull)
            np.testing.assert_allclose(r["errors"][i, :P], e64[j], rtol=1e-3)
        b64.close()
        m64.close()
    mass.close()
    kst.close()


def test_service_classes_share_one_budget(ctx):
    """two classes under one budget: the event data of both live in one arena of the budget's
    size (not one per class), and the admitted data (all columns) never exceed it"""
    mass = ctx.model(M.MASS, fp32=True)
    kst = ctx.model(M.KSTAR, fp32=True)
    n = [100000, 60000, 100000, 60000, 100000, 60000]
    cls = [0, 1, 0, 1, 0, 1]
    budget = 4 * (2 * 100000 + 60000 * 3) * 1.1  # about one dataset of each class at a time
    say("service shared budget")
    r = M.serve([mass, kst], [M.MASS_TRUTH, M.KSTAR_TRUTH], [M.MASS_TRUTH, M.KSTAR_TRUTH], cls, n,
                list(range(11, 17)), [0.0] * len(n), budget_bytes=budget, slots=8)
    assert np.all(r["status"] == 0)
    assert r["arena_bytes"] <= budget
    assert r["peak_bytes"] <= budget
    assert r["peak_bytes"] >= 4 * 60000 * 3  # an angular request counts its three columns
    mass.close()
    kst.close()

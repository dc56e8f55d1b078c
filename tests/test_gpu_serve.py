"""The fit service: requests of both classes arriving while others are on the GPU are
admitted under an HBM budget, their evaluations share launches, and every result is the
same alone and in the mix (bitwise: a segment's sums depend only on its own chunks)."""
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

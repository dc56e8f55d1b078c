"""Pins the FP64 oracle (oracle/) before anything is checked against it.

Sources of truth, none of them a retyped formula:
  * golden vectors written by the reference's own graph.hh (symbolic log p, d/dtheta,
    d2/dtheta2 of both configuration densities) and random.hh (PRNG streams, Poisson),
    committed in tests/golden/ by tests/golden/make_golden.py;
  * central differences (the reference's own check, fitter.hh:1012-1047, eps = 1e-6);
  * normalisation by numerical quadrature (integral of p over the range = 1);
  * the exact sum (math.fsum) for the summation semantics;
  * goodness of fit of the accept-reject sample to the density.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import integrate, stats

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _relmax(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


@pytest.mark.parametrize("name,model", [("mass", O.MASS), ("kstar", O.KSTAR), ("cb", O.CB), ("poly", O.POLY)])
def test_oracle_matches_reference_graph(name, model):
    """log p, gradient and Hessian of the oracle == the reference graph.hh's own symbolic
    derivatives evaluated by its interpreter (tests/golden/ref_graph.json)."""
    G = _load("ref_graph.json")
    P = O.NPAR[model]
    for pt in G[name]["points"]:
        lp, g, h = O.event_jet(model, pt["x"], pt["theta"])
        packed = [h[i, j] for i in range(P) for j in range(i + 1)]
        assert abs(lp - pt["logp"]) <= 1e-13 * max(1.0, abs(pt["logp"]))
        assert _relmax(g, pt["grad"]) < 1e-12
        assert _relmax(packed, pt["hess"]) < 1e-11
        lp2, g2 = O.event_grad(model, pt["x"], pt["theta"])
        assert abs(lp2 - pt["logp"]) <= 1e-13 * max(1.0, abs(pt["logp"]))
        assert _relmax(g2, pt["grad"]) < 1e-12


def test_mass_range_golden():
    G = _load("ref_graph.json")
    assert G["mass_range"] == [5.0, 5.6] or abs(G["mass_range"][1] - 5.6) < 1e-15


@pytest.mark.parametrize("model", [O.MASS, O.KSTAR])
def test_central_differences(model):
    """closed-form gradient vs central differences of log p; jet Hessian vs central
    differences of the gradient (eps = 1e-6 as fitter.hh:1012-1047)."""
    rs = np.random.default_rng(7)
    P = O.NPAR[model]
    th0 = np.array([5.28, 0.02, 0.3, -1.5]) if model == O.MASS else np.array([0.7, 0.05, -0.1, -0.2, 0.1, 0.02, -0.03, 0.01])
    for _ in range(10):
        if model == O.MASS:
            x = [5.0 + 0.6 * rs.random()]
        else:
            x = [2 * rs.random() - 1, 2 * rs.random() - 1, np.pi * (2 * rs.random() - 1)]
        lp, g = O.event_grad(model, x, th0)
        _, gj, H = O.event_jet(model, x, th0)
        for k in range(P):
            eps = 1e-6 * max(1.0, abs(th0[k]))
            tp, tm = th0.copy(), th0.copy()
            tp[k] += eps
            tm[k] -= eps
            fd = (O.event_grad(model, x, tp)[0] - O.event_grad(model, x, tm)[0]) / (2 * eps)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k]))
            fdg = (O.event_grad(model, x, tp)[1] - O.event_grad(model, x, tm)[1]) / (2 * eps)
            np.testing.assert_allclose(fdg, H[:, k], rtol=1e-5, atol=1e-5 * max(1.0, np.max(np.abs(H))))
        np.testing.assert_allclose(g, gj, rtol=1e-12, atol=1e-12 * max(1.0, np.max(np.abs(g))))


@pytest.mark.parametrize("theta", [[5.28, 0.02, 0.3, -1.5], [5.34, 0.022, 0.33, -1.65], [5.01, 0.1, 0.8, 2.0], [5.3, 0.05, 0.5, 0.0]])
def test_mass_normalisation(theta):
    """the normalised mass density integrates to one on [5.0, 5.6] (lambda = 0 included)."""
    val, err = integrate.quad(lambda m: O.prob(O.MASS, [m], theta), 5.0, 5.6, points=[theta[0]], limit=200, epsabs=1e-13)
    assert abs(val - 1.0) < 1e-10


def test_kstar_normalisation_and_positivity():
    """the angular density integrates to one over the full box (Gauss-Legendre in cos
    theta_l, cos theta_K; trapezoid in phi, exact for trigonometric polynomials) and is
    positive on a grid at the configuration values."""
    theta = [0.7, 0.0, -0.1, -0.2, 0.1, 0.0, 0.0, 0.0]
    theta2 = [0.6, 0.05, -0.05, -0.3, 0.2, 0.02, -0.03, 0.01]
    xg, wg = np.polynomial.legendre.leggauss(24)
    nphi = 16
    phis = -np.pi + 2 * np.pi * np.arange(nphi) / nphi
    for th in (theta, theta2):
        tot = 0.0
        for cl, wl in zip(xg, wg):
            for ck, wk in zip(xg, wg):
                for ph in phis:
                    tot += wl * wk * (2 * np.pi / nphi) * O.prob(O.KSTAR, [cl, ck, ph], th)
        # sqrt(1 - x^2) terms integrate to the S4/S5/S7/S8 parts only through odd functions
        # whose phi integral is zero, so the quadrature is exact to rounding
        assert abs(tot - 1.0) < 1e-12
    g = np.linspace(-0.999, 0.999, 21)
    for cl in g:
        for ck in g:
            for ph in np.linspace(-np.pi, np.pi, 21):
                assert O.prob(O.KSTAR, [cl, ck, ph], theta) > 0


def test_get_max_bounds_density():
    for th in ([5.28, 0.02, 0.3, -1.5], [5.34, 0.022, 0.33, -1.65], [5.3, 0.05, 0.5, 1.2]):
        mx = O.get_max(O.MASS, th)
        grid = np.linspace(5.0, 5.6, 6001)
        assert max(O.prob(O.MASS, [m], th) for m in grid) <= mx * (1 + 1e-12)
    th = [0.7, 0.0, -0.1, -0.2, 0.1, 0.0, 0.0, 0.0]
    mx = O.get_max(O.KSTAR, th)
    rs = np.random.default_rng(3)
    for _ in range(2000):
        x = [2 * rs.random() - 1, 2 * rs.random() - 1, np.pi * (2 * rs.random() - 1)]
        assert O.prob(O.KSTAR, x, th) <= mx


def test_prng_matches_reference_random():
    """Splitmix64, Xoshiro128++ seeding/stream/jump/float uniforms and Poisson == the
    reference random.hh (tests/golden/ref_random.json)."""
    R = _load("ref_random.json")
    st = __import__("ctypes").c_uint64(R["seed"])
    sm = [O.lib().orc_splitmix64_next(__import__("ctypes").byref(st)) for _ in range(16)]
    assert [str(v) for v in sm] == R["splitmix64"]
    s = O.xoshiro_seed(R["seed"])
    assert list(s) == R["xoshiro128pp_state"]
    assert [O.lib().orc_xoshiro128pp_next(s) for _ in range(32)] == R["xoshiro128pp_int32"]
    O.lib().orc_xoshiro128pp_jump(s)
    assert list(s) == R["xoshiro128pp_after_jump_state"]
    assert [O.lib().orc_xoshiro128pp_next(s) for _ in range(8)] == R["xoshiro128pp_after_jump_int32"]
    assert [O.lib().orc_xoshiro128pp_random(s) for _ in range(16)] == R["xoshiro128pp_random"]
    for k, (mean, vals) in enumerate(R["poisson"].items()):
        s = O.xoshiro_seed(R["seed"] + k)
        got = [O.lib().orc_poisson(s, float(mean)) for _ in range(64)]
        assert got == vals, mean


@pytest.mark.parametrize("model,n", [(O.MASS, 20000), (O.KSTAR, 5000)])
def test_sums_against_exact(model, n):
    """Kahan slices (1 and 5 threads) and the long-double sum vs math.fsum of per-event values."""
    theta = [5.3, 0.03, 0.4, -1.2] if model == O.MASS else [0.7, 0.0, -0.1, -0.2, 0.1, 0.0, 0.0, 0.0]
    data = O.generate(model, theta, n, seed=11, nstreams=4)
    per = []
    for i in range(n):
        lp, g = O.event_grad(model, data[:, i], theta)
        per.append([lp] + list(g))
    per = np.array(per)
    exact = np.array([math.fsum(per[:, k]) for k in range(per.shape[1])])
    scale = np.abs(per).sum(axis=0)
    for nt in (0, 1, 5):
        got = O.nll_sums(model, data, theta, mode=1, nthreads=nt)
        assert np.all(np.abs(got - exact) <= 1e-15 * scale + 1e-300), nt


@pytest.mark.parametrize("model", [O.MASS, O.KSTAR])
def test_generator_goodness_of_fit(model):
    """accept-reject sample follows the density: chi2 of a binned projection."""
    theta = [5.28, 0.02, 0.3, -1.5] if model == O.MASS else [0.7, 0.0, -0.1, -0.2, 0.1, 0.0, 0.0, 0.0]
    n = 100000
    data = O.generate(model, theta, n, seed=229387429, nstreams=8)
    lo, hi = O.RANGES[model][0], O.RANGES[model][1]
    edges = np.linspace(lo, hi, 41)
    h, _ = np.histogram(data[0], edges)
    if model == O.MASS:
        exp_ = np.array([integrate.quad(lambda m: O.prob(O.MASS, [m], theta), a, b, points=[theta[0]] if a < theta[0] < b else None)[0]
                         for a, b in zip(edges[:-1], edges[1:])]) * n
    else:
        # marginal in cos theta_l: 3/4 (1-FL)(1 - cl^2)... integrate numerically over the others
        xg, wg = np.polynomial.legendre.leggauss(12)
        phis = -np.pi + 2 * np.pi * np.arange(8) / 8

        def marg(cl):
            return sum(wk * (2 * np.pi / 8) * O.prob(O.KSTAR, [cl, ck, ph], theta) for ck, wk in zip(xg, wg) for ph in phis)

        exp_ = np.array([integrate.quad(marg, a, b)[0] for a, b in zip(edges[:-1], edges[1:])]) * n
    chi2 = np.sum((h - exp_) ** 2 / exp_)
    assert stats.chi2.sf(chi2, len(h) - 1) > 1e-3


def test_generator_streams_independent_of_partition():
    """the same streams give the same events whatever the number of streams per call"""
    a = O.generate(O.MASS, [5.28, 0.02, 0.3, -1.5], 1000, seed=5, events_per_stream=100)
    b = O.generate(O.MASS, [5.28, 0.02, 0.3, -1.5], 1000, seed=5, events_per_stream=100)
    assert np.array_equal(a, b)
    c = O.generate(O.MASS, [5.28, 0.02, 0.3, -1.5], 1000, seed=6, events_per_stream=100)
    assert not np.array_equal(a, c)


def test_stream_jump_ahead_matches_sequential_jumps():
    """generator streams are seeded on the device by jump-ahead (J^k from precomputed powers
    J^(2^i)); the matrix method equals k sequential jump()s of the oracle, for small k and for
    k with high bits set"""
    import ctypes
    import morefit_amd as M
    for seed in (1, 229387429, 2**61 + 12345):
        s = O.xoshiro_seed(seed)
        seq = []
        for k in range(70):
            seq.append([s[j] for j in range(4)])
            O.lib().orc_xoshiro128pp_jump(s)
        got = M.stream_states(seed, 0, 70)
        assert np.array_equal(got, np.array(seq, dtype=np.uint32)), seed
    # far streams: k = 1000..1002 from sequential jumps
    s = O.xoshiro_seed(7)
    for k in range(1000):
        O.lib().orc_xoshiro128pp_jump(s)
    far = []
    for k in range(3):
        far.append([s[j] for j in range(4)])
        O.lib().orc_xoshiro128pp_jump(s)
    assert np.array_equal(M.stream_states(7, 1000, 3), np.array(far, dtype=np.uint32))

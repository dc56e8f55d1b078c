"""Crystal Ball and polynomial densities (reference pdf.hh:195-419 and 716-892) and the box
integrals of all densities (definite_integral, used by the plotter's bin integrals).

Oracle pins (independent of the oracle's own formulas), beside the golden vectors of the
reference's own graph engine for both densities (tests/test_oracle.py, tests/golden/ref_graph.json):
  * Crystal Ball: scipy.stats.crystalball (beta = alpha, m = n, loc = mu, scale = sigma),
    truncated to the window, for the density, and quad for the normalisation; continuity of
    the density and of its slope at the join mu - alpha sigma (what fixes A and B);
  * polynomial: numpy.polynomial integration of 1 + c1 x + c2 x^2 + c3 x^3;
  * derivatives: central differences of log p (gradient) and of the gradient (Hessian);
  * envelope: the maximum over a dense grid.
Then the CPU backend's -2lnL, gradient and Hessian against the oracle's sums, and a fit
against an independent minimiser (scipy + Newton on the oracle).
"""
import numpy as np
import pytest
from scipy import integrate, optimize, stats

import morefit_amd as M
from oracle import oracle as O

CB_POINTS = [
    [5.28, 0.02, 1.5, 3.0],
    [5.30, 0.05, 0.5, 1.2],    # long tail, n close to 1
    [5.25, 0.015, 3.0, 10.0],  # join far below the peak
    [5.40, 0.08, 1.0, 1.0],  # n = 1: the tail integral's logarithmic branch
]
POLY_POINTS = [[0.3, -0.2, 0.1], [0.0, 0.0, 0.0], [-0.9, 0.5, 0.2], [0.2, 0.9, -0.3]]


def _cb_scipy(x, t, rng):
    mu, s, a, n = t
    d = stats.crystalball(a, n, loc=mu, scale=s)
    return d.pdf(x) / (d.cdf(rng[1]) - d.cdf(rng[0]))


@pytest.mark.parametrize("theta", CB_POINTS)
def test_crystalball_density_matches_scipy(theta):
    rng = O.MASS_RANGE
    xs = np.linspace(rng[0], rng[1], 301)
    p = np.array([O.prob(O.CB, [x], theta) for x in xs])
    # scipy needs n > 1: at n = 1 it is compared at 1 + 1e-9, where the densities differ by O(1e-9)
    t_sp = list(theta[:3]) + [max(theta[3], 1.0 + 1e-9)]
    np.testing.assert_allclose(p, _cb_scipy(xs, t_sp, rng), rtol=2e-6)
    norm, _ = integrate.quad(lambda x: O.prob(O.CB, [x], theta), rng[0], rng[1], points=[theta[0] - theta[1] * theta[2], theta[0]],
                             limit=400, epsabs=0, epsrel=1e-12)
    assert abs(norm - 1.0) < 1e-8


@pytest.mark.parametrize("theta", CB_POINTS[:3])
def test_crystalball_join_is_smooth(theta):
    mu, s, a, n = theta
    xj = mu - a * s
    h = 1e-7 * s
    lo, hi = O.prob(O.CB, [xj - h], theta), O.prob(O.CB, [xj + h], theta)
    assert abs(lo - hi) < 1e-5 * hi
    dlo = (O.prob(O.CB, [xj - h], theta) - O.prob(O.CB, [xj - 1e3 * h], theta)) / (999 * h)
    dhi = (O.prob(O.CB, [xj + 1e3 * h], theta) - O.prob(O.CB, [xj + h], theta)) / (999 * h)
    assert abs(dlo - dhi) < 2e-3 * abs(dhi)


@pytest.mark.parametrize("theta", POLY_POINTS)
def test_polynomial_density_matches_numpy(theta):
    rng = O.POLY_RANGE
    c = np.polynomial.Polynomial([1.0] + list(theta))
    integ = c.integ()
    norm = integ(rng[1]) - integ(rng[0])
    xs = np.linspace(rng[0], rng[1], 101)
    p = np.array([O.prob(O.POLY, [x], theta) for x in xs])
    np.testing.assert_allclose(p, c(xs) / norm, rtol=1e-13, atol=1e-15)
    # another window: the normalisation follows the range
    rng2 = np.array([-0.5, 0.8])
    norm2 = integ(rng2[1]) - integ(rng2[0])
    np.testing.assert_allclose(O.prob(O.POLY, [0.1], theta, rng2), c(0.1) / norm2, rtol=1e-13)


@pytest.mark.parametrize("model,theta", [(O.CB, t) for t in CB_POINTS[:3]] + [(O.POLY, t) for t in POLY_POINTS if any(t)])
def test_jets_match_central_differences(model, theta):
    rng = O.RANGES[model]
    theta = np.array(theta)
    for x in np.linspace(rng[0], rng[1], 9)[1:-1]:
        lp, g, H = O.event_jet(model, [x], theta)
        lp1, g1 = O.event_grad(model, [x], theta)
        assert abs(lp - lp1) < 1e-12 * max(1.0, abs(lp))
        np.testing.assert_allclose(g, g1, rtol=1e-10, atol=1e-10 * np.max(np.abs(g)))
        assert abs(lp - np.log(O.prob(model, [x], theta))) < 1e-12 * max(1.0, abs(lp))
        for k in range(len(theta)):
            h = 1e-5 * max(abs(theta[k]), 1e-2)
            tp, tm = theta.copy(), theta.copy()
            tp[k] += h
            tm[k] -= h
            fd = (np.log(O.prob(model, [x], tp)) - np.log(O.prob(model, [x], tm))) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (x, k, fd, g[k])
            gd = (O.event_grad(model, [x], tp)[1] - O.event_grad(model, [x], tm)[1]) / (2 * h)
            np.testing.assert_allclose(gd, H[k], rtol=1e-5, atol=1e-5 * max(1.0, np.max(np.abs(H[k]))))


@pytest.mark.parametrize("model,theta", [(O.CB, CB_POINTS[0]), (O.CB, CB_POINTS[1]), (O.POLY, POLY_POINTS[0]), (O.POLY, POLY_POINTS[2])])
def test_envelope_covers_the_density(model, theta):
    rng = O.RANGES[model]
    xs = np.linspace(rng[0], rng[1], 200001)
    pmax = max(O.prob(model, [x], theta) for x in xs[::50])
    pmax = max(pmax, max(O.prob(model, [x], theta) for x in xs[np.argmax([O.prob(model, [x], theta) for x in xs[::50]]) * 50 - 60:][:120]))
    m = O.get_max(model, theta)
    assert pmax <= m <= 1.01 * pmax


@pytest.mark.parametrize("model", [O.CB, O.POLY])
def test_generated_sample_follows_the_density(model):
    theta = M.TRUTH[model]
    rng = O.RANGES[model]
    data = O.generate(model, theta, 40000, seed=7, nstreams=4)[0]
    assert np.all((data >= rng[0]) & (data <= rng[1]))
    if model == O.CB:
        d = stats.crystalball(theta[2], theta[3], loc=theta[0], scale=theta[1])
        cdf = lambda x: (d.cdf(x) - d.cdf(rng[0])) / (d.cdf(rng[1]) - d.cdf(rng[0]))
    else:
        integ = np.polynomial.Polynomial([1.0] + list(theta)).integ()
        cdf = lambda x: (integ(x) - integ(rng[0])) / (integ(rng[1]) - integ(rng[0]))
    grid = np.linspace(rng[0], rng[1], 400)
    F = np.interp(data, grid, cdf(grid))
    assert stats.kstest(F, "uniform").pvalue > 1e-3


def _hess_full(P, packed):
    H = np.zeros((P, P))
    k = 0
    for i in range(P):
        for j in range(i + 1):
            H[i, j] = H[j, i] = packed[k]
            k += 1
    return H


@pytest.mark.parametrize("model,n", [(O.CB, 40000), (O.POLY, 40000)])
def test_cpu_backend_eval_matches_oracle(model, n):
    data = O.generate(model, M.TRUTH[model], n, seed=123, nstreams=8)
    theta = M.START[model]
    v, g, H = M.eval_cpu(model, data, theta, nthreads=3)
    ref = O.nll_sums(model, data, theta, mode=2, nthreads=0)
    absref = O.nll_sums(model, data, theta, mode=12, nthreads=1)
    P = O.NPAR[model]
    assert abs(v - (-2 * ref[0])) <= 1e-11 * absref[0] * 2
    assert np.all(np.abs(g - (-2 * ref[1:1 + P])) <= 1e-11 * absref[1:1 + P] + 1e-300)
    assert np.all(np.abs(H - (-2 * _hess_full(P, ref[1 + P:]))) <= 1e-10 * 2 * _hess_full(P, absref[1 + P:]) + 1e-300)


def _oracle_fit(model, data, start):
    P = O.NPAR[model]

    def f(t):
        s = O.nll_sums(model, data, t, mode=1, nthreads=4)
        return -2 * s[0], -2 * s[1:]

    t = np.array(start, dtype=float)
    if model == O.CB:
        r = optimize.minimize(f, t, jac=True, method="L-BFGS-B", bounds=[(5.0, 5.6), (0.005, 0.13), (0.2, 5.0), (1.05, 20.0)],
                              options={"ftol": 1e-15, "gtol": 1e-10, "maxiter": 3000})
        t = r.x.copy()
    for _ in range(60):
        s = O.nll_sums(model, data, t, mode=2, nthreads=4)
        g, H = -2 * s[1:1 + P], -2 * _hess_full(P, s[1 + P:])
        step = np.linalg.solve(H, g)
        f0, a = -2 * s[0], 1.0
        while a > 1e-8:
            f1 = f(t - a * step)[0]
            if np.isfinite(f1) and f1 <= f0 + 1e-12 * abs(f0):
                break
            a *= 0.5
        t = t - a * step
        if np.max(np.abs(a * step)) < 1e-14:
            break
    s = O.nll_sums(model, data, t, mode=2, nthreads=4)
    return t, np.sqrt(np.diag(2 * np.linalg.inv(-2 * _hess_full(P, s[1 + P:]))))


@pytest.mark.parametrize("model,n", [(O.CB, 60000), (O.POLY, 60000)])
def test_cpu_fit_matches_independent_minimiser(model, n):
    data = O.generate(model, M.TRUTH[model], n, seed=229387429, nstreams=16)
    vals, errs, out, ms = M.fit_cpu(model, data, M.START[model], nthreads=4)
    assert out.status == 0, out.status
    t, e = _oracle_fit(model, data, M.START[model])
    assert np.max(np.abs((vals - t) / e)) < 2e-3, (vals, t, e)
    np.testing.assert_allclose(errs, e, rtol=2e-3)
    assert np.max(np.abs((vals - M.TRUTH[model]) / errs)) < 5

"""The plotter (reference plotter.hh:44-640): the integral kernel over fine bins built from
PDF::definite_integral (pdf.hh:75-82, physicspdfs.hh:82-164), the data histogram, residuals,
pulls and the script files, on the CPU backend.

Oracle pins of the bin integrals (oracle.bin_integrals, Gauss-Legendre quadrature of the
density), each independent of the oracle's formulas:
  * mass: scipy.stats.norm / the exponential's closed-form CDF, truncated to the window;
  * Crystal Ball: scipy.stats.crystalball's CDF, truncated;
  * polynomial: numpy.polynomial's antiderivative;
  * angular: the textbook one-dimensional projections of the P-wave distribution,
      dG/dcos(theta_l) = 3/4 FL (1 - v^2) + 3/8 (1 - FL)(1 + v^2) + AFB v,
      dG/dcos(theta_K) = 3/2 FL u^2 + 3/4 (1 - FL)(1 - u^2),
      dG/dphi         = (1 + S3 cos 2phi + S9 sin 2phi) / (2 pi);
  * every density: the bins over the full range sum to 1.
Bars: FP64 kernels within 1e-13 of the oracle per bin; FP32 kernels within 1e-3 of the
largest bin (differences of FP32 antiderivatives at FP32 bin edges).
"""
import sys
import types

import numpy as np
import pytest
from scipy import stats

import morefit_amd as M
from oracle import oracle as O

ALL = [(M.MASS, 0), (M.CB, 0), (M.POLY, 0), (M.KSTAR, 0), (M.KSTAR, 1), (M.KSTAR, 2)]


def _edges(model, var, nfine):
    rng = O.RANGES[model][2 * var:2 * var + 2]
    return np.linspace(rng[0], rng[1], nfine + 1)


def _closed_form(model, var, theta, e):
    """bin integrals from library CDFs / textbook projections"""
    if model == M.MASS:
        mu, s, f, lam = theta
        a, b = O.MASS_RANGE
        N = stats.norm(mu, s)
        G = (N.cdf(e) - N.cdf(a)) / (N.cdf(b) - N.cdf(a))
        E = (np.exp(lam * e) - np.exp(lam * a)) / (np.exp(lam * b) - np.exp(lam * a))
        F = f * G + (1 - f) * E
    elif model == M.CB:
        mu, s, al, n = theta
        a, b = O.MASS_RANGE
        d = stats.crystalball(al, n, loc=mu, scale=s)
        F = (d.cdf(e) - d.cdf(a)) / (d.cdf(b) - d.cdf(a))
    elif model == M.POLY:
        P = np.polynomial.Polynomial([1.0] + list(theta)).integ()
        a, b = O.POLY_RANGE
        F = (P(e) - P(a)) / (P(b) - P(a))
    else:
        FL, S3, AFB, S9 = theta[0], theta[1], theta[4], theta[7]
        if var == 0:  # cos theta_l
            F = 0.75 * FL * (e - e ** 3 / 3) + 0.375 * (1 - FL) * (e + e ** 3 / 3) + AFB * e ** 2 / 2
        elif var == 1:  # cos theta_K
            F = 0.5 * FL * e ** 3 + 0.75 * (1 - FL) * (e - e ** 3 / 3)
        else:
            F = (e + 0.5 * S3 * np.sin(2 * e) - 0.5 * S9 * np.cos(2 * e)) / (2 * np.pi)
    return np.diff(F)


ANGULAR_POINTS = [M.KSTAR_TRUTH, np.array([0.4, 0.2, -0.2, 0.3, -0.25, 0.1, -0.1, 0.15])]


@pytest.mark.parametrize("model,var", ALL)
def test_oracle_bin_integrals_match_closed_forms(model, var):
    points = ANGULAR_POINTS if model == M.KSTAR else [M.TRUTH[model], M.START[model]]
    for theta in points:
        e = _edges(model, var, 200)
        got = O.bin_integrals(model, theta, var, e[:-1], e[1:])
        ref = _closed_form(model, var, theta, e)
        assert np.max(np.abs(got - ref)) < 1e-10 * max(1.0, ref.max()), np.max(np.abs(got - ref))
        assert abs(got.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("model,var", ALL)
@pytest.mark.parametrize("fp32", [False, True])
def test_cpu_integral_kernel_matches_oracle(model, var, fp32):
    theta = M.START[model]
    data = O.generate(model, M.TRUTH[model], 3000, seed=5, nstreams=4)
    nbins, mult = 40, 10
    r = M.plot(model, theta, data, var=var, nbins=nbins, multiplier=mult, fp32=fp32, nthreads=2)
    e = _edges(model, var, nbins * mult)
    ref = O.bin_integrals(model, theta, var, e[:-1], e[1:])
    tol = 1e-3 * ref.max() if fp32 else 1e-13
    assert np.max(np.abs(r["pdf_integrals"] - ref)) <= tol, np.max(np.abs(r["pdf_integrals"] - ref))


def test_histogram_pulls_and_residuals():
    model = M.MASS
    data = O.generate(model, M.MASS_TRUTH, 200000, seed=11, nstreams=8)
    nbins, mult = 60, 10
    r = M.plot(model, M.MASS_TRUTH, data, nbins=nbins, multiplier=mult)
    e = _edges(model, 0, nbins)
    h, _ = np.histogram(data[0], bins=e)
    np.testing.assert_array_equal(r["counts"], h)
    np.testing.assert_allclose(r["errors"], np.sqrt(h))
    # the density is scaled to the histogram: the residuals sum to zero over the full range
    scale = h.sum() / r["pdf_integrals"].sum() * mult
    resid = h - (r["pdf_integrals"] * scale).reshape(nbins, mult).mean(axis=1)
    assert abs(resid.sum()) < 1e-8 * h.sum()
    ok = h > 0
    np.testing.assert_allclose(r["pulls"][ok], resid[ok] / np.sqrt(h[ok]), rtol=1e-9, atol=1e-12)
    # data drawn from the plotted density: pulls are standard normal where bins are populated
    p = r["pulls"][h > 25]
    assert abs(p.mean()) < 0.5 and 0.6 < p.std() < 1.5


def test_weighted_histogram():
    data = O.generate(M.POLY, M.POLY_TRUTH, 5000, seed=2, nstreams=2)
    w = np.random.default_rng(1).uniform(0.2, 2.0, data.shape[1])
    r = M.plot(M.POLY, M.POLY_TRUTH, data, nbins=25, multiplier=4, weights=w)
    e = _edges(M.POLY, 0, 25)
    h, _ = np.histogram(data[0], bins=e, weights=w)
    h2, _ = np.histogram(data[0], bins=e, weights=w * w)
    np.testing.assert_allclose(r["counts"], h, rtol=1e-12)
    np.testing.assert_allclose(r["errors"], np.sqrt(h2), rtol=1e-12)


def test_sub_range_edges_and_replot():
    data = np.array([[-1.0, -0.5, -0.25, 0.0, 0.5, 0.5, 1.0]])
    r = M.plot(M.POLY, M.POLY_TRUTH, data, nbins=2, multiplier=3, range_min=-0.5, range_max=0.5)
    # [-0.5, 0) holds -0.5 and -0.25; [0, 0.5] holds 0 and both events at the upper edge
    np.testing.assert_array_equal(r["counts"], [2, 3])
    e = np.linspace(-0.5, 0.5, 7)
    np.testing.assert_allclose(r["pdf_integrals"], O.bin_integrals(M.POLY, M.POLY_TRUTH, 0, e[:-1], e[1:]), atol=1e-14)
    r2 = M.plot(M.POLY, M.POLY_TRUTH, data, nbins=2, multiplier=3, range_min=-0.5, range_max=0.5, replots=2)
    np.testing.assert_array_equal(r2["pdf_integrals"], r["pdf_integrals"])


def test_bad_arguments_are_refused():
    data = O.generate(M.MASS, M.MASS_TRUTH, 100, seed=1)
    with pytest.raises(M.MorefitError):
        M.plot(M.MASS, M.MASS_TRUTH, data, var=1)
    with pytest.raises(M.MorefitError):
        M.plot(M.MASS, M.MASS_TRUTH, data, nbins=0)
    with pytest.raises(M.MorefitError):
        M.plot(M.MASS, M.MASS_TRUTH, data, multiplier=0)


class _Recorder:
    """stands in for matplotlib.pyplot and records what the script draws"""

    def __init__(self):
        self.calls = []
        self.rcParams = {}

    def __getattr__(self, name):
        def call(*a, **k):
            self.calls.append((name, a, k))
            return self

        return call

    def subplots(self, *a, **k):
        self.calls.append(("subplots", a, k))
        return self, _Axes(self)


class _Axes:
    def __init__(self, rec):
        self.rec = rec

    def __getitem__(self, i):
        return self

    def __getattr__(self, name):
        def call(*a, **k):
            self.rec.calls.append((name, a, k))

        return call


@pytest.mark.parametrize("pulls", [False, True])
def test_matplotlib_script(tmp_path, pulls):
    data = O.generate(M.MASS, M.MASS_TRUTH, 20000, seed=3, nstreams=4)
    f = tmp_path / "mass_plot.py"
    r = M.plot(M.MASS, M.MASS_TRUTH, data, nbins=30, multiplier=10, filename=str(f), plot_type=M.PLOT_MATPLOTLIB, pulls=pulls)
    rec = _Recorder()
    mpl = types.ModuleType("matplotlib")
    mpl.pyplot = rec
    saved = {k: sys.modules.get(k) for k in ("matplotlib", "matplotlib.pyplot")}
    sys.modules["matplotlib"], sys.modules["matplotlib.pyplot"] = mpl, rec
    try:
        ns = {}
        exec(compile(f.read_text(), str(f), "exec"), ns)
    finally:
        for k, v in saved.items():
            if v is None:
                sys.modules.pop(k, None)
            else:
                sys.modules[k] = v
    np.testing.assert_allclose(ns["y"], r["counts"], rtol=1e-5)
    np.testing.assert_allclose(ns["dy"], 0.5 * r["errors"], rtol=1e-5)  # half-length bars, as the reference draws
    scale = r["counts"].sum() / r["pdf_integrals"].sum() * 10
    np.testing.assert_allclose(ns["y_pdf"], r["pdf_integrals"] * scale, rtol=1e-5, atol=1e-6 * scale)
    assert len(ns["x"]) == 30 and len(ns["x_pdf"]) == 300
    if pulls:
        np.testing.assert_allclose(ns["y_pull"], r["pulls"], rtol=1e-5, atol=1e-6)
    assert any(c[0] == "savefig" and c[1][0] == "mass_plot.pdf" for c in rec.calls)


def test_root_macro(tmp_path):
    data = O.generate(M.KSTAR, M.KSTAR_TRUTH, 5000, seed=3, nstreams=4)
    f = tmp_path / "phi_plot.C"
    r = M.plot(M.KSTAR, M.KSTAR_TRUTH, data, var=2, nbins=16, multiplier=5, filename=str(f), plot_type=M.PLOT_ROOT, pulls=True)
    text = f.read_text()
    assert text.startswith("int phi_plot() {") and text.rstrip().endswith("}")
    assert text.count("{") == text.count("}")

    def array(name):
        body = text.split("double " + name)[1].split("{", 1)[1].split("}", 1)[0]
        return np.array([float(v) for v in body.split(",")])

    np.testing.assert_allclose(array("hist_values"), r["counts"], rtol=1e-5)
    np.testing.assert_allclose(array("pdf_values"), r["pdf_integrals"], rtol=1e-5)
    np.testing.assert_allclose(array("pull_values"), r["pulls"], rtol=1e-5, atol=1e-6)
    assert 'c0->Print("phi_plot.eps", "eps");' in text

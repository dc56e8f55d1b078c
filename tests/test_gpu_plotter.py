"""GPU parity of the plotter's integral kernel (reference plotter.hh:138-216 over
PDF::definite_integral) through the C ABI: the fine-bin integrals of every density and
every angular projection against the oracle's quadrature of the same bins; the histogram,
residuals and pulls equal to the CPU backend's.
Bars: FP64 kernels within 1e-13 per bin; FP32 kernels within 1e-3 of the largest bin
(differences of FP32 antiderivatives at FP32 bin edges).
"""
import sys

import numpy as np
import pytest

import morefit_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ALL = [(M.MASS, 0), (M.CB, 0), (M.POLY, 0), (M.KSTAR, 0), (M.KSTAR, 1), (M.KSTAR, 2)]


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("model,var", ALL)
@pytest.mark.parametrize("fp32", [False, True])
def test_integral_kernel_matches_oracle(ctx, model, var, fp32):
    theta = M.START[model]
    data = O.generate(model, M.TRUTH[model], 5000, seed=9, nstreams=4)
    nbins, mult = 100, 10
    say("plot", model, var, "fp32", fp32)
    r = M.plot(model, theta, data, var=var, nbins=nbins, multiplier=mult, ctx=ctx, fp32=fp32)
    rng = O.RANGES[model][2 * var:2 * var + 2]
    e = np.linspace(rng[0], rng[1], nbins * mult + 1)
    ref = O.bin_integrals(model, theta, var, e[:-1], e[1:])
    tol = 1e-3 * ref.max() if fp32 else 1e-13
    assert np.max(np.abs(r["pdf_integrals"] - ref)) <= tol, np.max(np.abs(r["pdf_integrals"] - ref))
    c = M.plot(model, theta, data, var=var, nbins=nbins, multiplier=mult, fp32=False)
    np.testing.assert_array_equal(r["counts"], c["counts"])
    if not fp32:
        ok = c["counts"] > 0
        np.testing.assert_allclose(r["pulls"][ok], c["pulls"][ok], rtol=1e-9, atol=1e-12)


def test_replot_and_sub_range(ctx):
    data = O.generate(M.MASS, M.MASS_TRUTH, 20000, seed=4, nstreams=4)
    say("replot")
    r = M.plot(M.MASS, M.MASS_TRUTH, data, nbins=50, multiplier=20, range_min=5.2, range_max=5.4, ctx=ctx, replots=3)
    e = np.linspace(5.2, 5.4, 1001)
    np.testing.assert_allclose(r["pdf_integrals"], O.bin_integrals(M.MASS, M.MASS_TRUTH, 0, e[:-1], e[1:]), atol=1e-13)
    h, _ = np.histogram(data[0], bins=np.linspace(5.2, 5.4, 51))
    np.testing.assert_array_equal(r["counts"], h)

"""Failure detection on the CPU: non-finite sums (an event that is not a number, a density
undefined at an event outside its range, a zero width), events outside the
observables' ranges, and how each is reported with the fit (mfh_fit_out: nonfinite_evals,
nonfinite_events, first_nonfinite, out_of_range).  The reference checks its host sum for NaN
and stops the program (fitter.hh:1197-1204); here the fit returns a non-zero status with the
count and the first index of the offending events, and fits of other data are unaffected.
Compile errors, injected launch failures and over-budget requests are covered in test_abi.py
and test_gpu_serve.py.
"""
import numpy as np
import pytest

import morefit_amd as M
from oracle import oracle as O


def _mass(n=5000, seed=3):
    return O.generate(M.MASS, M.MASS_TRUTH, n, seed=seed, nstreams=2)


@pytest.mark.parametrize("flags", [M.FIT_DEFAULT, M.FIT_DEFAULT & ~M.FIT_KAHAN_ON_ACCELERATOR])
@pytest.mark.parametrize("fp32", [False, True])
def test_event_that_is_not_a_number(flags, fp32):
    d = _mass()
    d[0, 1234] = np.nan
    v, e, out, _ = M.fit_options(M.MASS, d, M.MASS_START, flags=flags, fp32=fp32)
    assert out.status != 0
    assert out.nonfinite_evals >= 1
    assert out.nonfinite_events == 1 and out.first_nonfinite == 1234
    assert out.out_of_range == 1


def test_event_outside_the_range_is_fitted_as_given_and_counted():
    d = _mass()
    d[0, 10] = 5.9  # outside [5.0, 5.6]: the Gaussian and exponential are finite there
    d[0, 20] = 4.0
    v, e, out, _ = M.fit_options(M.MASS, d, M.MASS_START)
    assert out.status == 0 and out.nonfinite_evals == 0
    assert out.out_of_range == 2
    # the same numbers as a fit that includes those two events in the sum (reference semantics)
    ref = O.nll_sums(M.MASS, d, v, mode=1, nthreads=1)
    assert abs(out.fval - (-2 * ref[0])) < 1e-8 * abs(ref[0])


def test_angular_event_where_the_density_is_undefined():
    d = O.generate(M.KSTAR, M.KSTAR_TRUTH, 4000, seed=5, nstreams=2)
    # cos(theta_l) = 3 (outside [-1, 1]): sin(theta_l) = sqrt(1 - cos^2) is not a number
    # there, so neither is log p for that event
    theta = np.array(M.KSTAR_START)
    x = np.array([3.0, 0.0, 0.0])
    assert not O.prob(M.KSTAR, x, theta) > 0
    d[:, 77] = x
    v, e, out, _ = M.fit_options(M.KSTAR, d, theta)
    assert out.status != 0
    assert out.nonfinite_events == 1 and out.first_nonfinite == 77
    assert out.out_of_range == 1


def test_zero_width_gives_non_finite_sums():
    d = _mass(2000)
    theta = np.array(M.MASS_TRUTH)
    theta[1] = 0.0
    v, g, H = M.eval_cpu(M.MASS, d, theta)
    assert not np.isfinite(v)
    # a fit starting there is moved inside the parameter's limits [0.005, 0.13] and converges
    vv, e, out, _ = M.fit_options(M.MASS, d, theta)
    assert out.status == 0 and np.all(np.isfinite(vv))


def test_session_reports_per_fit():
    """a plain session: a bad dataset does not leak into the next fit's report"""
    pf = M.PlainFitter(M.MASS)
    try:
        bad = _mass(3000, seed=8)
        bad[0, 5] = np.inf
        _, _, out, _ = pf.fit(bad, M.MASS_START)
        assert out.status != 0 and out.nonfinite_events == 1 and out.first_nonfinite == 5
        good = _mass(3000, seed=9)
        v, e, out2, _ = pf.fit(good, M.MASS_START)
        assert out2.status == 0 and out2.nonfinite_evals == 0 and out2.nonfinite_events == 0 and out2.out_of_range == 0
        vc, ec, _, _ = M.fit_cpu(M.MASS, good, M.MASS_START)
        np.testing.assert_allclose(v, vc, rtol=0, atol=1e-6 * np.max(ec))
    finally:
        pf.close()

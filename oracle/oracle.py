"""ctypes binding of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline.  The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MASS, KSTAR, CB, POLY = 0, 1, 2, 3
NPAR = {MASS: 4, KSTAR: 8, CB: 4, POLY: 3}
NOBS = {MASS: 1, KSTAR: 3, CB: 1, POLY: 1}
MASS_RANGE = np.array([5.0, 5.6])
KSTAR_RANGE = np.array([-1.0, 1.0, -1.0, 1.0, -np.pi, np.pi])
POLY_RANGE = np.array([-1.0, 1.0])
RANGES = {MASS: MASS_RANGE, KSTAR: KSTAR_RANGE, CB: MASS_RANGE, POLY: POLY_RANGE}

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        up = ctypes.POINTER(ctypes.c_uint32)
        L.orc_event_jet.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp]
        L.orc_event_grad.argtypes = [ctypes.c_int, dp, dp, dp, dp]
        L.orc_event_grad.restype = ctypes.c_double
        L.orc_prob.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_prob.restype = ctypes.c_double
        L.orc_get_max.argtypes = [ctypes.c_int, dp, dp]
        L.orc_get_max.restype = ctypes.c_double
        L.orc_bin_integral.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_bin_integral.restype = ctypes.c_double
        L.orc_nll.argtypes = [ctypes.c_int, dp, ctypes.c_long, ctypes.c_long, dp, dp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_splitmix64_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.orc_splitmix64_next.restype = ctypes.c_uint64
        L.orc_xoshiro128pp_seed.argtypes = [up, ctypes.c_uint64]
        L.orc_xoshiro128pp_next.argtypes = [up]
        L.orc_xoshiro128pp_next.restype = ctypes.c_uint32
        L.orc_xoshiro128pp_jump.argtypes = [up]
        L.orc_xoshiro128pp_random.argtypes = [up]
        L.orc_xoshiro128pp_random.restype = ctypes.c_double
        L.orc_poisson.argtypes = [up, ctypes.c_double]
        L.orc_poisson.restype = ctypes.c_uint
        L.orc_generate.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double, up, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_long), dp, ctypes.c_long]
        L.orc_generate.restype = ctypes.c_long
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def event_jet(model, x, theta, rng=None):
    """log p, gradient and full Hessian of one event (forward-mode jets)."""
    P = NPAR[model]
    x, theta = _f64(x), _f64(theta)
    rng = _f64(RANGES[model] if rng is None else rng)
    lp = np.zeros(1)
    g = np.zeros(P)
    h = np.zeros(P * P)
    lib().orc_event_jet(model, _dp(x), _dp(theta), _dp(rng), _dp(lp), _dp(g), _dp(h))
    return lp[0], g, h.reshape(P, P)


def event_grad(model, x, theta, rng=None):
    """log p and gradient of one event (closed form)."""
    P = NPAR[model]
    x, theta = _f64(x), _f64(theta)
    rng = _f64(RANGES[model] if rng is None else rng)
    g = np.zeros(P)
    lp = lib().orc_event_grad(model, _dp(x), _dp(theta), _dp(rng), _dp(g))
    return lp, g


def prob(model, x, theta, rng=None):
    x, theta = _f64(x), _f64(theta)
    rng = _f64(RANGES[model] if rng is None else rng)
    return lib().orc_prob(model, _dp(x), _dp(theta), _dp(rng))


def get_max(model, theta, rng=None):
    theta = _f64(theta)
    rng = _f64(RANGES[model] if rng is None else rng)
    return lib().orc_get_max(model, _dp(theta), _dp(rng))


def bin_integrals(model, theta, var, lo, hi, rng=None):
    """integral of the normalised density over each [lo[i], hi[i]] in observable var (the
    others over their full ranges), by quadrature"""
    theta = _f64(theta)
    rng = _f64(RANGES[model] if rng is None else rng)
    return np.array([lib().orc_bin_integral(model, _dp(theta), _dp(rng), int(var), float(a), float(b))
                     for a, b in zip(np.atleast_1d(lo), np.atleast_1d(hi))])


def nll_sums(model, data, theta, mode=1, nthreads=1, rng=None):
    """Per-output sums (sum log p, sum grad, sum Hessian packed i>=j) over SoA data
    of shape (nobs, n).  nthreads=0 selects the long-double precise sum; mode+10 sums
    absolute values (tolerance scales)."""
    data = _f64(np.atleast_2d(data))
    nobs, n = data.shape
    assert nobs == NOBS[model]
    theta = _f64(theta)
    rng = _f64(RANGES[model] if rng is None else rng)
    P = NPAR[model]
    m = mode % 10
    nout = 1 if m == 0 else (1 + P if m == 1 else 1 + P + P * (P + 1) // 2)
    out = np.zeros(nout)
    rc = lib().orc_nll(model, _dp(data), n, n, _dp(theta), _dp(rng), mode, nthreads, _dp(out))
    if rc != 0:
        raise ValueError("orc_nll failed")
    return out


def xoshiro_seed(seed):
    s = (ctypes.c_uint32 * 4)()
    lib().orc_xoshiro128pp_seed(s, ctypes.c_uint64(seed))
    return s


def generate(model, theta, n, seed, nstreams=1, rng=None, maxprob=None, events_per_stream=None):
    """Accept-reject sample of n events in `nstreams` streams separated by jump(); with
    events_per_stream, stream k owns events [k*eps, min(n, (k+1)*eps))."""
    theta = _f64(theta)
    rng = _f64(RANGES[model] if rng is None else rng)
    if maxprob is None:
        maxprob = get_max(model, theta, rng)
    if events_per_stream:
        nstreams = (n + events_per_stream - 1) // events_per_stream
    states = (ctypes.c_uint32 * (4 * nstreams))()
    s = xoshiro_seed(seed)
    for k in range(nstreams):
        for j in range(4):
            states[4 * k + j] = s[j]
        lib().orc_xoshiro128pp_jump(s)
    offs = (ctypes.c_long * (nstreams + 1))()
    for k in range(nstreams + 1):
        offs[k] = min(n, k * events_per_stream) if events_per_stream else (n * k) // nstreams
    out = np.zeros((NOBS[model], n))
    lib().orc_generate(model, _dp(theta), _dp(rng), maxprob, states, nstreams, offs, _dp(out), n)
    return out

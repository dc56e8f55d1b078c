"""The minimiser when the objective's value carries FP32 rounding (CPU only).

FP32 likelihood sums carry a coherent error: the kernel parameters are rounded to FP32 once
per call, so every event sees the same relative error and the value jumps by up to a few
n * 2^-24 as theta moves, far more than the random-walk rounding of the per-event terms.  Near
the minimum that exceeds the predicted decrease of a Newton step, so the value cannot decide
the step; the minimiser then follows the directional derivative.  Checked two ways against
the FP64 fit of the same data:

* tests/cpp/minimizer_noise.cc: the oracle's exact -2 lnL, gradient and Hessian with a
  piecewise-constant value perturbation of amplitude amp * n * 2^-24 (amp 4 and 8, and 64:
  far above the floor the minimiser assumes);
* the FP32 kernels through the CPU backend (the generated FP32 functor, FP32-rounded
  parameters, centred observables -- the device's FP32 semantics with host libm).

Bar: every fit converges (status 0) within 1e-3 sigma of the FP64 minimum.
"""
import os
import subprocess

import numpy as np
import pytest

import morefit_amd as M
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("mn") / "minimizer_noise")
    oracle_dir = os.path.join(ROOT, "oracle")
    if not os.path.exists(os.path.join(oracle_dir, "liboracle.so")):
        subprocess.check_call(["make", "-C", oracle_dir])
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "morefit_amd", "include"),
                           os.path.join(ROOT, "tests", "cpp", "minimizer_noise.cc"), "-L", oracle_dir, "-loracle",
                           "-Wl,-rpath," + oracle_dir, "-o", exe])
    return exe


def _toys(ntoys, seed):
    rs = np.random.default_rng(seed)
    for i in range(ntoys):
        n = int(rs.poisson(5000))
        yield i, n, O.generate(M.MASS, M.MASS_TRUTH, n, seed=seed * 1000 + i, nstreams=4), M.toy_starts(M.MASS, [n], seed=seed + i)[0]


@pytest.mark.parametrize("amp", [4.0, 8.0, 64.0])
def test_minimiser_converges_under_coherent_value_noise(probe, tmp_path, amp):
    path = str(tmp_path / "data.f64")
    checked = 0
    for i, n, data, start in _toys(30, 21):
        data[0].tofile(path)
        out = subprocess.check_output([probe, path, str(n), str(amp)] + [repr(float(x)) for x in start]).decode().split()
        s0, hesse_ok, s1, worst = int(out[0]), int(out[1]), int(out[2]), float(out[3])
        assert s0 == 0 and hesse_ok, (i, out)
        assert s1 == 0, (i, out)
        assert worst < 1e-3, (i, out)
        checked += 1
    assert checked == 30


def test_probe_fails_without_the_noise_floor(probe, tmp_path):
    """the probe is live: when the perturbed objective reports no rounding floor (as if it were
    FP64), value comparisons decide the steps near the minimum and fits stall"""
    path = str(tmp_path / "data.f64")
    stalled = 0
    for i, n, data, start in _toys(30, 21):
        data[0].tofile(path)
        out = subprocess.check_output([probe, path, str(n), "8", *[repr(float(x)) for x in start], "nofloor"]).decode().split()
        if int(out[2]) != 0 or float(out[3]) > 1e-3:
            stalled += 1
    assert stalled > 0


def test_fp32_fits_match_fp64_on_cpu_backend():
    for i, n, data, start in _toys(40, 23):
        v64, e64, o64, _ = M.fit_cpu(M.MASS, data, start)
        assert o64.status == 0 and np.all(e64 > 0), i
        v32, e32, o32, _ = M.fit_cpu(M.MASS, data, start, fp32=True)
        assert o32.status == 0, (i, o32.status)
        assert np.max(np.abs(v32 - v64) / e64) < 1e-3, (i, (v32 - v64) / e64)
        np.testing.assert_allclose(e32, e64, rtol=1e-3)


def test_host_fibres_switch_and_reuse_stacks(tmp_path):
    """the batched fits' host fibres (morefit/fiber.hh): 3000 fibres on 8 pinned pool threads,
    ~22 switches each, rounding modes kept per fibre, exceptions and deep recursion inside
    fibres, stacks reused by a second round; and the thread pool alone: 6000 runs of 1-16
    threads, back to back (workers polling) and after pauses (workers blocked), each w called
    once per run on its own OS thread (built at -O0 and -O2)"""
    for opt in ("-O0", "-O2"):
        exe = str(tmp_path / ("fiber_check" + opt))
        subprocess.check_call(["g++", opt, "-std=c++17", "-I", os.path.join(ROOT, "morefit_amd", "include"),
                               os.path.join(ROOT, "tests", "cpp", "fiber_check.cc"), "-lpthread", "-o", exe])
        out = subprocess.check_output([exe]).decode().split()
        assert out[0] == "ok", out

"""The reference's toy-study API on the HIP backend (tests/cpp/massfit_toys.cc built with the
HIP typedefs by build()): toystudy::toy with HIPBlock, drop-in, against the same toys fitted
together by the batched engine (toystudy::toy_batched), and the benchmark's numerical
derivatives against the analytic ones.  Bars: every toy converges; per-toy fits agree within
1e-3 sigma (FP32 batched vs FP32 sequential, FP64 numerical vs FP64 analytic), 1e-4 sigma
(FP64 batched vs FP64 sequential)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def _run(name, *args):
    exe = os.path.join(ROOT, "tests", "cpp", "_build", name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    say(name, *args)
    out = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout[out.stdout.index("{"):])


@pytest.mark.parametrize("name,lim", [("massfit_toys_hip", 1e-4), ("massfit_toys_hip_f32", 1e-3)])
def test_toystudy_sequential_equals_batched(name, lim):
    nruns, n = 24, 20000
    r = _run(name, nruns, n, 1, 1)
    s, b = r["sequential"], r["batched"]
    assert s["status"] == [0] * nruns and b["status"] == [0] * nruns
    assert s["events"] == b["events"] == [n] * nruns
    vs, es, vb = np.array(s["values"]), np.array(s["errors"]), np.array(b["values"])
    assert np.max(np.abs(vb - vs) / es) < lim, np.max(np.abs(vb - vs) / es)
    np.testing.assert_allclose(np.array(b["errors"]), es, rtol=10 * lim)


def test_toystudy_benchmark_settings_on_device():
    nruns, n = 10, 20000
    a = _run("massfit_toys_hip", nruns, n, 1)["sequential"]
    m = _run("massfit_toys_hip", nruns, n, 0)["sequential"]
    assert a["status"] == [0] * nruns and m["status"] == [0] * nruns
    va, ea, vm = np.array(a["values"]), np.array(a["errors"]), np.array(m["values"])
    assert np.max(np.abs(vm - va) / ea) < 1e-3
    np.testing.assert_allclose(np.array(m["errors"]), ea, rtol=1e-3)

"""FP64 exp, log, reciprocal, sine and cosine of the per-event kernels, checked on the host.

The runtime (morefit_amd/csrc/kernel_runtime.hip) writes these out instead of calling the
device library; the block from mf_exp to the MF_SINCOS_END marker is plain C++ apart from three
AMDGPU builtins. It is cut out, compiled with g++ into tests/cpp/fp64_math_check.cc and compared
with the C library: within 2 ulp (reciprocal 1 ulp) over the ranges a density meets and far
beyond, next to multiples of pi/2, and at the special arguments (NaN, infinities, zero, overflow
and underflow of exp). A wrong coefficient, sign or quadrant fails by many ulp.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RUNTIME = os.path.join(HERE, "..", "morefit_amd", "csrc", "kernel_runtime.hip")


def _block():
    src = open(RUNTIME).read()
    i = src.index("__device__ __forceinline__ double mf_exp(double x) {")
    j = src.index("// MF_SINCOS_END")
    return src[i:j]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fp64_math_within_two_ulp(tmp_path):
    (tmp_path / "fp64_math.inc").write_text(_block())
    exe = tmp_path / "fp64_math_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", str(tmp_path),
                    os.path.join(HERE, "cpp", "fp64_math_check.cc"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all ok" in r.stdout

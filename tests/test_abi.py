"""The C ABI libraries load, export every symbol their headers declare, and report errors
through status codes (no device needed: only calls that fail before touching a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

import morefit_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, prefix):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", txt)))


def _exported(lib):
    out = subprocess.check_output(["nm", "-D", "--defined-only", os.path.join(M.LIBDIR, lib)]).decode()
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


@pytest.mark.parametrize("header,lib,prefix,listed", [
    ("morefit_hip.h", "libmorefit_hip.so", "mf_", M.HIP_SYMBOLS),
    ("morefit_capi.h", "libmorefit_capi.so", "mfh_", M.CAPI_SYMBOLS),
])
def test_every_declared_symbol_is_exported(header, lib, prefix, listed):
    M.libs()
    declared = _declared(header, prefix)
    assert declared, header
    exp = _exported(lib)
    missing = [s for s in declared if s not in exp]
    assert not missing, missing
    assert sorted(listed) == declared


def test_compile_error_is_reported_not_fatal():
    hip, _ = M.libs()
    hip.mf_module_compile_only.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]
    hip.mf_last_error.restype = ctypes.c_char_p
    n = ctypes.c_size_t(0)
    rc = hip.mf_module_compile_only(b"#define MF_T double\n//@MF_RUNTIME@\nthis is not C++;\n", b"", ctypes.byref(n))
    assert rc == 3  # MF_ERR_COMPILE
    assert b"error" in hip.mf_last_error()


def test_injected_failure():
    hip, _ = M.libs()
    hip.mf_debug_fail_next.argtypes = [ctypes.c_char_p, ctypes.c_int]
    hip.mf_module_compile_only.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]
    assert hip.mf_debug_fail_next(b"mf_module_compile_only", 1) == 0
    n = ctypes.c_size_t(0)
    assert hip.mf_module_compile_only(b"x", b"", ctypes.byref(n)) == 6  # MF_ERR_INJECTED
    hip.mf_last_error.restype = ctypes.c_char_p
    assert b"injected" in hip.mf_last_error()


def test_null_arguments_rejected():
    hip, _ = M.libs()
    hip.mf_malloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    assert hip.mf_malloc(None, 16, None) == 1  # MF_ERR_INVALID
    hip.mf_launch_eval.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    assert hip.mf_launch_eval(None, b"k", None, 1, 256) == 1
    hip.mf_query.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    assert hip.mf_query(None, None) == 1


def test_generated_modules_compile_for_gfx950():
    """the four configuration modules (mass/angular x FP64/FP32) compile with hipRTC"""
    for kind in (M.MASS, M.KSTAR):
        for fp32 in (False, True):
            M.precompile(kind, fp32)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_open_without_gpu_raises():
    with pytest.raises(M.MorefitError):
        M.Context()


class _Res(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64)] + [(f, ctypes.c_int) for f in (
        "vgpr_count", "agpr_count", "sgpr_count", "vgpr_spill_count", "sgpr_spill_count",
        "private_segment_bytes", "group_segment_bytes", "pad")]


_LDS_MODULE = b"""#define MF_T float
#define MF_FP32 1
#define MF_NOBS 1
#define MF_NTHETA 1
#define MF_NHOIST 0
#define MF_BLOCK 256
#define MF_ACC 1
#define MF_FAST_MATH 1
#define MF_GEN_PROPOSE(x, s) do { } while (0)
//@MF_RUNTIME@
extern "C" __global__ void k_lds_probe(float* p) {
  __shared__ float s[1000];
  s[threadIdx.x] = p[threadIdx.x];
  __syncthreads();
  p[threadIdx.x] = s[999 - threadIdx.x];
}
"""


def test_module_resources_reads_code_object_metadata():
    """mf_module_resources parses the AMDGPU metadata note: a known kernel with a 4000-byte
    LDS array reports that size, its name, and registers within the architectural VGPRs"""
    hip, _ = M.libs()
    hip.mf_module_resources.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_Res), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int)]
    out = (_Res * 8)()
    n = ctypes.c_int(0)
    assert hip.mf_module_resources(_LDS_MODULE, b"", out, 8, ctypes.byref(n)) == 0
    res = {out[k].name.decode(): out[k] for k in range(min(n.value, 8))}
    assert "k_lds_probe" in res, list(res)
    r = res["k_lds_probe"]
    assert r.group_segment_bytes == 4000
    assert r.agpr_count == 0 and r.private_segment_bytes == 0
    assert 0 < r.vgpr_count <= 256


def test_kernels_fit_register_budget():
    """Every evaluation kernel of the configuration modules is fitted to the 256 architectural
    VGPRs: no AGPRs and no scratch.  At unroll 8 the FP64 angular Hessian kernel needs AGPRs
    (the condition under which the compiler left loop-carried address registers unset in
    lanes outside a divergent branch); the fit lowers its unroll until it needs none."""
    raw = M.kernel_resources(M.KSTAR, False, fit=False, unroll=8)
    assert raw[M.EVAL_HESS]["agpr"] > 0
    for kind in (M.MASS, M.KSTAR):
        for fp32 in (False, True):
            fit = M.kernel_resources(kind, fp32, fit=True, unroll=8)
            for k, r in fit.items():
                assert 1 <= r["unroll"] <= 8, (kind, fp32, k, r)
                assert r["agpr"] == 0 and r["scratch"] == 0, (kind, fp32, k, r)
    assert M.kernel_resources(M.KSTAR, False, fit=True, unroll=8)[M.EVAL_HESS]["unroll"] < 8


def test_autotune_candidates_by_workload_size():
    """The autotuner's launch shapes by workload: 2-16 events per thread below 4e6 events, the
    default 8-64 from there, and the default plus 16 events' loads in flight per thread from 1e8
    (where the kernel streams); every shape's kernels fit the register budget, an unroll of 16
    included"""
    small, mid, large = (M.autotune_candidates(n) for n in (10**6, 10**7, 5 * 10**8))
    assert {e for _, e, _ in small} == {2, 4, 8, 16} and all(u <= e for _, e, u in small)
    assert {e for _, e, _ in mid} == {8, 16, 32, 64} and {u for _, _, u in mid} == {4, 8}
    assert set(mid) < set(large) and {s for s in large if s not in mid} == {(b, e, 16) for b in (256, 512) for e in (32, 64)}
    for kind in (M.MASS, M.KSTAR):
        for k, r in M.kernel_resources(kind, True, fit=True, unroll=16).items():
            assert 1 <= r["unroll"] <= 16 and r["agpr"] == 0 and r["scratch"] == 0, (kind, k, r)
    assert M.kernel_resources(M.MASS, True, fit=True, unroll=16)[M.EVAL_GRAD]["unroll"] == 16


def test_fast_eval_extension_binds_and_checks_arguments():
    """morefit_amd._fast (Batch.eval through the CPython C API) is built, bound to the loaded
    library's mfh_batch_eval, and rejects a malformed call before reaching the library: wrong
    argument count, a theta of neither nseg x ntheta nor ntheta values, negative sizes"""
    fast = M._fast_eval()
    assert fast is not None, "morefit_amd/_fast.so not built"
    with pytest.raises(TypeError):
        fast(1, 2, [0.0])
    with pytest.raises(ValueError):
        fast(1, 2, [0.0] * 5, 0, 2, 4, 5)  # 5 values for 2 segments x 4 parameters
    with pytest.raises(ValueError):
        fast(1, 2, [0.0] * 4, 0, -1, 4, 5)
    with pytest.raises(ValueError):
        fast(0, 2, [0.0] * 4, 0, 1, 4, 5)  # null batch handle

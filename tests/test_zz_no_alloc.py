"""mf_launch_eval_noalloc (the evaluation launch that allocates nothing: its per-chunk partials
go to the library's static per-GPU scratch) through the device C ABI, against
mf_launch_eval with a partials buffer of the caller's: the totals are bitwise the same, on
two calls in a row with the second over fewer events; oversize launches and in-kernel
segment reduction are refused before anything is enqueued."""
import ctypes
import re
import sys

import numpy as np
import pytest

import morefit_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

vp, ip = ctypes.c_void_p, ctypes.c_int
MF_OK, MF_ERR_INVALID = 0, 1
SCRATCH_BYTES = 4 << 20  # MF_NOALLOC_SCRATCH_BYTES


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


class Segment(ctypes.Structure):
    _fields_ = [("data_off", ctypes.c_longlong), ("stride", ctypes.c_longlong), ("n", ip), ("param_off", ip),
                ("chunk_begin", ip), ("chunk_end", ip), ("out_off", ip), ("mode", ip)]


class Chunk(ctypes.Structure):
    _fields_ = [("seg", ip), ("begin", ip), ("end", ip), ("pad", ip)]


class EvalArgs(ctypes.Structure):
    _fields_ = [("data", vp), ("segs", vp), ("chunks", vp), ("nchunks", ip), ("nseg", ip), ("params", vp),
                ("partials", vp), ("results", vp), ("hoist64", vp), ("store", vp), ("store_stride", ctypes.c_longlong),
                ("flags", ip), ("pad", ip), ("seg_done", vp), ("done_flag", vp), ("done_value", ip),
                ("inline_params", ip), ("pinline", ctypes.c_uint32 * 128), ("hinline", ctypes.c_double * 64),
                ("moments", vp)]


def _hip():
    hip, _ = M.libs()
    hip.mf_device_open.argtypes = [ip, ctypes.POINTER(vp)]
    hip.mf_device_close.argtypes = [vp]
    hip.mf_malloc.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(vp)]
    hip.mf_free.argtypes = [vp, vp]
    hip.mf_memcpy_h2d.argtypes = [vp, vp, vp, ctypes.c_size_t]
    hip.mf_memcpy_d2h.argtypes = [vp, vp, vp, ctypes.c_size_t]
    hip.mf_sync.argtypes = [vp]
    hip.mf_module_load.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]
    hip.mf_module_unload.argtypes = [vp]
    hip.mf_launch_prep.argtypes = [vp, vp, vp, vp, vp, ip, vp, vp]
    hip.mf_launch_eval.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(EvalArgs), ip, ip]
    hip.mf_launch_eval_noalloc.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(EvalArgs), ip, ip]
    hip.mf_last_error.restype = ctypes.c_char_p
    return hip


def test_noalloc_launch_matches_launch_with_partials():
    hip = _hip()
    ctx = M.Context(device=0)
    mdl = ctx.model(M.MASS, fp32=True)
    src = mdl.source()
    block = int(re.search(r"#define MF_BLOCK (\d+)", src).group(1))
    nmom = int(re.search(r"#define MF_NMOM (\d+)", src).group(1))
    ntheta, nhoist, nout = mdl.ntheta, mdl.nhoist, mdl.nout[M.EVAL_GRAD]
    dev, mod = vp(), vp()
    assert hip.mf_device_open(0, ctypes.byref(dev)) == MF_OK, hip.mf_last_error()
    assert hip.mf_module_load(dev, src.encode(), b"", ctypes.byref(mod)) == MF_OK, hip.mf_last_error()
    bufs = []

    def alloc(nbytes):
        p = vp()
        assert hip.mf_malloc(dev, max(int(nbytes), 16), ctypes.byref(p)) == MF_OK, hip.mf_last_error()
        bufs.append(p)
        return p

    def h2d(dst, arr):
        assert hip.mf_memcpy_h2d(dev, dst, arr.ctypes.data, arr.nbytes) == MF_OK, hip.mf_last_error()

    nmax, events_per_chunk = 400000, block * 16
    data_d = alloc(nmax * 4)
    theta_d, params_d, hoist_d = alloc(8 * ntheta), alloc(4 * (ntheta + nhoist)), alloc(8 * max(1, nhoist))
    segs_d = alloc(ctypes.sizeof(Segment))
    chunks_d = alloc(ctypes.sizeof(Chunk) * (nmax // events_per_chunk + 1))
    partials_d = alloc(SCRATCH_BYTES)  # the caller's own buffer for mf_launch_eval
    res_d = alloc(8 * nout)
    mom_d = alloc(8 * max(1, nmom))  # the segment's data moments (mf_eval_mom), read by the outputs' combination
    theta = np.array(M.MASS_START, dtype=np.float64)
    h2d(theta_d, theta)
    assert hip.mf_launch_prep(mod, theta_d, params_d, hoist_d, None, 1, None, None) == MF_OK, hip.mf_last_error()

    def run(n, noalloc):
        """-2lnL + gradient sums of the first n stored events through one of the two entry points"""
        assert 0 < n <= nmax
        nch = (n + events_per_chunk - 1) // events_per_chunk
        chunks = (Chunk * nch)(*[Chunk(0, c * events_per_chunk, min(n, (c + 1) * events_per_chunk), 0) for c in range(nch)])
        assert all(0 <= c.begin < c.end <= n for c in chunks)
        seg = Segment(0, n, n, 0, 0, nch, 0, 1)
        assert hip.mf_memcpy_h2d(dev, segs_d, ctypes.addressof(seg), ctypes.sizeof(seg)) == MF_OK
        assert hip.mf_memcpy_h2d(dev, chunks_d, ctypes.addressof(chunks), ctypes.sizeof(chunks)) == MF_OK
        if nmom:
            m = EvalArgs()
            m.data, m.segs, m.chunks, m.nchunks, m.nseg = data_d, segs_d, chunks_d, nch, 1
            m.params, m.results, m.partials = params_d, mom_d, partials_d
            assert hip.mf_launch_eval(mod, b"mf_eval_mom", ctypes.byref(m), nch, block) == MF_OK, hip.mf_last_error()
        a = EvalArgs()
        a.data, a.segs, a.chunks, a.nchunks, a.nseg = data_d, segs_d, chunks_d, nch, 1
        a.params, a.results, a.hoist64 = params_d, res_d, hoist_d
        a.moments = mom_d if nmom else None
        a.partials = None if noalloc else partials_d
        say("eval", "noalloc" if noalloc else "with partials", "n", n, "chunks", nch)
        f = hip.mf_launch_eval_noalloc if noalloc else hip.mf_launch_eval
        assert f(mod, b"mf_eval_grad", ctypes.byref(a), nch, block) == MF_OK, hip.mf_last_error()
        out = np.zeros(nout)
        assert hip.mf_memcpy_d2h(dev, out.ctypes.data, res_d, out.nbytes) == MF_OK
        assert hip.mf_sync(dev) == MF_OK, hip.mf_last_error()
        return out

    data = O.generate(M.MASS, M.MASS_TRUTH, nmax, seed=77, nstreams=8)
    stored = (data[0] - 5.3).astype(np.float32)  # FP32 modules store offsets from the range centre
    h2d(data_d, stored)
    for n in (nmax, 70001):
        ref = run(n, noalloc=False)
        got = run(n, noalloc=True)
        assert np.array_equal(got, ref), (n, got - ref)
        # and both are the sums of the oracle on the same stored events (FP32 bar)
        st = (stored[:n].astype(np.float64) + 5.3)[None, :]
        o = O.nll_sums(M.MASS, st, theta, mode=1, nthreads=0)
        oabs = O.nll_sums(M.MASS, st, theta, mode=11, nthreads=1)
        assert np.max(np.abs(got - o) / oabs) <= 2e-6

    # refused before anything is enqueued: more partials than the scratch holds, and the
    # in-kernel segment reduction (its counters would have to start at zero)
    a = EvalArgs()
    a.data, a.segs, a.chunks, a.params, a.results, a.hoist64, a.nseg = data_d, segs_d, chunks_d, params_d, res_d, hoist_d, 1
    a.nchunks = SCRATCH_BYTES // 8 + 1
    assert hip.mf_launch_eval_noalloc(mod, b"mf_eval_grad", ctypes.byref(a), 1, block) == MF_ERR_INVALID
    assert b"scratch" in hip.mf_last_error()
    a.nchunks, a.seg_done = 1, res_d
    assert hip.mf_launch_eval_noalloc(mod, b"mf_eval_grad", ctypes.byref(a), 1, block) == MF_ERR_INVALID
    assert hip.mf_launch_eval_noalloc(mod, b"mf_prep", ctypes.byref(a), 1, block) != MF_OK
    # and a launch whose outputs need the data moments without them
    if nmom:
        a.seg_done, a.moments, a.partials = None, None, partials_d
        assert hip.mf_launch_eval(mod, b"mf_eval_grad", ctypes.byref(a), 1, block) == MF_ERR_INVALID
        assert b"moments" in hip.mf_last_error()

    for p in bufs:
        hip.mf_free(dev, p)
    hip.mf_module_unload(mod)
    hip.mf_device_close(dev)
    mdl.close()
    ctx.close()

"""morefit_amd -- Python bindings (ctypes) of the MI355X MoreFit likelihood backend.

The product is native: libmorefit_hip.so (device boundary, include/morefit_hip.h) and
libmorefit_capi.so (host C ABI over the morefit C++ headers, include/morefit_capi.h).  This
module only loads them and converts numpy arrays; it has no compute path of its own and
raises if the native libraries are missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
MASS, KSTAR, CB, POLY = 0, 1, 2, 3
EVAL_LH, EVAL_GRAD, EVAL_HESS = 1, 2, 3
NTHETA = {MASS: 4, KSTAR: 8, CB: 4, POLY: 3}
NOBS = {MASS: 1, KSTAR: 3, CB: 1, POLY: 1}
THETA_NAMES = {MASS: ["mu", "sigma", "f", "lambda"], KSTAR: ["FL", "S3", "S4", "S5", "AFB", "S7", "S8", "S9"],
               CB: ["mu", "sigma", "alpha", "n"], POLY: ["c1", "c2", "c3"]}
MASS_TRUTH = np.array([5.28, 0.02, 0.3, -1.5])
KSTAR_TRUTH = np.array([0.7, 0.0, -0.1, -0.2, 0.1, 0.0, 0.0, 0.0])
CB_TRUTH = np.array([5.28, 0.02, 1.5, 3.0])
POLY_TRUTH = np.array([0.3, -0.2, 0.1])
TRUTH = {MASS: MASS_TRUTH, KSTAR: KSTAR_TRUTH, CB: CB_TRUTH, POLY: POLY_TRUTH}
# start values: +10 % of the window width for mu, +10 % of the value otherwise, zeros stay zero
MASS_START = np.array([5.28 + 0.06, 0.022, 0.33, -1.65])
KSTAR_START = np.array([0.77, 0.0, -0.11, -0.22, 0.11, 0.0, 0.0, 0.0])
START = {MASS: MASS_START, KSTAR: KSTAR_START, CB: np.array([5.285, 0.021, 1.6, 3.3]), POLY: np.array([0.33, -0.22, 0.11])}
# statistical errors of one fit at the truth, per parameter, scaled as 1/sqrt(N) from N_REF events
SIGMA_REF = {MASS: (np.array([1.57e-4, 1.41e-4, 1.93e-3, 2.24e-2]), 100000),
             KSTAR: (np.array([2.9e-3, 3.8e-3, 4.7e-3, 4.7e-3, 2.6e-3, 4.8e-3, 4.8e-3, 3.8e-3]), 50000),
             CB: (np.array([7.9e-5, 6.3e-5, 1.7e-2, 8.2e-2]), 100000),
             POLY: (np.array([1.25e-2, 9.2e-3, 1.88e-2]), 100000)}


def toy_starts(kind, sizes, seed):
    """independent start values per toy: truth + N(0, 1) * sigma_stat(N), fixed by the seed"""
    sizes = np.asarray(sizes, dtype=float)
    sig, nref = SIGMA_REF[kind]
    rs = np.random.default_rng(seed)
    return TRUTH[kind] + rs.standard_normal((len(sizes), len(sig))) * sig * np.sqrt(nref / np.maximum(sizes, 1.0))[:, None]


class MorefitError(RuntimeError):
    pass


class Options(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("block_size", ctypes.c_int), ("events_per_thread", ctypes.c_int),
                ("grid", ctypes.c_int), ("fast_math", ctypes.c_int), ("accumulate", ctypes.c_int), ("centre", ctypes.c_int),
                ("unroll", ctypes.c_int), ("balance", ctypes.c_int), ("base2", ctypes.c_int),
                ("fused_final", ctypes.c_int), ("event_columns", ctypes.c_int), ("event_cost", ctypes.c_double),
                ("linear", ctypes.c_int), ("exp_pivot", ctypes.c_int),
                ("host_params", ctypes.c_int), ("prefetch", ctypes.c_int)]


class FitOut(ctypes.Structure):
    _fields_ = [("fval", ctypes.c_double), ("edm", ctypes.c_double), ("status", ctypes.c_int), ("ncalls", ctypes.c_int),
                ("iterations", ctypes.c_int), ("hesse_ok", ctypes.c_int), ("thread_moved", ctypes.c_int),
                ("nonfinite_evals", ctypes.c_int), ("nonfinite_events", ctypes.c_longlong), ("first_nonfinite", ctypes.c_longlong),
                ("out_of_range", ctypes.c_longlong)]


class BatchStats(ctypes.Structure):
    _fields_ = [("ticks", ctypes.c_int), ("launches_lh", ctypes.c_int), ("launches_grad", ctypes.c_int),
                ("launches_hess", ctypes.c_int), ("wall_ms", ctypes.c_double), ("gpu_wait_ms", ctypes.c_double),
                ("host_ms", ctypes.c_double), ("host_threads", ctypes.c_int), ("pad", ctypes.c_int),
                ("setup_ms", ctypes.c_double), ("step_ms", ctypes.c_double)]


_hip = None
_capi = None

# every symbol include/morefit_hip.h and include/morefit_capi.h declare
HIP_SYMBOLS = [
    "mf_last_error", "mf_device_open", "mf_device_close", "mf_device_query", "mf_malloc", "mf_free", "mf_host_alloc",
    "mf_host_free", "mf_memcpy_h2d", "mf_memcpy_d2h", "mf_memcpy_d2d", "mf_fill_zero", "mf_sync", "mf_query", "mf_module_load",
    "mf_module_compile_only", "mf_module_unload", "mf_module_has_kernel", "mf_module_compile_ms", "mf_launch_eval",
    "mf_launch_prep", "mf_launch_prep_inline", "mf_launch_generate", "mf_graph_begin", "mf_graph_end", "mf_graph_launch", "mf_graph_destroy",
    "mf_timer_begin", "mf_timer_end", "mf_bench_copy", "mf_debug_fail_next", "mf_module_resources", "mf_host_alloc_mapped", "mf_launch_seed_streams", "mf_module_occupancy",
    "mf_launch_eval_noalloc", "mf_launch_check", "mf_launch_expand_streams",
]
CAPI_SYMBOLS = [
    "mfh_last_error", "mfh_default_options", "mfh_open", "mfh_close", "mfh_device_name", "mfh_copy_ceiling",
    "mfh_model_create", "mfh_model_destroy", "mfh_model_info", "mfh_model_source", "mfh_model_precompile", "mfh_model_kernels", "mfh_model_host_outputs", "mfh_stream_states", "mfh_autotune", "mfh_autotune_candidates",
    "mfh_batch_create", "mfh_batch_destroy", "mfh_batch_upload", "mfh_batch_download", "mfh_batch_generate",
    "mfh_batch_eval", "mfh_batch_check", "mfh_batch_time_kernel", "mfh_batch_time_eval", "mfh_batch_time_graph", "mfh_batch_bytes", "mfh_batch_fit",
    "mfh_fit_plain", "mfh_fit_cpu", "mfh_eval_cpu", "mfh_serve_run", "mfh_fit_weighted", "mfh_fit_options",
    "mfh_plain_create", "mfh_plain_fit", "mfh_plain_eval", "mfh_plain_destroy", "mfh_plot",
]


_FAST = []


def _fast_eval():
    """morefit_amd._fast.eval (the per-call evaluation through the CPython C API), bound to the
    loaded library's mfh_batch_eval; None when the extension is not built (the ctypes path)"""
    if not _FAST:
        try:
            from . import _fast
            _fast.init(ctypes.cast(libs()[1].mfh_batch_eval, ctypes.c_void_p).value)
            _FAST.append(_fast.eval)
        except ImportError:
            _FAST.append(None)
    return _FAST[0]


def libs():
    """Load (once) and return (libmorefit_hip, libmorefit_capi); raise if not built."""
    global _hip, _capi
    if _capi is None:
        hp = os.path.join(LIBDIR, "libmorefit_hip.so")
        cp = os.path.join(LIBDIR, "libmorefit_capi.so")
        if not (os.path.exists(hp) and os.path.exists(cp)):
            raise MorefitError("native libraries not built: run __graft_entry__.build() (make -C morefit_amd/csrc)")
        _hip = ctypes.CDLL(hp, mode=ctypes.RTLD_GLOBAL)
        L = ctypes.CDLL(cp)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(ctypes.c_void_p)
        L.mfh_last_error.restype = ctypes.c_char_p
        L.mfh_open.argtypes = [ctypes.POINTER(Options), pp]
        L.mfh_close.argtypes = [vp]
        L.mfh_device_name.argtypes = [vp, ctypes.c_char_p, ip, ctypes.POINTER(ctypes.c_int)]
        L.mfh_copy_ceiling.argtypes = [vp, ctypes.c_double, ip, dp]
        L.mfh_model_create.argtypes = [vp, ip, ip, dp, ip, pp]
        L.mfh_model_destroy.argtypes = [vp]
        L.mfh_model_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int), dp]
        L.mfh_model_source.argtypes = [vp, ctypes.c_char_p, ip]
        L.mfh_model_precompile.argtypes = [ip, ip, ip, ctypes.POINTER(Options), dp]
        L.mfh_model_kernels.argtypes = [ip, ip, ip, ctypes.POINTER(Options), ip, ctypes.POINTER(ctypes.c_int)]
        L.mfh_model_host_outputs.argtypes = [ip, ip, ip, dp, dp, ctypes.c_longlong, ctypes.c_double, ip, dp]
        L.mfh_autotune_candidates.argtypes = [ctypes.c_longlong, ctypes.POINTER(ctypes.c_int), ip, ctypes.POINTER(ctypes.c_int)]
        L.mfh_stream_states.argtypes = [ctypes.c_ulonglong, ctypes.c_longlong, ip, ctypes.POINTER(ctypes.c_uint32)]
        L.mfh_autotune.argtypes = [vp, ip, ip, ctypes.POINTER(ctypes.c_longlong), ip, ip, ctypes.POINTER(ctypes.c_int), ip, ip,
                                   ctypes.POINTER(ctypes.c_int), dp]
        L.mfh_batch_create.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ip, pp]
        L.mfh_batch_destroy.argtypes = [vp]
        L.mfh_batch_upload.argtypes = [vp, ip, dp, ctypes.c_longlong]
        L.mfh_batch_download.argtypes = [vp, ip, dp]
        L.mfh_batch_generate.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_ulonglong), ip, ctypes.POINTER(ctypes.c_ulonglong)]
        L.mfh_batch_eval.argtypes = [vp, ip, dp, ip, dp]
        L.mfh_batch_check.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ip, dp, ctypes.POINTER(ctypes.c_longlong)]
        L.mfh_batch_time_kernel.argtypes = [vp, ip, ip, ip, dp]
        L.mfh_batch_time_graph.argtypes = [vp, ip, ip, dp]
        L.mfh_batch_time_eval.argtypes = [vp, ip, ip, ip, dp]
        L.mfh_batch_bytes.argtypes = [vp]
        L.mfh_batch_bytes.restype = ctypes.c_longlong
        L.mfh_batch_fit.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ip, dp, ip, dp, dp, ctypes.POINTER(FitOut), ctypes.POINTER(BatchStats)]
        L.mfh_fit_plain.argtypes = [vp, ip, ip, dp, ctypes.c_longlong, dp, dp, dp, ctypes.POINTER(FitOut), dp]
        L.mfh_fit_cpu.argtypes = [ip, ip, dp, ctypes.c_longlong, dp, ip, ip, dp, dp, ctypes.POINTER(FitOut), dp]
        L.mfh_fit_weighted.argtypes = [vp, ip, ip, dp, dp, ctypes.c_longlong, dp, ip, dp, dp, ctypes.POINTER(FitOut), dp]
        L.mfh_eval_cpu.argtypes = [ip, dp, ctypes.c_longlong, dp, ip, dp, dp, dp]
        L.mfh_fit_options.argtypes = [vp, ip, ip, dp, ctypes.c_longlong, dp, ip, ip, dp, dp, ctypes.POINTER(FitOut), dp]
        L.mfh_plain_create.argtypes = [vp, ip, ip, ip, ip, ctypes.c_char_p, pp]
        L.mfh_plain_fit.argtypes = [vp, dp, ctypes.c_longlong, dp, dp, dp, ctypes.POINTER(FitOut), dp]
        L.mfh_plain_eval.argtypes = [vp, dp, ctypes.c_longlong, dp, ip, dp, dp, dp]
        L.mfh_plain_destroy.argtypes = [vp]
        L.mfh_plot.argtypes = [vp, ip, ip, ip, dp, dp, dp, ctypes.c_longlong, ip, ip, ip, ctypes.c_double, ctypes.c_double, ip, ip,
                               ctypes.c_char_p, ip, dp, dp, dp, dp, dp]
        L.mfh_serve_run.argtypes = [ctypes.POINTER(ctypes.c_void_p), ip, dp, dp, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_ulonglong), dp, ip,
                                    ctypes.c_double, ip, dp, dp, dp, ctypes.POINTER(ctypes.c_int), dp, dp]
        _capi = L
    return _hip, _capi


def _check(rc):
    if rc != 0:
        raise MorefitError(libs()[1].mfh_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def options(**kw):
    o = Options()
    libs()[1].mfh_default_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, float(v) if k == "event_cost" else int(v))
    return o


def kernel_resources(kind, fp32, with_hessian=True, fit=True, **kw):
    """Without a device: per evaluation kind {EVAL_LH, EVAL_GRAD, EVAL_HESS} the kernel's
    unroll, VGPRs, AGPRs and scratch bytes, fitted to the register budget (fit=True, as a
    model is built) or at the requested unroll as given (fit=False)."""
    info = (ctypes.c_int * 15)()
    _check(libs()[1].mfh_model_kernels(kind, int(fp32), int(with_hessian), ctypes.byref(options(**kw)), int(fit), info))
    return {k: {"unroll": info[j], "vgpr": info[3 + j], "agpr": info[6 + j], "scratch": info[9 + j], "fused": bool(info[12 + j])}
            for j, k in enumerate((EVAL_LH, EVAL_GRAD, EVAL_HESS))}


def host_outputs(kind, fp32, which, theta, data, event_cost=-1.0, linear=1):
    """Per-event outputs of a configuration module's value/gradient/Hessian kernel (which = 1/2/3)
    as generated, evaluated on the host in FP64 (no device): array (n, nout).  event_cost >= 0:
    with the event-only terms above that cost taken from precomputed columns."""
    data = _f64(np.atleast_2d(data))
    n = data.shape[1]
    nout = {1: 1, 2: 1 + NTHETA[kind], 3: 1 + NTHETA[kind] + NTHETA[kind] * (NTHETA[kind] + 1) // 2}[which]
    out = np.zeros(n * nout)
    _check(libs()[1].mfh_model_host_outputs(kind, int(fp32), which, _dp(_f64(theta)), _dp(data), n, float(event_cost), int(linear), _dp(out)))
    return out.reshape(n, nout)


def stream_states(seed, first, count):
    """xoshiro128++ states of generator streams first .. first+count-1 of `seed`, by the
    jump-ahead matrices the device seeding uses (no device): array (count, 4) of uint32"""
    out = (ctypes.c_uint32 * (4 * count))()
    _check(libs()[1].mfh_stream_states(int(seed), int(first), int(count), out))
    return np.frombuffer(out, dtype=np.uint32).reshape(count, 4).copy()


def autotune_candidates(total_events):
    """the autotuner's built-in launch shapes (block, events per thread, unroll) for a workload
    of total_events events"""
    cand = (ctypes.c_int * 3 * 64)()
    n = ctypes.c_int(0)
    _check(libs()[1].mfh_autotune_candidates(int(total_events), ctypes.cast(cand, ctypes.POINTER(ctypes.c_int)), 64, ctypes.byref(n)))
    flat = list(ctypes.cast(cand, ctypes.POINTER(ctypes.c_int))[:3 * n.value])
    return [tuple(flat[3 * i:3 * i + 3]) for i in range(n.value)]


def precompile(kind, fp32, with_hessian=True, **kw):
    """Compile a configuration module for gfx950 into the on-disk cache (no device needed)."""
    ms = ctypes.c_double(0)
    _check(libs()[1].mfh_model_precompile(kind, int(fp32), int(with_hessian), ctypes.byref(options(**kw)), ctypes.byref(ms)))
    return ms.value


class Context:
    """One device (HIPBackend) with its stream."""

    def __init__(self, **kw):
        self.h = ctypes.c_void_p()
        _check(libs()[1].mfh_open(ctypes.byref(options(**kw)), ctypes.byref(self.h)))

    def close(self):
        if self.h:
            libs()[1].mfh_close(self.h)
            self.h = ctypes.c_void_p()

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        cu = ctypes.c_int(0)
        _check(libs()[1].mfh_device_name(self.h, buf, 256, ctypes.byref(cu)))
        return buf.value.decode(), cu.value

    def copy_ceiling(self, nbytes=2 << 30, iters=5):
        g = ctypes.c_double(0)
        _check(libs()[1].mfh_copy_ceiling(self.h, float(nbytes), iters, ctypes.byref(g)))
        return g.value

    def autotune(self, kind, fp32, sizes, eval_kind=2, candidates=None, iters=30):
        """Pick the launch shape (block, events per thread, unroll) of this context's evaluation
        kernels by timing `eval_kind` evaluations of the configuration density on a batch of
        the given segment sizes (generated on the device).  Models created afterwards use it.
        Returns {"best": (block, ept, unroll), "table": [((block, ept, unroll), us), ...]}."""
        cands = candidates or autotune_candidates(int(np.sum(np.asarray(sizes, dtype=np.int64))))
        flat = (ctypes.c_int * (3 * len(cands)))(*[int(v) for c in cands for v in c])
        sz = (ctypes.c_longlong * len(sizes))(*[int(s) for s in sizes])
        best = (ctypes.c_int * 3)()
        us = np.zeros(len(cands))
        _check(libs()[1].mfh_autotune(self.h, kind, int(fp32), sz, len(sizes), eval_kind, flat, len(cands), iters, best, _dp(us)))
        return {"best": tuple(best), "table": [(tuple(c), float(u)) for c, u in zip(cands, us)]}

    def model(self, kind, fp32=False, rng=None, with_hessian=True):
        return Model(self, kind, fp32, rng, with_hessian)

    def fit_plain(self, kind, data, start, fp32=False):
        """One fit through the drop-in HIPBlock (reference fitter flow)."""
        data = _f64(np.atleast_2d(data))
        P = NTHETA[kind]
        vals, errs = np.zeros(P), np.zeros(P)
        out, ms = FitOut(), ctypes.c_double(0)
        _check(libs()[1].mfh_fit_plain(self.h, kind, int(fp32), _dp(data), data.shape[1], _dp(_f64(start)), _dp(vals), _dp(errs),
                                       ctypes.byref(out), ctypes.byref(ms)))
        return vals, errs, out, ms.value


class Model:
    """A density compiled for the device: value, gradient and Hessian kernels."""

    def __init__(self, ctx, kind, fp32, rng, with_hessian):
        self.ctx, self.kind, self.fp32 = ctx, kind, fp32
        self.h = ctypes.c_void_p()
        r = None if rng is None else _dp(_f64(rng))
        self._rng = None if rng is None else _f64(rng)
        if self._rng is not None:
            r = _dp(self._rng)
        _check(libs()[1].mfh_model_create(ctx.h, kind, int(fp32), r, int(with_hessian), ctypes.byref(self.h)))
        info = (ctypes.c_int * 18)()
        ms = (ctypes.c_double * 2)()
        _check(libs()[1].mfh_model_info(self.h, info, ms))
        self.ntheta, self.nhoist, self.nobs = info[0], info[1], info[2]
        self.nout = {EVAL_LH: info[3], EVAL_GRAD: info[4], EVAL_HESS: info[5]}
        self.statements, self.graph_nodes = info[6], info[7]
        self.unroll = {EVAL_LH: info[8], EVAL_GRAD: info[9], EVAL_HESS: info[10]}
        self.vgprs = {EVAL_LH: info[11], EVAL_GRAD: info[12], EVAL_HESS: info[13]}
        self.fused = {EVAL_LH: bool(info[14]), EVAL_GRAD: bool(info[15]), EVAL_HESS: bool(info[16])}
        self.ncol = info[17]
        self.codegen_ms, self.compile_ms = ms[0], ms[1]

    def source(self):
        buf = ctypes.create_string_buffer(1 << 22)
        _check(libs()[1].mfh_model_source(self.h, buf, 1 << 22))
        return buf.value.decode()

    def batch(self, sizes):
        return Batch(self, sizes)

    def close(self):
        if self.h:
            libs()[1].mfh_model_destroy(self.h)
            self.h = ctypes.c_void_p()


class Batch:
    """Resident datasets (segments) of one model."""

    # mfh_batch_eval with plain addresses: the per-call path of eval() avoids ctypes pointer objects
    _EVAL_PROTO = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p)

    def __init__(self, model, sizes):
        self.model = model
        self.sizes = np.asarray(sizes, dtype=np.int64)
        self.S = len(self.sizes)
        self.h = ctypes.c_void_p()
        arr = (ctypes.c_longlong * self.S)(*[int(s) for s in self.sizes])
        _check(libs()[1].mfh_batch_create(model.h, arr, self.S, ctypes.byref(self.h)))
        self._eval = Batch._EVAL_PROTO(ctypes.cast(libs()[1].mfh_batch_eval, ctypes.c_void_p).value)
        self._hv = self.h.value

    def close(self):
        if self.h:
            libs()[1].mfh_batch_destroy(self.h)
            self.h = ctypes.c_void_p()
            self._hv = None

    def nbytes(self):
        return libs()[1].mfh_batch_bytes(self.h)

    def upload(self, seg, cols):
        cols = _f64(np.atleast_2d(cols))
        assert cols.shape == (self.model.nobs, self.sizes[seg])
        _check(libs()[1].mfh_batch_upload(self.h, seg, _dp(cols), cols.shape[1]))

    def download(self, seg):
        out = np.zeros((self.model.nobs, int(self.sizes[seg])))
        _check(libs()[1].mfh_batch_download(self.h, seg, _dp(out)))
        return out

    def generate(self, theta_gen, seeds, events_per_stream=64):
        th = _f64(np.broadcast_to(np.asarray(theta_gen, dtype=np.float64), (self.S, self.model.ntheta)))
        sd = (ctypes.c_ulonglong * self.S)(*[int(s) for s in seeds])
        props = ctypes.c_ulonglong(0)
        _check(libs()[1].mfh_batch_generate(self.h, _dp(th), sd, events_per_stream, ctypes.byref(props)))
        return props.value

    def eval(self, kind, theta, use_graph=False):
        """sums of `kind` for every segment at theta (one row per segment, or one row for all):
        array (S, nout)"""
        if not self._hv:
            raise MorefitError("batch is closed")
        fast = _fast_eval()
        if fast is not None:
            r = fast(self._hv, kind, theta, int(use_graph), self.S, self.model.ntheta, self.model.nout[kind])
            if type(r) is int:
                _check(r)
            return r
        th = np.asarray(theta, dtype=np.float64)
        if th.size != self.S * self.model.ntheta or not th.flags.c_contiguous:
            th = np.ascontiguousarray(np.broadcast_to(th, (self.S, self.model.ntheta)))
        out = np.empty((self.S, self.model.nout[kind]))
        rc = self._eval(self._hv, kind, th.__array_interface__["data"][0], int(use_graph), out.__array_interface__["data"][0])
        if rc:
            _check(rc)
        return out

    def time_kernel(self, kind, iters=50, main_only=False):
        """ms per launch of the evaluation kernels of `kind` (HIP events, back-to-back launches);
        main_only: the per-chunk kernel without its final pass"""
        ms = ctypes.c_double(0)
        _check(libs()[1].mfh_batch_time_kernel(self.h, kind, iters, int(main_only), ctypes.byref(ms)))
        return ms.value

    def time_graph(self, kind, iters=50):
        """us of device time per call: the call graph replayed back to back, no host sync between"""
        us = ctypes.c_double(0)
        _check(libs()[1].mfh_batch_time_graph(self.h, kind, iters, ctypes.byref(us)))
        return us.value

    def time_eval(self, kind, iters=50, use_graph=False):
        us = ctypes.c_double(0)
        _check(libs()[1].mfh_batch_time_eval(self.h, kind, iters, int(use_graph), ctypes.byref(us)))
        return us.value

    def check(self, theta, segs=None):
        """failure detection at theta (one row per listed segment): per segment the events with
        non-finite log-density terms, the first of them (-1: none), the events out of range"""
        P = self.model.ntheta
        lst = list(range(self.S)) if segs is None else list(segs)
        n = len(lst)
        th = _f64(np.broadcast_to(np.asarray(theta, dtype=np.float64), (n, P)))
        out = np.zeros(3 * n, dtype=np.int64)
        _check(libs()[1].mfh_batch_check(self.h, (ctypes.c_int * n)(*lst), n, _dp(th),
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
        return out.reshape(n, 3)

    def fit(self, start, segs=None, use_graph=False):
        P = self.model.ntheta
        lst = list(range(self.S)) if segs is None else list(segs)
        n = len(lst)
        st = _f64(np.broadcast_to(np.asarray(start, dtype=np.float64), (n, P)))
        vals, errs = np.zeros((n, P)), np.zeros((n, P))
        outs = (FitOut * n)()
        stats = BatchStats()
        sp = None if segs is None else (ctypes.c_int * n)(*lst)
        _check(libs()[1].mfh_batch_fit(self.h, sp, n, _dp(st), int(use_graph), _dp(vals), _dp(errs), outs, ctypes.byref(stats)))
        return vals, errs, list(outs), stats


def serve(models, theta_gen, start, cls, n, seeds, arrival_s, budget_bytes, slots=256, request_start=None):
    """Run the continuous-batching fit service over `models` (request class = index).
    Returns dict with per-request arrival/admit/done times, values, errors, status, stats."""
    nm = len(models)
    maxp = max(m.ntheta for m in models)
    tg = np.full((nm, maxp), np.nan)
    st = np.full((nm, maxp), np.nan)
    for i, m in enumerate(models):
        tg[i, :m.ntheta] = theta_gen[i]
        st[i, :m.ntheta] = start[i]
    nreq = len(cls)
    mh = (ctypes.c_void_p * nm)(*[m.h.value for m in models])
    ci = (ctypes.c_int * nreq)(*[int(c) for c in cls])
    ni = (ctypes.c_longlong * nreq)(*[int(x) for x in n])
    si = (ctypes.c_ulonglong * nreq)(*[int(x) for x in seeds])
    arr = _f64(arrival_s)
    timings = np.zeros((nreq, 3))
    vals = np.zeros((nreq, maxp))
    errs = np.zeros((nreq, maxp))
    status = (ctypes.c_int * nreq)()
    stats = np.zeros(8)
    tg, st = _f64(tg), _f64(st)
    rst = None
    if request_start is not None:
        rst = np.full((nreq, maxp), np.nan)
        for i, r in enumerate(request_start):
            rst[i, :len(r)] = r
        rst = _f64(rst)
    _check(libs()[1].mfh_serve_run(mh, nm, _dp(tg), _dp(st), ci, ni, si, _dp(arr), nreq, float(budget_bytes), int(slots),
                                   _dp(timings), _dp(vals), _dp(errs), status, _dp(stats), None if rst is None else _dp(rst)))
    return {"arrival": timings[:, 0], "admit": timings[:, 1], "done": timings[:, 2], "values": vals, "errors": errs,
            "status": np.array(list(status)), "ticks": int(stats[0]), "launches": int(stats[1]), "peak_live": int(stats[2]),
            "rejected": int(stats[3]), "wall_s": stats[4], "peak_bytes": stats[5], "generate_ms": stats[6],
            "arena_bytes": stats[7]}


def fit_cpu(kind, data, start, nthreads=1, fp32=False, optimize_dimensions=False):
    """One fit through the host-compiled CPU backend (reference LLVM-backend semantics);
    fp32=True runs the FP32 kernels the way the device's FP32 path does; optimize_dimensions
    precomputes event-only terms into extra data columns first (fitter option of that name)."""
    data = _f64(np.atleast_2d(data))
    P = NTHETA[kind]
    vals, errs = np.zeros(P), np.zeros(P)
    out, ms = FitOut(), ctypes.c_double(0)
    _check(libs()[1].mfh_fit_cpu(kind, int(fp32), _dp(data), data.shape[1], _dp(_f64(start)), nthreads, int(optimize_dimensions),
                                 _dp(vals), _dp(errs), ctypes.byref(out), ctypes.byref(ms)))
    return vals, errs, out, ms.value


def fit_weighted(kind, data, weights, start, ctx=None, fp32=False, nthreads=1):
    """Weighted fit with the covariance correction -2 H^-1 F H^-1: through the device's drop-in
    path when ctx is given, else the CPU backend."""
    data = _f64(np.atleast_2d(data))
    w = _f64(weights)
    P = NTHETA[kind]
    vals, errs = np.zeros(P), np.zeros(P)
    out, ms = FitOut(), ctypes.c_double(0)
    h = ctx.h if ctx is not None else None
    _check(libs()[1].mfh_fit_weighted(h, kind, int(fp32), _dp(data), _dp(w), data.shape[1], _dp(_f64(start)), nthreads,
                                      _dp(vals), _dp(errs), ctypes.byref(out), ctypes.byref(ms)))
    return vals, errs, out, ms.value


FIT_ANALYTIC_GRADIENT, FIT_ANALYTIC_HESSIAN, FIT_KAHAN_ON_ACCELERATOR, FIT_POSTRUN_HESSE = 1, 2, 4, 8
FIT_DEFAULT = 15


def fit_options(kind, data, start, flags=FIT_DEFAULT, ctx=None, fp32=False, nthreads=1):
    """fitter::fit with the fitter_options switches of the FIT_* bits (analytic gradient,
    analytic Hessian, sums on the accelerator, HESSE after MIGRAD): through the device's
    drop-in path when ctx is given, else the CPU backend."""
    data = _f64(np.atleast_2d(data))
    P = NTHETA[kind]
    vals, errs = np.zeros(P), np.zeros(P)
    out, ms = FitOut(), ctypes.c_double(0)
    h = ctx.h if ctx is not None else None
    _check(libs()[1].mfh_fit_options(h, kind, int(fp32), _dp(data), data.shape[1], _dp(_f64(start)), nthreads, int(flags),
                                     _dp(vals), _dp(errs), ctypes.byref(out), ctypes.byref(ms)))
    return vals, errs, out, ms.value


class PlainFitter:
    """fitter::fit one dataset at a time with the kernels built once (refit): on the device's
    drop-in path when ctx is given, else on the CPU backend (nthreads, cpu_flags)."""

    def __init__(self, kind, fp32=False, ctx=None, nthreads=1, flags=FIT_DEFAULT, cpu_flags=None):
        self.kind = kind
        self.h = ctypes.c_void_p()
        _check(libs()[1].mfh_plain_create(ctx.h if ctx is not None else None, kind, int(fp32), nthreads, int(flags),
                                          cpu_flags.encode() if cpu_flags else None, ctypes.byref(self.h)))

    def fit(self, data, start):
        data = _f64(np.atleast_2d(data))
        P = NTHETA[self.kind]
        vals, errs = np.zeros(P), np.zeros(P)
        out, ms = FitOut(), ctypes.c_double(0)
        _check(libs()[1].mfh_plain_fit(self.h, _dp(data), data.shape[1], _dp(_f64(start)), _dp(vals), _dp(errs),
                                       ctypes.byref(out), ctypes.byref(ms)))
        return vals, errs, out, ms.value

    def time_eval(self, data, theta, iters):
        """-2lnL and gradient at theta, and the seconds `iters` evaluations took"""
        data = _f64(np.atleast_2d(data))
        v, sec = ctypes.c_double(0), ctypes.c_double(0)
        g = np.zeros(NTHETA[self.kind])
        _check(libs()[1].mfh_plain_eval(self.h, _dp(data), data.shape[1], _dp(_f64(theta)), int(iters), ctypes.byref(v), _dp(g),
                                        ctypes.byref(sec)))
        return v.value, g, sec.value

    def close(self):
        if self.h:
            libs()[1].mfh_plain_destroy(self.h)
            self.h = ctypes.c_void_p()


def eval_cpu(kind, data, theta, nthreads=1):
    """-2lnL, its gradient and Hessian through the CPU backend."""
    data = _f64(np.atleast_2d(data))
    P = NTHETA[kind]
    v = ctypes.c_double(0)
    g, h = np.zeros(P), np.zeros(P * P)
    _check(libs()[1].mfh_eval_cpu(kind, _dp(data), data.shape[1], _dp(_f64(theta)), nthreads, ctypes.byref(v), _dp(g), _dp(h)))
    return v.value, g, h.reshape(P, P)


PLOT_ROOT, PLOT_MATPLOTLIB = 0, 1


def plot(kind, theta, data, var=0, nbins=100, multiplier=10, range_min=0.0, range_max=0.0, weights=None, ctx=None, fp32=False,
         nthreads=1, filename=None, plot_type=PLOT_MATPLOTLIB, pulls=False, replots=0):
    """plotter::plot: the density at theta integrated over nbins*multiplier fine bins of
    observable var (integral kernel on the device when ctx is given, else the CPU backend),
    the histogram of data, residuals and pulls; filename writes a matplotlib script (.py) or a
    ROOT macro (.C).  Returns a dict of numpy arrays and the last call's milliseconds."""
    data = _f64(np.atleast_2d(data))
    n = data.shape[1]
    fine = np.zeros(nbins * multiplier)
    counts, errors, pl = np.zeros(nbins), np.zeros(nbins), np.zeros(nbins)
    ms = ctypes.c_double(0)
    w = _f64(weights) if weights is not None else None
    _check(libs()[1].mfh_plot(ctx.h if ctx is not None else None, kind, int(fp32), nthreads, _dp(_f64(theta)), _dp(data),
                              _dp(w) if w is not None else None, n, var, nbins, multiplier, range_min, range_max, plot_type,
                              1 if pulls else 0, filename.encode() if filename else None, replots, _dp(fine), _dp(counts),
                              _dp(errors), _dp(pl), ctypes.byref(ms)))
    return {"pdf_integrals": fine, "counts": counts, "errors": errors, "pulls": pl, "ms": ms.value}

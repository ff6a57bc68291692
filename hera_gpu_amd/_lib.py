"""ctypes binding of ``libheravis.so`` (declarations in include/heravis.h).

The library is built in-tree by ``hera_gpu_amd.build``.  There is no CPU
fallback: if the library or a GPU is missing, every call raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import threading
import weakref
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libheravis.so"

HV_OK = 0
HV_ERR_ARG = -1
HV_ERR_HIP = -2
HV_ERR_NOMEM = -3
HV_ERR_STATE = -4
HV_ERR_BUSY = -5
HV_ERR_CANCELED = -6

HV_F32 = 0
HV_F64 = 1

MODES = {"fp32": 0, "split_bf16": 1, "split_fp16": 2, "fp64": 3}
MODE_NAMES = {v: k for k, v in MODES.items()}

# Every symbol the header declares, with its ctypes signature.
_vp = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_d = C.c_double
_sz = C.c_size_t


class StageTimes(C.Structure):
    _fields_ = [("coords_ms", _d), ("phasor_ms", _d), ("herk_ms", _d), ("reduce_ms", _d),
                ("total_ms", _d), ("coords_n", _i64), ("phasor_n", _i64), ("herk_n", _i64),
                ("reduce_n", _i64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class JobInfo(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int), ("slot", C.c_int), ("graph_hit", C.c_int),
                ("reserved", C.c_int), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("footprint", C.c_int64), ("flops", C.c_double),
                ("error", C.c_char * 160)]


class Job(C.Structure):
    _fields_ = [("mode", C.c_int), ("flags", C.c_int), ("sky", C.c_void_p), ("nsrc", C.c_int64),
                ("beam", C.c_void_p), ("antpos", C.c_void_p), ("nant", C.c_int), ("freqs", C.c_void_p),
                ("nfreq", C.c_int), ("eq2tops", C.c_void_p), ("ntimes", C.c_int), ("vis_out", C.c_void_p)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("graph_captures", C.c_int64), ("graph_launches", C.c_int64), ("direct_launches", C.c_int64),
                ("peak_running", C.c_int64), ("pending", C.c_int64), ("running", C.c_int64),
                ("workspace_bytes", C.c_size_t), ("budget_bytes", C.c_size_t), ("cancelled", C.c_int64),
                ("instances", C.c_int64), ("evictions", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SIGNATURES = {
    "hv_version": (_i, []),
    "hv_last_error": (C.c_char_p, []),
    "hv_device_count": (_i, [C.POINTER(_i)]),
    "hv_ctx_create": (_i, [_i, _sz, C.POINTER(_vp)]),
    "hv_ctx_destroy": (_i, [_vp]),
    "hv_vis": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _vp, _i64, _vp, _i, _vp]),
    "hv_vis_ex": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _vp, _i64, _vp, _i, _i, _i, _vp]),
    "hv_vis_device_noalloc": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _vp, _i64, _vp, _i, _vp, _vp]),
    "hv_vis_pol": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _vp, _i64, _vp, _i, _vp]),
    "hv_omnical": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _vp, _i64, C.c_double, _i, _i, _i, C.c_double,
                        _vp, _vp, _vp]),
    "hv_set_omnical_place": (_i, [_i]),
    "hv_bench_omnical": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _vp, _i64, C.c_double, _i, _i, _i, C.c_double,
                              _vp, _vp, _vp, _i, C.POINTER(_d)]),
    "hv_medfilt2d": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i]),
    "hv_bench_medfilt2d": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, C.POINTER(_d)]),
    "hv_k_rotate": (_i, [_vp, _i, _vp, _vp, _i64, _vp]),
    "hv_k_beam": (_i, [_vp, _i, _vp, _i, _i, _vp, _i64, _vp]),
    "hv_k_phasor": (_i, [_vp, _i, _i, _vp, _i, _i, _vp, _vp, _i64, _vp, _d, _vp]),
    "hv_k_herk": (_i, [_vp, _i, _vp, _i, _i64, _vp]),
    "hv_sky_upload": (_i, [_vp, _i, _vp, _vp, _i64, C.POINTER(_vp)]),
    "hv_sky_free": (_i, [_vp]),
    "hv_beam_upload": (_i, [_vp, _i, _vp, _i, _i, C.POINTER(_vp)]),
    "hv_beam_upload_freq": (_i, [_vp, _i, _vp, _i, _i, _i, C.POINTER(_vp)]),
    "hv_beam_free": (_i, [_vp]),
    "hv_simulate": (_i, [_vp, _i, _vp, _i64, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i,
                         C.POINTER(StageTimes)]),
    "hv_bench_copy": (_i, [_vp, _sz, _i, C.POINTER(_d)]),
    "hv_bench_mfma": (_i, [_vp, _i, _i, C.POINTER(_d)]),
    "hv_bench_herk": (_i, [_vp, _i, _i, _i64, _i, _i, _i, _i, C.POINTER(_d)]),
    "hv_set_herk_variant": (_i, [_i, _i]),
    "hv_herk_choice": (_i, [_i, _i, C.POINTER(_i), C.POINTER(_i)]),
    "hv_job_ksplit": (_i, [_i, _i, C.c_int64, C.c_int64, C.POINTER(_i)]),
    "hv_set_ksplit_items": (_i, [C.c_int64]),
    "hv_server_create": (_i, [_vp, _i, _sz, _i, C.POINTER(_vp)]),
    "hv_server_submit": (_i, [_vp, C.POINTER(Job), C.POINTER(_i64)]),
    "hv_server_poll": (_i, [_vp, C.POINTER(JobInfo), _i, C.POINTER(_i)]),
    "hv_server_wait": (_i, [_vp, _i64, _d, C.POINTER(JobInfo)]),
    "hv_server_stats": (_i, [_vp, C.POINTER(ServerStats)]),
    "hv_server_cancel": (_i, [_vp, _i64]),
    "hv_server_destroy": (_i, [_vp]),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(f"{LIB_PATH} is not built; run `python -m hera_gpu_amd.build`")
            lib = C.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                f = getattr(lib, name)
                f.restype = res
                f.argtypes = args
            _lib = lib
    return _lib


class HeravisError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    if rc == HV_OK:
        return
    msg = load().hv_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == HV_ERR_ARG:
        raise AssertionError(text)          # the reference raises AssertionError (vis.py:147,162-169)
    if rc == HV_ERR_NOMEM:
        raise MemoryError(text)
    raise HeravisError(f"[{rc}] {text}")


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


_PAR_BYTES = 1 << 24        # arrays below 16 MB convert in the calling thread
_NPAR = min(8, os.cpu_count() or 1)
_pool = None


def contiguous(a, dtype) -> np.ndarray:
    """np.ascontiguousarray(a, dtype), with a large conversion (complex128 ->
    complex64 data of a calibration call, say) split by rows over up to 8
    threads: numpy's cast loops release the GIL, and one thread converts at
    a fraction of the bandwidth that the later upload runs at."""
    a = np.asarray(a)
    dtype = np.dtype(dtype)
    if (a.dtype == dtype and a.flags.c_contiguous) or a.ndim == 0 or a.shape[0] < 16 \
            or a.size * max(a.itemsize, dtype.itemsize) < _PAR_BYTES:
        return np.ascontiguousarray(a, dtype=dtype)
    global _pool
    with _lock:
        if _pool is None:
            from concurrent.futures import ThreadPoolExecutor
            _pool = ThreadPoolExecutor(max_workers=_NPAR, thread_name_prefix="hv-cast")
        pool = _pool
    out = np.empty(a.shape, dtype=dtype)
    n = _NPAR
    edges = [a.shape[0] * i // n for i in range(n + 1)]

    def cast(i):
        out[edges[i]:edges[i + 1]] = a[edges[i]:edges[i + 1]]

    list(pool.map(cast, range(n)))
    return out


class Context:
    """One device, stream and workspace (wraps hv_ctx).

    Objects that hold handles into the context (resident catalogs and beams,
    servers) register with `adopt`; `close` closes the live ones, newest
    first, before destroying the context, so no handle outlives the hv_ctx
    its C object points to (at interpreter exit the atexit hook closes the
    contexts before module globals are collected)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        lib = load()
        h = C.c_void_p()
        check(lib.hv_ctx_create(device, hbm_budget, C.byref(h)), "hv_ctx_create")
        self.handle = h
        self.device = device
        self._deps: list[weakref.ref] = []
        self._deps_lock = threading.Lock()

    def adopt(self, obj) -> None:
        """Close `obj` (anything with close()) before this context is destroyed."""
        with self._deps_lock:
            if len(self._deps) >= 64:
                self._deps = [r for r in self._deps if r() is not None]
            self._deps.append(weakref.ref(obj))

    def close(self):
        if self.handle:
            with self._deps_lock:
                deps, self._deps = self._deps, []
            # newest first: a server closes before the catalogs and beams its jobs read
            for r in reversed(deps):
                o = r()
                if o is not None:
                    try:
                        o.close()
                    except Exception:
                        pass
            load().hv_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_contexts: dict[int, Context] = {}


def context(device: int = 0) -> Context:
    """Process-wide context of a device, created on first use."""
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        budget = int(os.environ.get("HERAVIS_HBM_BUDGET", "0"))
        ctx = Context(device, budget)
        with _lock:
            _contexts.setdefault(device, ctx)
            ctx = _contexts[device]
    return ctx


@atexit.register
def _close_all():
    for c in list(_contexts.values()):
        try:
            c.close()
        except Exception:
            pass
    _contexts.clear()


def device_count() -> int:
    n = C.c_int(0)
    check(load().hv_device_count(C.byref(n)))
    return n.value

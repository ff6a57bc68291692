"""Kernel-level entry points and resident inputs.

Each stage function runs one device stage on host arrays, the way the
reference's kernel tests launch one kernel each
(hera_gpu/tests/test_vis_gpu.py:17-169):

* ``rotate``      -- rotation kernel (replaces the Sgemm of vis.py:228-232)
* ``beam_interp`` -- LDS bilinear beam (replaces InterpolateBeam, vis.py:46-89)
* ``phasor``      -- fused phasor build (replaces the delay Sgemm and MeasEq,
                     vis.py:233-245)
* ``herk``        -- MFMA Hermitian product + K-split reduction (replaces the
                     Cgemm and host sum, vis.py:246-258)

``Sky`` / ``Beam`` keep inputs resident in HBM and ``simulate`` runs the whole
pipeline on them (optionally timed with per-stage HIP events).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _dtype_code(dtype):
    dt = np.dtype(dtype)
    if dt == np.float32:
        return _lib.HV_F32
    if dt == np.float64:
        return _lib.HV_F64
    raise AssertionError("dtype must be float32 or float64")


def _mode(gemm):
    if isinstance(gemm, int):
        return gemm
    assert gemm in _lib.MODES, f"unknown gemm {gemm!r}"
    return _lib.MODES[gemm]


def rotate(eq2top, crd_eq, dtype=np.float32, device=0):
    """crd_top (3, K) = eq2top (3, 3) @ crd_eq (3, K), FP64 inside."""
    eq = np.ascontiguousarray(eq2top, dtype=dtype)
    crd = np.ascontiguousarray(crd_eq, dtype=dtype)
    assert eq.shape == (3, 3) and crd.ndim == 2 and crd.shape[0] == 3
    out = np.empty_like(crd)
    rc = _lib.load().hv_k_rotate(_lib.context(device).handle, _dtype_code(dtype), _lib.ptr(eq), _lib.ptr(crd),
                                 crd.shape[1], _lib.ptr(out))
    _lib.check(rc, "hv_k_rotate")
    return out


def beam_interp(bm_cube, crd_top, dtype=np.float32, device=0):
    """A (N, K): clamp-to-edge bilinear beam at crd_top (x, y), 0 unless z > 0."""
    bm = np.ascontiguousarray(bm_cube, dtype=dtype)
    crd = np.ascontiguousarray(crd_top, dtype=dtype)
    n, p, p2 = bm.shape
    assert p == p2 and crd.ndim == 2 and crd.shape[0] == 3
    out = np.empty((n, crd.shape[1]), dtype=dtype)
    rc = _lib.load().hv_k_beam(_lib.context(device).handle, _dtype_code(dtype), _lib.ptr(bm), n, p,
                               _lib.ptr(crd), crd.shape[1], _lib.ptr(out))
    _lib.check(rc, "hv_k_beam")
    return out


def phasor(bm_cube, crd_top, isqrt, antpos, freq, gemm="fp32", dtype=np.float32, device=0):
    """v (N, K) = A sqrt(I) exp(+2 pi i freq antpos . crd_top / c) as stored for `gemm`.

    Split modes return hi + lo; complex128 for ``gemm="fp64"``.
    """
    mode = _mode(gemm)
    bm = np.ascontiguousarray(bm_cube, dtype=dtype)
    crd = np.ascontiguousarray(crd_top, dtype=dtype)
    isq = np.ascontiguousarray(isqrt, dtype=dtype)
    ap = np.ascontiguousarray(antpos, dtype=dtype)
    n, p, _ = bm.shape
    k = crd.shape[1]
    assert crd.shape == (3, k) and isq.shape == (k,) and ap.shape == (n, 3)
    out = np.empty((n, k), dtype=np.complex128 if mode == _lib.MODES["fp64"] else np.complex64)
    rc = _lib.load().hv_k_phasor(_lib.context(device).handle, mode, _dtype_code(dtype), _lib.ptr(bm), n, p,
                                 _lib.ptr(crd), _lib.ptr(isq), k, _lib.ptr(ap), float(freq), _lib.ptr(out))
    _lib.check(rc, "hv_k_phasor")
    return out


def herk(v, gemm="fp32", device=0):
    """V (N, N) = v @ v^H for caller phasors v (N, K)."""
    mode = _mode(gemm)
    cd = np.complex128 if mode == _lib.MODES["fp64"] else np.complex64
    v = np.ascontiguousarray(v, dtype=cd)
    n, k = v.shape
    out = np.empty((n, n), dtype=cd)
    rc = _lib.load().hv_k_herk(_lib.context(device).handle, mode, _lib.ptr(v), n, k, _lib.ptr(out))
    _lib.check(rc, "hv_k_herk")
    return out


class Sky:
    """Source catalog resident in HBM: crd_eq (3, K) and sqrt(I) (K,)."""

    def __init__(self, crd_eq, I_sky, device=0, dtype=np.float32):
        self.ctx = _lib.context(device)
        crd = np.ascontiguousarray(crd_eq, dtype=dtype)
        isq = np.ascontiguousarray(np.sqrt(np.asarray(I_sky)).astype(dtype))
        assert crd.ndim == 2 and crd.shape[0] == 3 and isq.shape == (crd.shape[1],)
        self.nsrc = crd.shape[1]
        h = C.c_void_p()
        _lib.check(_lib.load().hv_sky_upload(self.ctx.handle, _dtype_code(dtype), _lib.ptr(crd), _lib.ptr(isq),
                                             self.nsrc, C.byref(h)), "hv_sky_upload")
        self.handle = h
        self.ctx.adopt(self)

    def close(self):
        if self.handle:
            # a closed context has closed this catalog first (Context.adopt)
            if self.ctx.handle:
                _lib.load().hv_sky_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Beam:
    """Beam cube (N, P, P) resident in HBM, or (F, N, P, P): a beam per
    frequency, for calls and jobs with exactly F frequencies."""

    def __init__(self, bm_cube, device=0, dtype=np.float32):
        self.ctx = _lib.context(device)
        bm = np.ascontiguousarray(bm_cube, dtype=dtype)
        assert bm.ndim in (3, 4)
        self.nfreq = bm.shape[0] if bm.ndim == 4 else 1
        self.nant, self.beam_px, p2 = bm.shape[-3:]
        assert self.beam_px == p2
        self.dtype = np.dtype(dtype)
        h = C.c_void_p()
        _lib.check(_lib.load().hv_beam_upload_freq(self.ctx.handle, _dtype_code(dtype), _lib.ptr(bm), self.nfreq,
                                                   self.nant, self.beam_px, C.byref(h)), "hv_beam_upload")
        self.handle = h
        self.ctx.adopt(self)

    def close(self):
        if self.handle:
            if self.ctx.handle:
                _lib.load().hv_beam_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulate(sky: Sky, beam: Beam, antpos, freqs, eq2tops, gemm="fp32", nsrc=None, reps=1,
             timing=False, want_output=True, dtype=np.float64):
    """Run the pipeline on resident inputs; returns (vis or None, times dict or None)."""
    mode = _mode(gemm)
    ap = np.ascontiguousarray(antpos, dtype=dtype)
    eq = np.ascontiguousarray(eq2tops, dtype=dtype)
    fr = np.ascontiguousarray(np.atleast_1d(freqs), dtype=np.float64)
    n = ap.shape[0]
    t = eq.shape[0]
    nsrc = sky.nsrc if nsrc is None else int(nsrc)
    cd = np.complex128 if mode == _lib.MODES["fp64"] else np.complex64
    out = np.empty((fr.size, t, n, n), dtype=cd) if want_output else None
    st = _lib.StageTimes()
    rc = _lib.load().hv_simulate(sky.ctx.handle, mode, sky.handle, nsrc, beam.handle, _dtype_code(dtype),
                                 _lib.ptr(ap), n, _lib.ptr(fr), fr.size, _lib.ptr(eq), t,
                                 _lib.ptr(out) if out is not None else None, reps,
                                 C.byref(st) if timing else None)
    _lib.check(rc, "hv_simulate")
    return out, (st.as_dict() if timing else None)


def herk_choice(gemm: str, nant: int) -> tuple:
    """(tile edge, pf) of the contraction variant calls of this arithmetic and size use."""
    t, p = C.c_int(), C.c_int()
    _lib.check(_lib.load().hv_herk_choice(_mode(gemm), nant, C.byref(t), C.byref(p)), "hv_herk_choice")
    return t.value, p.value


def ksplit_choice(gemm: str, nant: int, units: int, nsrc: int) -> int:
    """Sources per K-split (8192 or 32768) of a job of `units` (time,
    frequency) units over nsrc sources: fixed by the job alone."""
    k = C.c_int()
    _lib.check(_lib.load().hv_job_ksplit(_mode(gemm), nant, units, nsrc, C.byref(k)), "hv_job_ksplit")
    return k.value


def bench_copy(nbytes: int = 1 << 30, reps: int = 20, device: int = 0) -> dict:
    """Plain device copy of nbytes: ms per copy and GB/s (read + write)."""
    ms = C.c_double()
    _lib.check(_lib.load().hv_bench_copy(_lib.context(device).handle, nbytes, reps, C.byref(ms)), "hv_bench_copy")
    return {"bytes": nbytes, "ms": ms.value, "GBps": 2 * nbytes / (ms.value * 1e-3) / 1e9}


def bench_mfma(gemm: str, iters: int = 20000, device: int = 0) -> float:
    """Measured matrix-core TFLOP/s of the MFMA instruction used by `gemm`."""
    t = C.c_double()
    _lib.check(_lib.load().hv_bench_mfma(_lib.context(device).handle, _mode(gemm), iters, C.byref(t)),
               "hv_bench_mfma")
    return t.value


def bench_herk(gemm: str, nant: int, nsrc: int, units: int = 1, tile: int = 0, pf: int = 0, reps: int = 5,
               device: int = 0) -> dict:
    """Contraction alone on synthetic operands: ms per launch and useful TFLOP/s."""
    ms = C.c_double()
    _lib.check(_lib.load().hv_bench_herk(_lib.context(device).handle, _mode(gemm), nant, nsrc, units, tile, pf,
                                         reps, C.byref(ms)), "hv_bench_herk")
    flops = 4.0 * nant * (nant + 1) * nsrc * units
    return {"ms": ms.value, "tflops": flops / (ms.value * 1e-3) / 1e12}

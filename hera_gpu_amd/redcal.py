"""Omnical redundant calibration on the GPU: ``omnical`` with the arguments and
returns of ``hera_gpu.redcal.omnical`` (hera_gpu/redcal.py:358-406, returns
:740-743).

Each pixel (a row of ``data``) is an independent system
data[b] ~ gains[i_b] conj(gains[j_b]) ubls[k_b]; one GPU workgroup solves one
pixel through all its iterations (see hv_omnical in include/heravis.h).  The
reference's hera_cal integration (``OmnicalSolverGPU``,
``RedundantCalibratorGPU``, redcal.py:746-883) subclasses hera_cal, which is
not part of this package.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

NTHREADS = 1024   # accepted for the reference's signature; one 256-thread workgroup per pixel is used


def omnical(ggu_indices, gains, ubls, data, wgts, conv_crit=1e-10, maxiter=50, check_every=4, check_after=1,
            gain=0.3, nthreads=NTHREADS, precision=1, verbose=False, *, device=0, _reps=0):
    """Solve every pixel; returns {'gains', 'ubls', 'chisq', 'iters', 'conv'}.

    Arguments as redcal.py:358-390: ggu_indices (nbls, 3) (i, j, k) per
    baseline; gains (ndata, nants), ubls (ndata, nubls) starting points;
    data, wgts (ndata, nbls); precision 1 (float32) or 2 (float64).  The
    input gains and ubls arrays are not modified.
    """
    ggu = np.asarray(ggu_indices)
    nbls = ggu.shape[0]
    assert ggu.shape == (nbls, 3)
    data = np.asarray(data)
    ndata = data.shape[0]
    assert data.shape == (ndata, nbls)
    wgts = np.asarray(wgts)
    assert wgts.shape == (ndata, nbls)
    gains = np.asarray(gains)
    nants = gains.shape[1]
    assert gains.shape == (ndata, nants)
    ubls = np.asarray(ubls)
    nubls = ubls.shape[1]
    assert ubls.shape == (ndata, nubls)
    assert precision in (1, 2)
    assert check_every > 1
    rd, cd = (np.float32, np.complex64) if precision == 1 else (np.float64, np.complex128)
    ggu = np.ascontiguousarray(ggu, dtype=np.uint32)
    data = _lib.contiguous(data, cd)
    wgts = _lib.contiguous(wgts, rd)
    g = np.array(gains, dtype=cd, order="C", copy=True)
    u = np.array(ubls, dtype=cd, order="C", copy=True)
    chisq = np.empty(ndata, dtype=rd)
    conv = np.empty(ndata, dtype=rd)
    iters = np.empty(ndata, dtype=np.uint32)
    if verbose:
        print("PRECISION:", precision, "NDATA:", ndata, "NANTS:", nants, "NUBLS:", nubls, "NBLS:", nbls)
    ctx = _lib.context(device)
    args = (ctx.handle, precision, _lib.ptr(ggu), nbls, _lib.ptr(g), nants, _lib.ptr(u), nubls, _lib.ptr(data),
            _lib.ptr(wgts), ndata, float(conv_crit), int(maxiter), int(check_every), int(check_after), float(gain),
            _lib.ptr(chisq), _lib.ptr(iters), _lib.ptr(conv))
    out = {"gains": g, "ubls": u, "chisq": chisq, "iters": iters, "conv": conv}
    if _reps:
        ms = C.c_double()
        _lib.check(_lib.load().hv_bench_omnical(*args, int(_reps), C.byref(ms)), "hv_bench_omnical")
        out["kernel_ms"] = ms.value
    else:
        _lib.check(_lib.load().hv_omnical(*args), "hv_omnical")
    return out


def bench_omnical(ggu_indices, gains, ubls, data, wgts, reps=5, **kw):
    """omnical with the inputs uploaded once and the solve repeated from the
    same start: the returned dict also holds 'kernel_ms', the mean kernel
    time of one solve of every pixel."""
    return omnical(ggu_indices, gains, ubls, data, wgts, _reps=reps, **kw)


def set_placement(place: int = -1) -> None:
    """Cap where the solver keeps its per-baseline arrays (tests and tuning):
    -1 automatic, 0 device memory, 1 weighted products in LDS, 2 products and
    data weights in LDS (lowered to what fits)."""
    _lib.check(_lib.load().hv_set_omnical_place(int(place)), "hv_set_omnical_place")

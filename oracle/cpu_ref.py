"""Restated reference CPU path ``vis_cpu`` -- the timed CPU baseline.

TEST INFRASTRUCTURE ONLY (and the ``cpu_baseline`` leg of bench.py).

Follows hera_gpu/tests/test_vis_gpu.py:278-357 step by step, written here
from that description (the reference module cannot be imported: it needs
pycuda, scikit-cuda and astropy):

* ``Isqrt = sqrt(I_sky)`` cast to the real dtype (:315)
* per time, ``crd_top = eq2top @ crd_eq`` (:331)
* per antenna, a degree-1 ``RectBivariateSpline`` over ``linspace(-1, 1, P)``
  in (y, x) evaluated at the source (y, x) (:326-335)
* horizon ``where(z > 0, A, 0)`` (:336)
* ``tau = antpos @ crd_top / c`` and ``v = exp(1j * 2 pi nu * tau) * A * Isqrt``
  (:339-344)
* upper triangle by per-row dot products of ``conj(v_i)`` with ``v[i:]``, then
  conjugate everything and fill the lower triangle (:347-355).

A caller with several frequencies calls it once per frequency, as a caller
of the reference does (``freq`` is a scalar there, vis.py:159).
"""
from __future__ import annotations

import numpy as np
from scipy.interpolate import RectBivariateSpline

C_LIGHT = 299792458.0


def vis_cpu(antpos, freq, eq2tops, crd_eq, I_sky, bm_cube, precision=1):
    """One frequency: complex (T, N, N) in complex64 (precision 1) or complex128."""
    assert precision in (1, 2)
    if precision == 1:
        real_dtype, complex_dtype = np.float32, np.complex64
    else:
        real_dtype, complex_dtype = np.float64, np.complex128
    nant, ncrd = antpos.shape
    assert ncrd == 3
    ntimes, n1, n2 = eq2tops.shape
    assert n1 == 3 and n2 == 3
    ncrd, npix = crd_eq.shape
    assert ncrd == 3
    assert I_sky.ndim == 1 and I_sky.shape[0] == npix
    bm_pix = bm_cube.shape[-1]
    assert bm_cube.shape == (nant, bm_pix, bm_pix)

    isqrt = np.sqrt(I_sky).astype(real_dtype)
    antpos = antpos.astype(real_dtype)
    ang_freq = real_dtype(2.0 * np.pi * freq)
    crd_eq = crd_eq.astype(real_dtype)
    grid = np.linspace(-1, 1, bm_pix)

    vis = np.empty((ntimes, nant, nant), dtype=complex_dtype)
    amp = np.empty((nant, npix), dtype=real_dtype)
    tau = np.empty((nant, npix), dtype=real_dtype)
    v = np.empty((nant, npix), dtype=complex_dtype)
    for t, eq2top in enumerate(eq2tops.astype(real_dtype)):
        tx, ty, tz = crd_top = np.dot(eq2top, crd_eq)
        for i in range(nant):
            spline = RectBivariateSpline(grid, grid, bm_cube[i], kx=1, ky=1)
            amp[i] = spline(ty, tx, grid=False)
        amp = np.where(tz > 0, amp, 0).astype(real_dtype)
        np.dot(antpos, crd_top, out=tau)
        tau *= real_dtype(1.0 / C_LIGHT)
        np.exp((1.0j * ang_freq) * tau, out=v)
        v *= amp * isqrt
        for i in range(nant):
            np.dot(v[i:i + 1].conj(), v[i:].T, out=vis[t, i:i + 1, i:])
    np.conj(vis, out=vis)
    for i in range(nant):
        vis[:, i + 1:, i] = vis[:, i, i + 1:].conj()
    return vis


def vis_cpu_multifreq(antpos, freqs, eq2tops, crd_eq, I_sky, bm_cube, precision=1):
    """One ``vis_cpu`` call per frequency; returns (F, T, N, N)."""
    freqs = np.atleast_1d(freqs)
    return np.stack([vis_cpu(antpos, float(nu), eq2tops, crd_eq, I_sky, bm_cube,
                             precision=precision) for nu in freqs])


# ---------------------------------------------------------------------------
# The same computation with the per-antenna rows spread over processes.
#
# scipy's spline evaluation holds the GIL, so threads do not help; the rows
# of amp and v are computed by `workers` spawned processes (fresh
# interpreters: nothing of the parent's state, GPU included, is inherited)
# into shared memory.  The parent forms crd_top and tau with the same numpy
# calls as vis_cpu and runs the same dot-product loop, so the result equals
# vis_cpu bit for bit (every step is elementwise per row, or the same BLAS
# call on the same arrays).

def _rows_worker(args):
    (shm_names, shapes, lo, hi, freq, precision) = args
    from multiprocessing import shared_memory
    real_dtype = np.float32 if precision == 1 else np.float64
    complex_dtype = np.complex64 if precision == 1 else np.complex128
    shms = [shared_memory.SharedMemory(name=n) for n in shm_names]
    try:
        crd = np.ndarray(shapes[0], dtype=real_dtype, buffer=shms[0].buf)
        tau = np.ndarray(shapes[1], dtype=real_dtype, buffer=shms[1].buf)
        bm = np.ndarray(shapes[2], dtype=np.float64, buffer=shms[2].buf)
        isqrt = np.ndarray(shapes[3], dtype=real_dtype, buffer=shms[3].buf)
        v = np.ndarray(shapes[4], dtype=complex_dtype, buffer=shms[4].buf)
        tx, ty, tz = crd
        grid = np.linspace(-1, 1, bm.shape[-1])
        ang_freq = real_dtype(2.0 * np.pi * freq)
        for i in range(lo, hi):
            spline = RectBivariateSpline(grid, grid, bm[i], kx=1, ky=1)
            a = spline(ty, tx, grid=False).astype(real_dtype)
            a = np.where(tz > 0, a, 0).astype(real_dtype)
            row = np.exp((1.0j * ang_freq) * tau[i])
            row *= a * isqrt
            v[i] = row
    finally:
        for s in shms:
            s.close()
    return hi - lo


def vis_cpu_parallel(antpos, freq, eq2tops, crd_eq, I_sky, bm_cube, precision=1, workers=8, pool=None):
    """vis_cpu with the per-antenna beam and phasor rows on `workers` processes."""
    from multiprocessing import get_context, shared_memory
    assert precision in (1, 2)
    real_dtype = np.float32 if precision == 1 else np.float64
    complex_dtype = np.complex64 if precision == 1 else np.complex128
    nant = antpos.shape[0]
    npix = crd_eq.shape[1]
    isqrt = np.sqrt(I_sky).astype(real_dtype)
    antpos = antpos.astype(real_dtype)
    crd_eq = crd_eq.astype(real_dtype)
    vis = np.empty((eq2tops.shape[0], nant, nant), dtype=complex_dtype)
    shapes = [(3, npix), (nant, npix), tuple(bm_cube.shape), (npix,), (nant, npix)]
    sizes = [3 * npix * isqrt.itemsize, nant * npix * isqrt.itemsize, bm_cube.size * 8, isqrt.nbytes,
             nant * npix * np.dtype(complex_dtype).itemsize]
    shms = [shared_memory.SharedMemory(create=True, size=max(1, s)) for s in sizes]
    own_pool = pool is None
    if own_pool:
        pool = get_context("spawn").Pool(workers)
    try:
        crd = np.ndarray(shapes[0], dtype=real_dtype, buffer=shms[0].buf)
        tau = np.ndarray(shapes[1], dtype=real_dtype, buffer=shms[1].buf)
        np.ndarray(shapes[2], dtype=np.float64, buffer=shms[2].buf)[...] = bm_cube
        np.ndarray(shapes[3], dtype=real_dtype, buffer=shms[3].buf)[...] = isqrt
        v = np.ndarray(shapes[4], dtype=complex_dtype, buffer=shms[4].buf)
        bounds = np.linspace(0, nant, workers + 1).astype(int)
        for t, eq2top in enumerate(eq2tops.astype(real_dtype)):
            crd[...] = np.dot(eq2top, crd_eq)
            np.dot(antpos, crd, out=tau)
            tau *= real_dtype(1.0 / C_LIGHT)
            jobs = [([s.name for s in shms], shapes, int(bounds[w]), int(bounds[w + 1]), freq, precision)
                    for w in range(workers) if bounds[w + 1] > bounds[w]]
            pool.map(_rows_worker, jobs)
            for i in range(nant):
                np.dot(v[i:i + 1].conj(), v[i:].T, out=vis[t, i:i + 1, i:])
        np.conj(vis, out=vis)
        for i in range(nant):
            vis[:, i + 1:, i] = vis[:, i, i + 1:].conj()
    finally:
        if own_pool:
            pool.close()
            pool.join()
        del crd, tau, v
        for s in shms:
            s.close()
            s.unlink()
    return vis

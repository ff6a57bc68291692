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

"""Drop-in ``vis_gpu`` on MI355X (mirrors hera_gpu/vis.py:144-262).

    from hera_gpu_amd import vis_gpu
    vis = vis_gpu(antpos, freq, eq2tops, crd_eq, I_sky, bm_cube)   # (ntimes, nant, nant)

Same positional arguments, same shape checks (``AssertionError`` on a bad
shape or precision, vis.py:147,161-169), same casts (vis.py:171-175, with
``sqrt(I_sky)`` taken on the host), same output shape and dtype
(complex64 for precision=1, complex128 for precision=2, vis.py:212,262).

Differences a caller may notice:

* every source is summed (the reference drops ``npix mod chunk`` tail
  sources, vis.py:176-177);
* ``freq`` may also be a 1-D array of frequencies; the result is then
  (nfreq, ntimes, nant, nant), and ``bm_cube`` may then be (nfreq, nant, P, P),
  a beam per frequency;
* the delay phase is reduced in FP64 (the reference evaluates it in FP32 for
  precision=1, vis.py:106-108), so realistic phases of ~10^3 rad keep their
  accuracy;
* ``real_dtype`` / ``complex_dtype`` (the spelling of the reference's
  ``speed_test.vis_cpu``, speed_test.py:132) select the precision too, and
  ``gemm`` selects the contraction arithmetic for precision=1:
  ``"fp32"`` (default), ``"split_bf16"`` or ``"split_fp16"``;
* ``nthreads``, ``max_memory`` and ``verbose`` are accepted and ignored (no
  per-call JIT or chunk plan exists to tune; the device workspace is bounded by
  the ``HERAVIS_HBM_BUDGET`` environment variable, default 90% of free HBM).
"""
from __future__ import annotations

import numpy as np

from . import _lib

NTHREADS = 1024
MAX_MEMORY = 2 ** 29
MIN_CHUNK = 8


def _resolve(precision, real_dtype, complex_dtype, gemm):
    if real_dtype is not None or complex_dtype is not None:
        rd = np.dtype(real_dtype) if real_dtype is not None else None
        cd = np.dtype(complex_dtype) if complex_dtype is not None else None
        if rd == np.float64 or cd == np.complex128:
            precision = 2
        elif rd in (None, np.float32) and cd in (None, np.complex64):
            precision = 1
        else:
            raise AssertionError("real_dtype/complex_dtype must be float32/complex64 or float64/complex128")
        if rd is not None and cd is not None and (rd == np.float64) != (cd == np.complex128):
            raise AssertionError("real_dtype and complex_dtype disagree on precision")
    assert precision in (1, 2)
    if precision == 1:
        gemm = "fp32" if gemm is None else gemm
        assert gemm in ("fp32", "split_bf16", "split_fp16"), f"unknown gemm {gemm!r} for precision=1"
        return np.float32, np.complex64, _lib.MODES[gemm], _lib.HV_F32
    gemm = "fp64" if gemm is None else gemm
    assert gemm == "fp64", "precision=2 computes the contraction in FP64"
    return np.float64, np.complex128, _lib.MODES["fp64"], _lib.HV_F64


LAYOUTS = {"full": 0, "ftb": 1, "tbf": 2}


def baseline_pairs(nant):
    """(ant1, ant2) of the baseline layouts: every a <= b, a-major, autos included
    (baseline index a N - a (a-1) / 2 + (b - a))."""
    ant1, ant2 = np.triu_indices(nant)
    return ant1.astype(np.int32), ant2.astype(np.int32)


def vis_gpu(antpos, freq, eq2tops, crd_eq, I_sky, bm_cube,
            nthreads=NTHREADS, max_memory=MAX_MEMORY, precision=1, verbose=False,
            *, real_dtype=None, complex_dtype=None, gemm=None, device=0):
    """Visibilities V[t, a, b] = sum_s v_a(s) conj(v_b(s)) on the GPU.

    v_a(s) = A_a(s) sqrt(I(s)) exp(+2 pi i nu antpos_a . crd_top(s) / c), with
    A the clamp-to-edge bilinear beam, zero unless crd_top z > 0.
    """
    out, scalar_freq = _vis(antpos, freq, eq2tops, crd_eq, I_sky, bm_cube, precision, real_dtype, complex_dtype,
                            gemm, device, "full")
    return out[0] if scalar_freq else out


def vis_gpu_baselines(antpos, freqs, eq2tops, crd_eq, I_sky, bm_cube, precision=1, *, order="tbf",
                      real_dtype=None, complex_dtype=None, gemm=None, device=0):
    """The visibilities of the baselines a <= b only, written on the device in
    baseline order (half the bytes of the (T, N, N) matrices).

    order "tbf": data (ntimes, nbl, nfreq) -- time-major baseline-times
    (Nblts = ntimes * nbl) by frequency, the order of a UVData data array;
    order "ftb": data (nfreq, ntimes, nbl).  Returns (data, ant1, ant2) with
    ant1 <= ant2 the antenna indices of the nbl = N (N+1) / 2 baselines.
    Same arguments and arithmetic as vis_gpu (data[..., k] equals
    vis_gpu(...)[..., ant1[k], ant2[k]] bit for bit).
    """
    assert order in ("tbf", "ftb"), "order must be 'tbf' or 'ftb'"
    out, _ = _vis(antpos, freqs, eq2tops, crd_eq, I_sky, bm_cube, precision, real_dtype, complex_dtype, gemm,
                  device, order)
    ant1, ant2 = baseline_pairs(np.asarray(antpos).shape[0])
    return out, ant1, ant2


def vis_gpu_polarized(antpos, freq, eq2tops, crd_eq, I_sky, jones, precision=1, *, real_dtype=None,
                      complex_dtype=None, gemm=None, device=0):
    """Visibilities of every feed pair for complex (Jones) beams.

    ``jones``: complex (nant, 2, 2, P, P) = [antenna, feed p, sky axis k, y, x]
    (a real array is taken as its complex value).  For an unpolarised sky,
    V[p, q, t, a, b] = sum_s sum_k J_a^{pk}(s) conj(J_b^{qk}(s)) I(s)
    exp(+2 pi i nu (antpos_a - antpos_b) . crd_top(s) / c), J interpolated
    bilinearly per real and imaginary part (zero below the horizon).  Returns
    (2, 2, ntimes, nant, nant), or (nfreq, 2, 2, ntimes, nant, nant) for an
    array of frequencies.  With J_a^{pk} = A_a delta_pk the p = q blocks are
    ``vis_gpu``'s result (to the arithmetic's rounding) and p != q are zero.
    The reference's beam is a real scalar (vis.py:162-175, "XXX complex?" at
    vis.py:175); this extends it.
    """
    real_dtype, complex_dtype, mode, dtype = _resolve(precision, real_dtype, complex_dtype, gemm)
    antpos = np.asarray(antpos)
    eq2tops = np.asarray(eq2tops)
    crd_eq = np.asarray(crd_eq)
    I_sky = np.asarray(I_sky)
    jones = np.asarray(jones)
    nant = antpos.shape[0]
    assert antpos.shape == (nant, 3)
    assert crd_eq.ndim == 2 and crd_eq.shape[0] == 3
    npix = crd_eq.shape[1]
    assert I_sky.shape == (npix,)
    assert eq2tops.ndim == 3 and eq2tops.shape[1:] == (3, 3)
    ntimes = eq2tops.shape[0]
    scalar_freq = np.ndim(freq) == 0
    freqs = np.ascontiguousarray(np.atleast_1d(np.asarray(freq, dtype=np.float64)))
    assert freqs.ndim == 1 and freqs.size >= 1
    assert jones.ndim == 5
    beam_px = jones.shape[-1]
    assert jones.shape == (nant, 2, 2, beam_px, beam_px)
    antpos = np.ascontiguousarray(antpos, dtype=real_dtype)
    eq2tops = np.ascontiguousarray(eq2tops, dtype=real_dtype)
    crd_eq = np.ascontiguousarray(crd_eq, dtype=real_dtype)
    isqrt = np.ascontiguousarray(np.sqrt(I_sky).astype(real_dtype))
    jones = np.ascontiguousarray(jones, dtype=np.complex128 if real_dtype == np.float64 else np.complex64)
    out = np.empty((freqs.size, 2, 2, ntimes, nant, nant), dtype=complex_dtype)
    ctx = _lib.context(device)
    rc = _lib.load().hv_vis_pol(ctx.handle, mode, dtype, _lib.ptr(antpos), nant, _lib.ptr(freqs), freqs.size,
                                _lib.ptr(eq2tops), ntimes, _lib.ptr(crd_eq), _lib.ptr(isqrt), npix,
                                _lib.ptr(jones), beam_px, _lib.ptr(out))
    _lib.check(rc, "hv_vis_pol")
    return out[0] if scalar_freq else out


def _vis(antpos, freq, eq2tops, crd_eq, I_sky, bm_cube, precision, real_dtype, complex_dtype, gemm, device,
         layout):
    real_dtype, complex_dtype, mode, dtype = _resolve(precision, real_dtype, complex_dtype, gemm)
    antpos = np.asarray(antpos)
    eq2tops = np.asarray(eq2tops)
    crd_eq = np.asarray(crd_eq)
    I_sky = np.asarray(I_sky)
    bm_cube = np.asarray(bm_cube)
    # shape checks of vis.py:161-169
    nant = antpos.shape[0]
    assert antpos.shape == (nant, 3)
    assert crd_eq.ndim == 2
    npix = crd_eq.shape[1]
    assert crd_eq.shape == (3, npix)
    assert I_sky.shape == (npix,)
    assert eq2tops.ndim == 3
    ntimes = eq2tops.shape[0]
    assert eq2tops.shape == (ntimes, 3, 3)
    scalar_freq = np.ndim(freq) == 0
    freqs = np.ascontiguousarray(np.atleast_1d(np.asarray(freq, dtype=np.float64)))
    assert freqs.ndim == 1 and freqs.size >= 1
    # (N, P, P) as the reference, or (F, N, P, P): a beam per frequency
    assert bm_cube.ndim in (3, 4)
    beam_px = bm_cube.shape[-1]
    if bm_cube.ndim == 3:
        assert bm_cube.shape == (nant, beam_px, beam_px)
    else:
        assert bm_cube.shape == (freqs.size, nant, beam_px, beam_px)
    beam_nfreq = 1 if bm_cube.ndim == 3 else freqs.size
    # casts of vis.py:171-175
    antpos = np.ascontiguousarray(antpos, dtype=real_dtype)
    eq2tops = np.ascontiguousarray(eq2tops, dtype=real_dtype)
    crd_eq = np.ascontiguousarray(crd_eq, dtype=real_dtype)
    isqrt = np.ascontiguousarray(np.sqrt(I_sky).astype(real_dtype))
    bm_cube = np.ascontiguousarray(bm_cube, dtype=real_dtype)

    nbl = nant * (nant + 1) // 2
    shape = {"full": (freqs.size, ntimes, nant, nant), "ftb": (freqs.size, ntimes, nbl),
             "tbf": (ntimes, nbl, freqs.size)}[layout]
    out = np.empty(shape, dtype=complex_dtype)
    ctx = _lib.context(device)
    rc = _lib.load().hv_vis_ex(ctx.handle, mode, dtype, _lib.ptr(antpos), nant, _lib.ptr(freqs), freqs.size,
                               _lib.ptr(eq2tops), ntimes, _lib.ptr(crd_eq), _lib.ptr(isqrt), npix,
                               _lib.ptr(bm_cube), beam_px, beam_nfreq, LAYOUTS[layout], _lib.ptr(out))
    _lib.check(rc, "hv_vis")
    return out, scalar_freq

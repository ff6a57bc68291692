"""FP64 NumPy oracle for the ``vis_gpu`` visibility equation.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path in
``hera_gpu_amd``; the product never imports it.

It restates, in float64 throughout, what the reference computes:

* rotation   ``crd_top = eq2top @ crd_eq``            (hera_gpu/vis.py:230-231,
                                                       test_vis_gpu.py:331)
* beam       clamp-to-edge bilinear interpolation of ``bm_cube[ant, y, x]``
             at texel position ``(P-1)*(0.5*x+0.5)``   (vis.py:63-72,80-84,
             the ``lerp`` of vis.py:29-32; test_vis_gpu.py:326-335)
* horizon    ``A = 0`` unless ``z > 0`` (strict)       (vis.py:75,85-87;
                                                       test_vis_gpu.py:336)
* delay      ``tau = antpos @ crd_top / c``            (vis.py:234-235;
                                                       test_vis_gpu.py:339-340)
* phasor     ``v = A * sqrt(I) * exp(+i 2 pi nu tau)`` (vis.py:106-108,159;
                                                       test_vis_gpu.py:341-344)
* visibility ``V[a, b] = sum_s v[a, s] * conj(v[b, s])`` (vis.py:248-249 in
             column-major; test_vis_gpu.py:347-355)

Deliberate deviation from the reference: every source is summed.  The
reference drops the ``K mod chunk`` tail sources (vis.py:176-177).

c = 299792458 m/s exactly (the reference takes it from astropy, vis.py:4,8).

Pinning (see tests/test_oracle.py): the bilinear beam is checked against
``scipy.interpolate.RectBivariateSpline(kx=1, ky=1)`` including out-of-range
points; the whole equation against a pure-Python brute-force triple loop on
tiny inputs and against the restated reference ``vis_cpu``
(``oracle.cpu_ref``) in float64; plus the reference's closed-form cases
(test_vis_gpu.py:31-139, 203-244) and invariants (Hermitian, real
non-negative autos equal to sum A^2 I, invariance under a common shift of all
antennas, Cauchy-Schwarz).
"""
from __future__ import annotations

import numpy as np

C_LIGHT = 299792458.0  # m/s, exact SI value


def bilinear_clamp(beam: np.ndarray, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Clamp-to-edge bilinear interpolation of ``beam[y, x]`` on a [-1, 1] grid.

    Follows hera_gpu/vis.py:63-72 (texel coordinate ``(P-1)*(0.5*x+0.5)``,
    integer part and fraction) and vis.py:80-84 (two lerps in x, one in y,
    out-of-range texel indices clamped to the edge, which is CUDA's default
    texture addressing).  ``beam`` is (P, P) or (N, P, P); the result has shape
    ``beam.shape[:-2] + x.shape``.
    """
    beam = np.asarray(beam, dtype=np.float64)
    P = beam.shape[-1]
    assert beam.shape[-2] == P
    bx = (P - 1) * (0.5 * np.asarray(x, dtype=np.float64) + 0.5)
    by = (P - 1) * (0.5 * np.asarray(y, dtype=np.float64) + 0.5)
    px = np.floor(bx)
    py = np.floor(by)
    fx = bx - px
    fy = by - py
    ix0 = np.clip(px, 0, P - 1).astype(np.int64)
    ix1 = np.clip(px + 1, 0, P - 1).astype(np.int64)
    iy0 = np.clip(py, 0, P - 1).astype(np.int64)
    iy1 = np.clip(py + 1, 0, P - 1).astype(np.int64)
    b00 = beam[..., iy0, ix0]
    b01 = beam[..., iy0, ix1]
    b10 = beam[..., iy1, ix0]
    b11 = beam[..., iy1, ix1]
    top = b00 + fx * (b01 - b00)
    bot = b10 + fx * (b11 - b10)
    return top + fy * (bot - top)


def beam_amplitude(bm_cube, crd_top, I_sky):
    """``A * sqrt(I)`` per (antenna, source) with the strict horizon cut.

    vis.py:75,85-87 and vis.py:106 (amp = A * Isqrt); test_vis_gpu.py:332-336.
    Returns (N, K) float64.
    """
    x, y, z = crd_top
    A = bilinear_clamp(bm_cube, x, y)
    A = np.where(z > 0, A, 0.0)
    return A * np.sqrt(np.asarray(I_sky, dtype=np.float64))[None, :]


def vis_oracle(antpos, freqs, eq2tops, crd_eq, I_sky, bm_cube) -> np.ndarray:
    """FP64 visibilities for every (freq, time): complex128 (F, T, N, N).

    ``freqs`` may be a scalar (then F = 1).  Inputs are used at the values
    they hold (the caller casts to float32 first to mirror precision=1).
    ``bm_cube`` is (N, P, P), the reference's one beam for every frequency
    (vis.py:167), or (F, N, P, P), a beam per frequency (the extension the
    reference leaves open at vis.py:175): frequency f then interpolates
    plane f with the same bilinear rule.
    """
    antpos = np.asarray(antpos, dtype=np.float64)
    eq2tops = np.asarray(eq2tops, dtype=np.float64)
    crd_eq = np.asarray(crd_eq, dtype=np.float64)
    I_sky = np.asarray(I_sky, dtype=np.float64)
    bm_cube = np.asarray(bm_cube, dtype=np.float64)
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    N = antpos.shape[0]
    T = eq2tops.shape[0]
    F = freqs.shape[0]
    out = np.empty((F, T, N, N), dtype=np.complex128)
    per_freq = bm_cube.ndim == 4
    if per_freq:
        assert bm_cube.shape[:2] == (F, N), "a per-frequency beam needs one plane per frequency"
    for t in range(T):
        crd_top = eq2tops[t] @ crd_eq
        amp = None if per_freq else beam_amplitude(bm_cube, crd_top, I_sky)
        dist = antpos @ crd_top  # metres of path difference, (N, K)
        for f, nu in enumerate(freqs):
            a_f = beam_amplitude(bm_cube[f], crd_top, I_sky) if per_freq else amp
            phase = (2.0 * np.pi * nu / C_LIGHT) * dist
            v = a_f * np.exp(1j * phase)
            out[f, t] = v @ v.conj().T
    return out


def vis_oracle_units(antpos, freqs, eq2tops, crd_eq, I_sky, bm_cube, units):
    """FP64 visibilities for a sample of (freq_index, time_index) units.

    Returns complex128 (len(units), N, N).  Used to check large
    configurations on a stated sample of units.  ``bm_cube`` is (N, P, P) or
    (F, N, P, P) as in :func:`vis_oracle`.
    """
    antpos = np.asarray(antpos, dtype=np.float64)
    eq2tops = np.asarray(eq2tops, dtype=np.float64)
    crd_eq = np.asarray(crd_eq, dtype=np.float64)
    I_sky = np.asarray(I_sky, dtype=np.float64)
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    N = antpos.shape[0]
    out = np.empty((len(units), N, N), dtype=np.complex128)
    cache = {}
    for i, (f, t) in enumerate(units):
        if t not in cache:
            crd_top = eq2tops[t] @ crd_eq
            cache = {t: (crd_top, None if bm_cube.ndim == 4 else beam_amplitude(bm_cube, crd_top, I_sky),
                         antpos @ crd_top)}
        crd_top, amp, dist = cache[t]
        if amp is None:
            amp = beam_amplitude(bm_cube[f], crd_top, I_sky)
        v = amp * np.exp(1j * (2.0 * np.pi * freqs[f] / C_LIGHT) * dist)
        out[i] = v @ v.conj().T
    return out


def vis_oracle_pol(antpos, freqs, eq2tops, crd_eq, I_sky, jones) -> np.ndarray:
    """FP64 polarised visibilities: complex128 (F, 2, 2, T, N, N).

    ``jones``: complex (N, 2, 2, P, P) = [antenna, feed p, sky axis k, y, x].
    V[f, p, q, t] = sum_k (J^{pk} sqrt(I) e^{i phase}) @ (J^{qk} sqrt(I) e^{i phase})^H
    with J^{pk} at each source from the bilinear rule of ``beam_amplitude``
    applied to the real and imaginary parts (the reference's beam is the
    real, one-component case, vis.py:46-89; "XXX complex?" at vis.py:175).
    """
    antpos = np.asarray(antpos, dtype=np.float64)
    eq2tops = np.asarray(eq2tops, dtype=np.float64)
    crd_eq = np.asarray(crd_eq, dtype=np.float64)
    I_sky = np.asarray(I_sky, dtype=np.float64)
    jones = np.asarray(jones, dtype=np.complex128)
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    N, F, T = antpos.shape[0], freqs.size, eq2tops.shape[0]
    assert jones.shape[:3] == (N, 2, 2)
    out = np.empty((F, 2, 2, T, N, N), dtype=np.complex128)
    for t in range(T):
        crd_top = eq2tops[t] @ crd_eq
        # J[p][k]: (N, K) complex at every source, times sqrt(I)
        J = [[beam_amplitude(jones[:, p, k].real, crd_top, I_sky) +
              1j * beam_amplitude(jones[:, p, k].imag, crd_top, I_sky) for k in range(2)] for p in range(2)]
        dist = antpos @ crd_top
        for f, nu in enumerate(freqs):
            ph = np.exp(1j * (2.0 * np.pi * nu / C_LIGHT) * dist)
            for p in range(2):
                for q in range(2):
                    out[f, p, q, t] = sum((J[p][k] * ph) @ (J[q][k] * ph).conj().T for k in range(2))
    return out


def rel_err(v, v_ref) -> float:
    """Frobenius relative error ``||v - v_ref|| / ||v_ref||`` (0 if both zero)."""
    v = np.asarray(v, dtype=np.complex128)
    v_ref = np.asarray(v_ref, dtype=np.complex128)
    den = np.linalg.norm(v_ref.ravel())
    num = np.linalg.norm((v - v_ref).ravel())
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return float(num / den)

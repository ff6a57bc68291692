"""Pins of the polarised (Jones-beam) oracle ``vis_oracle_pol``.

The reference simulates a real scalar beam only (hera_gpu/vis.py:162-175,
"XXX complex?" at vis.py:175), so there is no reference output to compare
with.  The polarised oracle is pinned by what the mathematics fixes:
  * a diagonal Jones beam J^{pk} = A delta_pk gives the scalar oracle on each
    feed and exact zeros across feeds;
  * a unitary mixing of the sky axes, the same for every antenna and
    direction, leaves the result of a diagonal beam unchanged;
  * a pure-Python brute force with scipy's degree-1 spline per real and
    imaginary plane on tiny inputs with random complex beams;
  * Hermitian symmetry V[p, q, a, b] = conj(V[q, p, b, a]);
and planted mistakes (conjugate on the wrong side, feed and sky axes
swapped) that the brute force must catch.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import RectBivariateSpline

from oracle import vis_oracle as vo

C = vo.C_LIGHT


def case(seed, N=3, K=7, T=2, F=2, P=9):
    rng = np.random.default_rng(seed)
    antpos = rng.uniform(-300, 300, (N, 3))
    antpos[:, 2] *= 0.01
    z = rng.uniform(-0.3, 1, K)
    phi = rng.uniform(0, 2 * np.pi, K)
    crd = np.stack([np.sqrt(1 - z * z) * np.cos(phi), np.sqrt(1 - z * z) * np.sin(phi), z])
    th = rng.uniform(0, 2 * np.pi, T)
    eq = np.stack([[[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]] for a in th])
    I = rng.lognormal(0, 1, K)
    jones = rng.uniform(-1, 1, (N, 2, 2, P, P)) + 1j * rng.uniform(-1, 1, (N, 2, 2, P, P))
    freqs = np.array([1.0e8, 1.9e8][:F])
    return antpos, freqs, eq, crd, I, jones


def brute_force(antpos, freqs, eq2tops, crd_eq, I_sky, jones, mistake=None):
    N, K = len(antpos), crd_eq.shape[1]
    P = jones.shape[-1]
    grid = np.linspace(-1, 1, P)
    spl = {}
    for a in range(N):
        for p in range(2):
            for k in range(2):
                spl[a, p, k] = (RectBivariateSpline(grid, grid, jones[a, p, k].real, kx=1, ky=1),
                                RectBivariateSpline(grid, grid, jones[a, p, k].imag, kx=1, ky=1))
    out = np.zeros((len(freqs), 2, 2, len(eq2tops), N, N), dtype=np.complex128)
    for f, nu in enumerate(freqs):
        for t, R in enumerate(eq2tops):
            for s in range(K):
                top = [sum(float(R[i][j]) * float(crd_eq[j][s]) for j in range(3)) for i in range(3)]
                if top[2] <= 0:
                    continue
                e = []
                for a in range(N):
                    tau = sum(float(antpos[a][j]) * top[j] for j in range(3)) / C
                    e.append(complex(math.cos(2 * math.pi * nu * tau), math.sin(2 * math.pi * nu * tau)))
                J = {}
                for a in range(N):
                    for p in range(2):
                        for k in range(2):
                            sr, si = spl[a, p, k] if mistake != "swap_axes" else spl[a, k, p]
                            J[a, p, k] = complex(float(sr(top[1], top[0], grid=False)),
                                                 float(si(top[1], top[0], grid=False)))
                for p in range(2):
                    for q in range(2):
                        for a in range(N):
                            for b in range(N):
                                acc = 0j
                                for k in range(2):
                                    va = J[a, p, k] * e[a]
                                    vb = J[b, q, k] * e[b]
                                    acc += (va.conjugate() * vb) if mistake == "conj_side" else (va * vb.conjugate())
                                out[f, p, q, t, a, b] += acc * float(I_sky[s])
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_pol_oracle_matches_brute_force(seed):
    c = case(seed)
    ours = vo.vis_oracle_pol(*c)
    ref = brute_force(*c)
    np.testing.assert_allclose(ours, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_diagonal_jones_is_scalar_beam_per_feed():
    antpos, freqs, eq, crd, I, _ = case(2, N=4, K=60, P=15)
    bm = np.random.default_rng(3).uniform(0, 1, (4, 15, 15))
    jones = np.zeros((4, 2, 2, 15, 15), dtype=np.complex128)
    jones[:, 0, 0] = bm
    jones[:, 1, 1] = bm
    v = vo.vis_oracle_pol(antpos, freqs, eq, crd, I, jones)
    scalar = vo.vis_oracle(antpos, freqs, eq, crd, I, bm)
    for p in range(2):
        np.testing.assert_allclose(v[:, p, p], scalar, rtol=0, atol=1e-12 * np.abs(scalar).max())
    assert np.all(v[:, 0, 1] == 0) and np.all(v[:, 1, 0] == 0)


def test_unitary_sky_axis_mixing_leaves_diagonal_beam_unchanged():
    antpos, freqs, eq, crd, I, _ = case(4, N=4, K=60, P=15)
    bm = np.random.default_rng(5).uniform(0, 1, (4, 15, 15))
    th, ph = 0.7, 1.9
    U = np.array([[math.cos(th), -math.sin(th) * np.exp(1j * ph)], [math.sin(th), math.cos(th) * np.exp(1j * ph)]])
    assert np.allclose(U @ U.conj().T, np.eye(2))
    jones = bm[:, None, None] * U[None, :, :, None, None]
    v = vo.vis_oracle_pol(antpos, freqs, eq, crd, I, jones)
    scalar = vo.vis_oracle(antpos, freqs, eq, crd, I, bm)
    tol = 1e-12 * np.abs(scalar).max()
    for p in range(2):
        for q in range(2):
            np.testing.assert_allclose(v[:, p, q], scalar if p == q else 0 * scalar, rtol=0, atol=tol)


def test_hermitian_across_feeds():
    c = case(6, N=4, K=40, P=11)
    v = vo.vis_oracle_pol(*c)
    np.testing.assert_allclose(v, np.conj(np.transpose(v, (0, 2, 1, 3, 5, 4))), rtol=0, atol=1e-12 * np.abs(v).max())


@pytest.mark.parametrize("mistake", ["conj_side", "swap_axes"])
def test_pins_catch_planted_mistakes(mistake):
    c = case(7, N=3, K=7, P=7)
    ref = brute_force(*c)
    bad = brute_force(*c, mistake=mistake)
    assert vo.rel_err(bad, ref) > 1e-3
    assert vo.rel_err(vo.vis_oracle_pol(*c), ref) < 1e-12

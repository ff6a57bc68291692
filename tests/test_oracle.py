"""Pins for the FP64 oracle (oracle/vis_oracle.py).

Each check is independent of the oracle's own formula:
  * the bilinear beam against scipy's degree-1 RectBivariateSpline, the
    library routine the reference CPU path uses (test_vis_gpu.py:334-335);
  * the whole equation against a pure-Python brute-force loop (math.cos /
    math.sin, explicit conjugation) on tiny inputs;
  * against the restated reference ``vis_cpu`` in float64 (oracle/cpu_ref.py,
    a different code path: spline objects, per-row dot products, conj-then-
    conj fill);
  * the reference's closed-form cases (test_vis_gpu.py:31-139, 203-244);
  * invariants: Hermitian, real non-negative autos equal to sum A^2 I,
    Cauchy-Schwarz, invariance under a common antenna shift.
Mutation checks plant a mistake in a copy of the oracle (wrong phase sign,
transposed beam axes, dropped horizon) and require a pin to catch it.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import RectBivariateSpline

from oracle import vis_oracle as vo
from oracle.cpu_ref import vis_cpu

C = vo.C_LIGHT


def brute_force(antpos, freqs, eq2tops, crd_eq, I_sky, bm_cube):
    """Pure-Python loops; the beam through scipy's linear spline (clamped)."""
    N = len(antpos)
    K = crd_eq.shape[1]
    P = bm_cube.shape[-1]
    grid = np.linspace(-1, 1, P)
    planes = bm_cube if bm_cube.ndim == 4 else [bm_cube] * len(freqs)   # a beam per frequency
    out = np.zeros((len(freqs), len(eq2tops), N, N), dtype=np.complex128)
    for f, nu in enumerate(freqs):
        splines = [RectBivariateSpline(grid, grid, np.asarray(planes[f][i], dtype=np.float64), kx=1, ky=1)
                   for i in range(N)]
        for t, R in enumerate(eq2tops):
            v = [[0j] * K for _ in range(N)]
            for s in range(K):
                top = [sum(float(R[i][j]) * float(crd_eq[j][s]) for j in range(3)) for i in range(3)]
                for a in range(N):
                    A = float(splines[a](top[1], top[0], grid=False)) if top[2] > 0 else 0.0
                    tau = sum(float(antpos[a][j]) * top[j] for j in range(3)) / C
                    ph = 2 * math.pi * float(nu) * tau
                    amp = A * math.sqrt(float(I_sky[s]))
                    v[a][s] = complex(amp * math.cos(ph), amp * math.sin(ph))
            for a in range(N):
                for b in range(N):
                    out[f, t, a, b] = sum(v[a][s] * v[b][s].conjugate() for s in range(K))
    return out


def tiny_case(seed=0, N=3, K=7, T=2, F=2, P=9, big_phase=True):
    rng = np.random.default_rng(seed)
    antpos = rng.uniform(-300, 300, (N, 3)) if big_phase else rng.uniform(0, 1, (N, 3))
    antpos[:, 2] *= 0.01
    z = rng.uniform(-0.3, 1, K)
    phi = rng.uniform(0, 2 * np.pi, K)
    crd = np.stack([np.sqrt(1 - z * z) * np.cos(phi), np.sqrt(1 - z * z) * np.sin(phi), z])
    th = rng.uniform(0, 2 * np.pi, T)
    eq = np.stack([[[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]] for a in th])
    I = rng.lognormal(0, 1, K)
    bm = rng.uniform(0, 1, (N, P, P))
    freqs = np.array([1.0e8, 1.9e8][:F])
    return antpos, freqs, eq, crd, I, bm


# ------------------------------------------------------------------ beam
def test_bilinear_matches_rectbivariatespline_inside_and_outside():
    rng = np.random.default_rng(1)
    for P in (2, 5, 31, 63):
        beam = rng.standard_normal((P, P))
        grid = np.linspace(-1, 1, P)
        spl = RectBivariateSpline(grid, grid, beam, kx=1, ky=1)
        x = rng.uniform(-1.6, 1.6, 4000)
        y = rng.uniform(-1.6, 1.6, 4000)
        x[:5] = [-1, 1, 0, -1.0, 1.0]
        y[:5] = [-1, 1, 0, 1.0, -1.0]
        ours = vo.bilinear_clamp(beam, x, y)
        ref = spl(y, x, grid=False)
        np.testing.assert_allclose(ours, ref, rtol=0, atol=1e-12 * np.abs(beam).max())


def test_bilinear_grid_nodes_exact():
    P = 31
    beam = np.arange(P * P, dtype=np.float64).reshape(P, P)
    grid = np.linspace(-1, 1, P)
    for iy in (0, 7, 15, 30):
        for ix in (0, 3, 15, 30):
            assert vo.bilinear_clamp(beam, grid[ix], grid[iy]) == pytest.approx(beam[iy, ix], abs=1e-9)


# ---------------------------------------------- reference closed forms
def test_flat_beam_is_one_above_and_zero_below_horizon():
    # test_vis_gpu.py:31-65
    bm = np.ones((8, 31, 31))
    crd = np.zeros((3, 128))
    crd[2] = 0.9
    amp = vo.beam_amplitude(bm, crd, np.ones(128))
    assert np.all(amp == 1)
    crd[2] = -0.9
    assert np.all(vo.beam_amplitude(bm, crd, np.ones(128)) == 0)


def test_delta_beam_centre_and_off_centre():
    # test_vis_gpu.py:66-102
    bm = np.zeros((8, 31, 31))
    bm[:, 15, 15] = 1
    crd = np.zeros((3, 128))
    crd[2] = 1.0
    assert np.all(vo.beam_amplitude(bm, crd, np.ones(128)) == 1)
    crd[2] = 0.5 ** 0.5
    crd[1] = 0.5 ** 0.5
    assert np.all(vo.beam_amplitude(bm, crd, np.ones(128)) == 0)


def test_values_closed_forms():
    # test_vis_gpu.py:203-232
    NANT, NTIMES, BM, NPIX = 16, 10, 31, 3072
    antpos = np.ones((NANT, 3))
    eq2tops = np.array([np.identity(3)] * NTIMES)
    crd = np.zeros((3, NPIX))
    crd[2] = 1
    I = np.ones(NPIX)
    bm = np.ones((NANT, BM, BM))
    assert np.all(vo.vis_oracle(antpos, 1.0, eq2tops, crd, I * 0, bm) == 0)
    assert np.all(vo.vis_oracle(antpos, 1.0, eq2tops, crd, I, bm * 0) == 0)
    v = vo.vis_oracle(antpos, 1.0, eq2tops, crd, I, bm)
    np.testing.assert_allclose(v, NPIX, rtol=1e-14)
    crd2 = np.zeros((3, 2))
    crd2[2, 0] = 1
    crd2[1, 1] = np.sqrt(0.5)
    crd2[2, 1] = np.sqrt(0.5)
    v = vo.vis_oracle(antpos, 1.0, eq2tops, crd2, np.ones(2), bm)
    np.testing.assert_almost_equal(v, 2, 2)


# ----------------------------------------------------- independent checks
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_matches_brute_force(seed):
    case = tiny_case(seed)
    ours = vo.vis_oracle(*case)
    ref = brute_force(*case)
    assert vo.rel_err(ours, ref) < 1e-12
    np.testing.assert_allclose(ours, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_oracle_matches_restated_vis_cpu_fp64():
    antpos, freqs, eq, crd, I, bm = tiny_case(3, N=6, K=200, T=3, F=2, P=31)
    ours = vo.vis_oracle(antpos, freqs, eq, crd, I, bm)
    for f, nu in enumerate(freqs):
        ref = vis_cpu(antpos, nu, eq, crd, I, bm, precision=2)
        assert vo.rel_err(ours[f], ref) < 1e-12


def test_vis_cpu_fp32_agrees_to_single_precision_small_phase():
    antpos, freqs, eq, crd, I, bm = tiny_case(4, N=5, K=300, T=2, F=1, P=31, big_phase=False)
    ref = vo.vis_oracle(antpos, freqs, eq, crd, I, bm)[0]
    v32 = vis_cpu(antpos.astype(np.float32), freqs[0], eq.astype(np.float32), crd.astype(np.float32),
                  I.astype(np.float32), bm.astype(np.float32), precision=1)
    assert v32.dtype == np.complex64
    assert vo.rel_err(v32, ref) < 1e-5


def test_per_frequency_beam_matches_brute_force():
    antpos, freqs, eq, crd, I, bm = tiny_case(7, N=3, K=9, T=2, F=2)
    bm4 = np.stack([bm, np.random.default_rng(8).uniform(0, 1, bm.shape)])
    ours = vo.vis_oracle(antpos, freqs, eq, crd, I, bm4)
    ref = brute_force(antpos, freqs, eq, crd, I, bm4)
    np.testing.assert_allclose(ours, ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    # frequency f uses plane f: each frequency equals a one-beam run with that plane
    for f in range(2):
        np.testing.assert_array_equal(ours[f], vo.vis_oracle(antpos, freqs[f:f + 1], eq, crd, I, bm4[f])[0])
    units = [(1, 0), (0, 1), (1, 1)]
    part = vo.vis_oracle_units(antpos, freqs, eq, crd, I, bm4, units)
    for i, (f, t) in enumerate(units):
        np.testing.assert_array_equal(part[i], ours[f, t])


def test_identical_frequency_planes_equal_one_beam():
    antpos, freqs, eq, crd, I, bm = tiny_case(9, N=4, K=40, T=2, F=2)
    np.testing.assert_array_equal(vo.vis_oracle(antpos, freqs, eq, crd, I, np.stack([bm, bm])),
                                  vo.vis_oracle(antpos, freqs, eq, crd, I, bm))


def test_units_sampler_matches_full():
    case = tiny_case(5, N=4, K=50, T=3, F=2)
    full = vo.vis_oracle(*case)
    units = [(1, 2), (0, 0), (1, 0)]
    part = vo.vis_oracle_units(*case, units)
    for i, (f, t) in enumerate(units):
        np.testing.assert_array_equal(part[i], full[f, t])


# ------------------------------------------------------------- invariants
def test_hermitian_real_autos_cauchy_schwarz():
    antpos, freqs, eq, crd, I, bm = tiny_case(6, N=8, K=500, T=2, F=2, P=15)
    v = vo.vis_oracle(antpos, freqs, eq, crd, I, bm)
    np.testing.assert_allclose(v, np.conj(np.swapaxes(v, -1, -2)), atol=1e-9)
    d = np.einsum("ftaa->fta", v)
    assert np.all(d.imag == 0) and np.all(d.real >= 0)
    for t in range(len(eq)):
        amp = vo.beam_amplitude(bm, eq[t] @ crd, I)
        np.testing.assert_allclose(d[0, t].real, (amp ** 2).sum(axis=1), rtol=1e-12)
    bound = np.sqrt(np.abs(d[..., :, None]) * np.abs(d[..., None, :]))
    assert np.all(np.abs(v) <= bound * (1 + 1e-12) + 1e-12)


def test_common_shift_invariance_realistic_phase():
    # Moving every antenna by the same vector multiplies each v_a(s) by the same
    # unit phasor per source, which cancels in v_a conj(v_b): a 1 km shift at
    # 200 MHz (~4e3 rad of extra phase) must leave V unchanged.
    antpos, freqs, eq, crd, I, bm = tiny_case(7, N=6, K=400, T=2, F=2, P=31)
    freqs = np.array([1.0e8, 2.0e8])
    shift = np.array([1000.0, -400.0, 3.0])
    a = vo.vis_oracle(antpos, freqs, eq, crd, I, bm)
    b = vo.vis_oracle(antpos + shift, freqs, eq, crd, I, bm)
    assert vo.rel_err(b, a) < 1e-10


# ------------------------------------------------ planted-mistake checks
def _mutant_vis(antpos, freqs, eq2tops, crd_eq, I_sky, bm_cube, mutation):
    antpos = np.asarray(antpos, dtype=np.float64)
    out = np.empty((len(freqs), len(eq2tops), len(antpos), len(antpos)), dtype=np.complex128)
    for t in range(len(eq2tops)):
        top = eq2tops[t] @ crd_eq
        x, y, z = top
        if mutation == "transposed_beam":
            A = vo.bilinear_clamp(np.swapaxes(bm_cube, -1, -2), x, y)
        else:
            A = vo.bilinear_clamp(bm_cube, x, y)
        if mutation != "no_horizon":
            A = np.where(z > 0, A, 0.0)
        amp = A * np.sqrt(I_sky)
        sign = -1.0 if mutation == "phase_sign" else 1.0
        for f, nu in enumerate(freqs):
            v = amp * np.exp(sign * 1j * 2 * np.pi * nu / C * (antpos @ top))
            out[f, t] = v @ v.conj().T
    return out


@pytest.mark.parametrize("mutation", ["phase_sign", "transposed_beam", "no_horizon"])
def test_pins_catch_planted_mistakes(mutation):
    case = tiny_case(8, N=3, K=9, T=2, F=1, P=7)
    ref = brute_force(*case)
    assert vo.rel_err(_mutant_vis(*case, mutation="none"), ref) < 1e-12
    assert vo.rel_err(_mutant_vis(*case, mutation=mutation), ref) > 1e-3


def test_pins_catch_wrong_frequency_plane():
    # a per-frequency beam read at plane 0 for every frequency (or planes in
    # reverse order) must fail the brute-force pin
    antpos, freqs, eq, crd, I, bm = tiny_case(10, N=3, K=9, T=2, F=2, P=7)
    bm4 = np.stack([bm, np.random.default_rng(11).uniform(0, 1, bm.shape)])
    ref = brute_force(antpos, freqs, eq, crd, I, bm4)
    for wrong in (np.stack([bm4[0], bm4[0]]), bm4[::-1]):
        assert vo.rel_err(vo.vis_oracle(antpos, freqs, eq, crd, I, wrong), ref) > 1e-3


@pytest.mark.parametrize("precision", [1, 2])
def test_vis_cpu_parallel_equals_vis_cpu_bitwise(precision):
    # the CPU baseline's process-parallel form (per-antenna rows on spawned
    # workers, shared memory) against the restated vis_cpu: same bits
    from hera_gpu_amd import configs
    from oracle.cpu_ref import vis_cpu_parallel

    w = configs.make_hex37()
    a = vis_cpu(w.antpos, 150e6, w.eq2tops[:3], w.crd_eq, w.I_sky, w.bm_cube, precision=precision)
    b = vis_cpu_parallel(w.antpos, 150e6, w.eq2tops[:3], w.crd_eq, w.I_sky, w.bm_cube, precision=precision,
                         workers=3)
    assert a.dtype == b.dtype
    np.testing.assert_array_equal(a, b)

"""Omnical on the GPU against the CPU restatement (oracle/omnical_oracle.py,
pinned in test_omnical_oracle.py): same iteration counts, and solutions, chi^2
and convergence measures equal to rounding, in both precisions; the
reference's own check (hera_gpu/tests/test_redcal.py:129-146) compares
against hera_cal, which is not installed."""
import numpy as np
import pytest

from hera_gpu_amd.redcal import bench_omnical, omnical, set_placement
from oracle import omnical_oracle as oo
from oracle.omnical_oracle import redundant_array

pytestmark = pytest.mark.gpu


def perturbed(seed, n=4, nd=6, noise=0.01):
    ggu, g, u, data, wgts = redundant_array(n=n, seed=seed, nd=nd, noise=noise)
    rng = np.random.default_rng(seed + 100)
    g0 = g * (1 + 0.05 * (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)))
    u0 = u * (1 + 0.05 * (rng.standard_normal(u.shape) + 1j * rng.standard_normal(u.shape)))
    wgts = wgts * rng.uniform(0.5, 1.5, wgts.shape)
    return ggu, g0, u0, data, wgts


def compare(gpu, ref, rtol):
    np.testing.assert_array_equal(gpu["iters"], ref["iters"])
    for key in ("gains", "ubls"):
        scale = np.abs(ref[key]).max()
        np.testing.assert_allclose(gpu[key], ref[key], rtol=0, atol=rtol * scale)
    np.testing.assert_allclose(gpu["chisq"], ref["chisq"], rtol=rtol, atol=rtol * np.abs(ref["chisq"]).max())
    np.testing.assert_allclose(gpu["conv"], ref["conv"], rtol=rtol, atol=0)


@pytest.mark.parametrize("precision", [1, 2])
def test_omnical_matches_restatement(precision):
    ggu, g0, u0, data, wgts = perturbed(1)
    kw = dict(maxiter=60, check_every=4, check_after=1, gain=0.3, precision=precision)
    gpu = omnical(ggu, g0, u0, data, wgts, **kw)
    ref = oo.omnical(ggu, g0, u0, data, wgts, **kw)
    compare(gpu, ref, 1e-12 if precision == 2 else 1e-5)
    assert gpu["gains"].dtype == (np.complex64 if precision == 1 else np.complex128)


def test_omnical_bitwise_equal_to_restatement_fp64():
    # the same operations in the same order: the FP64 solve is the
    # restatement's bit for bit
    ggu, g0, u0, data, wgts = perturbed(2)
    kw = dict(maxiter=40, precision=2)
    gpu = omnical(ggu, g0, u0, data, wgts, **kw)
    ref = oo.omnical(ggu, g0, u0, data, wgts, **kw)
    for key in ("gains", "ubls", "chisq", "iters", "conv"):
        np.testing.assert_array_equal(gpu[key], ref[key], err_msg=key)


@pytest.mark.parametrize("place", [0, 1, 2])
def test_omnical_every_placement_bitwise_fp64(place):
    # per-baseline arrays in device memory, products in LDS, or all in LDS:
    # the same operations in the same order whichever is used
    ggu, g0, u0, data, wgts = perturbed(5, n=5, nd=300)
    kw = dict(maxiter=24, check_every=4, check_after=1, precision=2)
    set_placement(place)
    try:
        gpu = omnical(ggu, g0, u0, data, wgts, **kw)
    finally:
        set_placement(-1)
    ref = oo.omnical(ggu, g0, u0, data, wgts, **kw)
    for key in ("gains", "ubls", "chisq", "iters", "conv"):
        np.testing.assert_array_equal(gpu[key], ref[key], err_msg=f"{key} (placement {place})")


def test_omnical_many_pixels_and_options():
    # more pixels than resident workgroups (each workgroup solves several),
    # convergence checks every 3 iterations from iteration 5, a zero-weight
    # baseline, and a pixel that starts at the exact solution
    ggu, g0, u0, data, wgts = perturbed(3, n=3, nd=2500, noise=0.02)
    wgts[:, 5] = 0.0
    ggu_e, g_e, u_e, data_e, w_e = redundant_array(n=3, seed=9, nd=1)
    g0[0], u0[0], data[0], wgts[0] = g_e[0], u_e[0], data_e[0], w_e[0]
    kw = dict(maxiter=30, check_every=3, check_after=5, gain=0.4, precision=2)
    gpu = omnical(ggu, g0, u0, data, wgts, **kw)
    ref = oo.omnical(ggu, g0, u0, data, wgts, **kw)
    compare(gpu, ref, 1e-12)
    assert gpu["iters"][0] == 0 and gpu["chisq"][0] < 1e-20


def test_omnical_argument_checks():
    ggu, g0, u0, data, wgts = perturbed(4, nd=2)
    bad = ggu.copy()
    bad[0, 0] = 99
    with pytest.raises(AssertionError):
        omnical(bad, g0, u0, data, wgts)
    with pytest.raises(AssertionError):
        omnical(ggu, g0, u0, data, wgts, check_every=1)
    with pytest.raises(AssertionError):
        omnical(ggu, g0[:, :3], u0, data, wgts)


def test_timed_solve_repeats_the_same_solve():
    ggu, g0, u0, data, wgts = perturbed(6, n=4, nd=40)
    kw = dict(maxiter=20, precision=2)
    one = omnical(ggu, g0, u0, data, wgts, **kw)
    timed = bench_omnical(ggu, g0, u0, data, wgts, reps=3, **kw)
    assert 0 < timed["kernel_ms"] < 1e4
    for key in ("gains", "ubls", "chisq", "iters", "conv"):
        np.testing.assert_array_equal(timed[key], one[key], err_msg=key)

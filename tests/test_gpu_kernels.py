"""Kernel-level parity on the GPU, one device stage per test.

Mirrors the reference's TestCuda (hera_gpu/tests/test_vis_gpu.py:17-169) with
the same closed-form cases, then checks each stage against the FP64 oracle
(or FP64 NumPy for the contraction) on random inputs.
"""
import numpy as np
import pytest

from hera_gpu_amd import stages
from oracle import vis_oracle as vo

pytestmark = pytest.mark.gpu
C = vo.C_LIGHT
MODES = ["fp32", "split_bf16", "split_fp16", "fp64"]
# relative Frobenius tolerance of V = v v^H against FP64, per arithmetic
HERK_TOL = {"fp32": 2e-6, "split_bf16": 2e-5, "split_fp16": 2e-6, "fp64": 1e-13}


# ------------------------------------------------------------ rotation
def test_rotate_identity_like_test_cublas():
    rng = np.random.default_rng(0)
    crd = rng.standard_normal((3, 128)).astype(np.float32)
    out = stages.rotate(np.eye(3, dtype=np.float32), crd)
    np.testing.assert_allclose(out, crd, rtol=1e-6)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rotate_matches_fp64(dtype):
    rng = np.random.default_rng(1)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    crd = rng.standard_normal((3, 5000))
    out = stages.rotate(q.astype(dtype), crd.astype(dtype), dtype=dtype)
    ref = q.astype(dtype).astype(np.float64) @ crd.astype(dtype).astype(np.float64)
    tol = 1e-14 if dtype == np.float64 else 1e-7
    np.testing.assert_allclose(out, ref, rtol=0, atol=tol * 4)


# ---------------------------------------------------------------- beam
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_beam_flat(dtype):
    bm = np.ones((8, 31, 31), dtype=dtype)
    crd = np.zeros((3, 128), dtype=dtype)
    crd[2] = 0.9
    assert np.all(stages.beam_interp(bm, crd, dtype=dtype) == 1)
    crd[2] = -0.9
    assert np.all(stages.beam_interp(bm, crd, dtype=dtype) == 0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_beam_middle(dtype):
    bm = np.zeros((8, 31, 31), dtype=dtype)
    bm[:, 15, 15] = 1
    crd = np.zeros((3, 128), dtype=dtype)
    crd[2] = 1.0
    assert np.all(stages.beam_interp(bm, crd, dtype=dtype) == 1)
    crd[2] = 0.5 ** 0.5
    crd[1] = 0.5 ** 0.5
    assert np.all(stages.beam_interp(bm, crd, dtype=dtype) == 0)


@pytest.mark.parametrize("dtype,P", [(np.float32, 31), (np.float32, 63), (np.float64, 63), (np.float32, 2)])
def test_beam_random_vs_oracle(dtype, P):
    rng = np.random.default_rng(2)
    N, K = 7, 3001
    bm = rng.uniform(-1, 1, (N, P, P)).astype(dtype)
    crd = rng.uniform(-1.4, 1.4, (3, K)).astype(dtype)   # includes out-of-range and z <= 0
    out = stages.beam_interp(bm, crd, dtype=dtype)
    ref = vo.beam_amplitude(bm.astype(np.float64), crd.astype(np.float64), np.ones(K))
    tol = 1e-12 if dtype == np.float64 else 2e-6 * P
    np.testing.assert_allclose(out, ref, rtol=0, atol=tol)


# --------------------------------------------------------------- phasor
@pytest.mark.parametrize("gemm", MODES)
def test_phasor_unit_amplitude_zero_delay(gemm):
    # test_meas_eq part 1: A = I = 1 and no delay give v == 1 exactly
    dtype = np.float64 if gemm == "fp64" else np.float32
    N, K = 8, 128
    crd = np.zeros((3, K), dtype=dtype)
    crd[2] = 1.0
    v = stages.phasor(np.ones((N, 31, 31), dtype), crd, np.ones(K, dtype), np.zeros((N, 3), dtype), 1.5e8,
                      gemm=gemm, dtype=dtype)
    assert np.all(v == 1)


@pytest.mark.parametrize("gemm", MODES)
def test_phasor_whole_turns(gemm):
    # test_meas_eq part 2: a delay of exactly 10 cycles gives v ~= 1
    dtype = np.float64 if gemm == "fp64" else np.float32
    N, K = 8, 128
    nu = 1.0e8
    crd = np.zeros((3, K), dtype=dtype)
    crd[2] = 1.0
    antpos = np.zeros((N, 3), dtype=dtype)
    antpos[:, 2] = 10 * C / nu  # 29.979... m, exact in float32? no: the path is FP64 after the cast
    ap64 = antpos.astype(np.float64)
    v = stages.phasor(np.ones((N, 31, 31), dtype), crd, np.ones(K, dtype), antpos, nu, gemm=gemm, dtype=dtype)
    ref = np.exp(2j * np.pi * nu / C * ap64[:, 2])[:, None] * np.ones((1, K))
    np.testing.assert_allclose(v, ref, rtol=1e-6 if gemm != "split_bf16" else 2e-5)
    np.testing.assert_allclose(v, 1, rtol=1e-5)


@pytest.mark.parametrize("gemm", MODES)
def test_phasor_realistic_vs_oracle(gemm):
    rng = np.random.default_rng(3)
    dtype = np.float64 if gemm == "fp64" else np.float32
    N, K, P = 11, 5000, 63
    antpos = np.zeros((N, 3))
    antpos[:, :2] = rng.uniform(-1000, 1000, (N, 2))
    antpos = antpos.astype(dtype)
    z = rng.uniform(-0.3, 1, K)
    phi = rng.uniform(0, 2 * np.pi, K)
    crd = np.stack([np.sqrt(1 - z * z) * np.cos(phi), np.sqrt(1 - z * z) * np.sin(phi), z]).astype(dtype)
    isq = rng.uniform(0.1, 3, K).astype(dtype)
    bm = rng.uniform(0, 1, (N, P, P)).astype(dtype)
    nu = 1.97e8
    v = stages.phasor(bm, crd, isq, antpos, nu, gemm=gemm, dtype=dtype)
    c64 = [a.astype(np.float64) for a in (bm, crd, isq, antpos)]
    amp = vo.beam_amplitude(c64[0], c64[1], c64[2] ** 2)
    ref = amp * np.exp(2j * np.pi * nu / C * (c64[3] @ c64[1]))
    err = np.abs(v - ref).max() / np.abs(ref).max()
    tol = {"fp32": 1e-6, "split_bf16": 2e-5, "split_fp16": 2e-6, "fp64": 1e-12}[gemm]
    assert err < tol, err


# ------------------------------------------------------- contraction
@pytest.mark.parametrize("gemm", MODES)
@pytest.mark.parametrize("N,K", [(1, 1), (5, 100), (37, 1000), (64, 8192), (65, 8193), (130, 20000),
                                 (300, 9000)])
def test_herk_vs_numpy(gemm, N, K):
    rng = np.random.default_rng(N * 100003 + K)
    v = (rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))) * rng.uniform(0.01, 3, K)
    cd = np.complex128 if gemm == "fp64" else np.complex64
    v = v.astype(cd)
    V = stages.herk(v, gemm=gemm)
    ref = v.astype(np.complex128) @ v.astype(np.complex128).conj().T
    assert V.dtype == cd and V.shape == (N, N)
    assert vo.rel_err(V, ref) < HERK_TOL[gemm]
    # exact Hermitian structure from the upper-triangle computation
    np.testing.assert_array_equal(V, V.conj().T)
    assert np.all(np.diag(V).imag == 0)


@pytest.mark.parametrize("gemm", MODES)
def test_herk_colocated_is_exact(gemm):
    # test_values' co-located case: every antenna and source gives the same
    # phasor exp(2 pi i * 1 Hz * 1 m / c), so Re = K and Im = 0 exactly
    K = 3072
    row = np.full(K, np.exp(2j * np.pi / C))
    cd = np.complex128 if gemm == "fp64" else np.complex64
    v = np.tile(row, (16, 1)).astype(cd)
    V = stages.herk(v, gemm=gemm)
    assert np.all(V.imag == 0)
    assert np.all(V.real == K)


# every contraction variant (tile x prefetch scheme) against FP64 NumPy
VARIANTS = [(64, 1), (64, 2), (64, 3), (64, 4), (128, 1), (128, 2), (128, 3), (128, 4)]


@pytest.mark.parametrize("tile,pf", VARIANTS)
@pytest.mark.parametrize("gemm", MODES)
def test_herk_variants(gemm, tile, pf):
    from hera_gpu_amd import _lib
    if tile == 128 and gemm in ("fp32", "fp64"):
        pytest.skip("128-tiles exist for the split modes only")
    if gemm == "fp64" and pf == 2:
        pytest.skip("FP64 has prefetch depths 1, 3 and 4")
    lib = _lib.load()
    _lib.check(lib.hv_set_herk_variant(tile, pf))
    try:
        for N, K in ((37, 1000), (130, 20000), (350, 8192 * 2 + 64)):
            rng = np.random.default_rng(N + K + tile + pf)
            cd = np.complex128 if gemm == "fp64" else np.complex64
            v = ((rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))) * rng.uniform(0.01, 3, K)).astype(cd)
            V = stages.herk(v, gemm=gemm)
            ref = v.astype(np.complex128) @ v.astype(np.complex128).conj().T
            assert vo.rel_err(V, ref) < HERK_TOL[gemm], (N, K)
            np.testing.assert_array_equal(V, V.conj().T)
    finally:
        _lib.check(lib.hv_set_herk_variant(0, 0))


@pytest.mark.parametrize("gemm,tile", [("split_bf16", 64), ("split_bf16", 128), ("fp32", 64), ("fp64", 64)])
def test_herk_persistent_grid_equals_one_workgroup_per_item(gemm, tile):
    # more work items than resident workgroups, so each persistent workgroup
    # runs several (unit, split, tile) items; same arithmetic per item
    from hera_gpu_amd import _lib
    lib = _lib.load()
    N, K = 350, 8192 * 48 + 64
    rng = np.random.default_rng(7)
    cd = np.complex128 if gemm == "fp64" else np.complex64
    v = (rng.standard_normal((N, K), dtype=np.float32) + 1j * rng.standard_normal((N, K), dtype=np.float32)).astype(cd)
    out = []
    try:
        for pf in (3, 4):
            _lib.check(lib.hv_set_herk_variant(tile, pf))
            out.append(stages.herk(v, gemm=gemm))
    finally:
        _lib.check(lib.hv_set_herk_variant(0, 0))
    np.testing.assert_array_equal(out[0], out[1])
    # spot check one row against FP64 products
    ref = v[5].astype(np.complex128) @ v.astype(np.complex128).conj().T
    assert vo.rel_err(out[1][5], ref) < 3 * HERK_TOL[gemm]


"""Beams larger than 128 pixels (the reference's texture path takes any P,
hera_gpu/vis.py:113-138).  Planes that fit in the 160 KB of LDS are staged
there (P = 200 float32: 156 KB, one antenna per workgroup); larger ones (P =
256 float32, P = 160 float64, four polarised planes of P = 128) are read from
device memory by the same bilinear arithmetic.  Each case is checked against
the FP64 oracle at the tolerances of the other vis tests."""
import numpy as np
import pytest

from hera_gpu_amd import configs, stages, vis_gpu, vis_gpu_polarized
from hera_gpu_amd.serve import Server
from oracle import vis_oracle as vo

pytestmark = pytest.mark.gpu

VIS_TOL = {"fp32": 2e-6, "split_bf16": 1e-5, "split_fp16": 2e-6, "fp64": 1e-12}


def wide_beams(nant, P, seed, dtype=np.float32):
    """Gaussian beams of ~40 degree width plus 1% noise on a P x P grid."""
    rng = np.random.default_rng(seed)
    g = np.linspace(-1, 1, P)
    X, Y = np.meshgrid(g, g)
    w = rng.uniform(0.3, 0.45, nant)
    bm = np.exp(-(X[None] ** 2 + Y[None] ** 2) / (2 * w[:, None, None] ** 2))
    bm = bm + 0.01 * rng.standard_normal(bm.shape)
    return bm.astype(dtype)


def jones_beams(nant, P, seed):
    """Gaussian amplitude times smooth complex feed/axis couplings."""
    rng = np.random.default_rng(seed)
    g = np.linspace(-1, 1, P)
    X, Y = np.meshgrid(g, g)
    w = rng.uniform(0.3, 0.5, nant)
    amp = np.exp(-(X[None] ** 2 + Y[None] ** 2) / (2 * w[:, None, None] ** 2))
    J = np.empty((nant, 2, 2, P, P), dtype=np.complex128)
    for p in range(2):
        for k in range(2):
            c = rng.uniform(0.2, 1.0, nant) * np.exp(1j * rng.uniform(0, 2 * np.pi, nant))
            tilt = np.exp(1j * (rng.uniform(-1, 1) * X + rng.uniform(-1, 1) * Y))
            J[:, p, k] = (1.0 if p == k else 0.3) * c[:, None, None] * amp * tilt[None]
    return J


def small_sky(nsrc=4000, ntimes=2):
    w = configs.make_hex37()
    h = configs.make_hera350(nsrc=nsrc, ntimes=ntimes, nfreq=1)
    return w.antpos, np.array([110e6, 190e6]), h.eq2tops, h.crd_eq, h.I_sky


@pytest.mark.parametrize("P", [200, 256])
@pytest.mark.parametrize("gemm", ["fp32", "split_bf16", "split_fp16"])
def test_large_beam_vs_oracle(gemm, P):
    antpos, freqs, eq, crd, I = small_sky()
    bm = wide_beams(antpos.shape[0], P, 3)
    v = vis_gpu(antpos, freqs, eq, crd, I, bm, gemm=gemm)
    ref = vo.vis_oracle(antpos, freqs, eq, crd, I, bm)
    assert v.shape == ref.shape
    err = vo.rel_err(v, ref)
    assert err < VIS_TOL[gemm], f"{gemm} P={P}: rel err {err:.2e}"


def test_large_beam_fp64_from_device_memory():
    # a float64 plane of 160 x 160 (200 KB) does not fit in LDS
    antpos, freqs, eq, crd, I = small_sky(2000)
    bm = wide_beams(antpos.shape[0], 160, 4, np.float64)
    v = vis_gpu(antpos, 150e6, eq, crd, I.astype(np.float64), bm, precision=2)
    ref = vo.vis_oracle(antpos, 150e6, eq, crd, I.astype(np.float64), bm)[0]
    assert vo.rel_err(v, ref) < VIS_TOL["fp64"]


def test_large_beam_kernel_stage_vs_oracle():
    rng = np.random.default_rng(5)
    N, K, P = 5, 3001, 300
    bm = rng.uniform(-1, 1, (N, P, P)).astype(np.float32)
    crd = rng.uniform(-1.4, 1.4, (3, K)).astype(np.float32)   # out of range and z <= 0 included
    out = stages.beam_interp(bm, crd, dtype=np.float32)
    ref = vo.beam_amplitude(bm.astype(np.float64), crd.astype(np.float64), np.ones(K))
    np.testing.assert_allclose(out, ref, rtol=0, atol=2e-6 * P)


@pytest.mark.parametrize("gemm", ["split_bf16", "fp32"])
def test_large_polarized_beam_vs_oracle(gemm):
    # four float32 planes of 128 x 128 (256 KB) per feed: taps from device memory
    antpos, freqs, eq, crd, I = small_sky(2500)
    J = jones_beams(antpos.shape[0], 128, 6).astype(np.complex64)
    v = vis_gpu_polarized(antpos, freqs, eq, crd, I, J, gemm=gemm)
    Ieff = np.sqrt(I).astype(np.float32).astype(np.float64) ** 2
    ref = vo.vis_oracle_pol(antpos, freqs, eq, crd, Ieff, J)
    for f in range(2):
        for a in range(2):
            for b in range(2):
                assert vo.rel_err(v[f, a, b], ref[f, a, b]) < VIS_TOL[gemm], (gemm, f, a, b)


def test_large_beam_served_equals_alone():
    antpos, freqs, eq, crd, I = small_sky(6000, 3)
    bm = wide_beams(antpos.shape[0], 256, 7)
    beam, sky = stages.Beam(bm), stages.Sky(crd, I)
    try:
        with Server(nslots=2) as srv:
            ids = [srv.submit(sky, beam, antpos, freqs, eq, nsrc=n) for n in (6000, 4100)]
            res = [srv.wait(i, 120) for i in ids]
        for r, n in zip(res, (6000, 4100)):
            assert r.status == 0, r.error
            alone, _ = stages.simulate(sky, beam, antpos, freqs, eq, nsrc=n, gemm="split_bf16")
            np.testing.assert_array_equal(r.vis, alone)
    finally:
        beam.close()
        sky.close()


def test_beam_beyond_limit_refused():
    antpos, freqs, eq, crd, I = small_sky(256, 1)
    bm = np.ones((antpos.shape[0], 2049, 2049), dtype=np.float32)
    with pytest.raises(AssertionError, match="beam_px"):
        vis_gpu(antpos, 150e6, eq, crd, I, bm)

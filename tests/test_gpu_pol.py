"""Polarised (Jones) beams on the GPU: ``vis_gpu_polarized`` against the FP64
polarised oracle (pinned in test_oracle_pol.py) for every arithmetic, and the
diagonal-beam case against the scalar path."""
import numpy as np
import pytest

from hera_gpu_amd import configs, vis_gpu, vis_gpu_polarized
from oracle.vis_oracle import rel_err, vis_oracle_pol

pytestmark = pytest.mark.gpu

VIS_TOL = {"fp32": 2e-6, "split_bf16": 1e-5, "split_fp16": 2e-6, "fp64": 1e-12}


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


def workload(nsrc=3000):
    w = configs.make_hex37()
    return w.antpos, np.array([120e6, 180e6]), w.eq2tops[:3], w.crd_eq[:, :nsrc], w.I_sky[:nsrc]


@pytest.mark.parametrize("gemm", ["fp32", "split_bf16", "split_fp16", "fp64"])
def test_polarized_vs_oracle(gemm):
    antpos, freqs, eq, crd, I = workload()
    J = jones_beams(antpos.shape[0], 31, 1)
    p = 2 if gemm == "fp64" else 1
    rd = np.float64 if p == 2 else np.float32
    Ju = J.astype(np.complex128 if p == 2 else np.complex64)
    Iu = I.astype(np.float64) if p == 2 else I
    v = vis_gpu_polarized(antpos, freqs, eq, crd, Iu, Ju, precision=p, gemm=gemm)
    assert v.shape == (2, 2, 2, 3, 37, 37)
    # the oracle on the values the call computes with
    Ieff = np.sqrt(Iu).astype(rd).astype(np.float64) ** 2
    ref = vis_oracle_pol(antpos.astype(rd), freqs, eq.astype(rd), crd.astype(rd), Ieff, Ju)
    for f in range(2):
        for a in range(2):
            for b in range(2):
                assert rel_err(v[f, a, b], ref[f, a, b]) < VIS_TOL[gemm], (gemm, f, a, b)


def test_polarized_scalar_freq_shape_and_errors():
    antpos, freqs, eq, crd, I = workload(500)
    J = jones_beams(37, 15, 2)
    v = vis_gpu_polarized(antpos, 150e6, eq, crd, I, J)
    assert v.shape == (2, 2, 3, 37, 37) and v.dtype == np.complex64
    with pytest.raises(AssertionError):
        vis_gpu_polarized(antpos, 150e6, eq, crd, I, J[:, :1])
    # four planes of 128 x 128 do not fit in LDS: taps from device memory
    z = vis_gpu_polarized(antpos, 150e6, eq, crd, I, np.zeros((37, 2, 2, 128, 128), np.complex64))
    assert z.shape == (2, 2, 3, 37, 37) and np.all(z == 0)
    with pytest.raises(AssertionError, match="beam_px"):   # one antenna: 134 MB of zeros
        vis_gpu_polarized(antpos[:1], 150e6, eq, crd, I, np.zeros((1, 2, 2, 2049, 2049), np.complex64))


@pytest.mark.parametrize("gemm", ["fp32", "split_bf16"])
def test_diagonal_jones_equals_scalar_beam(gemm):
    # J^{pk} = A delta_pk: each feed's block is the scalar visibility (the
    # columns of the zero axis add exact zeros), the cross-feed blocks are 0
    antpos, freqs, eq, crd, I = workload(4000)
    w = configs.make_hex37()
    bm = w.bm_cube
    J = np.zeros((37, 2, 2) + bm.shape[1:], dtype=np.complex64)
    J[:, 0, 0] = bm
    J[:, 1, 1] = bm
    v = vis_gpu_polarized(antpos, freqs, eq, crd, I, J, gemm=gemm)
    s = vis_gpu(antpos, freqs, eq, crd, I, bm, gemm=gemm)
    assert np.all(v[:, 0, 1] == 0) and np.all(v[:, 1, 0] == 0)
    for p in range(2):
        assert rel_err(v[:, p, p], s) < VIS_TOL[gemm]

"""API-level parity of the drop-in ``vis_gpu`` on the GPU.

Mirrors the reference's TestVisGpu (hera_gpu/tests/test_vis_gpu.py:171-276):
shapes, dtypes, closed-form values, comparison with the CPU ``vis_cpu`` and
its phases; then parity with the FP64 oracle on the committed fixtures
(including phases of ~10^3 rad) for every contraction arithmetic, and
invariants that hold at any size.
"""
from pathlib import Path

import numpy as np
import pytest

from hera_gpu_amd import configs, stages, vis_gpu
from oracle.cpu_ref import vis_cpu
from oracle.vis_oracle import rel_err, vis_oracle

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
NANT, NTIMES, BM_PIX, NPIX = 16, 10, 31, 12 * 16 ** 2
# relative Frobenius error of V against the FP64 oracle, per arithmetic
VIS_TOL = {"fp32": 2e-6, "split_bf16": 1e-5, "split_fp16": 2e-6, "fp64": 1e-12}
P1_MODES = ["fp32", "split_bf16", "split_fp16"]


def load(name):
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def run(d, gemm):
    precision = 2 if gemm == "fp64" else 1
    # vis_gpu takes sqrt(I_sky) in the dtype it is given, as vis.py:174 does;
    # for the FP64 path pass float64 intensities so it sees the oracle's values
    I_sky = d["I_sky"].astype(np.float64) if precision == 2 else d["I_sky"]
    return vis_gpu(d["antpos"], d["freqs"], d["eq2tops"], d["crd_eq"], I_sky, d["bm_cube"],
                   precision=precision, gemm=gemm)


def effective_intensity(I_sky, precision):
    """The intensity vis_gpu computes with: (sqrt(I) in I's dtype, cast like vis.py:174)^2."""
    rd = np.float32 if precision == 1 else np.float64
    return np.sqrt(I_sky).astype(rd).astype(np.float64) ** 2


# ------------------------------------------------ reference API tests
def test_shapes():
    antpos = np.zeros((NANT, 3))
    eq2tops = np.zeros((NTIMES, 3, 3))
    crd_eq = np.zeros((3, NPIX))
    I_sky = np.zeros(NPIX)
    bm_cube = np.zeros((NANT, BM_PIX, BM_PIX))
    v = vis_gpu(antpos, 0.15, eq2tops, crd_eq, I_sky, bm_cube)
    assert v.shape == (NTIMES, NANT, NANT)
    for bad in ("antpos", "eq2tops", "crd_eq", "bm_cube"):
        args = dict(antpos=antpos, eq2tops=eq2tops, crd_eq=crd_eq, bm_cube=bm_cube)
        args[bad] = args[bad].T
        with pytest.raises(AssertionError):
            vis_gpu(args["antpos"], 0.15, args["eq2tops"], args["crd_eq"], I_sky, args["bm_cube"])


def test_dtypes():
    z = (np.zeros((NANT, 3)), 0.15, np.zeros((NTIMES, 3, 3)), np.zeros((3, NPIX)), np.zeros(NPIX),
         np.zeros((NANT, BM_PIX, BM_PIX)))
    for precision, cdtype in zip((1, 2), (np.complex64, np.complex128)):
        assert vis_gpu(*z, precision=precision).dtype == cdtype
    assert vis_gpu(*z, real_dtype=np.float64, complex_dtype=np.complex128).dtype == np.complex128
    assert vis_gpu(*z, real_dtype=np.float32, complex_dtype=np.complex64).dtype == np.complex64


@pytest.mark.parametrize("precision", [1, 2])
def test_values(precision):
    antpos = np.ones((NANT, 3))
    eq2tops = np.array([np.identity(3)] * NTIMES)
    crd_eq = np.zeros((3, NPIX))
    crd_eq[2] = 1
    I_sky = np.ones(NPIX)
    bm_cube = np.ones((NANT, BM_PIX, BM_PIX))
    v = vis_gpu(antpos, 1.0, eq2tops, crd_eq, I_sky * 0, bm_cube, precision=precision)
    np.testing.assert_equal(v, 0)
    v = vis_gpu(antpos, 1.0, eq2tops, crd_eq, I_sky, bm_cube * 0, precision=precision)
    np.testing.assert_equal(v, 0)
    v = vis_gpu(antpos, 1.0, eq2tops, crd_eq, I_sky, bm_cube, precision=precision)
    np.testing.assert_almost_equal(v, NPIX, 22)
    crd2 = np.zeros((3, 2))
    crd2[2, 0] = 1
    crd2[1, 1] = np.sqrt(0.5)
    crd2[2, 1] = np.sqrt(0.5)
    I2 = np.ones(2)
    v = vis_gpu(antpos, 1.0, eq2tops, crd2, I2, bm_cube, precision=precision)
    np.testing.assert_almost_equal(v, 2, 2)
    antpos[0, 0] = 0
    antpos[0, 1] = 0
    v_cpu = vis_cpu(antpos, 1.0, eq2tops, crd2, I2, bm_cube, precision=precision)
    v = vis_gpu(antpos, 1.0, eq2tops, crd2, I2, bm_cube, precision=precision)
    np.testing.assert_almost_equal(v, v_cpu, 5 if precision == 1 else 15)


@pytest.mark.parametrize("precision", [1, 2])
def test_compare_cpu_and_phase(precision):
    rng = np.random.default_rng(0)
    for _ in range(3):
        antpos = rng.random((NANT, 3)).astype(np.float32)
        eq2tops = rng.random((NTIMES, 3, 3)).astype(np.float32)
        crd_eq = rng.random((3, NPIX)).astype(np.float32)
        I_sky = rng.random(NPIX).astype(np.float32)
        bm_cube = rng.random((NANT, BM_PIX, BM_PIX)).astype(np.float32)
        freq = float(rng.random())
        v = vis_gpu(antpos, freq, eq2tops, crd_eq, I_sky, bm_cube, precision=precision)
        v_cpu = vis_cpu(antpos, freq, eq2tops, crd_eq, I_sky, bm_cube, precision=precision)
        lim = 1e-4 if precision == 2 else 1e-2
        assert np.median(np.abs(v - v_cpu)) < lim
        assert np.median(np.abs(np.angle(v) - np.angle(v_cpu))) < lim
        # the reference's median bounds are loose; the oracle gives the real bar
        ref = vis_oracle(antpos, freq, eq2tops, crd_eq, effective_intensity(I_sky, precision), bm_cube)[0]
        assert rel_err(v, ref) < (1e-12 if precision == 2 else 2e-6)


# --------------------------------------------------- oracle fixtures
@pytest.mark.parametrize("gemm", P1_MODES + ["fp64"])
@pytest.mark.parametrize("name", ["hex37.npz", "realistic_small.npz"])
def test_fixture_parity(name, gemm):
    d = load(name)
    v = run(d, gemm)
    assert v.shape == d["vis"].shape
    for f in range(v.shape[0]):
        err = rel_err(v[f], d["vis"][f])
        assert err < VIS_TOL[gemm], (name, gemm, f, err)


def test_fixture_split_modes_close_to_fp32():
    d = load("realistic_small.npz")
    a = run(d, "fp32")
    for g in ("split_bf16", "split_fp16"):
        assert rel_err(run(d, g), a) < VIS_TOL[g] + VIS_TOL["fp32"]


# ------------------------------------------------------ invariants
@pytest.mark.parametrize("gemm", ["fp32", "split_bf16", "fp64"])
def test_common_shift_invariance_km_phases(gemm):
    w = configs.make_disk128(nsrc=20000, ntimes=2)
    freqs = np.array([100e6, 200e6])
    # positions on a 1/64 m grid so that both they and the shifted copies are
    # exact in float32 (the precision=1 cast must not change the baselines)
    antpos = (np.round(w.antpos.astype(np.float64) * 64) / 64).astype(np.float32)
    shift = np.array([1024.0, -704.0, 0.0], dtype=np.float32)
    assert np.array_equal((antpos + shift).astype(np.float64) - shift, antpos.astype(np.float64))
    p = 2 if gemm == "fp64" else 1
    a = vis_gpu(antpos, freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube, precision=p, gemm=gemm)
    b = vis_gpu(antpos + shift, freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube, precision=p, gemm=gemm)
    assert rel_err(b, a) < (1e-12 if gemm == "fp64" else 2 * VIS_TOL[gemm])


def test_multifreq_equals_single_freq_calls_bitwise():
    d = load("realistic_small.npz")
    allf = run(d, "split_bf16")
    for f, nu in enumerate(d["freqs"]):
        one = vis_gpu(d["antpos"], nu, d["eq2tops"], d["crd_eq"], d["I_sky"], d["bm_cube"], gemm="split_bf16")
        np.testing.assert_array_equal(one, allf[f])


def test_many_splits_and_tiles_vs_oracle_units():
    # K spans several 8192-source splits; N spans several 64-antenna tiles
    w = configs.make_disk128(nsrc=30000, ntimes=3)
    freqs = np.array([120e6, 180e6])
    ref = vis_oracle(w.antpos, freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube)
    for gemm in P1_MODES:
        v = vis_gpu(w.antpos, freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube, gemm=gemm)
        assert rel_err(v, ref) < VIS_TOL[gemm], gemm


def test_resident_simulate_matches_drop_in_bitwise():
    w = configs.make_hex37()
    sky = stages.Sky(w.crd_eq, w.I_sky)
    beam = stages.Beam(w.bm_cube)
    for gemm in P1_MODES:
        a, _ = stages.simulate(sky, beam, w.antpos, w.freqs, w.eq2tops, gemm=gemm, dtype=np.float32)
        b = vis_gpu(w.antpos, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube, gemm=gemm)
        np.testing.assert_array_equal(a, b)
    # a truncated catalog equals a call on the first nsrc sources
    a, t = stages.simulate(sky, beam, w.antpos, w.freqs, w.eq2tops, nsrc=333, timing=True, dtype=np.float32)
    b = vis_gpu(w.antpos, w.freqs, w.eq2tops, w.crd_eq[:, :333], w.I_sky[:333], w.bm_cube)
    np.testing.assert_array_equal(a, b)
    assert t["herk_n"] >= 1 and t["total_ms"] > 0


@pytest.mark.parametrize("gemm", ["split_bf16", "fp32"])
def test_batch_schedules_agree_bitwise(gemm, monkeypatch):
    # a small HBM budget splits the job into several batches; the one-stream
    # and the two-stream (HERAVIS_OVERLAP=1) schedules must give the same bits
    from hera_gpu_amd import _lib
    w = configs.make_disk128(nsrc=40000, ntimes=4)
    freqs = np.linspace(110e6, 190e6, 6)
    lib = _lib.load()
    ctx = _lib.Context(0, 160 << 20)
    try:
        outs = []
        for ov in ("0", "1"):
            monkeypatch.setenv("HERAVIS_OVERLAP", ov)
            out = np.empty((freqs.size, 4, 128, 128), dtype=np.complex64)
            isq = np.sqrt(w.I_sky).astype(np.float32)
            _lib.check(lib.hv_vis(ctx.handle, _lib.MODES[gemm], 0, _lib.ptr(w.antpos), 128, _lib.ptr(freqs),
                                  freqs.size, _lib.ptr(w.eq2tops), 4, _lib.ptr(w.crd_eq), _lib.ptr(isq),
                                  w.crd_eq.shape[1], _lib.ptr(w.bm_cube), w.bm_cube.shape[1], _lib.ptr(out)),
                       "hv_vis")
            outs.append(out)
    finally:
        ctx.close()
    np.testing.assert_array_equal(outs[0], outs[1])
    ref = vis_oracle(w.antpos, freqs[:2], w.eq2tops[:1], w.crd_eq, w.I_sky, w.bm_cube)
    assert rel_err(outs[0][:2, :1], ref) < VIS_TOL[gemm]


@pytest.mark.parametrize("gemm", ["split_bf16", "split_fp16"])
@pytest.mark.parametrize("tile,nant", [(64, 300), (128, 300), (192, 300), (192, 160), (192, 129), (192, 400), (192, 97), (192, 128), (192, 230), (128, 128), (128, 100), (128, 70), (192, 350), (192, 321), (192, 352), (192, 520)])
def test_operands_built_in_contraction_equal_phasor_build_bitwise(gemm, tile, nant):
    # pf 5 builds the operand stages inside the contraction from A sqrt(I);
    # pf 4 contracts the stages phasor_kernel wrote: same operand bits, same
    # MFMA sequence per 32x32 block, so the same visibilities bit for bit
    # (192-tiles exist only in place: they are checked against the 128-tile
    # ring, whose blocks run the same chains).  N = 300 leaves a partial
    # last tile, N = 160 and 129 one diagonal 192-tile with five block rows
    # (the default choice there); K spans several 8192-source splits; 3 times
    # x 4 frequencies
    from hera_gpu_amd import _lib
    lib = _lib.load()
    w = configs.make_hera350(nsrc=3 * 8192 + 1000, ntimes=3, nfreq=4)
    ap, bm = w.antpos, w.bm_cube
    if nant > len(ap):
        # N = 400: three 192-blocks, so full off-diagonal 192-tiles too; the
        # extra antennas are copies of the first ones 500 m east
        extra = nant - len(ap)
        ap = np.concatenate([ap, ap[:extra] + np.array([500.0, 0.0, 0.0], dtype=ap.dtype)])
        bm = np.concatenate([bm, bm[:extra]])
    ap, bm = ap[:nant], bm[:nant]
    outs = []
    try:
        for pf in (4, 5):
            _lib.check(lib.hv_set_herk_variant(128 if (tile == 192 and pf == 4) else tile, pf))
            outs.append(vis_gpu(ap, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, bm, gemm=gemm))
    finally:
        _lib.check(lib.hv_set_herk_variant(0, 0))
    np.testing.assert_array_equal(outs[0], outs[1])
    ref = vis_oracle(ap, w.freqs[-1:], w.eq2tops[:1], w.crd_eq, w.I_sky, bm)[0, 0]
    assert rel_err(outs[1][-1, 0], ref) < VIS_TOL[gemm]


@pytest.mark.parametrize("gemm", ["split_bf16", "split_fp16"])
@pytest.mark.parametrize("nant", [129, 160, 161, 192, 200, 230, 256, 257, 288, 289, 300, 320, 321, 350, 352])
def test_triangle_strip_items_equal_tile_pairs_bitwise(gemm, nant):
    # arrays of 129..352 antennas (5..11 blocks of 32 rows) run the items of
    # the triangle decomposition: one item holding every block up to eight
    # blocks, else a triangle item (the first 6, 7 or 8 block columns) and a
    # strip item (the last three); pf 6 runs the same array as 192-tiles
    # (the paired diagonal tiles and the off-diagonal tile).  Every block
    # runs the same MFMA chain on the same operand bits, so the visibilities
    # are equal bit for bit; the sizes cover each block count, one antenna
    # past a block edge (161, 257, 289, 321) and full last blocks (160, 192,
    # 256, 288, 320, 352).  K spans several 8192-source splits; 2 times x 3
    # frequencies.
    from hera_gpu_amd import _lib
    lib = _lib.load()
    w = configs.make_hera350(nsrc=2 * 8192 + 700, ntimes=2, nfreq=3)
    ap, bm = w.antpos, w.bm_cube
    if nant > len(ap):
        extra = nant - len(ap)
        ap = np.concatenate([ap, ap[:extra] + np.array([0.0, 400.0, 0.0], dtype=ap.dtype)])
        bm = np.concatenate([bm, bm[:extra]])
    ap, bm = ap[:nant], bm[:nant]
    outs = []
    try:
        for pf in (5, 6):
            _lib.check(lib.hv_set_herk_variant(192, pf))
            outs.append(vis_gpu(ap, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, bm, gemm=gemm))
    finally:
        _lib.check(lib.hv_set_herk_variant(0, 0))
    np.testing.assert_array_equal(outs[0], outs[1])
    ref = vis_oracle(ap, w.freqs[:1], w.eq2tops[-1:], w.crd_eq, w.I_sky, bm)[0, 0]
    assert rel_err(outs[0][0, -1], ref) < VIS_TOL[gemm]


@pytest.mark.parametrize("gemm", ["split_bf16", "split_fp16"])
@pytest.mark.parametrize("nant,nfreq", [(33, 7), (37, 13), (64, 13), (65, 5), (96, 9), (97, 3), (128, 7), (128, 4)])
def test_multi_unit_items_equal_tile_kernels_bitwise(gemm, nant, nfreq):
    # arrays of 33..128 antennas run items holding several frequency units
    # of one time (6, 4 or 3 units at 2, 3 or 4 blocks of 32 rows) that share
    # the time's amplitude rows.  Every block runs the MFMA chain of the
    # 128-tile LDS ring on the same operand bits (tile 128, pf 4: one unit
    # per work item, operands written by phasor_kernel), so the visibilities
    # are equal bit for bit (the 64-tile kernels fold their accumulators
    # every 512 sources and differ in the last bits).  The frequency counts
    # leave the last group of each time short; K spans several 8192-source
    # splits; 2 times.
    from hera_gpu_amd import _lib
    lib = _lib.load()
    w = configs.make_hera350(nsrc=2 * 8192 + 700, ntimes=2, nfreq=nfreq)
    ap, bm = w.antpos[:nant], w.bm_cube[:nant]
    outs = []
    try:
        for tile, pf in ((0, 0), (128, 4)):
            _lib.check(lib.hv_set_herk_variant(tile, pf))
            outs.append(vis_gpu(ap, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, bm, gemm=gemm))
    finally:
        _lib.check(lib.hv_set_herk_variant(0, 0))
    np.testing.assert_array_equal(outs[0], outs[1])
    ref = vis_oracle(ap, w.freqs[-1:], w.eq2tops[-1:], w.crd_eq, w.I_sky, bm)[0, 0]
    assert rel_err(outs[0][-1, -1], ref) < VIS_TOL[gemm]


@pytest.mark.parametrize("gemm", ["split_bf16", "split_fp16"])
@pytest.mark.parametrize("nant,nfreq,tile", [(37, 1, 64), (37, 5, 64), (64, 5, 64), (65, 3, 128), (100, 2, 128),
                                             (128, 1, 128)])
def test_few_frequency_jobs_take_tile_kernels_bitwise(gemm, nant, nfreq, tile):
    # an item of several units computes all its 6, 4 or 3 unit slots, so a
    # job with fewer frequencies per time than that runs one unit per item
    # (the 64- or 128-tile in-place kernels): the bits of the LDS ring of
    # that tile (pf 4), and at N <= 64 not those of the 128-tile chains the
    # multi-unit items share (the 64-tiles fold every 512 sources)
    from hera_gpu_amd import _lib
    lib = _lib.load()
    w = configs.make_hera350(nsrc=2 * 8192 + 700, ntimes=2, nfreq=nfreq)
    ap, bm = w.antpos[:nant], w.bm_cube[:nant]
    outs = []
    try:
        for t, pf in ((0, 0), (tile, 4), (128, 4)):
            _lib.check(lib.hv_set_herk_variant(t, pf))
            outs.append(vis_gpu(ap, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, bm, gemm=gemm))
    finally:
        _lib.check(lib.hv_set_herk_variant(0, 0))
    np.testing.assert_array_equal(outs[0], outs[1])
    if tile == 64:
        assert not np.array_equal(outs[0], outs[2])
    ref = vis_oracle(ap, w.freqs[-1:], w.eq2tops[-1:], w.crd_eq, w.I_sky, bm)[0, 0]
    assert rel_err(outs[0][-1, -1], ref) < VIS_TOL[gemm]


LONG_K = 5 * 32768 + 700   # six 32768-source splits, the last one short


@pytest.fixture
def long_splits():
    # the K-split rule's work-item threshold lowered from 32768 to 2048, so
    # jobs of test size take 32768-source splits as the large ones do
    from hera_gpu_amd import _lib
    lib = _lib.load()
    _lib.check(lib.hv_set_ksplit_items(2048))
    try:
        yield
    finally:
        _lib.check(lib.hv_set_ksplit_items(0))


@pytest.mark.parametrize("gemm", ["split_bf16", "split_fp16"])
@pytest.mark.parametrize("nant,ntimes,variants", [(37, 32, ((0, 0), (128, 4))), (128, 16, ((0, 0), (128, 4))),
                                                  (300, 3, ((192, 5), (192, 6))), (350, 3, ((192, 5), (192, 6)))])
def test_long_splits_equal_across_kernels_bitwise(gemm, nant, ntimes, variants, long_splits):
    # a job with 32768-source K-splits (one FP32 chain per split and block):
    # the built-in items (multi-unit items at N = 37, 128; triangle/strip
    # items at N = 300, 350) equal the 128-tile LDS ring / the 192-tile
    # pairs bit for bit, as with 8192-source splits, and match the oracle.
    # 64 frequencies.
    from hera_gpu_amd import _lib
    lib = _lib.load()
    assert stages.ksplit_choice(gemm, nant, ntimes * 64, LONG_K) == 32768
    w = configs.make_hera350(nsrc=LONG_K, ntimes=ntimes, nfreq=64)
    ap, bm = w.antpos[:nant], w.bm_cube[:nant]
    outs = []
    try:
        for tile, pf in variants:
            _lib.check(lib.hv_set_herk_variant(tile, pf))
            outs.append(vis_gpu(ap, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, bm, gemm=gemm))
    finally:
        _lib.check(lib.hv_set_herk_variant(0, 0))
    np.testing.assert_array_equal(outs[0], outs[1])
    ref = vis_oracle(ap, w.freqs[::63], w.eq2tops[-1:], w.crd_eq, w.I_sky, bm)
    assert rel_err(outs[0][::63, -1:], ref) < VIS_TOL[gemm]


def test_long_split_job_batches_agree_bitwise(long_splits):
    # the split length follows the job, not the launch: a job of 16 x 64
    # units cut into one launch per time by a small HBM budget gives the
    # bits of the job run in one launch
    from hera_gpu_amd import _lib
    assert stages.ksplit_choice("split_bf16", 128, 1024, LONG_K) == 32768
    w = configs.make_disk128(nsrc=LONG_K, ntimes=16)
    freqs = np.linspace(110e6, 190e6, 64)
    lib = _lib.load()
    isq = np.sqrt(w.I_sky).astype(np.float32)
    outs = []
    for budget in (0, 600 << 20):
        ctx = _lib.Context(0, budget)
        try:
            out = np.empty((freqs.size, 16, 128, 128), dtype=np.complex64)
            _lib.check(lib.hv_vis(ctx.handle, _lib.MODES["split_bf16"], 0, _lib.ptr(w.antpos), 128, _lib.ptr(freqs),
                                  freqs.size, _lib.ptr(w.eq2tops), 16, _lib.ptr(w.crd_eq), _lib.ptr(isq),
                                  w.crd_eq.shape[1], _lib.ptr(w.bm_cube), w.bm_cube.shape[1], _lib.ptr(out)),
                       "hv_vis")
            outs.append(out)
        finally:
            ctx.close()
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("gemm", ["fp32", "split_bf16", "fp64"])
def test_baseline_layouts_equal_full_matrix_bitwise(gemm):
    # the reduce writes only a <= b, in (T, Nbl, F) or (F, T, Nbl) order; the
    # values are those of the full (F, T, N, N) output bit for bit
    from hera_gpu_amd import vis_gpu_baselines
    w = configs.make_disk128(nsrc=20000, ntimes=3)
    freqs = np.array([110e6, 150e6, 190e6])
    p = 2 if gemm == "fp64" else 1
    I = w.I_sky.astype(np.float64) if p == 2 else w.I_sky
    full = vis_gpu(w.antpos, freqs, w.eq2tops, w.crd_eq, I, w.bm_cube, precision=p, gemm=gemm)
    tbf, a1, a2 = vis_gpu_baselines(w.antpos, freqs, w.eq2tops, w.crd_eq, I, w.bm_cube, precision=p, gemm=gemm)
    ftb, b1, b2 = vis_gpu_baselines(w.antpos, freqs, w.eq2tops, w.crd_eq, I, w.bm_cube, precision=p, gemm=gemm,
                                    order="ftb")
    nbl = 128 * 129 // 2
    assert tbf.shape == (3, nbl, 3) and ftb.shape == (3, 3, nbl) and a1.shape == (nbl,)
    assert np.array_equal(a1, b1) and np.array_equal(a2, b2) and np.all(a1 <= a2)
    np.testing.assert_array_equal(ftb, full[:, :, a1, a2])
    np.testing.assert_array_equal(tbf, full[:, :, a1, a2].transpose(1, 2, 0))


def _freq_beams(nant, nfreq, P, seed):
    """Gaussian beams whose width scales as 1/nu: one (nant, P, P) plane per frequency."""
    rng = np.random.default_rng(seed)
    g = np.linspace(-1, 1, P)
    r2 = g[None, :] ** 2 + g[:, None] ** 2
    w0 = rng.uniform(0.3, 0.5, nant)
    return np.stack([np.exp(-r2[None] / (2 * (w0[:, None, None] * (1.0 - 0.2 * f)) ** 2))
                     for f in range(nfreq)]).astype(np.float32)


@pytest.mark.parametrize("gemm", P1_MODES + ["fp64"])
def test_per_frequency_beams_vs_oracle(gemm):
    # one beam plane per frequency (F, N, P, P): each frequency interpolates
    # its own plane; N = 128 and K over several splits
    w = configs.make_disk128(nsrc=20000, ntimes=2)
    freqs = np.array([110e6, 150e6, 190e6])
    bm4 = _freq_beams(128, 3, 31, 5)
    p = 2 if gemm == "fp64" else 1
    I = w.I_sky.astype(np.float64) if p == 2 else w.I_sky
    bm = bm4.astype(np.float64) if p == 2 else bm4
    v = vis_gpu(w.antpos, freqs, w.eq2tops, w.crd_eq, I, bm, precision=p, gemm=gemm)
    ref = vis_oracle(w.antpos, freqs, w.eq2tops, w.crd_eq, effective_intensity(I, p), bm4)
    assert rel_err(v, ref) < VIS_TOL[gemm]
    # and each frequency's plane is the one used: plane 0 for every frequency is far off
    wrong = vis_oracle(w.antpos, freqs, w.eq2tops, w.crd_eq, effective_intensity(I, p), bm4[[0, 0, 0]])
    assert rel_err(v, wrong) > 1e-2


@pytest.mark.parametrize("gemm", ["fp32", "split_bf16", "fp64"])
@pytest.mark.parametrize("pf", [0, 4])
def test_identical_frequency_planes_equal_one_beam_bitwise(gemm, pf):
    # F copies of the one beam give the one-beam result bit for bit, through
    # the built-in contraction (pf 0: operands built in place for split modes
    # at N > 64) and through the phasor build (pf 4); N = 300 leaves a partial
    # last tile
    from hera_gpu_amd import _lib
    if pf == 4 and gemm == "fp64":
        pytest.skip("pf 4 is a split / FP32 variant")
    lib = _lib.load()
    w = configs.make_hera350(nsrc=2 * 8192 + 500, ntimes=2, nfreq=3)
    ap, bm = w.antpos[:300], w.bm_cube[:300]
    p = 2 if gemm == "fp64" else 1
    if p == 2:
        bm = bm.astype(np.float64)
    bm4 = np.stack([bm] * 3)
    try:
        _lib.check(lib.hv_set_herk_variant(128 if pf else 0, pf))
        one = vis_gpu(ap, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, bm, precision=p, gemm=gemm)
        per = vis_gpu(ap, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, bm4, precision=p, gemm=gemm)
    finally:
        _lib.check(lib.hv_set_herk_variant(0, 0))
    np.testing.assert_array_equal(per, one)


def test_per_frequency_beam_resident_and_checks():
    w = configs.make_hex37()
    freqs = np.array([120e6, 160e6])
    bm4 = _freq_beams(37, 2, w.bm_cube.shape[-1], 6)
    sky = stages.Sky(w.crd_eq, w.I_sky)
    beam = stages.Beam(bm4)
    try:
        a, _ = stages.simulate(sky, beam, w.antpos, freqs, w.eq2tops, gemm="split_bf16", dtype=np.float32)
        b = vis_gpu(w.antpos, freqs, w.eq2tops, w.crd_eq, w.I_sky, bm4, gemm="split_bf16")
        np.testing.assert_array_equal(a, b)
        # a per-frequency beam needs exactly one plane per frequency
        with pytest.raises(AssertionError):
            stages.simulate(sky, beam, w.antpos, freqs[:1], w.eq2tops, gemm="split_bf16", dtype=np.float32)
    finally:
        beam.close()
        sky.close()
    with pytest.raises(AssertionError):
        vis_gpu(w.antpos, freqs[:1], w.eq2tops, w.crd_eq, w.I_sky, bm4)


def test_diffuse_sky_config_vs_oracle():
    # the diffuse-sky configuration (equal-area ring pixels, exponentiated
    # Gaussian random field, HERA-350, frequencies across 100-200 MHz) at
    # nside 16 (3072 pixels); every frequency of a time goes in one
    # contraction launch (fewer launches than times when times batch too)
    w = configs.make_diffuse(nside=16, ntimes=2, nfreq=4, lmax=24)
    ref = vis_oracle(w.antpos, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube)
    for gemm in ("split_bf16", "fp32"):
        v = vis_gpu(w.antpos, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube, gemm=gemm)
        errs = [rel_err(v[f], ref[f]) for f in range(len(w.freqs))]
        assert max(errs) < VIS_TOL[gemm], (gemm, errs)
    sky = stages.Sky(w.crd_eq, w.I_sky)
    beam = stages.Beam(w.bm_cube)
    try:
        a, st = stages.simulate(sky, beam, w.antpos, w.freqs, w.eq2tops, gemm="split_bf16", timing=True,
                                dtype=np.float32)
    finally:
        beam.close()
        sky.close()
    assert 1 <= st["herk_n"] <= len(w.eq2tops)
    np.testing.assert_array_equal(a, vis_gpu(w.antpos, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube,
                                             gemm="split_bf16"))


@pytest.mark.parametrize("gemm", P1_MODES + ["fp64"])
def test_empty_and_unseen_skies_give_zero(gemm):
    # no sources, and sources only at or below the horizon (the horizon test
    # is strict z > 0, vis.py:75): every visibility is exactly zero
    p = 2 if gemm == "fp64" else 1
    antpos = np.random.default_rng(1).uniform(-50, 50, (NANT, 3))
    eq2tops = np.array([np.identity(3)] * 3)
    bm_cube = np.ones((NANT, BM_PIX, BM_PIX))
    v = vis_gpu(antpos, 1.5e8, eq2tops, np.zeros((3, 0)), np.zeros(0), bm_cube, precision=p, gemm=gemm)
    assert v.shape == (3, NANT, NANT)
    np.testing.assert_array_equal(v, 0)
    crd = np.zeros((3, 100))
    crd[0] = np.linspace(-1, 1, 100)
    crd[2, 50:] = -np.sqrt(1 - crd[0, 50:] ** 2)       # half on the horizon (z = 0), half below
    v = vis_gpu(antpos, 1.5e8, eq2tops, crd, np.ones(100), bm_cube, precision=p, gemm=gemm)
    np.testing.assert_array_equal(v, 0)


def test_amplitude_pass_chunking_bitwise():
    # 16 times x 128 antennas x 40000 sources in one launch: the beam
    # amplitude pass gives each workgroup 32768 sources (the launch keeps
    # >= 1024 workgroups); one time alone gives it 8192.  The amplitudes,
    # and so the visibilities, are the same bit for bit.
    w = configs.make_disk128(nsrc=40000, ntimes=16)
    freqs = np.array([150e6])
    allt = vis_gpu(w.antpos, freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube, gemm="split_bf16")
    for t in (0, 7, 15):
        one = vis_gpu(w.antpos, freqs, w.eq2tops[t:t + 1], w.crd_eq, w.I_sky, w.bm_cube, gemm="split_bf16")
        np.testing.assert_array_equal(one[:, 0], allt[:, t])


@pytest.mark.parametrize("gemm", P1_MODES + ["fp64"])
def test_single_antenna_and_negative_intensity(gemm):
    # N = 1 (autos only) against the oracle; a negative intensity gives NaN
    # everywhere, as the reference's host-side sqrt does (vis.py:174), also
    # through sources below the horizon (0 * NaN)
    p = 2 if gemm == "fp64" else 1
    d = load("realistic_small.npz")
    I_sky = d["I_sky"].astype(np.float64) if p == 2 else d["I_sky"]
    ap, bm = d["antpos"][:1], d["bm_cube"][:1]
    v = vis_gpu(ap, d["freqs"], d["eq2tops"], d["crd_eq"], I_sky, bm, precision=p, gemm=gemm)
    ref = vis_oracle(ap, d["freqs"], d["eq2tops"], d["crd_eq"], d["I_sky"], bm)
    assert v.shape == ref.shape and rel_err(v, ref) < VIS_TOL[gemm]
    assert np.all(v.imag == 0) and np.all(v.real > 0)
    bad = I_sky.copy()
    bad[3] = -1.0
    with pytest.warns(RuntimeWarning):           # numpy's sqrt of a negative, as in the reference
        v = vis_gpu(d["antpos"], d["freqs"], d["eq2tops"], d["crd_eq"], bad, d["bm_cube"], precision=p, gemm=gemm)
    assert np.isnan(v).all()

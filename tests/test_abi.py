"""The C-ABI library loads and exports every symbol include/heravis.h declares.

No compute calls here: this runs without a GPU.
"""
import re
from pathlib import Path

import numpy as np
import pytest

from hera_gpu_amd import _lib, vis_gpu

HEADER = Path(__file__).resolve().parents[1] / "include" / "heravis.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hv_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_surface():
    names = declared_functions()
    for n in ("hv_vis", "hv_k_rotate", "hv_k_beam", "hv_k_phasor", "hv_k_herk", "hv_simulate",
              "hv_ctx_create", "hv_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_version_and_device_count_without_compute():
    lib = _lib.load()
    assert lib.hv_version() >= 100
    assert _lib.device_count() >= 0


def test_null_context_is_an_argument_error():
    lib = _lib.load()
    rc = lib.hv_ctx_destroy(None)
    assert rc == _lib.HV_OK
    out = np.zeros(4, dtype=np.complex64)
    rc = lib.hv_vis(None, 0, 0, None, 1, None, 1, None, 1, None, None, 1, None, 1, _lib.ptr(out))
    assert rc == _lib.HV_ERR_ARG
    assert b"ctx" in lib.hv_last_error()


# Shape checks of vis.py:161-169 fire before any device work (test_vis_gpu.py:172-191).
NANT, NTIMES, BM_PIX, NPIX = 16, 10, 31, 3072


@pytest.mark.parametrize("which", ["antpos", "eq2tops", "crd_eq", "bm_cube"])
def test_transposed_inputs_raise_assertion(which):
    args = dict(antpos=np.zeros((NANT, 3)), eq2tops=np.zeros((NTIMES, 3, 3)), crd_eq=np.zeros((3, NPIX)),
                I_sky=np.zeros(NPIX), bm_cube=np.zeros((NANT, BM_PIX, BM_PIX)))
    args[which] = args[which].T
    with pytest.raises(AssertionError):
        vis_gpu(args["antpos"], 0.15, args["eq2tops"], args["crd_eq"], args["I_sky"], args["bm_cube"])


def test_bad_precision_raises_assertion():
    with pytest.raises(AssertionError):
        vis_gpu(np.zeros((2, 3)), 0.15, np.zeros((1, 3, 3)), np.zeros((3, 4)), np.zeros(4),
                np.zeros((2, 3, 3)), precision=3)
    with pytest.raises(AssertionError):
        vis_gpu(np.zeros((2, 3)), 0.15, np.zeros((1, 3, 3)), np.zeros((3, 4)), np.zeros(4),
                np.zeros((2, 3, 3)), precision=2, gemm="split_bf16")


@pytest.mark.parametrize("nant,tile", [(1, 64), (37, 64), (64, 64), (65, 128), (128, 128), (129, 192), (160, 192),
                                       (192, 192), (193, 192), (256, 192), (257, 192), (350, 192), (520, 192)])
def test_contraction_tile_choice(nant, tile):
    # host-side choice, no GPU call: split modes store their partials in
    # 64-tiles up to 64 antennas, 128-tiles up to 128 (items of several
    # frequency units), 192-tiles above (the triangle decomposition up to
    # 352 antennas, 192-tiles beyond); FP32 keeps 64-tiles
    import ctypes as C
    lib = _lib.load()
    t, p = C.c_int(), C.c_int()
    assert lib.hv_herk_choice(_lib.MODES["split_bf16"], nant, C.byref(t), C.byref(p)) == 0
    assert (t.value, p.value) == (tile, 5)
    assert lib.hv_herk_choice(_lib.MODES["fp32"], nant, C.byref(t), C.byref(p)) == 0
    assert t.value == 64


@pytest.mark.parametrize("gemm,nant,units,nsrc,ksplit", [
    ("split_bf16", 350, 2048, 1 << 20, 32768),   # the HERA-350 job: 2 items per (unit, split)
    ("split_bf16", 350, 512, 1 << 20, 32768),    # 512 x 32 splits x 2 = 32768 items: just enough
    ("split_bf16", 350, 511, 1 << 20, 8192),
    ("split_bf16", 350, 128, 1 << 20, 8192),     # two times of it, as a job of their own
    ("split_bf16", 350, 1, 1 << 20, 8192),       # a lone unit keeps 128 splits x 2 items
    ("split_bf16", 256, 1024, 1 << 20, 32768),   # one triangle item per (unit, split)
    ("split_bf16", 256, 1023, 1 << 20, 8192),
    ("split_fp16", 37, 6144, 1 << 20, 32768),    # items of 6 units
    ("split_fp16", 37, 6143, 1 << 20, 8192),
    ("split_bf16", 128, 3072, 1 << 20, 32768),   # items of 3 units
    ("split_bf16", 520, 171, 1 << 20, 32768),    # 192-tiles: 6 per unit
    ("split_bf16", 520, 170, 1 << 20, 8192),
    ("fp32", 350, 2048, 1 << 20, 8192),          # FP32 and FP64 always keep short splits
    ("fp64", 128, 100000, 1 << 20, 8192),
    ("split_bf16", 37, 10 ** 6, 0, 32768),       # no sources: one split either way
])
def test_ksplit_rule(gemm, nant, units, nsrc, ksplit):
    # host-side rule, no GPU call: 32768-source K-splits where a split-mode
    # job still holds 32768 contraction work items with them, else 8192
    from hera_gpu_amd import stages
    assert stages.ksplit_choice(gemm, nant, units, nsrc) == ksplit


def test_ksplit_threshold_knob():
    # the rule's work-item threshold is process-wide (tests and tuning); 0
    # restores 32768
    from hera_gpu_amd import stages
    lib = _lib.load()
    assert stages.ksplit_choice("split_bf16", 350, 32, 1 << 20) == 8192
    try:
        assert lib.hv_set_ksplit_items(2048) == 0
        assert stages.ksplit_choice("split_bf16", 350, 32, 1 << 20) == 32768    # 2048 items
        assert stages.ksplit_choice("split_bf16", 350, 31, 1 << 20) == 8192
        assert stages.ksplit_choice("fp32", 350, 2048, 1 << 20) == 8192
    finally:
        assert lib.hv_set_ksplit_items(0) == 0
    assert stages.ksplit_choice("split_bf16", 350, 32, 1 << 20) == 8192
    assert lib.hv_set_ksplit_items(-1) != 0

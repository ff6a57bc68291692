"""hv_vis_device_noalloc (device arrays in and out, no allocation) against
hv_vis (host arrays) on the same values: bitwise equal, on two calls in a row
with the second smaller, for split-BF16 and FP32; device memory unchanged by
the calls; a problem beyond the scratch refused with HV_ERR_NOMEM."""
import ctypes as C

import numpy as np
import pytest

from hera_gpu_amd import _lib, configs, vis_gpu

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to("cuda:0")


def _noalloc(lib, ctx, mode, w, nsrc):
    """Inputs as device float64 (the float32 values of the workload), beam float32."""
    N, T, F = w.antpos.shape[0], w.eq2tops.shape[0], len(w.freqs)
    ap = _dev(w.antpos, np.float64)
    fr = _dev(w.freqs, np.float64)
    eq = _dev(w.eq2tops, np.float64)
    crd = _dev(w.crd_eq[:, :nsrc], np.float64)
    isq = _dev(np.sqrt(w.I_sky[:nsrc].astype(np.float32)), np.float64)
    bm = _dev(w.bm_cube, np.float32)
    out = torch.empty((F, T, N, N), dtype=torch.complex64, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(0)[0]
    rc = lib.hv_vis_device_noalloc(ctx.handle, mode, ap.data_ptr(), N, fr.data_ptr(), F, eq.data_ptr(), T,
                                   crd.data_ptr(), isq.data_ptr(), nsrc, bm.data_ptr(), w.bm_cube.shape[-1],
                                   out.data_ptr(), C.c_void_p(stream))
    _lib.check(rc, "hv_vis_device_noalloc")
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info(0)[0]
    return out.cpu().numpy(), free0 - free1


@pytest.mark.parametrize("gemm", ["split_bf16", "fp32"])
def test_noalloc_equals_hv_vis_twice(gemm):
    lib = _lib.load()
    ctx = _lib.context(0)
    mode = _lib.MODES[gemm]
    big = configs.make_hera350(nsrc=24576, ntimes=3, nfreq=4)          # 350 antennas, 192-tiles
    small = configs.make_disk128(nsrc=4096, ntimes=2)                  # 128 antennas, smaller
    small.freqs = np.array([110e6, 190e6])
    # warm-up of each case: the first call on the device resolves the scratch
    # symbol, and the first dispatch of a kernel that spills registers makes
    # the HIP runtime size its per-queue private-segment memory (not a buffer
    # of the library); the measured calls follow, the second smaller
    for w, nsrc in ((big, 24576), (small, 4096)):
        _noalloc(lib, ctx, mode, w, nsrc)
    for w, nsrc in ((big, 24576), (small, 4096)):
        got, grew = _noalloc(lib, ctx, mode, w, nsrc)
        ref = vis_gpu(w.antpos, w.freqs, w.eq2tops, w.crd_eq[:, :nsrc], w.I_sky[:nsrc], w.bm_cube, gemm=gemm)
        assert got.shape == ref.shape
        np.testing.assert_array_equal(got, ref, err_msg=f"{gemm} N={w.antpos.shape[0]} K={nsrc}")
        assert grew == 0, f"device memory fell by {grew} bytes during the call"


def test_noalloc_refuses_beyond_scratch_and_bad_args():
    lib = _lib.load()
    ctx = _lib.context(0)
    w = configs.make_hera350(nsrc=1 << 19, ntimes=1, nfreq=1)   # A sqrt(I) alone is 733 MB
    with pytest.raises(MemoryError, match="scratch"):
        _noalloc(lib, ctx, _lib.MODES["split_bf16"], w, 1 << 19)
    with pytest.raises(AssertionError, match="multiple of 64"):
        _noalloc(lib, ctx, _lib.MODES["split_bf16"], w, 1000)
    with pytest.raises(AssertionError, match="HV_MODE"):
        _noalloc(lib, ctx, _lib.MODES["split_fp16"], w, 1024)


def test_noalloc_long_split_job_in_time_batches():
    # 16 times x 64 frequencies of 128 antennas over 164544 sources need more
    # than the scratch in one batch, so the call runs one time per batch; the
    # K-splits follow the whole job (32768 sources: 2048 work items, with the
    # rule's threshold lowered to that), so the results equal hv_vis bit for
    # bit
    from hera_gpu_amd import stages
    lib = _lib.load()
    ctx = _lib.context(0)
    nsrc = 164544
    _lib.check(lib.hv_set_ksplit_items(2048))
    try:
        assert stages.ksplit_choice("split_bf16", 128, 1024, nsrc) == 32768
        w = configs.make_disk128(nsrc=nsrc, ntimes=16)
        w.freqs = np.linspace(110e6, 190e6, 64)
        mode = _lib.MODES["split_bf16"]
        _noalloc(lib, ctx, mode, w, nsrc)
        got, grew = _noalloc(lib, ctx, mode, w, nsrc)
        ref = vis_gpu(w.antpos, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube, gemm="split_bf16")
    finally:
        _lib.check(lib.hv_set_ksplit_items(0))
    np.testing.assert_array_equal(got, ref)
    assert grew == 0

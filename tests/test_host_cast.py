"""Host-side conversion of large caller arrays (no GPU needed)."""
import numpy as np

from hera_gpu_amd import _lib


def test_parallel_cast_equals_numpy():
    rng = np.random.default_rng(3)
    data = rng.standard_normal((1001, 2100)) + 1j * rng.standard_normal((1001, 2100))   # 33.6 MB
    for src, dtype in ((data, np.complex64), (data[:, ::2], np.complex64), (data.real, np.float32),
                       (data.real.T, np.float32)):
        out = _lib.contiguous(src, dtype)
        ref = np.ascontiguousarray(src, dtype=dtype)
        assert out.dtype == ref.dtype and out.shape == ref.shape and out.flags.c_contiguous
        np.testing.assert_array_equal(out, ref)


def test_cast_passes_through_and_handles_small_arrays():
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    assert _lib.contiguous(a, np.float32) is a
    np.testing.assert_array_equal(_lib.contiguous(a[:, ::2], np.float64), a[:, ::2].astype(np.float64))
    np.testing.assert_array_equal(_lib.contiguous([[1, 2], [3, 4]], np.complex64), np.array([[1, 2], [3, 4]]))
    assert _lib.contiguous(np.float64(2.5), np.float32).shape == np.ascontiguousarray(2.5, np.float32).shape

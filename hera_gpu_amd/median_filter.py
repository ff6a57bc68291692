"""2-D median filter on the GPU: ``MedianFilter`` with the arguments and
return of ``hera_gpu.median_filter.MedianFilter`` (hera_gpu/median_filter.py:8-365).

Every pixel becomes the element of rank (rows*cols)//2 of its rows x cols
window, zero outside the image (scipy.signal.medfilt2d for odd windows).
All images of a batch go to the device in one copy and are filtered by one
launch (see hv_medfilt2d in include/heravis.h).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _window(kernel_size) -> tuple[int, int]:
    # median_filter.py:19-22: an int, or (rows, cols)
    if isinstance(kernel_size, (int, np.integer)):
        return int(kernel_size), int(kernel_size)
    wr, wc = kernel_size
    return int(wr), int(wc)


def medfilt2d(images, kernel_size=3, *, device=0) -> np.ndarray:
    """Filter a (N, M) image or an (B, N, M) batch; returns float32 of the same shape."""
    x = _lib.contiguous(images, np.float32)     # median_filter.py:318 casts to float32
    one = x.ndim == 2
    if one:
        x = x[None]
    assert x.ndim == 3, "images must be (N, M) or (B, N, M)"
    wr, wc = _window(kernel_size)
    out = np.empty_like(x)
    if x.size:
        ctx = _lib.context(device)
        rc = _lib.load().hv_medfilt2d(ctx.handle, _lib.ptr(x), _lib.ptr(out), x.shape[0], x.shape[1], x.shape[2],
                                      wr, wc)
        _lib.check(rc, "hv_medfilt2d")
    return out[0] if one else out


def MedianFilter(input=None, kernel_size=3, bw=32, bh=32):
    """Median-filter one image or a batch of equally shaped images; returns a
    list of float32 images (median_filter.py:8-365).  bw and bh (the
    reference's thread-block shape) are accepted and ignored."""
    imgs = np.asarray(input)
    if imgs.ndim == 2:
        imgs = imgs[None]
    return list(medfilt2d(imgs, kernel_size))


def bench_medfilt2d(images, kernel_size=3, reps=10, *, device=0) -> float:
    """Milliseconds per launch of the filter over a batch already in HBM."""
    x = _lib.contiguous(images, np.float32)
    if x.ndim == 2:
        x = x[None]
    wr, wc = _window(kernel_size)
    ms = C.c_double()
    ctx = _lib.context(device)
    rc = _lib.load().hv_bench_medfilt2d(ctx.handle, _lib.ptr(x), x.shape[0], x.shape[1], x.shape[2], wr, wc,
                                        int(reps), C.byref(ms))
    _lib.check(rc, "hv_bench_medfilt2d")
    return ms.value

"""CPU restatement of the reference's 2-D median filter (test infrastructure:
only tests, smoke() and bench.py's CPU leg use it).

Follows hera_gpu/median_filter.py:8-365:
  * kernel_size is an int or (rows, cols) (:19-22); the images are cast to
    float32 (:318) and zero padded by rows//2 and cols//2 (:318);
  * output (r, c) is the element of rank (rows*cols)//2 (0-based,
    median_filter.py:213) of the window of padded-image rows
    r - rows//2 ... r - rows//2 + rows - 1 and columns c - cols//2 ...
    c - cols//2 + cols - 1 (the loops of the mf kernel, :196-211), which is
    the centred window for odd sizes and the one the reference uses for even
    ones;
  * a 2-D image gives a list of one output (:27-31); a 3-D array a list of
    its images' outputs.
Ranks are taken in IEEE 754 totalOrder on float32: -0 < +0 and NaNs sort to
the ends by sign (the reference's comparisons leave both cases to the order
its selection happens to visit; for ordinary finite data without signed
zeros the two agree, which tests/test_medfilt_oracle.py pins against
scipy.signal.medfilt2d, the routine the reference's tests compare with).
"""
from __future__ import annotations

import numpy as np
from numpy.lib.stride_tricks import sliding_window_view


def to_keys(x: np.ndarray) -> np.ndarray:
    """float32 -> int32 keys whose signed order is IEEE totalOrder."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.int32)
    return u ^ ((u >> 31) & np.int32(0x7FFFFFFF))


def from_keys(k: np.ndarray) -> np.ndarray:
    """Inverse of to_keys (the map is an involution)."""
    k = np.ascontiguousarray(k, dtype=np.int32)
    return (k ^ ((k >> 31) & np.int32(0x7FFFFFFF))).view(np.float32)


def window_shape(kernel_size) -> tuple[int, int]:
    """(rows, cols) of the window (median_filter.py:19-22)."""
    if isinstance(kernel_size, (int, np.integer)):
        return int(kernel_size), int(kernel_size)
    wr, wc = kernel_size
    return int(wr), int(wc)


def _filter(img: np.ndarray, wr: int, wc: int, rank: int, row_off: int, col_off: int) -> np.ndarray:
    """Rank-`rank` element of each (wr, wc) window whose top-left corner sits
    row_off rows and col_off columns before the output pixel; zero outside."""
    n, m = img.shape
    keys = to_keys(img)
    pad = np.zeros((n + wr - 1, m + wc - 1), dtype=np.int32)   # to_keys(+0.0) == 0
    pad[row_off:row_off + n, col_off:col_off + m] = keys
    out = np.empty((n, m), dtype=np.int32)
    step = max(1, (1 << 22) // max(1, m * wr * wc))              # bound the window copy
    for r0 in range(0, n, step):
        r1 = min(n, r0 + step)
        win = sliding_window_view(pad[r0:r1 + wr - 1], (wr, wc)).reshape(r1 - r0, m, wr * wc)
        out[r0:r1] = np.partition(win, rank, axis=-1)[..., rank]
    return from_keys(out)


def medfilt2d_one(img: np.ndarray, kernel_size=3) -> np.ndarray:
    """One image (median_filter.py:196-213 for every output pixel)."""
    wr, wc = window_shape(kernel_size)
    img = np.asarray(img, dtype=np.float32)
    assert img.ndim == 2 and wr >= 1 and wc >= 1
    return _filter(img, wr, wc, (wr * wc) // 2, wr // 2, wc // 2)


def median_filter(input, kernel_size=3) -> list:
    """MedianFilter(input, kernel_size) of median_filter.py:8 (a list of
    float32 images)."""
    imgs = np.asarray(input)
    if imgs.ndim == 2:
        imgs = imgs[None]
    assert imgs.ndim == 3
    return [medfilt2d_one(im, kernel_size) for im in imgs]


def brute_force(img, kernel_size) -> np.ndarray:
    """Pure-Python loops over an explicitly padded image (tiny inputs only):
    pad by rows//2, cols//2 on every side as median_filter.py:318 does, then
    for each output sort its window in totalOrder and take rank rows*cols//2."""
    wr, wc = window_shape(kernel_size)
    a = np.asarray(img, dtype=np.float32)
    n, m = a.shape
    hr, hc = wr // 2, wc // 2
    padded = [[np.float32(0.0)] * (m + 2 * hc) for _ in range(n + 2 * hr)]
    for r in range(n):
        for c in range(m):
            padded[r + hr][c + hc] = a[r, c]
    out = np.empty((n, m), dtype=np.float32)
    for r in range(n):
        for c in range(m):
            vals = [padded[r + i][c + j] for i in range(wr) for j in range(wc)]
            vals.sort(key=lambda v: int(to_keys(np.array([v], dtype=np.float32))[0]))
            out[r, c] = vals[(wr * wc) // 2]
    return out

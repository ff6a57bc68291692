"""The GPU median filter against the CPU restatement (oracle/medfilt_oracle.py,
pinned to scipy.signal.medfilt2d in test_medfilt_oracle.py): bit for bit,
for every compiled window shape, the radix-selection path, batches, special
values, and the reference's own test cases (hera_gpu/tests/test_median_filter.py)
at their full sizes."""
import numpy as np
import pytest
import scipy.signal as sps

from hera_gpu_amd import MedianFilter
from hera_gpu_amd.median_filter import bench_medfilt2d, medfilt2d
from oracle import medfilt_oracle as mo

pytestmark = pytest.mark.gpu

ODD = (1, 3, 5, 7, 9, 11)


def same_bits(a, b):
    np.testing.assert_array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


@pytest.mark.parametrize("wr", ODD)
def test_every_compiled_window(wr):
    # 37 x 131: neither side a multiple of the 8 x 64 tile
    rng = np.random.default_rng(wr)
    x = (rng.standard_normal((37, 131)) * 10).astype(np.float32)
    for wc in ODD:
        print(f"window {wr}x{wc}", flush=True)
        same_bits(medfilt2d(x, (wr, wc)), mo.medfilt2d_one(x, (wr, wc)))


RADIX = [2, (4, 2), (2, 5), (13, 13), (1, 13), (13, 1), (15, 3), (6, 10), (31, 31), (127, 1), (1, 127), (40, 40)]


@pytest.mark.parametrize("case", range(len(RADIX)))
def test_radix_selection_windows(case):
    ks = RADIX[case]
    rng = np.random.default_rng(case)
    x = rng.integers(-50, 50, (45, 150)).astype(np.float32) * np.float32(0.25)   # ties
    print(f"window {ks}", flush=True)
    same_bits(medfilt2d(x, ks), mo.medfilt2d_one(x, ks))


def test_window_larger_than_image():
    x = np.random.default_rng(5).random((3, 4)).astype(np.float32)
    for ks in (9, (7, 3), 13):
        same_bits(medfilt2d(x, ks), mo.medfilt2d_one(x, ks))
    one = np.array([[2.5]], dtype=np.float32)
    for ks in (1, 3, 9, 11):
        same_bits(medfilt2d(one, ks), mo.medfilt2d_one(one, ks))


def test_batch_in_one_launch():
    x = np.random.default_rng(6).random((5, 70, 90)).astype(np.float32)
    for ks in ((3, 3), (5, 9), (4, 4)):
        got = medfilt2d(x, ks)
        for i in range(5):
            same_bits(got[i], mo.medfilt2d_one(x[i], ks))


def test_special_values_total_order():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((40, 70)).astype(np.float32)
    x[rng.random(x.shape) < 0.05] = np.nan
    x[rng.random(x.shape) < 0.05] = -np.float32(np.nan)
    x[rng.random(x.shape) < 0.05] = np.inf
    x[rng.random(x.shape) < 0.05] = -np.inf
    x[rng.random(x.shape) < 0.1] = -0.0
    x[rng.random(x.shape) < 0.1] = 0.0
    x[rng.random(x.shape) < 0.05] = np.float32(1e-42)    # subnormal
    for ks in (3, (5, 7), 9, (2, 2), 11):
        same_bits(medfilt2d(x, ks), mo.medfilt2d_one(x, ks))


def test_reference_api_and_cases_full_size():
    # test_median_filter.py:8-98 at the reference's sizes, against
    # scipy.signal.medfilt2d as there (scipy pins the oracle on these windows)
    rng = np.random.default_rng(8)
    cases = [((7000, 7000), 1), ((7000, 7000), 3), ((4000, 4000), 5), ((997, 997), 9), ((8013, 700), 5),
             ((8013, 700), (9, 5)), ((1, 73), (1, 11)), ((5, 3), (11, 1)), ((2, 3), (3, 5)), ((1, 1), 9)]
    for shape, ks in cases:
        print(f"reference case {shape} {ks}", flush=True)
        x = rng.random(shape)
        got = MedianFilter(input=x, kernel_size=ks)
        assert isinstance(got, list) and len(got) == 1 and got[0].dtype == np.float32
        np.testing.assert_array_equal(got[0], sps.medfilt2d(x.astype(np.float32), ks))
    imgs = [rng.random((29, 29)) for _ in range(3)]
    got = MedianFilter(input=imgs, kernel_size=3, bw=16, bh=16)
    assert len(got) == 3
    for g, x in zip(got, imgs):
        np.testing.assert_array_equal(g, sps.medfilt2d(x.astype(np.float32), 3))


def test_argument_checks():
    x = np.zeros((8, 8), dtype=np.float32)
    with pytest.raises(AssertionError):
        medfilt2d(x, 128)
    with pytest.raises(AssertionError):
        medfilt2d(x, (0, 3))


def test_bench_entry_point_times_the_same_filter():
    x = np.random.default_rng(9).random((512, 512)).astype(np.float32)
    ms = bench_medfilt2d(x, 5, reps=3)
    assert 0 < ms < 1000

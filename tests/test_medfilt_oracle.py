"""Pins for the CPU restatement of the 2-D median filter (oracle/medfilt_oracle.py):
scipy.signal.medfilt2d (what hera_gpu/tests/test_median_filter.py:14-98
compares against) on the reference's own shapes and windows, a pure-Python
brute force for even windows (which scipy refuses), totalOrder special
cases, and planted mistakes the pins must catch."""
import numpy as np
import pytest
import scipy.signal as sps

from oracle import medfilt_oracle as mo


# (image shape, kernel_size) from test_median_filter.py:8-98, large images cut down
REF_CASES = [((1, 1), 1), ((29, 29), 1), ((93, 93), 1), ((1, 1), 3), ((29, 29), 3), ((93, 93), 3),
             ((1, 1), 5), ((400, 300), 5), ((1, 1), 9), ((197, 197), 9), ((1, 73), 1), ((5, 3), 1),
             ((2, 3), 3), ((801, 70), 5), ((1, 73), (1, 11)), ((5, 3), (11, 1)), ((2, 3), (3, 5)),
             ((801, 70), (9, 5))]


@pytest.mark.parametrize("case", range(len(REF_CASES)))
def test_matches_scipy_medfilt2d(case):
    shape, ks = REF_CASES[case]
    rng = np.random.default_rng(case)
    x = rng.random(shape).astype(np.float32)
    got = mo.median_filter(x, kernel_size=ks)
    assert len(got) == 1 and got[0].dtype == np.float32 and got[0].shape == shape
    np.testing.assert_array_equal(got[0], sps.medfilt2d(x, ks))


def test_matches_scipy_with_ties_and_negatives():
    rng = np.random.default_rng(7)
    x = rng.integers(-3, 4, (60, 45)).astype(np.float32)          # many ties, negatives
    for ks in (3, (5, 3), (3, 7), 7):
        np.testing.assert_array_equal(mo.median_filter(x, ks)[0], sps.medfilt2d(x, ks))


def test_float64_input_is_cast_to_float32():
    x = np.random.default_rng(8).random((20, 30))
    got = mo.median_filter(x, 3)[0]
    np.testing.assert_array_equal(got, sps.medfilt2d(x.astype(np.float32), 3))


def test_batch_is_per_image():
    x = np.random.default_rng(9).random((3, 17, 23)).astype(np.float32)
    got = mo.median_filter(x, (3, 5))
    assert len(got) == 3
    for i in range(3):
        np.testing.assert_array_equal(got[i], sps.medfilt2d(x[i], (3, 5)))


@pytest.mark.parametrize("shape,ks", [((6, 7), 2), ((5, 9), (4, 2)), ((7, 4), (2, 3)), ((1, 5), (1, 4)),
                                      ((4, 4), (6, 6)), ((3, 8), (5, 3))])
def test_matches_brute_force(shape, ks):
    rng = np.random.default_rng(shape[0] * 31 + shape[1])
    x = (rng.random(shape).astype(np.float32) - 0.3).astype(np.float32)
    np.testing.assert_array_equal(mo.median_filter(x, ks)[0], mo.brute_force(x, ks))


def test_total_order_signed_zero_and_nan():
    # an all -0.0 image: interior windows hold only -0.0; an edge window of a
    # 3x3 holds 6 (-0.0) and 3 padding +0.0, so rank 4 is -0.0; a corner
    # window holds 4 (-0.0) and 5 (+0.0), so rank 4 is +0.0
    x = np.full((4, 5), -0.0, dtype=np.float32)
    got = mo.median_filter(x, 3)[0]
    sign = np.signbit(got)
    assert sign[1:-1, 1:-1].all() and sign[0, 1:-1].all() and sign[1:-1, 0].all()
    assert not sign[0, 0] and not sign[0, -1] and not sign[-1, 0] and not sign[-1, -1]
    np.testing.assert_array_equal(np.signbit(mo.brute_force(x, 3)), sign)
    # +NaN sorts above +inf, -NaN below -inf
    y = np.random.default_rng(3).random((6, 6)).astype(np.float32)
    y[2, 2] = np.nan
    y[3, 4] = -np.float32(np.nan)
    y[1, 1] = np.inf
    a, b = mo.median_filter(y, 3)[0], mo.brute_force(y, 3)
    np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))
    z = np.full((3, 3), np.nan, dtype=np.float32)
    assert np.isnan(mo.median_filter(z, 1)[0]).all()


def test_keys_are_total_order():
    v = np.array([-np.inf, -2.5, -1e-40, -0.0, 0.0, 1e-40, 3.0, np.inf, np.nan], dtype=np.float32)
    k = mo.to_keys(v)
    assert (np.diff(k.astype(np.int64)) > 0).all()
    np.testing.assert_array_equal(mo.from_keys(k).view(np.int32), v.view(np.int32))


@pytest.mark.parametrize("mistake", ["rank", "transpose", "offset"])
def test_planted_mistakes_fail_the_pins(mistake):
    rng = np.random.default_rng(11)
    x = rng.random((40, 33)).astype(np.float32)
    wr, wc = 5, 3
    if mistake == "rank":
        bad = mo._filter(x, wr, wc, (wr * wc) // 2 - 1, wr // 2, wc // 2)
    elif mistake == "transpose":
        bad = mo._filter(x, wc, wr, (wr * wc) // 2, wc // 2, wr // 2)
    else:
        bad = mo._filter(x, wr, wc, (wr * wc) // 2, wr // 2 + 1, wc // 2)
    assert not np.array_equal(bad, sps.medfilt2d(x, (wr, wc)))
    # an even-window offset mistake shows against the brute force
    y = rng.random((9, 8)).astype(np.float32)
    bad_even = mo._filter(y, 4, 4, 8, 1, 1)
    assert not np.array_equal(bad_even, mo.brute_force(y, 4))

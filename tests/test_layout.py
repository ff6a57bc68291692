"""Baseline order of the upper-triangle output layouts (host side)."""
import numpy as np

from hera_gpu_amd import baseline_pairs


def test_baseline_index_formula():
    # the reduce kernel writes baseline (a, b), a <= b, at a N - a (a-1)/2 + (b - a)
    for n in (1, 2, 16, 37, 350):
        a1, a2 = baseline_pairs(n)
        assert a1.shape == (n * (n + 1) // 2,)
        k = a1.astype(np.int64) * n - a1.astype(np.int64) * (a1 - 1) // 2 + (a2 - a1)
        np.testing.assert_array_equal(k, np.arange(a1.size))
        assert np.all(a1 <= a2)


def test_pairs_cover_the_upper_triangle_once():
    n = 9
    a1, a2 = baseline_pairs(n)
    m = np.zeros((n, n), int)
    np.add.at(m, (a1, a2), 1)
    np.testing.assert_array_equal(m, np.triu(np.ones((n, n), int)))

"""Pins of the Omnical restatement (oracle/omnical_oracle.py).

The reference's own test (hera_gpu/tests/test_redcal.py:129-146) compares
against hera_cal's CPU omnical, which is not installed here, so the
restatement is pinned by what the algorithm fixes:
  * closed form of one update on a one-baseline system (redcal.py:274-275,
    296-297: wc = (1 - g) + g buf / wgt, applied to gains and ubls);
  * the exact model is a fixed point: data = g_i conj(g_j) u_k at the start
    stays there and converges at the first check;
  * from a perturbed start on noiseless redundant data, chi^2 falls by many
    orders of magnitude and the model reproduces the data;
  * the fixed-order pixel sum equals the exact sum to rounding;
and a planted mistake (the second antenna's term not conjugated) must fail.
"""
import numpy as np
import pytest

from oracle import omnical_oracle as oo
from oracle.omnical_oracle import redundant_array


def test_one_update_closed_form():
    ggu = np.array([[0, 1, 0]], dtype=np.uint32)
    gains = np.ones((1, 2), complex)
    ubls = np.ones((1, 1), complex)
    data = np.array([[2.0 + 0j]])
    wgts = np.ones((1, 1))
    r = oo.omnical(ggu, gains, ubls, data, wgts, maxiter=1, precision=2)
    gs = float(np.float32(0.3))
    wc = float(np.float32(1) - np.float32(0.3)) + gs * 2.0 / 1.0
    np.testing.assert_allclose(r["gains"], wc * np.ones((1, 2)), rtol=1e-15)
    np.testing.assert_allclose(r["ubls"], [[wc]], rtol=1e-15)
    assert r["iters"][0] == 1
    np.testing.assert_allclose(r["chisq"], [(2 - wc ** 3) ** 2], rtol=1e-14)
    d2 = 3 * (wc - 1) ** 2
    np.testing.assert_allclose(r["conv"], [np.sqrt(d2 / (3 * wc ** 2))], rtol=1e-14)


def test_exact_model_is_a_fixed_point():
    ggu, g, u, data, wgts = redundant_array(seed=1)
    r = oo.omnical(ggu, g, u, data, wgts, maxiter=20, precision=2)
    np.testing.assert_allclose(r["gains"], g, rtol=1e-12)
    np.testing.assert_allclose(r["ubls"], u, rtol=1e-12)
    assert np.all(r["chisq"] < 1e-20 * np.sum(np.abs(data) ** 2, axis=1))
    assert np.all(r["conv"] < 1e-10)


@pytest.mark.parametrize("precision", [1, 2])
def test_converges_from_perturbed_start(precision):
    ggu, g, u, data, wgts = redundant_array(seed=2)
    rng = np.random.default_rng(3)
    g0 = g * (1 + 0.05 * (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)))
    u0 = u * (1 + 0.05 * (rng.standard_normal(u.shape) + 1j * rng.standard_normal(u.shape)))
    i, j, k = ggu[:, 0], ggu[:, 1], ggu[:, 2]
    chisq0 = np.sum(np.abs(data - g0[:, i] * np.conj(g0[:, j]) * u0[:, k]) ** 2, axis=1)
    r = oo.omnical(ggu, g0, u0, data, wgts, maxiter=300, precision=precision)
    fall = 1e-10 if precision == 2 else 1e-6
    assert np.all(r["chisq"] < fall * chisq0)
    model = r["gains"][:, i] * np.conj(r["gains"][:, j]) * r["ubls"][:, k]
    tol = 1e-5 if precision == 2 else 1e-3
    assert np.max(np.abs(model - data)) < tol * np.max(np.abs(data))
    assert np.all(r["iters"] > 0) and np.all(r["iters"] <= 300)


def test_fixed_sum_is_a_sum():
    x = np.random.default_rng(4).uniform(0, 1, (3, 1000))
    np.testing.assert_allclose(oo.fixed_sum(x), x.sum(axis=1), rtol=1e-13)
    assert oo.fixed_sum(np.ones((1, 1))) == 1.0


def test_planted_unconjugated_second_antenna_fails(monkeypatch):
    ggu, g, u, data, wgts = redundant_array(seed=2)
    rng = np.random.default_rng(3)
    g0 = g * (1 + 0.05 * (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)))
    good = oo.omnical(ggu, g0, u, data, wgts, maxiter=200, precision=2)

    def bad_gen(ggu_, gr, gi, ur, ui):
        i, j, k = ggu_[:, 0], ggu_[:, 1], ggu_[:, 2]
        tr, ti = oo.cmul(gr[:, i], gi[:, i], gr[:, j], gi[:, j])     # conj dropped
        return oo.cmul(tr, ti, ur[:, k], ui[:, k])

    monkeypatch.setattr(oo, "gen_dmdl", bad_gen)
    bad = oo.omnical(ggu, g0, u, data, wgts, maxiter=200, precision=2)
    assert np.all(bad["chisq"] > 1e3 * good["chisq"] + 1e-12)

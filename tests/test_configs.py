"""Synthetic workload generators: shapes, dtypes, determinism, layout facts."""
import numpy as np

from hera_gpu_amd import configs as c


def test_hex37_shape_and_spacing():
    w = c.make_hex37()
    assert w.shape == dict(N=37, K=1000, T=10, F=1, P=31)
    assert w.antpos.dtype == np.float32 and w.crd_eq.dtype == np.float32
    d = np.linalg.norm(w.antpos[:, None, :2] - w.antpos[None, :, :2], axis=-1)
    d[d == 0] = np.inf
    np.testing.assert_allclose(d.min(), 14.6, rtol=1e-6)
    assert np.all(w.antpos[:, 2] == 0)
    np.testing.assert_allclose(np.linalg.norm(w.crd_eq, axis=0), 1, atol=1e-6)
    assert w.freqs[0] == 150e6


def test_determinism():
    a, b = c.make_hex37(), c.make_hex37()
    for k in ("antpos", "crd_eq", "I_sky", "bm_cube", "eq2tops"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k))


def test_eq2tops_are_rotations():
    eq = c.eq2tops_lst(5).astype(np.float64)
    for R in eq:
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)
        assert abs(np.linalg.det(R) - 1) < 1e-6
    # zenith of the first time is at declination = latitude
    up = eq[0][2]
    np.testing.assert_allclose(np.degrees(np.arcsin(up[2])), -30.7, atol=1e-4)


def test_hera350_layout():
    rng = np.random.default_rng(0)
    a = c.hera350_layout(rng)
    assert a.shape == (350, 3)
    r = np.hypot(a[:, 0], a[:, 1])
    assert (r[:320] < 160).all() and (r[320:] > 140).all() and (r <= 1000).all()
    assert len(c.hex_layout(10)) == 331


def test_ring_pixels_equal_area_count():
    crd = c.ring_pixel_centres(8)
    assert crd.shape == (3, 12 * 64)
    np.testing.assert_allclose(np.linalg.norm(crd, axis=0), 1, atol=1e-12)
    # equal-area rings: mean z over all pixels is 0
    assert abs(crd[2].mean()) < 1e-12


def test_powerlaw_flux_range_and_slope():
    s = c.powerlaw_flux(np.random.default_rng(0), 200000)
    assert s.min() >= 0.01 * (1 - 1e-6) and s.max() <= 100 * (1 + 1e-6)
    # dN/dS ~ S^-2.5 => N(>S) ~ S^-1.5: the fraction above 1 Jy is ~ (1/0.01)^-1.5
    frac = (s > 1).mean()
    assert abs(frac - 1e-3) < 3e-4


def test_grf_unit_variance():
    crd = c.ring_pixel_centres(16)
    f = c.gaussian_random_field(np.random.default_rng(0), crd, lmax=24)
    assert abs(f.mean()) < 1e-9 and abs(f.std() - 1) < 1e-9


def test_serving_stream_draws():
    jobs = c.make_serving_stream(300, rate_hz=10.0)
    assert len(jobs) == 300
    assert {j.layout for j in jobs} == {"hex37", "disk128", "hera350"}
    assert all(1000 <= j.nsrc <= 2 ** 20 for j in jobs)
    assert {j.ntimes for j in jobs} <= {1, 8, 32} and {j.nfreq for j in jobs} <= {1, 16, 64}
    arr = np.array([j.arrival_s for j in jobs])
    assert np.all(np.diff(arr) >= 0)
    assert abs(np.diff(arr).mean() - 0.1) < 0.02

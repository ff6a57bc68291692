"""The oracle and the workload generator reproduce the committed fixtures."""
from pathlib import Path

import numpy as np
import pytest

from hera_gpu_amd import configs
from oracle.vis_oracle import vis_oracle, rel_err

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_hex37_inputs_match_generator():
    g = load("hex37.npz")
    w = configs.make_hex37()
    for k in ("antpos", "eq2tops", "crd_eq", "I_sky", "bm_cube", "freqs"):
        np.testing.assert_array_equal(g[k], getattr(w, k))


@pytest.mark.parametrize("name", ["hex37.npz", "realistic_small.npz"])
def test_oracle_reproduces_fixture(name):
    g = load(name)
    v = vis_oracle(g["antpos"], g["freqs"], g["eq2tops"], g["crd_eq"], g["I_sky"], g["bm_cube"])
    assert rel_err(v, g["vis"]) < 1e-13


def test_realistic_fixture_has_large_phases():
    g = load("realistic_small.npz")
    d = np.abs(g["antpos"].astype(np.float64) @ (g["eq2tops"][0].astype(np.float64) @ g["crd_eq"]))
    assert (2 * np.pi * g["freqs"].max() / 299792458.0 * d).max() > 1000

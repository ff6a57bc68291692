"""Write the golden fixtures of this directory from the FP64 oracle.

    python tests/golden/make_golden.py

* hex37.npz          -- the 37-antenna workload (inputs + oracle V, complex128
                        (T, N, N)), float32 inputs as the precision=1 path sees them;
* realistic_small.npz -- 6 antennas spread over 1 km, 300 sources, 3 times,
                        3 frequencies over 100-200 MHz: phases of ~10^3 rad.
The reference ships no stored vectors (hera_gpu/tests/test_vis_gpu.py holds
only inline closed-form cases), so these fixtures pin regressions of the
oracle and give the GPU tests fixed inputs; the oracle itself is pinned by
tests/test_oracle.py.
"""
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from hera_gpu_amd import configs  # noqa: E402
from oracle.vis_oracle import vis_oracle  # noqa: E402


def realistic_small():
    rng = np.random.default_rng(20240607)
    N, K, T = 6, 300, 3
    antpos = np.zeros((N, 3))
    antpos[:, :2] = rng.uniform(-500, 500, (N, 2))
    antpos = antpos.astype(np.float32)
    z = rng.uniform(-0.2, 1, K)
    phi = rng.uniform(0, 2 * np.pi, K)
    crd = np.stack([np.sqrt(1 - z * z) * np.cos(phi), np.sqrt(1 - z * z) * np.sin(phi), z]).astype(np.float32)
    I = rng.lognormal(0, 1, K).astype(np.float32)
    bm = configs.gaussian_beams(rng, N, 31, fwhm_deg=60.0)
    eq = configs.eq2tops_lst(T, step_deg=1.0, lst0_deg=30.0)
    freqs = np.array([100e6, 150e6, 200e6])
    return dict(antpos=antpos, freqs=freqs, eq2tops=eq, crd_eq=crd, I_sky=I, bm_cube=bm)


def main():
    w = configs.make_hex37()
    d = dict(antpos=w.antpos, freqs=w.freqs, eq2tops=w.eq2tops, crd_eq=w.crd_eq, I_sky=w.I_sky, bm_cube=w.bm_cube)
    d["vis"] = vis_oracle(d["antpos"], d["freqs"], d["eq2tops"], d["crd_eq"], d["I_sky"], d["bm_cube"])
    np.savez_compressed(HERE / "hex37.npz", **d)
    r = realistic_small()
    r["vis"] = vis_oracle(r["antpos"], r["freqs"], r["eq2tops"], r["crd_eq"], r["I_sky"], r["bm_cube"])
    np.savez_compressed(HERE / "realistic_small.npz", **r)
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


if __name__ == "__main__":
    main()

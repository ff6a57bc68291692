"""Contraction variant sweep on the GPU (tuning aid for the built-in choice).

    python -m hera_gpu_amd.tune [--nant 350 128] [--nsrc 262144] [--units 8]

For every arithmetic and every kernel variant (tile edge x pf, see
``hv_set_herk_variant``) it times the contraction alone on synthetic operands
(``hv_bench_herk``) and, for the precision-1 arithmetics, the relative error of
a small HERA-350 simulation run with that variant against the FP64 path.  One
JSON line per measurement on stdout.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from . import _lib, configs, stages
from .vis import vis_gpu

VARIANTS = {
    "split_bf16": ((64, 1), (64, 3), (64, 4), (128, 2), (128, 3), (128, 4)),
    "split_fp16": ((64, 3), (128, 3), (128, 4)),
    "fp32": ((64, 1), (64, 3), (64, 4)),
    "fp64": ((64, 1), (64, 3), (64, 4)),
}


def rel_err(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def sweep_speed(nants, nsrc, units, gemms):
    for gemm in gemms:
        for nant in nants:
            for tile, pf in VARIANTS[gemm]:
                k = nsrc if gemm != "fp64" else nsrc // 4
                try:
                    r = stages.bench_herk(gemm, nant, k, units=units, tile=tile, pf=pf, reps=3)
                except Exception as e:      # a variant that does not exist for this size
                    r = {"error": str(e)}
                r.update(gemm=gemm, nant=nant, nsrc=k, units=units, tile=tile, pf=pf)
                print(json.dumps(r), flush=True)


def sweep_accuracy(gemms):
    lib = _lib.load()
    w = configs.make_hera350(nsrc=2 ** 17, ntimes=2, nfreq=2)
    # the FP64 path on the intensities a precision-1 call computes with
    Ieff = np.sqrt(w.I_sky).astype(np.float32).astype(np.float64) ** 2
    ref = vis_gpu(w.antpos, w.freqs, w.eq2tops, w.crd_eq, Ieff, w.bm_cube, precision=2)
    for gemm in gemms:
        if gemm == "fp64":
            continue
        for tile, pf in VARIANTS[gemm]:
            _lib.check(lib.hv_set_herk_variant(tile, pf), "hv_set_herk_variant")
            try:
                v = vis_gpu(w.antpos, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube, gemm=gemm)
            finally:
                _lib.check(lib.hv_set_herk_variant(0, 0), "hv_set_herk_variant")
            err = max(rel_err(v[f, t], ref[f, t]) for f in range(2) for t in range(2))
            print(json.dumps({"accuracy": gemm, "tile": tile, "pf": pf, "rel_err": err}), flush=True)


PIPELINE_VARIANTS = ((64, 4), (128, 4), (128, 5), (64, 5))


def sweep_pipeline(nants, nsrc, ntimes, nfreq, gemm):
    """Whole simulation (rotation, operand build, contraction, reduce) per
    variant on a synthetic sky: ms per (time, frequency) unit."""
    lib = _lib.load()
    rng = np.random.default_rng(1)
    z = rng.uniform(-1, 1, nsrc)
    ph = rng.uniform(0, 2 * np.pi, nsrc)
    crd = np.stack([np.sqrt(1 - z * z) * np.cos(ph), np.sqrt(1 - z * z) * np.sin(ph), z]).astype(np.float32)
    flux = rng.lognormal(0, 1, nsrc).astype(np.float32)
    sky = stages.Sky(crd, flux)
    eq = configs.eq2tops_lst(ntimes)
    freqs = np.linspace(100e6, 200e6, nfreq)
    for nant in nants:
        ap = configs.hera350_layout(np.random.default_rng(3))[:nant].astype(np.float32)
        beam = stages.Beam(configs.gaussian_beams(np.random.default_rng(4), nant, 63))
        for tile, pf in PIPELINE_VARIANTS:
            _lib.check(lib.hv_set_herk_variant(tile, pf), "hv_set_herk_variant")
            try:
                stages.simulate(sky, beam, ap, freqs, eq, gemm=gemm, want_output=False, dtype=np.float32)
                _, t = stages.simulate(sky, beam, ap, freqs, eq, gemm=gemm, want_output=False, timing=True,
                                       dtype=np.float32)
            finally:
                _lib.check(lib.hv_set_herk_variant(0, 0), "hv_set_herk_variant")
            units = ntimes * nfreq
            print(json.dumps({"pipeline": gemm, "nant": nant, "nsrc": nsrc, "tile": tile, "pf": pf,
                              "ms_per_unit": t["total_ms"] / units,
                              "tflops": 4.0 * nant * (nant + 1) * nsrc * units / (t["total_ms"] * 1e-3) / 1e12}),
                  flush=True)
        beam.close()
    sky.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nant", type=int, nargs="+", default=[350, 128])
    ap.add_argument("--nsrc", type=int, default=2 ** 18)
    ap.add_argument("--units", type=int, default=8)
    ap.add_argument("--gemm", nargs="+", default=list(VARIANTS))
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--pipeline", action="store_true",
                    help="time whole simulations per variant (split-BF16) instead of the contraction alone")
    args = ap.parse_args()
    if args.pipeline:
        sweep_pipeline(args.nant, args.nsrc, 2, 16, "split_bf16")
        return
    sweep_speed(args.nant, args.nsrc, args.units, args.gemm)
    if not args.no_accuracy:
        sweep_accuracy(args.gemm)


if __name__ == "__main__":
    main()

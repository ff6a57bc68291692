#!/usr/bin/env python3
"""Benchmark of the visibility simulator on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload hera350] [--gemm split_bf16]

A step is one whole simulation of the workload with its inputs resident in
HBM: for the default ``hera350`` workload, 350 antennas x 2^20 sources x 32
times x 64 frequencies (2048 (time, frequency) units).  The headline value is
the useful Hermitian GEMM rate 4 N (N+1) K T F / step time in TFLOP/s of the
split-BF16 path; the plain FP32 path is timed on the same workload beside it,
and both are compared with the FP64 path on a sample of units.  One JSON line
is printed on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import contextlib
import json
import os
import sys
import time

import numpy as np

METRIC = "vis GEMM TFLOP/s & p99 job latency, split-BF16 vs FP32 at ≤1e-5 rel err"
PEAK_TFLOPS = {"split_bf16": 2500.0, "split_fp16": 2500.0, "fp32": 157.3, "fp64": 78.6}
HBM_PEAK_GBS = 8000.0
# HBM traffic of one contraction launch at the bench's launch shape (workload,
# arithmetic, units per launch, tile edge), from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE passes over `bench.py --kernels-only` (KiB per dispatch), scaled
# by what the same counters read for a 1 GiB device copy: FETCH_SIZE x2.0,
# WRITE_SIZE x1.0.  `src` is the hash of the contraction's sources
# (KERNEL_SOURCES) when the counters were read: once they change, the figure
# is reported as null (stale) until the passes are run again.
MEASURED_TRAFFIC = {
    # in-place build (herk_t192_kernel, triangle and strip items), two times
    # (128 units) per launch, 32768-source K-splits (a job of 2048 units)
    ("hera350", "split_bf16", 128, 192): {"fetch_kib": 5124967, "write_kib": 2173440, "fetch_scale": 2.0,
                                          "write_scale": 1.0, "src": "f025ed689749092f"},
}
KERNEL_SOURCES = ("hera_gpu_amd/csrc/hv_herk.hip", "hera_gpu_amd/csrc/hv_phasor_math.h",
                  "hera_gpu_amd/csrc/hv_internal.h")


def kernel_source_hash():
    import hashlib
    h = hashlib.sha256()
    root = os.path.dirname(os.path.abspath(__file__))
    for f in KERNEL_SOURCES:
        with open(os.path.join(root, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]
PF_NAMES = {1: "register-staged, 1 stage ahead", 2: "register-staged, 2 stages ahead",
            3: "LDS ring, one workgroup per work item", 4: "LDS ring, persistent grid",
            5: "operands built in place from A sqrt(I), 8 waves, persistent grid"}


def variant_name(tile, pf):
    if tile == 192:
        return ("operands built in place from A sqrt(I), 8 waves (Re and Im wave pair per SIMD), "
                "at 321..352 antennas a triangle item (32-row blocks r <= c < 8) and a strip item "
                "(block columns 8-10) per (unit, split), persistent XCD-grouped grid")
    return PF_NAMES[pf]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def herk_flops(N, K, units):
    return 4.0 * N * (N + 1) * K * units


def t192_mfma_blocks(N):
    """32x32 MFMA blocks the 192-tile kernel runs per unit and 16-source stage
    for each component: per tile (bi <= bj) 21 (diagonal) or 36 (off-diagonal)
    blocks, 15 and 30 when at most five 32-row blocks of the last block row or
    column hold antennas (at 321..352 antennas the triangle and strip items
    run the same 66 blocks: 36 + 30)."""
    nb = -(-N // 192)
    total = 0
    for bi in range(nb):
        for bj in range(bi, nb):
            five = N - bj * 192 <= 160
            total += (15 if five else 21) if bi == bj else (30 if five else 36)
    return total


@contextlib.contextmanager
def job_splits(gemm, N, units, K):
    """Run a part of the workload (a sample of its times, one unit) with the
    K-split length of the whole job, so its arithmetic is the headline's:
    a part of a job that takes 32768-source splits would take 8192 alone."""
    from hera_gpu_amd import _lib, stages
    lib = _lib.load()
    long = stages.ksplit_choice(gemm, N, units, K) == 32768
    if long:
        _lib.check(lib.hv_set_ksplit_items(1), "hv_set_ksplit_items")
    try:
        yield long
    finally:
        if long:
            _lib.check(lib.hv_set_ksplit_items(0), "hv_set_ksplit_items")


def rel_err(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def run_timed(stages, sky, beam, w, gemm, steps, warmup):
    for _ in range(warmup):
        stages.simulate(sky, beam, w.antpos, w.freqs, w.eq2tops, gemm=gemm, want_output=False,
                        dtype=np.float32)
    t0 = time.perf_counter()
    _, st = stages.simulate(sky, beam, w.antpos, w.freqs, w.eq2tops, gemm=gemm, reps=steps, timing=True,
                            want_output=False, dtype=np.float32)
    wall = time.perf_counter() - t0
    return st, wall


def cpu_baseline(w, nsrc, gpu_unit, workers):
    """vis_cpu restated (float32, one call per frequency) on one unit of the
    workload at full N and the first `nsrc` sources (all K by default), with
    the per-antenna beam and phasor rows on `workers` processes and the dot
    products on the host BLAS; the FP64 oracle on the same unit checks the
    CPU and GPU results.  Also the small 37-antenna workload in full."""
    from hera_gpu_amd import configs
    from oracle.cpu_ref import vis_cpu, vis_cpu_parallel
    from oracle.vis_oracle import vis_oracle, vis_oracle_units

    N = w.antpos.shape[0]
    T, F = w.eq2tops.shape[0], len(w.freqs)
    crd = w.crd_eq[:, :nsrc]
    I = w.I_sky[:nsrc]
    t0 = time.perf_counter()
    v = vis_cpu_parallel(w.antpos, float(w.freqs[-1]), w.eq2tops[:1], crd, I, w.bm_cube, precision=1,
                         workers=workers)
    dt = time.perf_counter() - t0
    ref = vis_oracle_units(w.antpos, w.freqs[-1:], w.eq2tops[:1], crd, I, w.bm_cube, [(0, 0)])[0]
    per_unit = dt * w.crd_eq.shape[1] / nsrc
    res = {
        "value": herk_flops(N, nsrc, 1) / dt / 1e12, "unit": "TFLOP/s", "cores": workers, "kind": "port",
        "sample": f"1 unit (time 0, {w.freqs[-1] / 1e6:.0f} MHz) of the workload, N={N}, K={nsrc}"
                  f"{'' if nsrc == w.crd_eq.shape[1] else ' (first sources)'}: {dt:.1f} s; vis_cpu restated in "
                  f"float32 (scipy linear spline beam), its per-antenna beam and phasor rows on {workers} "
                  f"processes, the per-row dot products on the host BLAS ({workers} threads)",
        "seconds": dt,
        "whole_job_extrapolated_s": per_unit * T * F,
        "extrapolation": f"x {T * F} units (T x F), linear in K: {per_unit * T * F / 3600:.1f} h for the whole "
                         f"workload (extrapolated, not run)",
        "cpu_fp32_rel_err_vs_oracle": rel_err(v[0], ref),
    }
    for gemm, vg in gpu_unit.items():
        res[f"gpu_{gemm}_rel_err_vs_oracle"] = rel_err(vg, ref)
    # the 37-antenna workload in full (10 times, 1000 sources, one frequency)
    h = configs.make_hex37()
    t1 = time.perf_counter()
    vh = vis_cpu(h.antpos, float(h.freqs[0]), h.eq2tops, h.crd_eq, h.I_sky, h.bm_cube, precision=1)
    dh = time.perf_counter() - t1
    sh = h.shape
    res["hex37_full"] = {"seconds": dh, "TFLOP/s": herk_flops(sh["N"], sh["K"], sh["T"] * sh["F"]) / dh / 1e12,
                         "rel_err_vs_oracle": rel_err(vh, vis_oracle(h.antpos, h.freqs, h.eq2tops, h.crd_eq,
                                                                      h.I_sky, h.bm_cube)[0]),
                         "sample": "the whole 37-antenna workload (N=37, K=1000, T=10, F=1), vis_cpu restated, "
                                   "one process"}
    return res


def omnical_leg(n=11, ndata=2000, maxiter=50, cpu_pixels=8):
    """Omnical (redcal.omnical) on a synthetic redundant array: an n x n grid
    (n = 11: 121 antennas, 7260 baselines, 220 unique baselines; the
    reference's profiling sizes are 2000 pixels of 54-180 antennas,
    hera_gpu/redcal.py:426-455), noisy data, starts 5% off the truth.  GPU
    wall time of the whole call (host arrays in, uploads and downloads
    included) per precision; the CPU restatement on a few pixels beside it,
    also the parity check of those pixels."""
    from hera_gpu_amd.redcal import bench_omnical, omnical
    from oracle.omnical_oracle import omnical as omnical_cpu, redundant_array

    ggu, g, u, data, wgts = redundant_array(n=n, seed=5, nd=ndata, noise=0.01)
    rng = np.random.default_rng(6)
    g0 = g * (1 + 0.05 * (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)))
    u0 = u * (1 + 0.05 * (rng.standard_normal(u.shape) + 1j * rng.standard_normal(u.shape)))
    out = {"problem": f"{ndata} pixels x {ggu.shape[0]} baselines, {g.shape[1]} antennas, {u.shape[1]} unique "
                      f"baselines; maxiter {maxiter}, check_every 4, gain 0.3",
           "time": "pixels_per_s: wall time of the call, host arrays in and out (uploads included); "
                   "kernel_pixels_per_s: the solve alone with the inputs in HBM"}
    for precision in (1, 2):
        kw = dict(maxiter=maxiter, precision=precision)
        omnical(ggu, g0[:64], u0[:64], data[:64], wgts[:64], **kw)          # warm-up
        t0 = time.perf_counter()
        r = omnical(ggu, g0, u0, data, wgts, **kw)
        dt = time.perf_counter() - t0
        kms = bench_omnical(ggu, g0, u0, data, wgts, reps=5, **kw)["kernel_ms"]
        t1 = time.perf_counter()
        c = omnical_cpu(ggu, g0[:cpu_pixels], u0[:cpu_pixels], data[:cpu_pixels], wgts[:cpu_pixels], **kw)
        dc = time.perf_counter() - t1
        same_iters = bool(np.array_equal(r["iters"][:cpu_pixels], c["iters"]))
        dg = float(np.max(np.abs(r["gains"][:cpu_pixels] - c["gains"])) / np.max(np.abs(c["gains"])))
        key = "fp32" if precision == 1 else "fp64"
        out[key] = {"seconds": dt, "pixels_per_s": ndata / dt, "kernel_ms": kms,
                    "kernel_pixels_per_s": ndata / (kms / 1e3), "mean_iters": float(np.mean(r["iters"])),
                    "cpu_pixels_per_s": cpu_pixels / dc, "cpu_seconds": dc, "cpu_pixels": cpu_pixels,
                    "cpu_kind": "port (oracle/omnical_oracle.py, numpy, one thread per array op)",
                    "gpu_vs_cpu_same_iters": same_iters, "gpu_vs_cpu_max_gain_diff": dg}
        log(f"[bench] omnical {key}: {ndata / dt:.0f} pixels/s ({dt * 1e3:.1f} ms; kernel {kms:.1f} ms, "
            f"{ndata / (kms / 1e3):.0f} pixels/s), mean iters "
            f"{np.mean(r['iters']):.1f}; cpu restatement {cpu_pixels / dc:.2f} pixels/s; same iters {same_iters}")
    return out


def variants_leg(w, gemm, ntimes=4):
    """The simulator's extensions on the same workload: a beam per frequency
    (resident inputs, the first `ntimes` times against the one-beam run of
    the same units), the baseline-major (T, Nbl, F) output against the full
    matrices (host buffers in and out), and complex 2x2 Jones beams (host
    buffers; 2N rows x 2K columns, the first two times)."""
    from hera_gpu_amd import stages, vis_gpu, vis_gpu_baselines, vis_gpu_polarized

    N, K, F = w.antpos.shape[0], w.crd_eq.shape[1], len(w.freqs)
    T = min(ntimes, w.eq2tops.shape[0])
    eq = w.eq2tops[:T]
    flops = herk_flops(N, K, T * F)
    out = {"times": T, "freqs": F}
    sky = stages.Sky(w.crd_eq, w.I_sky)
    one = stages.Beam(w.bm_cube)
    scale = (1.0 + 0.2 * np.arange(F) / F).astype(np.float32)[:, None, None, None]
    per_f = stages.Beam(w.bm_cube[None] * scale)
    res = {}
    for name, beam in (("one_beam", one), ("beam_per_freq", per_f)):
        stages.simulate(sky, beam, w.antpos, w.freqs, eq[:1], gemm=gemm, want_output=False, dtype=np.float32)
        _, st = stages.simulate(sky, beam, w.antpos, w.freqs, eq, gemm=gemm, timing=True, want_output=False,
                                dtype=np.float32)
        res[name] = st["total_ms"]
    out["beam_per_freq"] = {"ms": res["beam_per_freq"], "tflops": flops / (res["beam_per_freq"] / 1e3) / 1e12,
                            "one_beam_ms": res["one_beam"],
                            "one_beam_tflops": flops / (res["one_beam"] / 1e3) / 1e12,
                            "time": "resident inputs, whole pipeline"}
    log(f"[bench] beam per frequency: {res['beam_per_freq']:.1f} ms ({out['beam_per_freq']['tflops']:.1f} "
        f"TFLOP/s) vs one beam {res['one_beam']:.1f} ms, {T * F} units")
    sky.close()
    one.close()
    per_f.close()

    wall = {}
    for name, fn in (("full", vis_gpu), ("baselines_tbf", vis_gpu_baselines)):
        fn(w.antpos, w.freqs[:1], eq[:1], w.crd_eq, w.I_sky, w.bm_cube, gemm=gemm)
        t0 = time.perf_counter()
        r = fn(w.antpos, w.freqs, eq, w.crd_eq, w.I_sky, w.bm_cube, gemm=gemm)
        wall[name] = (time.perf_counter() - t0, (r[0] if isinstance(r, tuple) else r).nbytes)
    out["baseline_layout"] = {"full_s": wall["full"][0], "full_bytes": wall["full"][1],
                              "tbf_s": wall["baselines_tbf"][0], "tbf_bytes": wall["baselines_tbf"][1],
                              "tbf_tflops": flops / wall["baselines_tbf"][0] / 1e12,
                              "time": "host buffers in and out (uploads and result copy included)"}
    log(f"[bench] baseline-major output: {wall['baselines_tbf'][0]:.2f} s ({wall['baselines_tbf'][1] / 1e9:.2f} GB) "
        f"vs full matrices {wall['full'][0]:.2f} s ({wall['full'][1] / 1e9:.2f} GB)")

    Tp = min(2, T)
    rng = np.random.default_rng(13)
    P = w.bm_cube.shape[-1]
    jones = np.zeros((N, 2, 2, P, P), dtype=np.complex64)
    jones[:, 0, 0] = w.bm_cube
    jones[:, 1, 1] = w.bm_cube
    leak = (0.05 * np.exp(2j * np.pi * rng.random((N, 1, 1)))).astype(np.complex64)
    jones[:, 0, 1] = leak * w.bm_cube
    jones[:, 1, 0] = np.conj(leak) * w.bm_cube
    vis_gpu_polarized(w.antpos, w.freqs[:1], eq[:1], w.crd_eq, w.I_sky, jones, gemm=gemm)
    t0 = time.perf_counter()
    vp = vis_gpu_polarized(w.antpos, w.freqs, eq[:Tp], w.crd_eq, w.I_sky, jones, gemm=gemm)
    dt = time.perf_counter() - t0
    pflops = 4.0 * (2 * N) * (2 * N + 1) * (2 * K) * Tp * F
    out["polarized"] = {"seconds": dt, "tflops": pflops / dt / 1e12, "times": Tp, "output_bytes": vp.nbytes,
                        "flop_count": "useful 4 (2N)(2N+1)(2K) per unit (feed-antenna rows x source-axis columns)",
                        "time": "host buffers in and out; phasor build + LDS-ring contraction"}
    log(f"[bench] polarized (Jones) beams: {dt:.2f} s for {Tp * F} units, {pflops / dt / 1e12:.1f} TFLOP/s")
    return out


MEDFILT_CASES = (((7000, 7000), 3), ((4000, 4000), 5), ((997, 997), 9))   # test_median_filter.py:8-63


def medfilt_leg(reps=20, cpu_rows=256):
    """The 2-D median filter (MedianFilter) on the reference's own test images
    (random float32, hera_gpu/tests/test_median_filter.py): the kernel with
    the image resident in HBM (ms per launch, algorithmic bytes 4 in + 4 out
    per pixel), the whole call with host arrays, and the CPU restatement on
    the first `cpu_rows` rows beside it, also their parity check."""
    from hera_gpu_amd.median_filter import MedianFilter, bench_medfilt2d
    from oracle.medfilt_oracle import medfilt2d_one

    rng = np.random.default_rng(12)
    out = {"time": "kernel: ms per launch, image resident in HBM; call: MedianFilter with host arrays"}
    for shape, ks in MEDFILT_CASES:
        x = rng.random(shape).astype(np.float32)
        ms = bench_medfilt2d(x, ks, reps=reps)
        MedianFilter(input=x[:64], kernel_size=ks)
        t0 = time.perf_counter()
        got = MedianFilter(input=x, kernel_size=ks)[0]
        call = time.perf_counter() - t0
        t1 = time.perf_counter()
        ref = medfilt2d_one(x[:cpu_rows], ks)
        dc = time.perf_counter() - t1
        h = ks // 2    # rows whose window stays inside the CPU sample
        same = bool(np.array_equal(got[:cpu_rows - h].view(np.int32), ref[:cpu_rows - h].view(np.int32)))
        px = x.size
        gbs = 8 * px / ms / 1e6
        key = f"{shape[0]}x{shape[1]}_w{ks}"
        out[key] = {"kernel_ms": ms, "gpixels_per_s": px / ms / 1e6, "GB_per_s": gbs, "hbm_frac": gbs / 8000.0,
                    "call_ms": call * 1e3, "cpu_pixels_per_s": cpu_rows * shape[1] / dc,
                    "cpu_sample": f"first {cpu_rows} rows", "cpu_kind": "port (oracle/medfilt_oracle.py, numpy)",
                    "gpu_equals_cpu_on_sample": same}
        log(f"[bench] medfilt {key}: kernel {ms:.3f} ms ({px / ms / 1e6:.1f} Gpixel/s, {gbs:.0f} GB/s), call "
            f"{call * 1e3:.1f} ms; cpu restatement {cpu_rows * shape[1] / dc / 1e6:.2f} Mpixel/s; equal {same}")
    return out


SERVE_BUDGET = 200 << 30   # device bytes for the server's instances (clamped to what the context has left)


def serving_leg(njobs, gemm, nslots, rate=None, burst_jobs=200, fp32_jobs=200, graphs=True):
    """The fifth workload's serving load: a Poisson stream of jobs with N in
    {37,128,350}, Nsrc log-uniform in [1e3, 2^20], ntimes in {1,8,32}, Nfreq
    in {1,16,64}, beams and catalogs resident per layout.  The capacity of
    each arithmetic is measured on the same t=0 burst (the stream's first
    `burst_jobs` jobs, jobs / makespan), which also gives the like-for-like
    split-BF16 / FP32 rate ratio; the split-BF16 stream then brings all
    `njobs` jobs at 70% of its capacity (or `rate`), the FP32 stream its first
    `fp32_jobs` at 70% of the FP32 capacity, so the latencies are those of a
    loaded, not an overloaded, server."""
    import dataclasses

    from hera_gpu_amd import configs, serve, stages

    layouts = {}
    for lay in ("hex37", "disk128", "hera350"):
        antpos, beam, crd, flux = configs.serving_catalog(lay)
        layouts[lay] = (antpos, stages.Beam(beam), stages.Sky(crd, flux))
    shapes = configs.make_serving_stream(njobs, rate_hz=1.0)
    nant = {lay: len(v[0]) for lay, v in layouts.items()}

    def flops(js):
        return float(sum(herk_flops(nant[j.layout], j.nsrc, j.ntimes * j.nfreq) for j in js))

    def make_args(j):
        antpos, beam, sky = layouts[j.layout]
        eq, freqs = configs.job_times_freqs(j)
        return sky, beam, antpos, freqs, eq, j.nsrc

    burst = [dataclasses.replace(j, arrival_s=0.0) for j in shapes[:burst_jobs]]
    out = {"jobs": njobs, "slots": nslots, "graphs": graphs, "stream_flops": flops(shapes),
           "latency": "scheduled arrival to result in host memory",
           "load": f"Poisson arrivals at 70% of the capacity measured on a t=0 burst of the stream's first "
                   f"{burst_jobs} jobs",
           "burst": {"jobs": len(burst), "flops": flops(burst)}}
    gemms = [gemm] + (["fp32"] if fp32_jobs and gemm != "fp32" else [])
    cap = {}
    for g in gemms:
        with serve.Server(nslots=nslots, budget_bytes=SERVE_BUDGET, graphs=graphs) as srv:
            bsum, _ = serve.run_load(srv, burst, make_args, gemm=g, progress=log)
        cap[g] = bsum
        out["burst"][g] = {"jobs_per_s": len(burst) / bsum["makespan_s"], "tflops": bsum["tflops"],
                           "makespan_s": bsum["makespan_s"], "graph_hit_rate": bsum["graph_hit_rate"]}
        log(f"[bench] serving {g}: capacity {len(burst) / bsum['makespan_s']:.2f} jobs/s "
            f"({bsum['tflops']:.1f} TFLOP/s) on the {len(burst)}-job burst")
    if "fp32" in cap:
        out["burst"]["ratio_tflops"] = cap[gemm]["tflops"] / cap["fp32"]["tflops"]
    for g in gemms:
        n = njobs if g == gemm else min(fp32_jobs, njobs)
        r = rate if rate else 0.7 * len(burst) / cap[g]["makespan_s"]
        jobs = configs.make_serving_stream(n, rate_hz=r)
        log(f"[bench] serving {g}: stream of {n} jobs at {r:.2f}/s")
        with serve.Server(nslots=nslots, budget_bytes=SERVE_BUDGET, graphs=graphs) as srv:
            summ, _ = serve.run_load(srv, jobs, make_args, gemm=g, progress=log)
            summ["stats"] = srv.stats()
        summ["rate_hz"] = r
        summ["stream"] = "all jobs of the stream" if n == njobs else f"the stream's first {n} jobs (labelled)"
        summ["flops"] = flops(jobs)
        out[g] = summ
        log(f"[bench] serving {g}: p50 {summ['p50_ms']:.0f} ms p99 {summ['p99_ms']:.0f} ms "
            f"{summ['tflops']:.1f} TFLOP/s, graph hit rate {summ['graph_hit_rate']:.2f}")
    # the largest job of the mix (HERA-350 x 2^20 x 32 x 64) through the
    # server alone: arrival to result in host memory, against the same job as
    # a direct call with the result left on the device
    big = configs.JobSpec(job_id=0, arrival_s=0.0, layout="hera350", nsrc=2 ** 20, ntimes=32, nfreq=64)
    sky_b, beam_b, antpos_b, freqs_b, eq_b, _ = make_args(big)
    with serve.Server(nslots=nslots, budget_bytes=SERVE_BUDGET, graphs=graphs) as srv:
        lat = []
        for _ in range(2):   # the second run replays the graph captured by the first
            t0 = time.perf_counter()
            d = srv.wait(srv.submit(sky_b, beam_b, antpos_b, freqs_b, eq_b, nsrc=big.nsrc, gemm=gemm))
            lat.append(time.perf_counter() - t0)
            assert d.status == 0, d.error
            del d
    _, stb = stages.simulate(sky_b, beam_b, antpos_b, freqs_b, eq_b, nsrc=big.nsrc, gemm=gemm, timing=True,
                             want_output=False)
    fl = flops([big])
    out["largest_job_alone"] = {
        "job": "HERA-350, 2^20 sources, 32 times, 64 frequencies", "flops": fl,
        "served_s": lat[-1], "served_first_s": lat[0], "served_tflops": fl / lat[-1] / 1e12,
        "direct_device_s": stb["total_ms"] / 1e3, "direct_tflops": fl / (stb["total_ms"] / 1e3) / 1e12,
        "note": "served: submit to the 2 GB result in host memory (second run: graph replay); direct: "
                "hv_simulate with the result left in HBM"}
    log(f"[bench] largest job alone: served {lat[-1]:.2f} s ({fl / lat[-1] / 1e12:.1f} TFLOP/s), direct "
        f"{stb['total_ms'] / 1e3:.2f} s on the device")
    for _, b, s_ in layouts.values():
        b.close()
        s_.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hera350", choices=["hera350", "disk128", "diffuse", "hex37"])
    ap.add_argument("--gemm", default="split_bf16", choices=["split_bf16", "split_fp16", "fp32"])
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 comparison run")
    ap.add_argument("--no-fp16", action="store_true", help="skip the split-FP16 timing run")
    ap.add_argument("--cmp-steps", type=int, default=2, help="timed steps of the FP32 and split-FP16 runs")
    ap.add_argument("--cmp-warmup", type=int, default=1, help="warm-up steps of the FP32 and split-FP16 runs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 sample comparison")
    ap.add_argument("--cpu-nsrc", type=int, default=0, help="sources of the CPU baseline unit (0 = all)")
    ap.add_argument("--cpu-workers", type=int, default=0, help="CPU baseline processes (0 = cores, at most 16)")
    ap.add_argument("--no-dropin", action="store_true", help="skip the host-buffer vis_gpu timing")
    ap.add_argument("--no-omnical", action="store_true", help="skip the Omnical calibration leg")
    ap.add_argument("--no-medfilt", action="store_true", help="skip the median filter leg")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the per-frequency beam, baseline-layout and polarised-beam leg")
    ap.add_argument("--serve-jobs", type=int, default=2000, help="serving stream length (0 = skip)")
    ap.add_argument("--serve-burst", type=int, default=200, help="jobs of the capacity burst")
    ap.add_argument("--serve-fp32-jobs", type=int, default=200, help="jobs of the FP32 stream (0 = none)")
    ap.add_argument("--serve-no-graphs", action="store_true", help="serve with direct launches (A/B)")
    ap.add_argument("--serve-only", action="store_true",
                    help="profiling target: only the serving leg (prints its JSON object)")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="arrivals per second (default: 70%% of the measured serving capacity)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--herk-variant", default=None, metavar="TILE,PF",
                    help="contraction variant for every run (as hv_set_herk_variant; default: built-in)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="profiling target: copy calibration, then the workload's first time (all "
                         "frequencies) through the production kernels; prints per-kernel figures")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no "
              "multi-GPU path (time/frequency units would shard without communication); "
              "run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from hera_gpu_amd import _lib, configs, stages

    if args.herk_variant:
        tile, pf = (int(x) for x in args.herk_variant.split(","))
        _lib.check(_lib.load().hv_set_herk_variant(tile, pf), "hv_set_herk_variant")

    t0 = time.perf_counter()
    w = configs.WORKLOADS[args.workload]()
    s = w.shape
    N, K, T, F = s["N"], s["K"], s["T"], s["F"]
    units = T * F
    log(f"[bench] workload {w.name}: N={N} K={K} T={T} F={F} P={s['P']} ({time.perf_counter() - t0:.1f}s)")
    sky = stages.Sky(w.crd_eq, w.I_sky)
    beam = stages.Beam(w.bm_cube)
    flops_job = herk_flops(N, K, units)

    if args.serve_only:
        sky.close()
        beam.close()
        res = serving_leg(args.serve_jobs, args.gemm, args.serve_slots, args.serve_rate,
                          burst_jobs=min(args.serve_burst, args.serve_jobs),
                          fp32_jobs=0 if args.no_fp32 else args.serve_fp32_jobs, graphs=not args.serve_no_graphs)
        print(json.dumps({"serving": res}), flush=True)
        return

    if args.kernels_only:
        # a short, fixed sequence of dispatches for rocprofv3 counter runs:
        # 3 copies of 1 GiB (the yardstick for FETCH_SIZE / WRITE_SIZE), then
        # the workload's first two times (2 F units: one contraction launch at
        # the production batching of HERA-350) through the production kernels
        cp = stages.bench_copy(1 << 30, 3)
        nt = min(2, T)
        with job_splits(args.gemm, N, T * F, K):
            _, st = stages.simulate(sky, beam, w.antpos, w.freqs, w.eq2tops[:nt], gemm=args.gemm, timing=True,
                                    want_output=False, dtype=np.float32)
        u_launch = nt * F / st["herk_n"]
        print(json.dumps({"copy": cp, "gemm": args.gemm, "units": nt * F, "herk_launches": st["herk_n"],
                          "units_per_launch": u_launch, "herk_avg_ms": st["herk_ms"] / st["herk_n"],
                          "phasor_avg_ms": st["phasor_ms"] / st["phasor_n"],
                          "reduce_avg_ms": st["reduce_ms"] / st["reduce_n"],
                          "operand_bytes_per_launch": (16 if args.gemm == "fp64" else 8) * N * K * u_launch}),
              flush=True)
        return

    log(f"[bench] {args.gemm}: warmup {args.warmup}, timed {args.steps}")
    st, wall = run_timed(stages, sky, beam, w, args.gemm, args.steps, args.warmup)
    ms_step = st["total_ms"] / args.steps
    value = flops_job / (ms_step / 1e3) / 1e12
    herk_avg_ms = st["herk_ms"] / st["herk_n"]
    flops_per_launch = flops_job * args.steps / st["herk_n"]
    herk_tflops = flops_per_launch / (herk_avg_ms / 1e3) / 1e12
    log(f"[bench] {args.gemm}: {ms_step:.1f} ms/step, {value:.1f} TFLOP/s whole job, "
        f"herk {herk_tflops:.1f} TFLOP/s ({st['herk_n']} launches, {herk_avg_ms:.3f} ms avg)")
    phasor_avg_ms = st["phasor_ms"] / st["phasor_n"]
    elem = 16 if args.gemm == "fp64" else 8
    tile, pf = stages.herk_choice(args.gemm, N)
    if pf == 5:
        # stage 1 is the beam amplitude A sqrt(I): 4 B per (antenna, source, time)
        stage1 = "amp_kernel: A sqrt(I) per time, FP32"
        phasor_bytes = 4 * N * K * T * args.steps / st["phasor_n"]
    else:
        # phasor launch writes units_per_launch * N * K phasors of `elem` bytes
        stage1 = "phasor_kernel: operand stages"
        phasor_bytes = elem * N * K * units * args.steps / st["phasor_n"]
    stage_ms = {k: st[k] / args.steps for k in ("coords_ms", "phasor_ms", "herk_ms", "reduce_ms")}

    def compare_leg(gemm):
        log(f"[bench] {gemm} comparison: warmup {args.cmp_warmup}, timed {args.cmp_steps}")
        stc, _ = run_timed(stages, sky, beam, w, gemm, args.cmp_steps, args.cmp_warmup)
        ms = stc["total_ms"] / args.cmp_steps
        res = {"value": flops_job / (ms / 1e3) / 1e12, "ms_per_step": ms, "steps": args.cmp_steps,
               "warmup": args.cmp_warmup,
               "herk_tflops": flops_job * args.cmp_steps / stc["herk_n"] / (stc["herk_ms"] / stc["herk_n"] / 1e3)
               / 1e12, "herk_avg_ms": stc["herk_ms"] / stc["herk_n"]}
        log(f"[bench] {gemm}: {ms:.1f} ms/step, {res['value']:.1f} TFLOP/s")
        return res

    fp32 = compare_leg("fp32") if not args.no_fp32 and args.gemm != "fp32" else None
    fp16 = compare_leg("split_fp16") if not args.no_fp16 and args.gemm != "split_fp16" else None

    dropin = None
    if not args.no_dropin:
        # the drop-in call with host buffers: uploads, the same pipeline and
        # the (F, T, N, N) result back over PCIe (never the headline value)
        from hera_gpu_amd import vis_gpu
        vis_gpu(w.antpos, w.freqs[:1], w.eq2tops[:1], w.crd_eq, w.I_sky, w.bm_cube, gemm=args.gemm)
        t0 = time.perf_counter()
        vis_gpu(w.antpos, w.freqs, w.eq2tops, w.crd_eq, w.I_sky, w.bm_cube, gemm=args.gemm)
        dt = time.perf_counter() - t0
        dropin = {"seconds": dt, "TFLOP/s": flops_job / dt / 1e12,
                  "output_bytes": units * N * N * 8,
                  "note": "vis_gpu with host arrays: inputs uploaded and the result copied back inside the timing"}
        log(f"[bench] drop-in vis_gpu (host buffers): {dt:.2f} s, {dropin['TFLOP/s']:.1f} TFLOP/s")

    accuracy = None
    if not args.no_accuracy:
        # FP64 path (pinned to the oracle by the tests) on every frequency of
        # the first and the last time, all sources
        tsel = sorted({0, T - 1})
        beam64 = stages.Beam(w.bm_cube.astype(np.float64), dtype=np.float64)
        ref, _ = stages.simulate(sky, beam64, w.antpos, w.freqs, w.eq2tops[tsel], gemm="fp64", dtype=np.float32)
        accuracy = {"sample": {"times": tsel, "freqs": "all", "freqs_mhz": [round(float(x) / 1e6, 3) for x in
                                                                         w.freqs]},
                    "metric": "per frequency: ||V - V_fp64||_F / ||V_fp64||_F over the sampled times and all "
                              "baselines",
                    "reference": "FP64 GPU path (pinned to the FP64 oracle by tests)"}
        for gemm in ("split_bf16", "split_fp16", "fp32"):
            with job_splits(gemm, N, T * F, K):
                v, _ = stages.simulate(sky, beam, w.antpos, w.freqs, w.eq2tops[tsel], gemm=gemm, dtype=np.float32)
            per_f = [rel_err(v[i], ref[i]) for i in range(F)]
            accuracy[gemm] = {"max_rel_err": max(per_f), "per_freq": per_f}
            log(f"[bench] rel err {gemm}: max {max(per_f):.2e} over {F} frequencies")
            del v
        del ref
        beam64.close()

    cpu = None
    if not args.no_cpu:
        nsrc = min(args.cpu_nsrc, K) if args.cpu_nsrc else K
        workers = args.cpu_workers or min(16, len(os.sched_getaffinity(0)))
        sub = stages.Sky(w.crd_eq[:, :nsrc], w.I_sky[:nsrc])
        gpu_unit = {}
        for gemm in ("split_bf16", "fp32"):
            with job_splits(gemm, N, T * F, K):
                v, _ = stages.simulate(sub, beam, w.antpos, w.freqs[-1:], w.eq2tops[:1], gemm=gemm,
                                       dtype=np.float32)
            gpu_unit[gemm] = v[0, 0]
        sub.close()
        log(f"[bench] cpu baseline: 1 unit, {nsrc} sources, {workers} processes")
        cpu = cpu_baseline(w, nsrc, gpu_unit, workers)
        log(f"[bench] cpu: {cpu['seconds']:.1f} s, {cpu['value']:.4f} TFLOP/s")

    variants = None if args.no_variants else variants_leg(w, args.gemm)
    omni = None if args.no_omnical else omnical_leg()
    medf = None if args.no_medfilt else medfilt_leg()

    serving = None
    if args.serve_jobs > 0:
        serving = serving_leg(args.serve_jobs, args.gemm, args.serve_slots, args.serve_rate,
                              burst_jobs=min(args.serve_burst, args.serve_jobs),
                              fp32_jobs=0 if args.no_fp32 else args.serve_fp32_jobs,
                              graphs=not args.serve_no_graphs)

    peak = PEAK_TFLOPS[args.gemm]
    executed = {}
    if tile == 192 and args.gemm in ("split_bf16", "split_fp16"):
        # matrix-core flops the kernel executes: 3 products x 2 components x
        # 32x32x16 per block and 16-source step pair, against the useful ones
        per_useful = (t192_mfma_blocks(N) * 2 * 3 * 2 * 2 * 32 * 32 * 16) / (4.0 * N * (N + 1) * 16)
        executed = {"mfma_flops_per_useful": per_useful, "executed_tflops": herk_tflops * per_useful,
                    "executed_frac": herk_tflops * per_useful / peak}
    upl = units * args.steps / st["herk_n"]
    tr = MEASURED_TRAFFIC.get((args.workload, args.gemm, int(round(upl)), tile))
    traffic_note = "FETCH_SIZE x2.0 + WRITE_SIZE x1.0 of one launch at this shape, from rocprofv3 --pmc passes"
    if tr and tr["src"] != kernel_source_hash():
        tr, traffic_note = None, "stale: the contraction's sources changed since the counter passes"
    traffic = ((tr["fetch_kib"] * tr["fetch_scale"] + tr["write_kib"] * tr["write_scale"]) * 1024
               if tr and abs(upl - round(upl)) < 1e-9 else None)
    out = {
        "metric": METRIC, "value": value, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16" if args.gemm == "split_bf16" else args.gemm,
        "data": "synthetic",
        "config": {"workload": f"{w.name}: N={N} antennas, K={K} sources, T={T} times, F={F} freqs, "
                               f"beam {s['P']}x{s['P']}; one step = the whole {units}-unit simulation",
                   "gemm": args.gemm, "flops_per_step": flops_job,
                   "flop_count": "useful Hermitian 4 N (N+1) K per unit"},
        "roofline": {"bound": "mfma", "achieved": herk_tflops, "peak": peak, "unit": "TFLOP/s",
                     "frac": herk_tflops / peak, "traffic": traffic, "traffic_note": traffic_note,
                     "kernel": f"Hermitian contraction, {args.gemm}, {tile}x{tile} tiles, {variant_name(tile, pf)}",
                     "avg_launch_ms": herk_avg_ms, "launches": st["herk_n"], "units_per_launch": upl,
                     "flops_per_launch": flops_per_launch,
                     "operand_bytes_per_launch": (elem * N * K * upl) if pf != 5 else 0,
                     "flop_count": "useful 4 N (N+1) K per unit; peak = published dense MFMA rate of the "
                                   "arithmetic (split-BF16 spends 3 BF16 products per useful product)",
                     **executed},
        "stages_ms_per_step": stage_ms,
        "phasor": {"kernel": stage1, "avg_launch_ms": phasor_avg_ms, "bytes_per_launch": phasor_bytes,
                   "GBps": phasor_bytes / (phasor_avg_ms / 1e3) / 1e9, "hbm_frac": phasor_bytes /
                   (phasor_avg_ms / 1e3) / 1e9 / HBM_PEAK_GBS},
        "fp32": fp32,
        "split_fp16": fp16,
        "speedup_vs_fp32": (value / fp32["value"]) if fp32 else None,
        "accuracy": accuracy,
        "dropin_pcie_inclusive": dropin,
        "cpu_baseline": cpu,
        "serving": serving,
        "variants": variants,
        "omnical": omni,
        "medfilt": medf,
        "p99_latency_ms": serving[args.gemm]["p99_ms"] if serving else None,
        "p50_latency_ms": serving[args.gemm]["p50_ms"] if serving else None,
        "serving_tflops": serving[args.gemm]["tflops"] if serving else None,
        "serving_ratio_vs_fp32": serving["burst"].get("ratio_tflops") if serving else None,
        "wall_s_timed": wall,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

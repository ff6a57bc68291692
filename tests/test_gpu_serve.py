"""Serving loop on the GPU: continuous batching, graphs per shape bucket,
footprint admission, failure isolation, and results identical alone and in a
mix."""
import numpy as np
import pytest

from hera_gpu_amd import configs, stages
from hera_gpu_amd.serve import FAIL_TEST, Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def catalogs():
    out = {}
    for lay in ("hex37", "disk128"):
        antpos, beam, crd, flux = configs.serving_catalog(lay, nsrc_max=40000)
        out[lay] = (antpos, stages.Beam(beam), stages.Sky(crd, flux))
    yield out
    for _, b, s in out.values():
        b.close()
        s.close()


def mixed_jobs(n, seed=0):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        lay = ("hex37", "disk128")[int(rng.integers(0, 2))]
        nsrc = int(rng.integers(1, 40000))
        t = int(rng.choice([1, 2, 3]))
        f = int(rng.choice([1, 2, 5]))
        gemm = ("split_bf16", "fp32", "split_fp16")[int(rng.integers(0, 3))]
        jobs.append((i, lay, nsrc, t, f, gemm))
    return jobs


def job_inputs(cat, i, lay, t, f):
    antpos, beam, sky = cat[lay]
    eq = configs.eq2tops_lst(t, lst0_deg=13.0 * i).astype(np.float64)
    freqs = np.linspace(100e6, 200e6, f) if f > 1 else np.array([150e6 + i])
    return antpos, beam, sky, eq, freqs


def test_mixed_results_equal_alone_bitwise(catalogs):
    jobs = mixed_jobs(24)
    with Server(nslots=4) as srv:
        ids = {}
        for (i, lay, nsrc, t, f, gemm) in jobs:
            antpos, beam, sky, eq, freqs = job_inputs(catalogs, i, lay, t, f)
            ids[srv.submit(sky, beam, antpos, freqs, eq, nsrc=nsrc, gemm=gemm)] = (i, lay, nsrc, t, f, gemm)
        done = {}
        for jid in ids:
            d = srv.wait(jid, timeout_s=120)
            assert d.status == 0, d.error
            done[jid] = d
        st = srv.stats()
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert st["peak_running"] >= 2, "jobs never overlapped on the device"
    for jid, (i, lay, nsrc, t, f, gemm) in ids.items():
        antpos, beam, sky, eq, freqs = job_inputs(catalogs, i, lay, t, f)
        alone, _ = stages.simulate(sky, beam, antpos, freqs, eq, gemm=gemm, nsrc=nsrc)
        np.testing.assert_array_equal(done[jid].vis, alone, err_msg=f"job {i} {lay} {nsrc} {gemm}")


def test_graph_replay_same_bucket(catalogs):
    antpos, beam, sky, eq, freqs = job_inputs(catalogs, 3, "disk128", 2, 2)
    eq2 = configs.eq2tops_lst(2, lst0_deg=77.0).astype(np.float64)
    eq3 = configs.eq2tops_lst(2, lst0_deg=151.0).astype(np.float64)
    with Server(nslots=2) as srv:
        # one bucket (16384 < Kpad <= 32768): the first job captures the
        # graph of its instance; the next ones, one at a time so they find the
        # instance idle, replay it on whichever slot takes them
        a = srv.wait(srv.submit(sky, beam, antpos, freqs, eq, nsrc=20000), 120)
        b = srv.wait(srv.submit(sky, beam, antpos, freqs, eq2, nsrc=31000), 120)
        c = srv.wait(srv.submit(sky, beam, antpos, freqs, eq3, nsrc=17000), 120)
        st = srv.stats()
    assert a.status == 0 and b.status == 0 and c.status == 0
    assert not a.graph_hit and b.graph_hit and c.graph_hit
    assert st["direct_launches"] == 0 and st["graph_captures"] == 1 and st["graph_launches"] == 3
    assert st["instances"] == 1
    for d, e, n in ((a, eq, 20000), (b, eq2, 31000), (c, eq3, 17000)):
        alone, _ = stages.simulate(sky, beam, antpos, freqs, e, nsrc=n, gemm="split_bf16")
        np.testing.assert_array_equal(d.vis, alone)


def test_admission_by_footprint(catalogs):
    antpos, beam, sky, eq, freqs = job_inputs(catalogs, 5, "disk128", 3, 5)
    probe = Server(nslots=4)
    jid = probe.submit(sky, beam, antpos, freqs, eq, nsrc=40000)
    need = probe.wait(jid, 120).footprint
    probe.close()
    # a budget that fits one such job but not two: four slots, one job at a time
    with Server(nslots=4, budget_bytes=int(need * 1.5)) as srv:
        ids = [srv.submit(sky, beam, antpos, freqs, eq, nsrc=40000 - k) for k in range(5)]
        res = [srv.wait(i, 120) for i in ids]
        st = srv.stats()
        assert all(r.status == 0 for r in res)
        assert st["peak_running"] == 1
        # a job that can never fit is refused at submission
        big_ap, big_beam, big_sky = catalogs["disk128"]
        with pytest.raises(MemoryError):
            srv.submit(big_sky, big_beam, big_ap, np.linspace(1e8, 2e8, 64), configs.eq2tops_lst(32), nsrc=40000)


def test_failed_job_is_isolated(catalogs):
    antpos, beam, sky, eq, freqs = job_inputs(catalogs, 7, "hex37", 1, 1)
    with Server(nslots=2) as srv:
        ok1 = srv.submit(sky, beam, antpos, freqs, eq, nsrc=5000)
        bad = srv.submit(sky, beam, antpos, freqs, eq, nsrc=5000, flags=FAIL_TEST)
        ok2 = srv.submit(sky, beam, antpos, freqs, eq, nsrc=6000)
        r = {i: srv.wait(i, 60) for i in (ok1, bad, ok2)}
    assert r[ok1].status == 0 and r[ok2].status == 0
    assert r[bad].status != 0 and "injected" in r[bad].error
    with Server(nslots=1) as srv:
        with pytest.raises(AssertionError):
            srv.submit(sky, beam, antpos, freqs, eq, nsrc=10 ** 7)


def test_baseline_layout_jobs_equal_full_jobs(catalogs):
    from hera_gpu_amd import baseline_pairs
    from hera_gpu_amd.serve import FTB, TBF
    antpos, beam, sky, eq, freqs = job_inputs(catalogs, 7, "disk128", 2, 5)
    with Server(nslots=2) as srv:
        ids = [srv.submit(sky, beam, antpos, freqs, eq, nsrc=12345, flags=fl) for fl in (0, FTB, TBF)]
        full, ftb, tbf = (srv.wait(i, 120) for i in ids)
    assert full.status == 0 and ftb.status == 0 and tbf.status == 0
    a1, a2 = baseline_pairs(128)
    np.testing.assert_array_equal(ftb.vis, full.vis[:, :, a1, a2])
    np.testing.assert_array_equal(tbf.vis, full.vis[:, :, a1, a2].transpose(1, 2, 0))


def test_per_frequency_beam_jobs_equal_alone(catalogs):
    # a resident (F, N, P, P) beam: served jobs (two of one bucket, so the
    # second replays a graph) equal the direct simulation bit for bit, and a
    # job whose frequency count differs from the beam's is refused at submit
    antpos, beam3, sky, eq, freqs = job_inputs(catalogs, 9, "disk128", 2, 5)
    P = beam3.beam_px
    g = np.linspace(-1, 1, P)
    r2 = g[None, :] ** 2 + g[:, None] ** 2
    bm4 = np.stack([np.broadcast_to(np.exp(-r2 / (2 * (0.45 - 0.04 * f) ** 2)), (128, P, P))
                    for f in range(5)]).astype(np.float32)
    beam = stages.Beam(bm4)
    try:
        with Server(nslots=2) as srv:
            ids = [srv.submit(sky, beam, antpos, freqs, e, nsrc=n)
                   for e, n in ((eq, 15000), (configs.eq2tops_lst(2, lst0_deg=40.0), 16000))]
            res = [srv.wait(i, 120) for i in ids]
            with pytest.raises(AssertionError):
                srv.submit(sky, beam, antpos, freqs[:3], eq, nsrc=1000)
        for r, (e, n) in zip(res, ((eq, 15000), (configs.eq2tops_lst(2, lst0_deg=40.0), 16000))):
            assert r.status == 0
            alone, _ = stages.simulate(sky, beam, antpos, freqs, e, nsrc=n, gemm="split_bf16")
            np.testing.assert_array_equal(r.vis, alone)
    finally:
        beam.close()


def test_large_outputs_prefaulted_equal_alone_bitwise(catalogs):
    # outputs of 42 MB (N = 128, 8 times x 40 frequencies, few sources): the
    # server pre-faults their pages from submission on, in 32 MB chunks, while
    # earlier jobs run; one job fails at admission with its output still being
    # pre-faulted, and the others are unaffected
    antpos, beam, sky = catalogs["disk128"]
    freqs = np.linspace(100e6, 200e6, 40)
    specs = [(i, 500 + 700 * i, configs.eq2tops_lst(8, lst0_deg=7.0 * i).astype(np.float64)) for i in range(6)]
    with Server(nslots=2) as srv:
        ids = {srv.submit(sky, beam, antpos, freqs, eq, nsrc=nsrc, gemm="split_bf16"): (i, nsrc, eq)
               for i, nsrc, eq in specs}
        bad = srv.submit(sky, beam, antpos, freqs, specs[0][2], nsrc=100, gemm="split_bf16", flags=FAIL_TEST)
        assert srv.wait(bad, timeout_s=120).status != 0
        done = {jid: srv.wait(jid, timeout_s=120) for jid in ids}
    for jid, (i, nsrc, eq) in ids.items():
        assert done[jid].status == 0, done[jid].error
        alone, _ = stages.simulate(sky, beam, antpos, freqs, eq, gemm="split_bf16", nsrc=nsrc)
        np.testing.assert_array_equal(done[jid].vis, alone, err_msg=f"job {i}")


def test_empty_source_set_job_gives_zero(catalogs):
    # a job over the first 0 sources of the catalog: the padded directions are
    # zero (not above the horizon), so every visibility is exactly zero, also
    # after a job whose sources filled the same slot
    antpos, beam, sky, eq, freqs = job_inputs(catalogs, 3, "hex37", 2, 2)
    with Server(nslots=1) as srv:
        full = srv.wait(srv.submit(sky, beam, antpos, freqs, eq, nsrc=30000), 60)
        empty = srv.wait(srv.submit(sky, beam, antpos, freqs, eq, nsrc=0), 60)
    assert full.status == 0 and empty.status == 0, (full.error, empty.error)
    assert np.abs(full.vis).max() > 0
    np.testing.assert_array_equal(empty.vis, 0)


def test_cancel_queued_and_running(catalogs):
    # one slot: the first job runs, the others queue behind it; cancelling a
    # queued job finishes it at once, cancelling the running one drops its
    # result when its kernels end; the rest equal their stand-alone runs
    antpos, beam, sky, eq, freqs = job_inputs(catalogs, 4, "disk128", 3, 5)
    with Server(nslots=1) as srv:
        ids = [srv.submit(sky, beam, antpos, freqs, eq, nsrc=39000 - 1000 * k) for k in range(4)]
        srv.cancel(ids[2])
        srv.cancel(ids[0])
        res = {i: srv.wait(i, 120) for i in ids}
        with pytest.raises(AssertionError):
            srv.cancel(ids[1])            # finished and collected
        st = srv.stats()
    assert res[ids[2]].status == -6 and "before it started" in res[ids[2]].error
    assert res[ids[0]].status in (0, -6)  # it may have finished before the request
    assert st["cancelled"] >= 1
    for k in (1, 3):
        d = res[ids[k]]
        assert d.status == 0, d.error
        alone, _ = stages.simulate(sky, beam, antpos, freqs, eq, nsrc=39000 - 1000 * k, gemm="split_bf16")
        np.testing.assert_array_equal(d.vis, alone)


# ---------------------------------------------------------------------------
# The serving load of the fifth workload: jobs drawn from
# configs.make_serving_stream (N in {37, 128, 350}, Nsrc log-uniform in
# [10^3, 2^20], T in {1, 8, 32}, F in {1, 16, 64}, beams shared per layout).
# Selection from the first 400 jobs of the seed-5 stream, in stream order:
# the first HERA-350 job with Nsrc >= 2^19, the first HERA-350 job with
# T = 32, the first with F = 64, the first with T = 8 and F = 16, then the
# remaining jobs in stream order whose output is at most 256 MB, up to 18.
SERVE_TOL = {"split_bf16": 1e-5}   # relative Frobenius error vs the FP64 oracle
SAMPLE_ANTS = {37: np.arange(37), 128: np.r_[0:20, 108:128],
               350: np.r_[0:10, 160:170, 330:350]}   # 350: core, centre and the 20 outermost (outriggers)


def _config5_jobs():
    stream = configs.make_serving_stream(400, rate_hz=50.0)
    nant = {"hex37": 37, "disk128": 128, "hera350": 350}
    picks = []
    rules = [lambda j: j.layout == "hera350" and j.nsrc >= 2 ** 19,
             lambda j: j.layout == "hera350" and j.ntimes == 32,
             lambda j: j.nfreq == 64,
             lambda j: j.ntimes == 8 and j.nfreq == 16]
    for rule in rules:
        j = next(j for j in stream if rule(j) and j not in picks)
        picks.append(j)
    for j in stream:
        if len(picks) >= 18:
            break
        out = nant[j.layout] ** 2 * j.ntimes * j.nfreq * 8
        if j not in picks and out <= 256 << 20:
            picks.append(j)
    return sorted(picks, key=lambda j: j.job_id), nant


@pytest.fixture(scope="module")
def full_catalogs():
    out = {}
    for lay in ("hex37", "disk128", "hera350"):
        antpos, beam, crd, flux = configs.serving_catalog(lay)
        out[lay] = (antpos, beam, crd, flux, stages.Beam(beam), stages.Sky(crd, flux))
    yield out
    for v in out.values():
        v[4].close()
        v[5].close()


def test_config5_mix_equals_alone_and_oracle(full_catalogs):
    from oracle.vis_oracle import rel_err, vis_oracle_units

    jobs, nant = _config5_jobs()
    assert sum(j.layout == "hera350" for j in jobs) >= 3
    assert any(j.layout == "hera350" and j.nsrc >= 2 ** 19 for j in jobs)
    assert {8, 32} <= {j.ntimes for j in jobs} and {16, 64} <= {j.nfreq for j in jobs}
    with Server(nslots=4) as srv:
        ids = {}
        for j in jobs:
            antpos, _, _, _, beam, sky = full_catalogs[j.layout]
            eq, freqs = configs.job_times_freqs(j)
            ids[srv.submit(sky, beam, antpos, freqs, eq, nsrc=j.nsrc, gemm="split_bf16")] = j
        done = {jid: srv.wait(jid, 300) for jid in ids}
        st = srv.stats()
    assert st["peak_running"] >= 2 and st["failed"] == 0
    for jid, j in ids.items():
        d = done[jid]
        assert d.status == 0, d.error
        antpos, bm, crd, flux, beam, sky = full_catalogs[j.layout]
        eq, freqs = configs.job_times_freqs(j)
        alone, _ = stages.simulate(sky, beam, antpos, freqs, eq, nsrc=j.nsrc, gemm="split_bf16")
        np.testing.assert_array_equal(d.vis, alone, err_msg=f"job {j.job_id}")
        # first and last (frequency, time) unit on the antenna sample
        S = SAMPLE_ANTS[nant[j.layout]]
        units = sorted({(0, 0), (j.nfreq - 1, j.ntimes - 1)})
        ref = vis_oracle_units(antpos[S], freqs, eq, crd[:, :j.nsrc], flux[:j.nsrc], bm[S], units)
        for k, (f, t) in enumerate(units):
            err = rel_err(d.vis[f, t][np.ix_(S, S)], ref[k])
            assert err < SERVE_TOL["split_bf16"], f"job {j.job_id} unit ({f}, {t}): rel err {err:.2e}"
        del d.vis, alone


@pytest.fixture
def ksplit_items_1000():
    # the K-split rule's work-item threshold lowered from 32768 to 1000, so
    # HERA-350 jobs of two times take 32768-source splits from 98304 sources
    from hera_gpu_amd import _lib
    lib = _lib.load()
    _lib.check(lib.hv_set_ksplit_items(1000))
    try:
        yield
    finally:
        _lib.check(lib.hv_set_ksplit_items(0))


def test_long_split_jobs_served_equal_alone(ksplit_items_1000):
    # HERA-350 jobs of 2 times x 64 frequencies: 70000 sources keep
    # 8192-source splits (768 work items with long ones), 100000 and 120000
    # take 32768-source splits (1024).  All three share the bucket of padded
    # sizes up to 131072, so the split length in the bucket key keeps the
    # first apart; the third replays the second's graph.  Served results
    # equal the stand-alone runs bit for bit and match the oracle.
    from oracle.vis_oracle import rel_err, vis_oracle_units
    antpos, bm, crd, flux = configs.serving_catalog("hera350", nsrc_max=131072)
    beam, sky = stages.Beam(bm), stages.Sky(crd, flux)
    freqs = np.linspace(100e6, 200e6, 64)
    eq = configs.eq2tops_lst(2, lst0_deg=40.0).astype(np.float64)
    eq2 = configs.eq2tops_lst(2, lst0_deg=95.0).astype(np.float64)
    assert stages.ksplit_choice("split_bf16", 350, 128, 70000) == 8192
    assert stages.ksplit_choice("split_bf16", 350, 128, 100000) == 32768
    try:
        with Server(nslots=2) as srv:
            a = srv.wait(srv.submit(sky, beam, antpos, freqs, eq, nsrc=70000), 300)
            b = srv.wait(srv.submit(sky, beam, antpos, freqs, eq, nsrc=100000), 300)
            c = srv.wait(srv.submit(sky, beam, antpos, freqs, eq2, nsrc=120000), 300)
            st = srv.stats()
        assert a.status == 0 and b.status == 0 and c.status == 0
        assert not b.graph_hit and c.graph_hit
        assert st["instances"] >= 2
        for d, e, n in ((a, eq, 70000), (b, eq, 100000), (c, eq2, 120000)):
            alone, _ = stages.simulate(sky, beam, antpos, freqs, e, nsrc=n, gemm="split_bf16")
            np.testing.assert_array_equal(d.vis, alone, err_msg=f"{n} sources")
            del alone
        S = SAMPLE_ANTS[350]
        ref = vis_oracle_units(antpos[S], freqs, eq2, crd[:, :120000], flux[:120000], bm[S], [(63, 1)])
        assert rel_err(c.vis[63, 1][np.ix_(S, S)], ref[0]) < SERVE_TOL["split_bf16"]
    finally:
        beam.close()
        sky.close()

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
    with Server(nslots=1) as srv:
        # one bucket (16384 < Kpad <= 32768): the first job runs direct, the
        # second is captured, the third replays the graph
        a = srv.wait(srv.submit(sky, beam, antpos, freqs, eq, nsrc=20000), 120)
        b = srv.wait(srv.submit(sky, beam, antpos, freqs, eq2, nsrc=31000), 120)
        c = srv.wait(srv.submit(sky, beam, antpos, freqs, eq3, nsrc=17000), 120)
        st = srv.stats()
    assert a.status == 0 and b.status == 0 and c.status == 0
    assert not a.graph_hit and not b.graph_hit and c.graph_hit
    assert st["direct_launches"] == 1 and st["graph_captures"] == 1 and st["graph_launches"] == 2
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

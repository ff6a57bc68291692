"""Host-side serving logic: latency summary and percentiles (no GPU)."""
import numpy as np

from hera_gpu_amd.serve import Done, percentile, summarize


def test_percentile_matches_numpy():
    x = np.random.default_rng(0).exponential(size=1001)
    assert percentile(x, 50) == np.percentile(x, 50)
    assert percentile(x, 99) == np.percentile(x, 99)
    assert np.isnan(percentile([], 50))


def test_summary_uses_arrival_to_result_latency():
    t0 = 100.0
    arrivals = {1: 0.0, 2: 1.0, 3: 2.0}
    done = [Done(1, 0, 0, False, 100.0, 100.1, 100.5, 10, 1e12, ""),
            Done(2, 0, 1, True, 101.0, 101.0, 101.2, 10, 1e12, ""),
            Done(3, -4, 0, False, 102.0, 102.0, 102.1, 10, 1e12, "injected")]
    s = summarize(done, arrivals, t0, 0.0)
    assert s["ok"] == 2 and s["failed"] == 1
    assert abs(s["p50_ms"] - 350.0) < 1e-6
    assert abs(s["makespan_s"] - 1.2) < 1e-9
    assert abs(s["tflops"] - 2e12 / 1.2 / 1e12) < 1e-9
    assert s["graph_hit_rate"] == 0.5

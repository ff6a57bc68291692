"""Continuous-batching simulation server and a Poisson load driver.

``Server`` wraps the native server of libheravis (hv_server_*): jobs name a
resident sky catalog and beam (``stages.Sky`` / ``stages.Beam``) and bring
their own antenna positions, times and frequencies.  The native worker admits
queued jobs into free slots as slots and device memory free up, runs them
concurrently on per-slot streams.  Each job runs in an instance of its shape
bucket (device buffers at fixed addresses plus the hipGraph captured from the
first job that ran there); later jobs of the bucket replay that graph on
whichever slot admits them.

``run_load`` drives a timed stream of jobs (arrival times in seconds) through
a server and reports latency percentiles (arrival to result on the host) and
the aggregate useful GEMM rate.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, stages


JobInfo = _lib.JobInfo
Job = _lib.Job
ServerStats = _lib.ServerStats
FAIL_TEST = 1     # HV_JOB_FAIL_TEST
FTB = 2           # HV_JOB_FTB: result (F, T, Nbl), baselines a <= b
TBF = 4           # HV_JOB_TBF: result (T, Nbl, F)


def _bind():
    return _lib.load()


@dataclass
class Done:
    id: int
    status: int
    slot: int
    graph_hit: bool
    t_submit: float
    t_admit: float
    t_done: float
    footprint: int
    flops: float
    error: str
    vis: np.ndarray = field(repr=False, default=None)


class Server:
    def __init__(self, nslots: int = 4, budget_bytes: int = 0, graphs: bool = True, device: int = 0):
        self.lib = _bind()
        self.ctx = _lib.context(device)
        h = C.c_void_p()
        _lib.check(self.lib.hv_server_create(self.ctx.handle, nslots, budget_bytes, int(graphs), C.byref(h)),
                   "hv_server_create")
        self.handle = h
        self._keep = {}        # id -> (host arrays kept alive until the job finishes)
        self.ctx.adopt(self)

    def submit(self, sky: stages.Sky, beam: stages.Beam, antpos, freqs, eq2tops, nsrc=None, gemm="split_bf16",
               flags=0) -> int:
        mode = _lib.MODES[gemm] if isinstance(gemm, str) else int(gemm)
        ap = np.ascontiguousarray(antpos, dtype=np.float64)
        fr = np.ascontiguousarray(np.atleast_1d(freqs), dtype=np.float64)
        eq = np.ascontiguousarray(eq2tops, dtype=np.float64)
        n, t, f = ap.shape[0], eq.shape[0], fr.shape[0]
        cd = np.complex128 if mode == _lib.MODES["fp64"] else np.complex64
        nbl = n * (n + 1) // 2
        if flags & TBF:
            shape = (t, nbl, f)
        elif flags & FTB:
            shape = (f, t, nbl)
        else:
            shape = (f, t, n, n)
        out = np.empty(shape, dtype=cd)
        job = Job(mode, flags, sky.handle, sky.nsrc if nsrc is None else int(nsrc), beam.handle,
                  ap.ctypes.data, n, fr.ctypes.data, f, eq.ctypes.data, t, out.ctypes.data)
        jid = C.c_int64()
        _lib.check(self.lib.hv_server_submit(self.handle, C.byref(job), C.byref(jid)), "hv_server_submit")
        self._keep[jid.value] = (ap, fr, eq, out)
        return jid.value

    def _done(self, info: JobInfo) -> Done:
        arrs = self._keep.pop(info.id, None)
        return Done(info.id, info.status, info.slot, bool(info.graph_hit), info.t_submit, info.t_admit,
                    info.t_done, info.footprint, info.flops, info.error.decode(errors="replace"),
                    arrs[3] if arrs is not None else None)

    def poll(self, max_jobs: int = 256) -> list[Done]:
        buf = (JobInfo * max_jobs)()
        n = C.c_int()
        _lib.check(self.lib.hv_server_poll(self.handle, buf, max_jobs, C.byref(n)), "hv_server_poll")
        return [self._done(buf[i]) for i in range(n.value)]

    def wait(self, jid: int, timeout_s: float = -1.0) -> Done:
        info = JobInfo()
        _lib.check(self.lib.hv_server_wait(self.handle, jid, timeout_s, C.byref(info)), "hv_server_wait")
        return self._done(info)

    def cancel(self, jid: int) -> None:
        """Cancel a queued or running job (it then finishes with HV_ERR_CANCELED)."""
        _lib.check(self.lib.hv_server_cancel(self.handle, jid), "hv_server_cancel")

    def stats(self) -> dict:
        st = ServerStats()
        _lib.check(self.lib.hv_server_stats(self.handle, C.byref(st)), "hv_server_stats")
        return st.as_dict()

    def close(self):
        if self.handle:
            if self.ctx.handle:
                self.lib.hv_server_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def percentile(xs, q):
    xs = np.sort(np.asarray(xs, dtype=np.float64))
    if xs.size == 0:
        return float("nan")
    return float(np.percentile(xs, q, method="linear"))


def summarize(done: list[Done], arrivals: dict, t0: float, flops_total: float):
    """Latency = result on host minus scheduled arrival; rate = useful flops / makespan."""
    ok = [d for d in done if d.status == 0]
    lat = [d.t_done - (t0 + arrivals[d.id]) for d in ok]
    queue = [d.t_admit - (t0 + arrivals[d.id]) for d in ok]
    end = max((d.t_done for d in ok), default=t0)
    makespan = end - t0
    return {
        "jobs": len(done), "ok": len(ok), "failed": len(done) - len(ok),
        "p50_ms": 1e3 * percentile(lat, 50), "p99_ms": 1e3 * percentile(lat, 99),
        "mean_ms": 1e3 * float(np.mean(lat)) if lat else float("nan"),
        "max_ms": 1e3 * float(np.max(lat)) if lat else float("nan"),
        "p50_queue_ms": 1e3 * percentile(queue, 50), "p99_queue_ms": 1e3 * percentile(queue, 99),
        "makespan_s": makespan,
        "tflops": sum(d.flops for d in ok) / makespan / 1e12 if makespan > 0 else float("nan"),
        "graph_hit_rate": (sum(d.graph_hit for d in ok) / len(ok)) if ok else float("nan"),
    }


def run_load(server: Server, jobs, make_args, gemm="split_bf16", keep_results=False, progress=None):
    """Submit ``jobs`` at their ``arrival_s`` offsets and collect every result.

    ``make_args(job)`` returns (sky, beam, antpos, freqs, eq2tops, nsrc).
    Returns (summary dict, list of Done).
    """
    order = sorted(jobs, key=lambda j: j.arrival_s)
    arrivals = {}
    done: list[Done] = []
    flops_total = 0.0
    t0 = time.monotonic()
    i = 0
    last_print = t0
    while i < len(order) or len(done) < len(order):
        now = time.monotonic()
        while i < len(order) and t0 + order[i].arrival_s <= now:
            j = order[i]
            sky, beam, antpos, freqs, eq2tops, nsrc = make_args(j)
            jid = server.submit(sky, beam, antpos, freqs, eq2tops, nsrc=nsrc, gemm=gemm)
            arrivals[jid] = j.arrival_s
            i += 1
        for d in server.poll():
            if not keep_results:
                d.vis = None
            done.append(d)
        if progress and now - last_print > 10.0:
            progress(f"[serve] t={now - t0:.0f}s submitted {i}/{len(order)} done {len(done)}")
            last_print = now
        nxt = t0 + order[i].arrival_s if i < len(order) else now + 1e-3
        time.sleep(max(0.0, min(nxt - time.monotonic(), 1e-3)))
    return summarize(done, arrivals, t0, flops_total), done

"""Teardown order of the Python handles: catalogs, beams and servers are
closed before the context whose hv_ctx their C objects point to, whatever
order the interpreter collects them in at exit (Context.adopt).  Runs on the
CPU with a stand-in for the library's entry points that records the calls."""
import gc

import numpy as np
import pytest

from hera_gpu_amd import _lib, serve, stages


class RecordingLib:
    """The entry points Context, Sky, Beam and Server call, recording each."""

    def __init__(self):
        self.calls = []
        self.next = 100

    def _handle(self, out, what):
        self.next += 1
        out._obj.value = self.next
        self.calls.append((what, self.next))
        return _lib.HV_OK

    def hv_ctx_create(self, device, budget, out):
        return self._handle(out, "ctx_create")

    def hv_ctx_destroy(self, h):
        self.calls.append(("ctx_destroy", h.value))
        return _lib.HV_OK

    def hv_sky_upload(self, ctx, code, crd, isq, n, out):
        return self._handle(out, "sky_upload")

    def hv_sky_free(self, h):
        self.calls.append(("sky_free", h.value))
        return _lib.HV_OK

    def hv_beam_upload_freq(self, ctx, code, bm, nf, nant, px, out):
        return self._handle(out, "beam_upload")

    def hv_beam_free(self, h):
        self.calls.append(("beam_free", h.value))
        return _lib.HV_OK

    def hv_server_create(self, ctx, nslots, budget, graphs, out):
        return self._handle(out, "server_create")

    def hv_server_destroy(self, h):
        self.calls.append(("server_destroy", h.value))
        return _lib.HV_OK

    def names(self):
        return [c[0] for c in self.calls]


@pytest.fixture
def fake(monkeypatch):
    lib = RecordingLib()
    monkeypatch.setattr(_lib, "_lib", lib)
    monkeypatch.setattr(_lib, "_contexts", {})
    return lib


def make_objects():
    sky = stages.Sky(np.zeros((3, 64)), np.ones(64))
    beam = stages.Beam(np.ones((2, 5, 5)))
    srv = serve.Server(nslots=1)
    return sky, beam, srv


def test_context_close_closes_live_handles_first(fake):
    sky, beam, srv = make_objects()
    _lib._close_all()                       # what the atexit hook runs
    assert fake.names() == ["ctx_create", "sky_upload", "beam_upload", "server_create",
                            "server_destroy", "beam_free", "sky_free", "ctx_destroy"]
    n = len(fake.calls)
    # later close() calls and collection touch nothing
    sky.close()
    srv.close()
    del sky, beam, srv
    gc.collect()
    assert len(fake.calls) == n


def test_handles_collected_before_context_free_once(fake):
    sky, beam, srv = make_objects()
    del srv
    gc.collect()                             # Server.__del__ destroys it
    sky.close()
    del beam
    gc.collect()
    _lib._close_all()
    assert fake.names() == ["ctx_create", "sky_upload", "beam_upload", "server_create",
                            "server_destroy", "sky_free", "beam_free", "ctx_destroy"]
    assert fake.names().count("ctx_destroy") == 1


def test_closed_context_frees_nothing_later(fake):
    sky, beam, _ = make_objects()
    ctx = sky.ctx
    ctx.close()
    n = len(fake.calls)
    assert ctx.handle is None and sky.handle is None and beam.handle is None
    sky.close()
    beam.close()
    assert len(fake.calls) == n

"""Bounded device waits and edge-case batches on the GPU.

Every search, and every wait inside it, has a wall-clock limit; with
GCL_TIME_LIMIT_MS set to 1 ms the limit is forced from the host, so that
the workers give up early: every kernel still exits normally, the call
reports GCL_E_TIMEOUT, and the next call on the same device runs normally."""
import sys

import numpy as np
import pytest

import gcl

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _batch():
    # graphs whose searches take well over 1 ms and publish root children
    gs = [gcl.random_regular(1200 + 2 * i, 3, 40 + i) for i in range(6)]
    gs += [gcl.torus(20, 30), gcl.cfi(60, True, 4)]
    return gcl.Batch(gs)


def test_forced_time_limit_reports_timeout_then_recovers(monkeypatch):
    b = _batch()
    _log("limits: reference run without a limit")
    ref = gcl.batch_canon(b, engine="gpu")
    monkeypatch.setenv("GCL_TIME_LIMIT_MS", "1")
    _log("limits: forced 1 ms limit")
    with pytest.raises(gcl.GclError) as ei:
        gcl.batch_canon(b, engine="gpu")
    assert ei.value.status == -9
    with pytest.raises(gcl.GclError) as ei:
        gcl.canon_label(gcl.random_regular(3000, 3, 1), engine="gpu")
    assert ei.value.status == -9
    monkeypatch.delenv("GCL_TIME_LIMIT_MS")
    _log("limits: run again without a limit")
    again = gcl.batch_canon(b, engine="gpu")
    assert np.array_equal(again.labels, ref.labels)
    assert np.array_equal(again.nodes, ref.nodes)


def test_empty_batch_on_device():
    import ctypes as C
    b = gcl._Batch(0, None, None, None, None)
    h = C.c_void_p()
    st = gcl._BatchStats()
    assert gcl.lib().gcl_device_batch_create(C.byref(b), 10, C.byref(h)) == 0
    assert gcl.lib().gcl_device_batch_run(h, 0, C.byref(st)) == 0
    assert gcl.lib().gcl_device_batch_run(h, 1, C.byref(st)) == 0  # captured graph of an empty run
    assert gcl.lib().gcl_device_batch_results(h, None, None, C.byref(st)) == 0
    gcl.lib().gcl_device_batch_destroy(h)
    lab = (C.c_int32 * 1)()
    assert gcl.lib().gcl_batch_canon(C.byref(b), 10, 0, lab, None, C.byref(st)) == 0

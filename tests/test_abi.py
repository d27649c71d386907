"""The C-ABI library loads and exports every entry point include/gcl.h
declares (no device calls)."""
import os
import re
import subprocess

import gcl
from conftest import REPO


def declared():
    text = open(os.path.join(REPO, "include", "gcl.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gcl_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("gcl_nauty_parallel", "gcl_nauty_sequential", "gcl_canon_label", "gcl_is_isomorphic",
                 "gcl_batch_canon", "gcl_read_dimacs"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = gcl.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", gcl.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    for name in declared():
        assert name in exported, name
        assert hasattr(lib, name)
        assert name in gcl.exported_symbols(), name


def test_status_strings_and_errors():
    lib = gcl.lib()
    assert lib.gcl_status_string(0) == b"ok"
    bad = gcl.Graph(3, [0, 1, 1, 1], [1])  # asymmetric
    try:
        gcl.validate(bad)
        raise AssertionError("expected GclError")
    except gcl.GclError as e:
        assert e.status == -1 and "symmetric" in str(e)


def test_invalid_arguments_rejected_without_device():
    g = gcl.torus(3, 3)
    for kw in ({"max_aut": 0},):
        try:
            gcl.canon_label(g, engine="cpu", **kw)
            raise AssertionError("expected GclError")
        except gcl.GclError as e:
            assert e.status == -1
    empty = gcl.Graph(0, [0], [])
    try:
        gcl.canon_label(empty, engine="cpu")
        raise AssertionError("expected GclError")
    except gcl.GclError as e:
        assert e.status == -1


def test_server_config_defaults_and_no_device():
    c = gcl.server_config()
    assert (c.slots, c.max_n, c.max_aut, c.lifetime_ms) == (0, 4096, 10000, 0)
    assert c.max_graph_nnz > 0 and c.queue_bytes > 0
    if gcl.device_count() == 0:
        try:
            gcl.Server()
            raise AssertionError("expected GclError")
        except gcl.GclError as e:
            assert e.status == -4  # no device


def test_batch_tuning_arguments_rejected_without_device(monkeypatch):
    monkeypatch.delenv("GCL_SPEC_MIN_N", raising=False)
    assert gcl.batch_tuning_default() == {"spec_min_n": 128, "min_class_n": 64}
    batch = gcl.Batch([gcl.torus(3, 3), gcl.torus(4, 4)])
    bad = ({"spec_min_n": -1, "min_class_n": 64}, {"spec_min_n": 128, "min_class_n": 100})
    for t in bad:
        try:
            gcl.DeviceBatch(batch, tuning=t)
            raise AssertionError("expected GclError")
        except gcl.GclError as e:
            assert e.status == -1, t
    try:
        gcl.autotune(batch, sample=0)
        raise AssertionError("expected GclError")
    except gcl.GclError as e:
        assert e.status == -1


def test_dimacs_vertex_ids_below_one(tmp_path):
    """'n' lines with id 0 or a negative id are ignored (the reference reads
    the id as unsigned, so a negative one falls outside 1..n); an edge with a
    negative id is an error."""
    f = tmp_path / "g.dimacs"
    f.write_text("p edge 3 2\nn 0 9\nn -5 7\nn 2 3\ne 1 2\ne 2 3\n")
    g = gcl.read_dimacs(str(f))
    assert g.n == 3 and list(g.color) == [0, 3, 0]
    f.write_text("p edge 3 1\ne -1 2\n")
    try:
        gcl.read_dimacs(str(f))
        raise AssertionError("expected GclError")
    except gcl.GclError as e:
        assert e.status == -1


def test_empty_batch_without_arrays():
    """A batch of zero graphs may come with null arrays: no entry point reads
    them (on a machine without a GPU the device entry points report no
    device)."""
    import ctypes as C
    b = gcl._Batch(0, None, None, None, None)
    h = C.c_void_p()
    st = gcl._BatchStats()
    r = gcl.lib().gcl_device_batch_create(C.byref(b), 10, C.byref(h))
    if gcl.device_count() == 0:
        assert r == -4
    else:
        assert r == 0
        assert gcl.lib().gcl_device_batch_run(h, 0, C.byref(st)) == 0
        gcl.lib().gcl_device_batch_destroy(h)
    assert gcl.lib().gcl_batch_canon(C.byref(b), 10, 1, None, None, C.byref(st)) == -1  # null labels
    lab = (C.c_int32 * 1)()
    assert gcl.lib().gcl_batch_canon(C.byref(b), 10, 1, lab, None, C.byref(st)) == 0

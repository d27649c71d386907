"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke() entry and
bench.py's cpu_baseline leg, as the checker.  No product code uses it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_BIN = os.path.join(_HERE, "_ref", "gpu_canon_label_ref")


class Stats(C.Structure):
    _fields_ = [("nodes", C.c_uint64), ("automorphisms", C.c_uint64), ("refinements", C.c_uint64),
                ("splitter_steps", C.c_uint64), ("stale_classify", C.c_uint64), ("alg_bytes", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.oracle_canon.restype = C.c_int
        L.oracle_canon.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        L.oracle_refine.restype = C.c_int
        L.oracle_refine.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.oracle_format_adjacency.restype = C.c_int64
        L.oracle_format_adjacency.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]
        L.oracle_compare_upper.restype = C.c_int
        L.oracle_compare_upper.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def canon(n, row_ptr, col, color=None, max_aut=10000, want_dense=True, aut_cap=0):
    """Runs the oracle.  Returns dict(dense, label, nodes, automorphisms, auts, ...)."""
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    color = None if color is None else np.ascontiguousarray(color, np.int32)
    dense = np.zeros(n * n, np.uint8) if want_dense else None
    label = np.zeros(max(n, 1), np.int32)
    auts = np.zeros(max(aut_cap, 1) * max(n, 1), np.int32) if aut_cap else None
    st = Stats()
    rc = lib().oracle_canon(n, _p(row_ptr), _p(col) if col.size else None, _p(color), max_aut, _p(dense),
                            _p(label), _p(auts), aut_cap, C.byref(st))
    if rc != 0:
        raise RuntimeError("oracle_canon failed")
    out = dict(label=label[:n], nodes=st.nodes, automorphisms=st.automorphisms, refinements=st.refinements,
               splitter_steps=st.splitter_steps, stale_classify=st.stale_classify, alg_bytes=st.alg_bytes)
    if want_dense:
        out["dense"] = dense.reshape(n, n)
        out["keys"] = np.nonzero(dense)[0].astype(np.uint32)
    if aut_cap:
        stored = min(int(st.automorphisms), max_aut, aut_cap)
        out["auts"] = auts[:stored * n].reshape(stored, n)
    return out


def canon_graph(g, max_aut=10000, **kw):
    return canon(g.n, g.row_ptr, g.col, g.color, max_aut, **kw)


def refine(n, row_ptr, col, elem, cells, level, active=None):
    """One refinement call; cells = list of (first, length, in_level)."""
    elem = np.ascontiguousarray(elem, np.uint32).copy()
    cf = np.zeros(n + 1, np.uint32)
    cl = np.zeros(n + 1, np.uint32)
    ci = np.zeros(n + 1, np.uint32)
    for i, (f, l, lv) in enumerate(cells):
        cf[i], cl[i], ci[i] = f, l, lv
    nc = C.c_uint32(len(cells))
    lev = C.c_uint32(level)
    act = None if active is None else np.ascontiguousarray(active, np.uint32)
    code = lib().oracle_refine(n, _p(np.ascontiguousarray(row_ptr, np.int32)), _p(np.ascontiguousarray(col, np.int32)),
                               _p(elem), _p(cf), _p(cl), _p(ci), C.byref(nc), C.byref(lev), _p(act),
                               -1 if act is None else len(act))
    cells_out = [(int(cf[i]), int(cl[i]), int(ci[i])) for i in range(nc.value)]
    return code, elem, cells_out, lev.value


def format_adjacency(dense):
    n = dense.shape[0]
    d = np.ascontiguousarray(dense, np.uint8)
    ln = lib().oracle_format_adjacency(_p(d), n, None, 0)
    buf = C.create_string_buffer(int(ln) + 1)
    lib().oracle_format_adjacency(_p(d), n, buf, ln + 1)
    return buf.raw[:ln].decode()


def compare_upper(a, b):
    n = a.shape[0]
    return lib().oracle_compare_upper(_p(np.ascontiguousarray(a, np.uint8)), _p(np.ascontiguousarray(b, np.uint8)), n)

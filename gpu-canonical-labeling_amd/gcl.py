"""Python binding of libgcl (ctypes over the C ABI of include/gcl.h).

The library is the product: this module only marshals numpy arrays.  It
raises GclError when libgcl.so has not been built or a call fails; there is no
Python fallback for any computation.

    import gcl
    g = gcl.erdos_renyi(128, 0.08, seed=0)
    res = gcl.canon_label(g, engine="gpu")        # reference --parallel
    res.keys      # sorted i*n+j keys of the canonical adjacency
    res.label     # label[i] = input vertex at canonical position i
"""
import ctypes as C
import os
import time
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgcl.so")
CLI_PATH = os.path.join(_HERE, "gpu_canon_label")

GCL_OK = 0
_STATUS = {0: "ok", -1: "invalid argument", -2: "out of memory", -3: "device error", -4: "no device",
           -5: "i/o error", -6: "capacity exceeded", -7: "internal error", -8: "busy",
           -9: "time limit exceeded"}
ENGINES = {"gpu": 0, "parallel": 0, "cpu": 1, "sequential": 1}


class GclError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status


class _CSR(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("row_ptr", C.c_void_p), ("col", C.c_void_p),
                ("color", C.c_void_p)]


class _Graph(C.Structure):
    _fields_ = [("csr", _CSR), ("row_ptr_buf", C.c_void_p), ("col_buf", C.c_void_p), ("color_buf", C.c_void_p)]


class _Stats(C.Structure):
    _fields_ = [("nodes", C.c_uint64), ("automorphisms", C.c_uint64), ("refinements", C.c_uint64),
                ("splitter_steps", C.c_uint64), ("csr_bytes", C.c_uint64), ("kernel_ms", C.c_double),
                ("total_ms", C.c_double), ("alg_bytes", C.c_uint64)]


class _Batch(C.Structure):
    _fields_ = [("num_graphs", C.c_int32), ("vtx_off", C.c_void_p), ("row_ptr", C.c_void_p), ("col", C.c_void_p),
                ("color", C.c_void_p)]


class _BatchStats(C.Structure):
    _fields_ = [("nodes", C.c_uint64), ("automorphisms", C.c_uint64), ("splitter_steps", C.c_uint64),
                ("csr_bytes", C.c_uint64), ("kernel_ms", C.c_double), ("total_ms", C.c_double),
                ("workers", C.c_int32), ("pad", C.c_int32), ("alg_bytes", C.c_uint64)]


class _ClassStats(C.Structure):
    _fields_ = [("max_n", C.c_int32), ("graphs", C.c_int32), ("workers", C.c_int32), ("pad", C.c_int32),
                ("kernel_ms", C.c_double), ("nodes", C.c_uint64), ("splitter_steps", C.c_uint64),
                ("csr_bytes", C.c_uint64), ("alg_bytes", C.c_uint64)]


class _ServerConfig(C.Structure):
    _fields_ = [("max_n", C.c_int32), ("max_graph_nnz", C.c_int64), ("max_aut", C.c_uint32), ("slots", C.c_int32),
                ("queue_bytes", C.c_uint64), ("lifetime_ms", C.c_int32), ("large_share_pct", C.c_int32),
                ("max_search_cost", C.c_uint64)]


class _ServerResult(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("status", C.c_int32), ("verdict", C.c_int32), ("n", C.c_int32),
                ("pad", C.c_int32), ("latency_ms", C.c_double), ("queue_ms", C.c_double), ("nodes", C.c_uint64),
                ("device_ms", C.c_double)]


class _ServerStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("submitted", "completed", "refused_busy", "launches", "graphs_launched",
                                           "pending", "in_flight", "queued_bytes")] + \
               [("sched_publish_ms", C.c_double), ("sched_collect_ms", C.c_double), ("refused_capacity", C.c_uint64)]


_lib = None

# exported symbols and their signatures (restype, argtypes)
_SIGS = {
    "gcl_last_error": (C.c_char_p, []),
    "gcl_status_string": (C.c_char_p, [C.c_int]),
    "gcl_csr_validate": (C.c_int, [C.c_void_p]),
    "gcl_nauty_parallel": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "gcl_nauty_sequential": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "gcl_canon_label": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    "gcl_is_isomorphic": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]),
    "gcl_batch_canon": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gcl_batch_canon_plain": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gcl_canonical_keys": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gcl_read_dimacs": (C.c_int, [C.c_char_p, C.c_void_p]),
    "gcl_write_dimacs": (C.c_int, [C.c_char_p, C.c_void_p]),
    "gcl_graph_free": (None, [C.c_void_p]),
    "gcl_format_adjacency_keys": (C.c_int64, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int64]),
    "gcl_gen_erdos_renyi": (C.c_int, [C.c_int32, C.c_double, C.c_uint64, C.c_void_p]),
    "gcl_gen_erdos_renyi_mean_degree": (C.c_int, [C.c_int32, C.c_double, C.c_uint64, C.c_void_p]),
    "gcl_gen_random_regular": (C.c_int, [C.c_int32, C.c_int32, C.c_uint64, C.c_void_p]),
    "gcl_gen_torus": (C.c_int, [C.c_int32, C.c_int32, C.c_void_p]),
    "gcl_gen_cfi": (C.c_int, [C.c_int32, C.c_int32, C.c_uint64, C.c_void_p]),
    "gcl_gen_permute": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "gcl_gen_edge_swap": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "gcl_device_count": (C.c_int32, []),
    "gcl_set_side_files": (C.c_int, [C.c_char_p]),
    "gcl_device_batch_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "gcl_batch_tuning_default": (C.c_int, [C.c_void_p]),
    "gcl_canon_automorphisms": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                          C.c_void_p]),
    "gcl_device_batch_create_tuned": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "gcl_batch_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gcl_device_batch_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "gcl_device_batch_results": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gcl_device_batch_destroy": (None, [C.c_void_p]),
    "gcl_device_batch_class_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "gcl_device_batch_graph_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gcl_gen_mix_batch": (C.c_int, [C.c_int32, C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "gcl_canon_label_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gcl_server_config_default": (None, [C.c_void_p]),
    "gcl_server_create": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gcl_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gcl_server_poll": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]),
    "gcl_server_stats_get": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gcl_server_replay": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "gcl_server_destroy": (None, [C.c_void_p]),
    "gcl_owned_batch_free": (None, [C.c_void_p]),
}


def lib():
    """Loads libgcl.so (built by build.py); raises GclError if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GclError(-7, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def _check(status):
    if status != GCL_OK:
        raise GclError(status, lib().gcl_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class Graph:
    """Undirected simple graph in CSR form (rows sorted, symmetric)."""
    n: int
    row_ptr: np.ndarray
    col: np.ndarray
    color: Optional[np.ndarray] = None

    @property
    def nnz(self):
        return int(self.row_ptr[self.n]) if self.n >= 0 else 0

    @property
    def m(self):
        return self.nnz // 2

    def _csr(self):
        self.row_ptr = np.ascontiguousarray(self.row_ptr, dtype=np.int32)
        self.col = np.ascontiguousarray(self.col, dtype=np.int32)
        if self.color is not None:
            self.color = np.ascontiguousarray(self.color, dtype=np.int32)
        c = _CSR(self.n, self.nnz, _ptr(self.row_ptr), _ptr(self.col) if self.col.size else None,
                 _ptr(self.color))
        return c

    def edges(self):
        """Undirected edges (u < v) as an (m, 2) array."""
        src = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.row_ptr))
        sel = src < self.col
        return np.stack([src[sel], self.col[sel].astype(np.int64)], axis=1)

    @staticmethod
    def from_edges(n, edges, color=None):
        edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        if edges.size:
            u = np.concatenate([edges[:, 0], edges[:, 1]])
            v = np.concatenate([edges[:, 1], edges[:, 0]])
        else:
            u = v = np.zeros(0, np.int64)
        order = np.lexsort((v, u))
        u, v = u[order], v[order]
        row_ptr = np.zeros(n + 1, np.int32)
        np.add.at(row_ptr, u + 1, 1)
        row_ptr = np.cumsum(row_ptr).astype(np.int32)
        return Graph(n, row_ptr, v.astype(np.int32), None if color is None else np.asarray(color, np.int32))


def _from_owned(gp):
    g = gp.contents.csr
    n = g.n
    rp = np.ctypeslib.as_array(C.cast(g.row_ptr, C.POINTER(C.c_int32)), (n + 1,)).copy()
    col = (np.ctypeslib.as_array(C.cast(g.col, C.POINTER(C.c_int32)), (g.nnz,)).copy()
           if g.nnz else np.zeros(0, np.int32))
    color = (np.ctypeslib.as_array(C.cast(g.color, C.POINTER(C.c_int32)), (n,)).copy()
             if g.color else None)
    lib().gcl_graph_free(gp)
    return Graph(n, rp, col, color)


def _gen(name, *args):
    gp = C.POINTER(_Graph)()
    _check(getattr(lib(), name)(*args, C.byref(gp)))
    return _from_owned(gp)


def erdos_renyi(n, p, seed):
    return _gen("gcl_gen_erdos_renyi", n, float(p), seed)


def erdos_renyi_mean_degree(n, mean_degree, seed):
    return _gen("gcl_gen_erdos_renyi_mean_degree", n, float(mean_degree), seed)


def random_regular(n, d, seed):
    return _gen("gcl_gen_random_regular", n, d, seed)


def torus(a, b):
    return _gen("gcl_gen_torus", a, b)


def cfi(base_n, twisted, seed):
    return _gen("gcl_gen_cfi", base_n, 1 if twisted else 0, seed)


def permute(g: Graph, seed):
    c = g._csr()
    return _gen("gcl_gen_permute", C.byref(c), seed)


def edge_swap(g: Graph, seed):
    c = g._csr()
    return _gen("gcl_gen_edge_swap", C.byref(c), seed)


def read_dimacs(path):
    gp = C.POINTER(_Graph)()
    _check(lib().gcl_read_dimacs(os.fsencode(path), C.byref(gp)))
    return _from_owned(gp)


def write_dimacs(path, g: Graph):
    c = g._csr()
    _check(lib().gcl_write_dimacs(os.fsencode(path), C.byref(c)))


def validate(g: Graph):
    c = g._csr()
    _check(lib().gcl_csr_validate(C.byref(c)))


@dataclass
class CanonResult:
    label: np.ndarray
    keys: np.ndarray
    nodes: int
    automorphisms: int
    splitter_steps: int
    csr_bytes: int
    kernel_ms: float
    total_ms: float
    alg_bytes: int = 0

    def dense(self, n):
        d = np.zeros(n * n, np.uint8)
        d[self.keys.astype(np.int64)] = 1
        return d.reshape(n, n)


def canon_label(g: Graph, max_aut=10000, engine="gpu") -> CanonResult:
    """Canonical labeling (reference Nauty_Parallel / Nauty_Sequential)."""
    c = g._csr()
    label = np.zeros(max(g.n, 1), np.int32)
    keys = np.zeros(max(g.nnz, 1), np.uint32)
    st = _Stats()
    _check(lib().gcl_canon_label(C.byref(c), max_aut, ENGINES[engine], _ptr(label), _ptr(keys), None,
                                 C.byref(st)))
    return CanonResult(label[:g.n], keys[:g.nnz], st.nodes, st.automorphisms, st.splitter_steps, st.csr_bytes,
                       st.kernel_ms, st.total_ms, st.alg_bytes)


def nauty_parallel(g: Graph, max_aut=10000):
    """Reference-shaped call: returns (n*n canonical matrix bytes, node count)."""
    c = g._csr()
    out = np.zeros(g.n * g.n, np.uint8)
    nodes = C.c_uint32()
    _check(lib().gcl_nauty_parallel(C.byref(c), max_aut, _ptr(out), C.byref(nodes)))
    return out, nodes.value


def nauty_sequential(g: Graph, max_aut=10000):
    c = g._csr()
    out = np.zeros(g.n * g.n, np.uint8)
    nodes = C.c_uint32()
    _check(lib().gcl_nauty_sequential(C.byref(c), max_aut, _ptr(out), C.byref(nodes)))
    return out, nodes.value


def is_isomorphic(a: Graph, b: Graph, max_aut=10000, engine="gpu") -> bool:
    ca, cb = a._csr(), b._csr()
    v = C.c_int()
    _check(lib().gcl_is_isomorphic(C.byref(ca), C.byref(cb), max_aut, ENGINES[engine], C.byref(v)))
    return bool(v.value)


def canonical_keys(g: Graph, label: np.ndarray) -> np.ndarray:
    c = g._csr()
    label = np.ascontiguousarray(label, dtype=np.int32)
    keys = np.zeros(max(g.nnz, 1), np.uint32)
    _check(lib().gcl_canonical_keys(C.byref(c), _ptr(label), _ptr(keys)))
    return keys[:g.nnz]


def format_adjacency(keys: np.ndarray, n: int) -> str:
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    ln = lib().gcl_format_adjacency_keys(_ptr(keys), keys.size, n, None, 0)
    buf = C.create_string_buffer(int(ln) + 1)
    lib().gcl_format_adjacency_keys(_ptr(keys), keys.size, n, buf, ln + 1)
    return buf.raw[:ln].decode()


class _OwnedBatch(C.Structure):
    _fields_ = [("batch", _Batch), ("vtx_buf", C.c_void_p), ("row_buf", C.c_void_p), ("col_buf", C.c_void_p),
                ("kind", C.c_void_p)]


class Batch:
    """A flat batch of graphs (the layout of gcl_batch in include/gcl.h)."""

    @staticmethod
    def from_arrays(vtx_off, row_ptr, col, color=None, kind=None):
        b = Batch.__new__(Batch)
        b.graphs = None
        b.vtx_off = np.ascontiguousarray(vtx_off, np.int64)
        b.row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        b.col = np.ascontiguousarray(col, np.int32)
        b.color = None if color is None else np.ascontiguousarray(color, np.int32)
        b.kind = kind
        return b

    @property
    def num_graphs(self):
        return len(self.vtx_off) - 1

    def graph(self, i) -> Graph:
        v0, v1 = int(self.vtx_off[i]), int(self.vtx_off[i + 1])
        e0 = int(self.row_ptr[v0])
        rp = (self.row_ptr[v0:v1 + 1] - e0).astype(np.int32)
        col = self.col[e0:int(self.row_ptr[v1])].copy()
        color = None if self.color is None else self.color[v0:v1].copy()
        return Graph(v1 - v0, rp, col, color)

    def __init__(self, graphs: Sequence[Graph]):
        self.graphs = list(graphs)
        ns = np.array([g.n for g in self.graphs], np.int64)
        self.vtx_off = np.zeros(len(ns) + 1, np.int64)
        self.vtx_off[1:] = np.cumsum(ns)
        rows, cols, colors = [], [], []
        base = 0
        any_color = any(g.color is not None for g in self.graphs)
        for g in self.graphs:
            rows.append(g.row_ptr[:-1].astype(np.int64) + base)
            cols.append(np.asarray(g.col, np.int32))
            base += g.nnz
            if any_color:
                colors.append(g.color if g.color is not None else np.zeros(g.n, np.int32))
        rows.append(np.array([base], np.int64))
        self.row_ptr = np.ascontiguousarray(np.concatenate(rows))
        self.col = np.ascontiguousarray(np.concatenate(cols) if cols else np.zeros(0, np.int32))
        self.color = np.ascontiguousarray(np.concatenate(colors)) if any_color else None

    def _c(self):
        return _Batch(self.num_graphs, _ptr(self.vtx_off), _ptr(self.row_ptr),
                      _ptr(self.col) if self.col.size else None, _ptr(self.color))

    def labels_of(self, labels, i):
        return labels[self.vtx_off[i]:self.vtx_off[i + 1]]


@dataclass
class BatchResult:
    labels: np.ndarray
    nodes: np.ndarray
    total_nodes: int
    automorphisms: int
    splitter_steps: int
    csr_bytes: int
    kernel_ms: float
    total_ms: float
    workers: int
    alg_bytes: int = 0


def mix_batch(count, seed, n_min=16, n_max=1024, threads=0) -> Batch:
    """The batched benchmark mix (gcl_gen_mix_batch)."""
    ob = C.POINTER(_OwnedBatch)()
    _check(lib().gcl_gen_mix_batch(count, seed, n_min, n_max, threads, C.byref(ob)))
    o = ob.contents
    G = o.batch.num_graphs
    vtx = np.ctypeslib.as_array(C.cast(o.vtx_buf, C.POINTER(C.c_int64)), (G + 1,)).copy()
    tn = int(vtx[-1])
    rows = np.ctypeslib.as_array(C.cast(o.row_buf, C.POINTER(C.c_int64)), (tn + 1,)).copy()
    nnz = int(rows[-1])
    col = (np.ctypeslib.as_array(C.cast(o.col_buf, C.POINTER(C.c_int32)), (nnz,)).copy()
           if nnz else np.zeros(0, np.int32))
    kind = np.ctypeslib.as_array(C.cast(o.kind, C.POINTER(C.c_int32)), (max(G, 1),)).copy()[:G]
    lib().gcl_owned_batch_free(ob)
    return Batch.from_arrays(vtx, rows, col, None, kind)


class _Tuning(C.Structure):
    _fields_ = [("spec_min_n", C.c_int32), ("min_class_n", C.c_int32)]


def canon_label_noalloc(g: Graph, max_aut=10000) -> CanonResult:
    """gcl_canon_label_noalloc: the GPU search of one graph (n <= 1024,
    nnz <= 32768) with no device allocation (a static device array)."""
    c = g._csr()
    label = np.zeros(max(g.n, 1), np.int32)
    keys = np.zeros(max(g.nnz, 1), np.uint32)
    st = _Stats()
    _check(lib().gcl_canon_label_noalloc(C.byref(c), max_aut, _ptr(label), _ptr(keys), C.byref(st)))
    return CanonResult(label[:g.n], keys[:g.nnz], st.nodes, st.automorphisms, st.splitter_steps, st.csr_bytes,
                       st.kernel_ms, st.total_ms, st.alg_bytes)


def canon_automorphisms(g: "Graph", max_aut=10000):
    """gcl_canon_automorphisms: (labels, automorphisms as an array of rows
    a with a[v] the image of v, automorphisms found) from the GPU search."""
    c = g._csr()
    label = np.zeros(max(g.n, 1), np.int32)
    cap = max_aut
    auts = np.zeros((cap, max(g.n, 1)), np.int32)
    rows = C.c_int64(0)
    st = _Stats()
    _check(lib().gcl_canon_automorphisms(C.byref(c), max_aut, _ptr(label), _ptr(auts), cap, C.byref(rows),
                                         C.byref(st)))
    return label[:g.n], auts[:rows.value, :g.n], st.automorphisms


def batch_tuning_default():
    """The default launch tuning (gcl_batch_tuning_default) as a dict."""
    t = _Tuning()
    _check(lib().gcl_batch_tuning_default(C.byref(t)))
    return {"spec_min_n": t.spec_min_n, "min_class_n": t.min_class_n}


def autotune(batch: "Batch", max_aut=10000, sample=4096):
    """gcl_batch_autotune: the fastest launch tuning on a strided sample of
    the batch; returns (tuning dict, its kernel ms, candidates tried)."""
    t = _Tuning()
    ms = C.c_double(0.0)
    tried = C.c_int32(0)
    c = batch._c()
    _check(lib().gcl_batch_autotune(C.byref(c), max_aut, sample, C.byref(t), C.byref(ms), C.byref(tried)))
    return {"spec_min_n": t.spec_min_n, "min_class_n": t.min_class_n}, ms.value, tried.value


class DeviceBatch:
    """A batch resident in device memory (gcl_device_batch_*); `tuning` is a
    dict with spec_min_n and min_class_n (gcl_device_batch_create_tuned)."""

    def __init__(self, batch: Batch, max_aut=10000, tuning=None):
        self.batch = batch
        self._c = batch._c()
        self.h = C.c_void_p()
        if tuning is None:
            _check(lib().gcl_device_batch_create(C.byref(self._c), max_aut, C.byref(self.h)))
        else:
            t = _Tuning(int(tuning["spec_min_n"]), int(tuning["min_class_n"]))
            _check(lib().gcl_device_batch_create_tuned(C.byref(self._c), max_aut, C.byref(t), C.byref(self.h)))

    def run(self, use_graph=False):
        st = _BatchStats()
        _check(lib().gcl_device_batch_run(self.h, 1 if use_graph else 0, C.byref(st)))
        return st.kernel_ms, st.total_ms

    def results(self):
        b = self.batch
        labels = np.zeros(max(int(b.vtx_off[-1]), 1), np.int32)
        nodes = np.zeros(max(b.num_graphs, 1), np.uint64)
        st = _BatchStats()
        _check(lib().gcl_device_batch_results(self.h, _ptr(labels), _ptr(nodes), C.byref(st)))
        return BatchResult(labels[:int(b.vtx_off[-1])], nodes[:b.num_graphs], st.nodes, st.automorphisms,
                           st.splitter_steps, st.csr_bytes, 0.0, 0.0, st.workers, st.alg_bytes)

    def graph_stats(self):
        """Per-graph (nodes, automorphisms, splitter_steps, device_ns) of the last run."""
        G = max(self.batch.num_graphs, 1)
        a = [np.zeros(G, np.uint64) for _ in range(4)]
        _check(lib().gcl_device_batch_graph_stats(self.h, *[_ptr(x) for x in a]))
        return tuple(x[:self.batch.num_graphs] for x in a)

    def class_stats(self):
        """Per size class of the last run: dicts with max_n, graphs, workers,
        kernel_ms (HIP events on the class's stream), nodes, splitter_steps,
        csr_bytes and alg_bytes (all algorithmic bytes of refinement)."""
        cnt = C.c_int32(0)
        _check(lib().gcl_device_batch_class_stats(self.h, None, 0, C.byref(cnt)))
        arr = (_ClassStats * max(cnt.value, 1))()
        _check(lib().gcl_device_batch_class_stats(self.h, arr, cnt.value, None))
        return [{f: getattr(arr[i], f) for f, _ in _ClassStats._fields_ if f != "pad"} for i in range(cnt.value)]

    def close(self):
        if self.h:
            lib().gcl_device_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_canon(batch: Batch, max_aut=10000, engine="gpu", plain=False) -> BatchResult:
    c = batch._c()
    labels = np.zeros(max(int(batch.vtx_off[-1]), 1), np.int32)
    nodes = np.zeros(max(batch.num_graphs, 1), np.uint64)
    st = _BatchStats()
    if plain:
        _check(lib().gcl_batch_canon_plain(C.byref(c), max_aut, _ptr(labels), _ptr(nodes), C.byref(st)))
    else:
        _check(lib().gcl_batch_canon(C.byref(c), max_aut, ENGINES[engine], _ptr(labels), _ptr(nodes),
                                     C.byref(st)))
    return BatchResult(labels[:int(batch.vtx_off[-1])], nodes[:batch.num_graphs], st.nodes, st.automorphisms,
                       st.splitter_steps, st.csr_bytes, st.kernel_ms, st.total_ms, st.workers, st.alg_bytes)


def set_side_files(prefix):
    """Side files after each GPU search launch (gcl_set_side_files): <prefix>_counters.csv and
    <prefix>_nodes_per_cu.csv; None or "" turns them off."""
    _check(lib().gcl_set_side_files(prefix.encode() if prefix else None))


def device_count():
    return int(lib().gcl_device_count())


def server_config(**kw):
    """Default server configuration (gcl_server_config_default) with overrides."""
    c = _ServerConfig()
    lib().gcl_server_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Server:
    """Request server with continuous batching (gcl_server_*).

    submit() queues a canonicalize request (graph) or an isomorphism test
    (graph, other) and returns a ticket; poll() returns finished results as
    dicts; labels of canonicalize requests are kept per ticket."""

    def __init__(self, **config):
        self._cfg = server_config(**config)
        self.h = C.c_void_p()
        _check(lib().gcl_server_create(C.byref(self._cfg), C.byref(self.h)))
        self._keep = {}

    def submit(self, g: Graph, other: Graph = None, want_labels=True):
        ca = g._csr()
        cb = other._csr() if other is not None else None
        lab = np.zeros(max(g.n, 1), np.int32) if want_labels else None
        t = C.c_uint64()
        _check(lib().gcl_server_submit(self.h, C.byref(ca), C.byref(cb) if cb is not None else None,
                                       _ptr(lab) if lab is not None else None, C.byref(t)))
        self._keep[t.value] = (g, other, lab)
        return t.value

    def poll(self, cap=1024, timeout_ms=0):
        buf = (_ServerResult * cap)()
        k = lib().gcl_server_poll(self.h, buf, cap, timeout_ms)
        if k < 0:
            _check(k)
        out = []
        for i in range(k):
            r = {f: getattr(buf[i], f) for f, _ in _ServerResult._fields_}
            g, other, lab = self._keep.pop(r["ticket"], (None, None, None))
            if lab is not None and g is not None:
                r["labels"] = lab[:g.n]
            out.append(r)
        return out

    def wait_all(self, tickets, timeout_s=120.0):
        """Results of the given tickets (dict ticket -> result), polling until all are done."""
        want, got = set(tickets), {}
        t_end = time.time() + timeout_s
        while want - set(got):
            if time.time() > t_end:
                raise GclError(-7, "server results did not arrive in time")
            for r in self.poll(timeout_ms=50):
                got[r["ticket"]] = r
        return got

    def stats(self):
        st = _ServerStats()
        _check(lib().gcl_server_stats_get(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _ServerStats._fields_}

    def replay(self, batch: Batch, arrival_s, pair_with=None):
        """Submits graph i at arrival_s[i] (paired with graph pair_with[i] when >= 0); returns per-request
        arrays: status, verdict, latency_ms, nodes, launch_graphs."""
        count = len(arrival_s)
        arr = np.ascontiguousarray(arrival_s, np.float64)
        pw = None if pair_with is None else np.ascontiguousarray(pair_with, np.int32)
        out = (_ServerResult * max(count, 1))()
        c = batch._c()
        _check(lib().gcl_server_replay(self.h, C.byref(c), _ptr(pw) if pw is not None else None, _ptr(arr), count,
                                       out))
        res = {f: np.array([getattr(out[i], f) for i in range(count)]) for f, _ in _ServerResult._fields_}
        return res

    def close(self):
        if self.h:
            lib().gcl_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

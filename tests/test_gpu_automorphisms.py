"""Every automorphism the GPU search stores is an automorphism of the graph
(gcl_canon_automorphisms): each returned row is a permutation that maps the
edge set onto itself and keeps vertex colours; as many rows come back as the
search reports (up to max_aut), and the labels are those of gcl_canon_label.
The check is independent of the search: it uses only the CSR."""
import sys

import numpy as np
import pytest

import gcl

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _edges(g):
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(np.asarray(g.row_ptr, np.int64)))
    dst = np.asarray(g.col, np.int64)
    return set((src * g.n + dst).tolist())


def _check(g, max_aut, colors=None):
    label, auts, found = gcl.canon_automorphisms(g, max_aut=max_aut)
    r = gcl.canon_label(g, max_aut=max_aut, engine="gpu")
    assert np.array_equal(label, r.label)
    assert found == r.automorphisms
    assert auts.shape[0] == min(found, max_aut)
    e = _edges(g)
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(np.asarray(g.row_ptr, np.int64)))
    dst = np.asarray(g.col, np.int64)
    for a in auts:
        assert np.array_equal(np.sort(a), np.arange(g.n)), "not a permutation"
        img = set((a[src].astype(np.int64) * g.n + a[dst]).tolist())
        assert img == e, "edge set not preserved"
        assert not np.array_equal(a, np.arange(g.n)), "identity stored"
        if colors is not None:
            assert np.array_equal(colors[a], colors)
    return found


@pytest.mark.parametrize("name,make", [
    ("torus12x17", lambda: gcl.permute(gcl.torus(12, 17), 3)),
    ("cfi40", lambda: gcl.cfi(40, False, 5)),
    ("rr3_200", lambda: gcl.random_regular(200, 3, 7)),
    ("er300", lambda: gcl.erdos_renyi_mean_degree(300, 2.0, 9)),
    ("torus64x64", lambda: gcl.torus(64, 64)),
])
def test_gpu_stored_automorphisms_are_automorphisms(name, make):
    g = make()
    _log(f"automorphisms {name} n={g.n}")
    found = _check(g, 10000)
    _log(f"  {found} automorphisms checked")


def test_gpu_stored_automorphisms_small_ring():
    # with max_aut below the number found the ring keeps the latest ones
    g = gcl.permute(gcl.torus(10, 10), 11)
    _log("automorphisms torus10x10, max_aut=2")
    assert _check(g, 2) >= 2


def test_gpu_stored_automorphisms_keep_colours():
    g = gcl.torus(8, 8)
    colors = (np.arange(g.n) % 8 == 0).astype(np.int32)
    gc = gcl.Graph(g.n, g.row_ptr, g.col, colors)
    _log("automorphisms coloured torus")
    _check(gc, 10000, colors)

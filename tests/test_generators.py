"""Seeded generators of the benchmark workloads."""
import numpy as np

import gcl
import oracle


def degrees(g):
    return np.diff(g.row_ptr)


def test_random_regular():
    for n, d, s in [(10, 3, 1), (100, 4, 2), (50, 5, 3), (4096, 3, 0)]:
        g = gcl.random_regular(n, d, s)
        gcl.validate(g)
        assert g.n == n and np.all(degrees(g) == d)
        h = gcl.random_regular(n, d, s)
        assert np.array_equal(g.col, h.col)
    assert not np.array_equal(gcl.random_regular(100, 3, 1).col, gcl.random_regular(100, 3, 2).col)


def test_torus():
    g = gcl.torus(64, 64)
    gcl.validate(g)
    assert g.n == 4096 and np.all(degrees(g) == 4) and g.m == 2 * 4096


def test_erdos_renyi_density():
    g = gcl.erdos_renyi(2000, 0.01, 5)
    gcl.validate(g)
    expect = 0.01 * 2000 * 1999 / 2
    assert abs(g.m - expect) < 5 * np.sqrt(expect)
    h = gcl.erdos_renyi_mean_degree(5000, 8.0, 3)
    assert abs(2 * h.m / h.n - 8.0) < 0.3
    assert np.array_equal(gcl.erdos_renyi(128, 0.08, 0).col, gcl.erdos_renyi(128, 0.08, 0).col)


def test_cfi_structure_and_pair():
    a = gcl.cfi(10, False, 4)
    b = gcl.cfi(10, True, 4)
    for g in (a, b):
        gcl.validate(g)
        assert g.n == 100 and np.all(degrees(g) == 3)
    oa, ob = oracle.canon_graph(a), oracle.canon_graph(b)
    assert not np.array_equal(oa["keys"], ob["keys"])  # CFI pair is not isomorphic
    assert np.array_equal(oracle.canon_graph(gcl.permute(a, 3))["keys"], oa["keys"])


def test_permute_and_edge_swap():
    g = gcl.erdos_renyi_mean_degree(60, 4.0, 9)
    p = gcl.permute(g, 2)
    gcl.validate(p)
    assert sorted(degrees(p)) == sorted(degrees(g))
    assert np.array_equal(oracle.canon_graph(p)["keys"], oracle.canon_graph(g)["keys"])
    s = gcl.edge_swap(g, 5)
    gcl.validate(s)
    assert np.array_equal(degrees(s), degrees(g))
    eg = {tuple(e) for e in g.edges()}
    es = {tuple(e) for e in s.edges()}
    assert len(eg - es) == 2 and len(es - eg) == 2


def test_dimacs_roundtrip(tmp_path):
    g = gcl.cfi(6, True, 1)
    g.color = np.arange(g.n, dtype=np.int32) % 3
    f = str(tmp_path / "g.dimacs")
    gcl.write_dimacs(f, g)
    h = gcl.read_dimacs(f)
    assert h.n == g.n and np.array_equal(h.row_ptr, g.row_ptr) and np.array_equal(h.col, g.col)
    assert np.array_equal(h.color, g.color)

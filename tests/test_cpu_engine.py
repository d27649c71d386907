"""The product's host engine (--sequential) through the C ABI: canonical
adjacency bytes, node and automorphism counts equal to the reference binary
(golden fixtures) and to the oracle on seeded generated graphs."""
import numpy as np

import gcl
import oracle
from conftest import golden_graph, golden_output


def test_cpu_golden_singles(golden):
    for e in golden["singles"]:
        g = golden_graph(e["name"])
        r = gcl.canon_label(g, max_aut=e["max_aut"], engine="cpu")
        assert gcl.format_adjacency(r.keys, g.n).encode() == golden_output(e), e["name"]
        assert (r.nodes, r.automorphisms) == (e["nodes"], e["automorphisms"]), e["name"]


def test_cpu_golden_pairs(golden):
    for p in golden["pairs"]:
        v = gcl.is_isomorphic(golden_graph(p["a"]), golden_graph(p["b"]), engine="cpu")
        assert ("ISOMORPHIC" if v else "NOT ISOMORPHIC") == p["verdict"], p


def _cases():
    rs = np.random.RandomState(1234)
    out = []
    for s in range(25):
        out.append(gcl.erdos_renyi_mean_degree(int(rs.randint(5, 200)), float(rs.uniform(0.5, 9)), s))
    for s in range(12):
        d = int(rs.choice([3, 4, 5]))
        n = int(rs.randint(6, 120))
        n += (n * d) % 2
        out.append(gcl.random_regular(n, d, 50 + s))
    for a, b in [(3, 4), (5, 5), (4, 11), (9, 9), (12, 7)]:
        out.append(gcl.torus(a, b))
    for base, s in [(4, 1), (6, 2), (8, 3), (12, 4), (16, 5)]:
        out.append(gcl.cfi(base, s % 2 == 0, s))
    return out


def test_cpu_engine_vs_oracle():
    for g in _cases():
        for max_aut in (10000, 2):
            o = oracle.canon_graph(g, max_aut=max_aut)
            r = gcl.canon_label(g, max_aut=max_aut, engine="cpu")
            assert np.array_equal(r.keys, o["keys"]), g.n
            assert np.array_equal(r.label, o["label"]), g.n
            assert (r.nodes, r.automorphisms) == (o["nodes"], o["automorphisms"]), g.n


def test_cpu_batch_loop_matches_single():
    graphs = _cases()[:20]
    b = gcl.Batch(graphs)
    res = gcl.batch_canon(b, engine="cpu")
    for i, g in enumerate(graphs):
        keys = gcl.canonical_keys(g, b.labels_of(res.labels, i))
        assert np.array_equal(keys, gcl.canon_label(g, engine="cpu").keys)


def test_nauty_sequential_dense_matches_oracle():
    g = golden_graph("cube")
    dense, nodes = gcl.nauty_sequential(g, 10000)
    o = oracle.canon_graph(g)
    assert np.array_equal(dense.reshape(g.n, g.n), o["dense"]) and nodes == o["nodes"]

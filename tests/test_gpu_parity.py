"""GPU engine parity through the C ABI: canonical forms bit-identical to the
oracle / reference, and (single-graph path, full automorphism ring) the same
search-tree node and automorphism counts."""
import sys

import numpy as np
import pytest

import gcl
import oracle
from conftest import golden_graph, golden_output

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gpu_golden_singles(golden):
    for e in golden["singles"]:
        _log(f"golden {e['name']} max_aut={e['max_aut']}")
        g = golden_graph(e["name"])
        r = gcl.canon_label(g, max_aut=e["max_aut"], engine="gpu")
        assert gcl.format_adjacency(r.keys, g.n).encode() == golden_output(e), e["name"]
        assert r.nodes == e["nodes"], (e["name"], r.nodes, e["nodes"])
        assert r.automorphisms == e["automorphisms"], (e["name"], r.automorphisms, e["automorphisms"])
        o = oracle.canon_graph(g, max_aut=e["max_aut"], want_dense=False)
        assert r.alg_bytes == o["alg_bytes"], (e["name"], r.alg_bytes, o["alg_bytes"])


def test_gpu_golden_pairs(golden):
    for p in golden["pairs"]:
        _log(f"pair {p['a']} {p['b']}")
        a, b = golden_graph(p["a"]), golden_graph(p["b"])
        v = gcl.is_isomorphic(a, b, engine="gpu")
        assert ("ISOMORPHIC" if v else "NOT ISOMORPHIC") == p["verdict"], p


def _generated_cases():
    cases = []
    rs = np.random.RandomState(7)
    for s in range(12):
        cases.append(("er", gcl.erdos_renyi_mean_degree(int(rs.randint(10, 400)), float(rs.uniform(1.0, 10.0)), s)))
    for s in range(6):
        cases.append(("rr3", gcl.random_regular(int(2 * rs.randint(10, 100)), 3, 100 + s)))
    for s in range(3):
        cases.append(("rr5", gcl.random_regular(int(2 * rs.randint(10, 60)), 5, 200 + s)))
    for a, b in [(4, 9), (7, 7), (10, 12), (16, 16)]:
        cases.append(("torus", gcl.torus(a, b)))
    for base, s in [(12, 1), (20, 2), (30, 3)]:
        cases.append(("cfi", gcl.cfi(base, s % 2 == 1, s)))
    cases.append(("er_big", gcl.erdos_renyi_mean_degree(3000, 8.0, 5)))  # global-memory class
    return cases


def test_gpu_generated_vs_oracle():
    for kind, g in _generated_cases():
        for max_aut in (10000, 1):
            _log(f"generated {kind} n={g.n} m={g.m} max_aut={max_aut}")
            o = oracle.canon_graph(g, max_aut=max_aut)
            r = gcl.canon_label(g, max_aut=max_aut, engine="gpu")
            assert np.array_equal(r.keys, o["keys"]), (kind, g.n)
            assert r.nodes == o["nodes"], (kind, g.n, r.nodes, o["nodes"])
            assert r.automorphisms == o["automorphisms"], (kind, g.n)
            # the same traversal, so the same refinement byte count
            assert r.alg_bytes == o["alg_bytes"], (kind, g.n, r.alg_bytes, o["alg_bytes"])


def test_gpu_relabel_invariance():
    for g in [gcl.random_regular(512, 3, 3), gcl.torus(20, 24), gcl.cfi(40, True, 9)]:
        _log(f"invariance n={g.n}")
        k0 = gcl.canon_label(g, engine="gpu").keys
        for s in range(3):
            assert np.array_equal(gcl.canon_label(gcl.permute(g, 50 + s), engine="gpu").keys, k0)


def _mixed_graphs(count, seed):
    rs = np.random.RandomState(seed)
    graphs = []
    for i in range(count):
        n = int(np.exp(rs.uniform(np.log(16), np.log(300))))
        kind = rs.randint(3)
        if kind == 0:
            g = gcl.erdos_renyi_mean_degree(n, float(rs.uniform(3, 12)), seed * 1000 + i)
        elif kind == 1:
            d = int(rs.choice([3, 4, 5]))
            n += (n * d) % 2
            g = gcl.random_regular(n, d, seed * 1000 + i)
        else:
            g = gcl.torus(int(rs.randint(4, 17)), int(rs.randint(4, 17)))
        graphs.append(gcl.permute(g, seed * 7919 + i))
    return graphs


def test_gpu_batch_matches_oracle_and_plain():
    graphs = _mixed_graphs(120, 3)
    batch = gcl.Batch(graphs)
    _log("batch launch")
    res = gcl.batch_canon(batch, engine="gpu")
    _log("plain loop")
    plain = gcl.batch_canon(batch, plain=True)
    for i, g in enumerate(graphs):
        lab = batch.labels_of(res.labels, i)
        keys = gcl.canonical_keys(g, lab)
        o = oracle.canon_graph(g, max_aut=64)
        assert np.array_equal(keys, o["keys"]), i
        assert int(res.nodes[i]) == o["nodes"], (i, int(res.nodes[i]), o["nodes"])
        assert np.array_equal(batch.labels_of(plain.labels, i), lab), i


@pytest.mark.parametrize("lo,hi,count,seed", [(257, 512, 48, 22), (513, 1024, 36, 21)])
def test_gpu_benchmark_mix_large_classes_match_oracle(lo, hi, count, seed):
    """The benchmark mix's two largest LDS classes as a batch (G(n,p),
    random regular and torus graphs, root children refined by helpers):
    every graph's canonical form and search-tree node count equal the
    oracle's, and the batch's algorithmic byte total is the oracle's sum."""
    batch = gcl.mix_batch(count, seed, lo, hi)
    _log(f"mix batch n in [{lo}, {hi}]: {batch.num_graphs} graphs")
    res = gcl.batch_canon(batch, max_aut=10000, engine="gpu")
    total = 0
    for i in range(batch.num_graphs):
        g = batch.graph(i)
        o = oracle.canon_graph(g, max_aut=10000)
        keys = gcl.canonical_keys(g, batch.labels_of(res.labels, i))
        assert np.array_equal(keys, o["keys"]), (i, g.n, int(batch.kind[i]))
        assert int(res.nodes[i]) == o["nodes"], (i, g.n, int(res.nodes[i]), o["nodes"])
        total += o["alg_bytes"]
    assert res.alg_bytes == total, (res.alg_bytes, total)


def test_gpu_bench_batch_sample_matches_oracle():
    """The 65,536-graph batch that bench.py times, run whole on the device:
    every 64th graph (1,024 of them, all sizes and families) has the
    oracle's canonical form, search-tree node count and automorphism count."""
    batch = gcl.mix_batch(65536, 1)
    db = gcl.DeviceBatch(batch, max_aut=10000)
    try:
        _log("bench batch: one device step")
        db.run()
        res = db.results()
        nodes, auts, _, _ = db.graph_stats()
    finally:
        db.close()
    for i in range(0, batch.num_graphs, 64):
        g = batch.graph(i)
        o = oracle.canon_graph(g, max_aut=10000)
        keys = gcl.canonical_keys(g, batch.labels_of(res.labels, i))
        assert np.array_equal(keys, o["keys"]), (i, g.n, int(batch.kind[i]))
        assert int(nodes[i]) == o["nodes"] and int(auts[i]) == o["automorphisms"], (i, g.n, int(nodes[i]), o["nodes"])
        if i % 8192 == 0:
            _log(f"checked through graph {i}")


def test_gpu_batch_coloured_mix_matches_oracle():
    # coloured graphs (an initial partition of several cells, vertices in
    # (colour, id) order) next to uncoloured ones in one batch, some large
    # enough to publish root children
    rs = np.random.RandomState(11)
    graphs = []
    for i, g in enumerate(_mixed_graphs(24, 9) + [gcl.random_regular(600, 3, 4), gcl.torus(20, 30)]):
        if i % 2 == 0:
            g = gcl.Graph(g.n, g.row_ptr, g.col, rs.randint(0, 3, size=g.n).astype(np.int32))
        graphs.append(g)
    _log(f"coloured batch of {len(graphs)}")
    res = gcl.batch_canon(gcl.Batch(graphs), engine="gpu")
    batch = gcl.Batch(graphs)
    for i, g in enumerate(graphs):
        o = oracle.canon_graph(g, max_aut=64)
        keys = gcl.canonical_keys(g, batch.labels_of(res.labels, i))
        assert np.array_equal(keys, o["keys"]), (i, g.n, g.color is not None)


def test_gpu_batch_result_same_alone_and_in_mix():
    graphs = _mixed_graphs(60, 11)
    batch = gcl.Batch(graphs)
    res = gcl.batch_canon(batch, engine="gpu")
    for i in (0, 17, 42):
        alone = gcl.batch_canon(gcl.Batch([graphs[i]]), engine="gpu")
        assert np.array_equal(alone.labels, batch.labels_of(res.labels, i))


def test_gpu_nauty_parallel_dense():
    g = golden_graph("petersen")
    dense, nodes = gcl.nauty_parallel(g, 10000)
    o = oracle.canon_graph(g)
    assert np.array_equal(dense.reshape(g.n, g.n), o["dense"])
    assert nodes == o["nodes"]


def test_gpu_speculation_keeps_traversal_exact(monkeypatch):
    # root children refined by idle workers, and descents below the best path
    # recorded by them (single graphs of n > 512), must not change the
    # traversal: same canonical form, node and automorphism counts as the
    # oracle, with speculation forced on for small graphs and switched off
    rs = np.random.RandomState(7)
    er393 = None
    for s in range(12):
        n, d = int(rs.randint(10, 400)), float(rs.uniform(1.0, 10.0))
        g = gcl.erdos_renyi_mean_degree(n, d, s)
        if n == 393:
            er393 = g
    graphs = [er393, gcl.random_regular(300, 3, 21), gcl.torus(12, 12), gcl.cfi(20, True, 5),
              gcl.erdos_renyi_mean_degree(150, 2.0, 3), gcl.cfi(64, False, 9), gcl.cfi(70, True, 4),
              gcl.torus(30, 27)]
    for spec in ("64", "0"):
        monkeypatch.setenv("GCL_SPEC_MIN_N", spec)
        for g in graphs:
            for max_aut in (10000, 1):
                _log(f"speculation min_n={spec} n={g.n} m={g.m} max_aut={max_aut}")
                o = oracle.canon_graph(g, max_aut=max_aut)
                r = gcl.canon_label(g, max_aut=max_aut, engine="gpu")
                assert np.array_equal(r.keys, o["keys"]), g.n
                assert (r.nodes, r.automorphisms) == (o["nodes"], o["automorphisms"]), (g.n, spec, max_aut)


def test_gpu_device_batch_counters_are_deterministic():
    # work done by idle workers on a graph's root children is credited to
    # that graph, so per-graph steps and nodes do not depend on who did it
    graphs = _mixed_graphs(200, 5) + [gcl.random_regular(600, 3, 8), gcl.torus(24, 30)]
    db = gcl.DeviceBatch(gcl.Batch(graphs), max_aut=64)
    runs = []
    # eager runs, then several replays of the captured hipGraph: each replay
    # must agree with the eager result
    for rep, use_graph in enumerate([False, False, True, True, True]):
        _log(f"device batch run {rep} hipgraph={use_graph}")
        db.run(use_graph=use_graph)
        nodes, auts, steps, _ = db.graph_stats()
        runs.append((nodes.copy(), auts.copy(), steps.copy(), db.results().labels.copy()))
    cs = db.class_stats()
    assert sum(c["graphs"] for c in cs) == len(graphs)
    assert sum(c["splitter_steps"] for c in cs) == int(runs[-1][2].sum())
    db.close()
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            assert np.array_equal(a, b)

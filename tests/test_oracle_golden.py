"""The oracle against the reference binary's own outputs (golden fixtures):
canonical adjacency text bytes, node counts and automorphism counts of the
reference --sequential run, and the pair verdicts."""
import pytest

import oracle
from conftest import golden_graph, golden_output


def test_golden_present(golden):
    assert len(golden["singles"]) >= 50 and len(golden["pairs"]) >= 10


def test_oracle_matches_reference_outputs(golden):
    for e in golden["singles"]:
        g = golden_graph(e["name"])
        r = oracle.canon_graph(g, max_aut=e["max_aut"])
        text = oracle.format_adjacency(r["dense"]).encode()
        assert text == golden_output(e), e["name"]
        assert r["nodes"] == e["nodes"], (e["name"], e["max_aut"])
        assert r["automorphisms"] == e["automorphisms"], (e["name"], e["max_aut"])
        assert r["stale_classify"] == 0


def test_oracle_pair_verdicts(golden):
    for p in golden["pairs"]:
        a, b = golden_graph(p["a"]), golden_graph(p["b"])
        da = oracle.canon_graph(a)["dense"]
        db = oracle.canon_graph(b)["dense"]
        # main.cpp:189-193 compares the upper triangles with graph 1's n
        same = a.n == b.n and oracle.compare_upper(da, db) == 0
        assert ("ISOMORPHIC" if same else "NOT ISOMORPHIC") == p["verdict"], p


def test_max_aut_does_not_change_output(golden):
    by = {}
    for e in golden["singles"]:
        by.setdefault(e["name"], []).append(e)
    multi = [v for v in by.values() if len(v) > 1]
    assert multi
    for runs in multi:
        outs = {golden_output(e) for e in runs}
        assert len(outs) == 1, runs[0]["name"]

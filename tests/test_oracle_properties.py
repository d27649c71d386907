"""Pins of the oracle that do not come from the reference binary:
the canonical form is the maximum leaf key of the UNPRUNED search tree
(trail codes with the 32767 terminator, then the upper triangle), stored
automorphisms are automorphisms, and relabeling does not change the form."""
import itertools

import numpy as np

import gcl
import oracle

SENTINEL = 0o77777


def initial(g):
    n = g.n
    color = g.color if g.color is not None else np.zeros(n, np.int32)
    elem = sorted(range(n), key=lambda v: (color[v], v))
    cells, start = [], 0
    while start < n:
        end = start + 1
        while end < n and color[elem[end]] == color[elem[start]]:
            end += 1
        cells.append((start, end - start, 0))
        start = end
    return np.array(elem, np.uint32), cells


def leaf_key(g, elem, trail):
    n = g.n
    perm = np.empty(n, np.int64)
    perm[elem.astype(np.int64)] = np.arange(n)
    m = np.zeros((n, n), np.uint8)
    e = g.edges()
    m[perm[e[:, 0]], perm[e[:, 1]]] = 1
    m[perm[e[:, 1]], perm[e[:, 0]]] = 1
    upper = tuple(m[np.triu_indices(n, 1)])
    return (tuple(trail) + (SENTINEL,), upper), m


def brute_force(g):
    """Exhaustive individualization-refinement tree, no pruning."""
    elem, cells = initial(g)
    code, elem, cells, lev = oracle.refine(g.n, g.row_ptr, g.col, elem, cells, 0, None)
    best = [None, None]

    def rec(elem, cells, lev, trail):
        if all(c[1] == 1 for c in cells):
            k, m = leaf_key(g, elem, trail)
            if best[0] is None or k > best[0]:
                best[0], best[1] = k, m
            return
        ti = next(i for i, c in enumerate(cells) if c[1] > 1)
        f, L, inl = cells[ti]
        for v in sorted(int(x) for x in elem[f:f + L]):
            e2 = elem.copy()
            seg = list(e2[f:f + L])
            seg.remove(v)
            e2[f:f + L] = [v] + seg
            c2 = cells[:ti] + [(f, 1, lev + 1), (f + 1, L - 1, inl)] + cells[ti + 1:]
            code2, e3, c3, lev2 = oracle.refine(g.n, g.row_ptr, g.col, e2, c2, lev, [f])
            rec(e3, c3, lev2, trail + [code2])

    rec(elem, cells, lev, [code])
    return best[1]


def tiny_graphs():
    out = [gcl.Graph.from_edges(6, [(i, (i + 1) % 6) for i in range(6)]),
           gcl.Graph.from_edges(6, [(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5)]),
           gcl.Graph.from_edges(5, []),
           gcl.Graph.from_edges(6, [(i, j) for i in range(6) for j in range(i + 1, 6)]),
           gcl.Graph.from_edges(6, [(i, j) for i in range(3) for j in range(3, 6)]),
           gcl.Graph.from_edges(7, [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6)]),
           gcl.Graph.from_edges(8, [(i, i ^ b) for i in range(8) for b in (1, 2, 4) if i < i ^ b]),
           gcl.Graph.from_edges(6, [(i, (i + 1) % 6) for i in range(6)], color=[0, 1, 0, 1, 0, 1]),
           gcl.torus(3, 3)]
    rs = np.random.RandomState(3)
    for s in range(8):
        n = int(rs.randint(4, 9))
        pairs = [(i, j) for i, j in itertools.combinations(range(n), 2) if rs.rand() < 0.45]
        out.append(gcl.Graph.from_edges(n, pairs))
    return out


def test_oracle_is_max_leaf_of_unpruned_tree():
    for g in tiny_graphs():
        for max_aut in (10000, 1):
            o = oracle.canon_graph(g, max_aut=max_aut)
            assert np.array_equal(o["dense"], brute_force(g)), (g.n, g.m)


def test_oracle_automorphisms_are_automorphisms():
    for g in [gcl.torus(6, 6), gcl.cfi(6, False, 2), gcl.Graph.from_edges(7, []), gcl.random_regular(20, 3, 1)]:
        o = oracle.canon_graph(g, aut_cap=10000)
        edges = {tuple(e) for e in g.edges()}
        for a in o["auts"]:
            assert sorted(a) == list(range(g.n))
            mapped = {tuple(sorted((int(a[u]), int(a[v])))) for u, v in edges}
            assert mapped == edges


def test_oracle_canonical_form_is_a_relabeling():
    for g in [gcl.erdos_renyi_mean_degree(80, 5, 1), gcl.cfi(8, True, 3)]:
        o = oracle.canon_graph(g)
        lab = o["label"]
        n = g.n
        perm = np.empty(n, np.int64)
        perm[lab] = np.arange(n)
        m = np.zeros((n, n), np.uint8)
        e = g.edges()
        m[perm[e[:, 0]], perm[e[:, 1]]] = 1
        m[perm[e[:, 1]], perm[e[:, 0]]] = 1
        assert np.array_equal(m, o["dense"])
        for s in range(4):
            assert np.array_equal(oracle.canon_graph(gcl.permute(g, s))["dense"], o["dense"])


def test_oracle_algorithmic_bytes_hand_counted():
    """The refinement byte count B of the oracle on two graphs small enough to
    trace by hand: 12 per splitter vertex, 4 per neighbour id, 8 per
    non-singleton cell hashed under a non-singleton splitter, 8 per vertex of
    a split cell, 8n per refinement call.

    Path 0-1-2 (n=3): root: splitter {0,1,2} 12*3 + 4*4, one cell hashed 8,
    the cell splits into {0,2}{1} 8*3, then splitter {1} 12 + 4*2 leaves
    {0,2} whole; + 8*3 = 128.  Two root children, each discrete right after
    the individualization: 8*3 each.  Total 176, 3 nodes.

    Star K_{1,3} (centre 0): root: splitter {0..3} 12*4 + 4*6, one cell 8,
    split into {1,2,3}{0} 8*4, splitter {0} 12 + 4*3 leaves {1,2,3}; + 8*4 =
    168.  Child 1: splitter {1} 12 + 4 splits nothing, + 32 = 48; its
    children 2 and 3: 32 each (discrete).  Child 2 (not in the orbit of 1 under
    the automorphism (2 3) found below 1): 48, its first child 32, which gives
    the automorphism (1 2) and prunes child 3.  Total 168 + 48 + 32 + 32 + 48
    + 32 = 360, 6 nodes."""
    path = oracle.canon(3, np.array([0, 1, 3, 4]), np.array([1, 0, 2, 1]))
    assert (path["alg_bytes"], path["nodes"]) == (176, 3)
    star = oracle.canon(4, np.array([0, 3, 4, 5, 6]), np.array([1, 2, 3, 0, 0, 0]))
    assert (star["alg_bytes"], star["nodes"]) == (360, 6)


def test_host_engine_bytes_equal_oracle(golden):
    """The host engine (--sequential, traversal-exact) counts the same
    refinement bytes as the oracle, graph by graph."""
    from conftest import golden_graph
    for e in golden["singles"]:
        g = golden_graph(e["name"])
        o = oracle.canon_graph(g, max_aut=e["max_aut"], want_dense=False)
        r = gcl.canon_label(g, max_aut=e["max_aut"], engine="cpu")
        assert r.alg_bytes == o["alg_bytes"], (e["name"], r.alg_bytes, o["alg_bytes"])
        assert r.alg_bytes > r.csr_bytes, e["name"]

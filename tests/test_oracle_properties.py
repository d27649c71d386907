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

"""Large graphs (n 10k-50k) on the global-memory size class, through the C
ABI and the CLI: canonical forms equal the host engine's (the --sequential
path, itself pinned to the oracle on smaller graphs), and the CLI writes the
`output` file of a 50k-vertex graph from the sparse canonical edge list."""
import os
import subprocess
import sys

import numpy as np
import pytest

import gcl

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,seed", [(6000, 7), (20000, 3), (50000, 5)])
def test_gpu_large_er_matches_host_engine(n, seed):
    g = gcl.permute(gcl.erdos_renyi_mean_degree(n, 8.0, seed), seed + 100)
    _log(f"large er8 n={n} m={g.m}")
    r = gcl.canon_label(g, engine="gpu")
    c = gcl.canon_label(g, engine="cpu")
    assert np.array_equal(r.keys, c.keys)
    assert (r.nodes, r.automorphisms) == (c.nodes, c.automorphisms)


def test_gpu_large_torus_matches_host_engine():
    # 80 x 80 torus (n = 6400, the 8192 LDS class): a vertex-transitive graph
    # whose search prunes by automorphisms
    g = gcl.permute(gcl.torus(80, 80), 31)
    _log(f"large torus80x80 n={g.n}")
    r = gcl.canon_label(g, engine="gpu")
    c = gcl.canon_label(g, engine="cpu")
    assert np.array_equal(r.keys, c.keys)
    assert (r.nodes, r.automorphisms) == (c.nodes, c.automorphisms)


@pytest.mark.parametrize("n", [4096, 10000])
def test_gpu_large_regular_invariance(n):
    # 3-regular graphs (n + 1 search nodes, n^2 splitter steps: beyond the
    # host engine in test time), in an LDS size class and in the global one:
    # the canonical form and the search (node count) do not depend on the
    # vertex names, and the labeling is a permutation
    g = gcl.random_regular(n, 3, 4)
    _log(f"large rr3 n={n}")
    r0 = gcl.canon_label(g, engine="gpu")
    assert np.array_equal(np.sort(r0.label), np.arange(g.n))
    assert r0.nodes == g.n + 1
    h = gcl.permute(g, 77)
    _log(f"large rr3 n={n} relabeled")
    r1 = gcl.canon_label(h, engine="gpu")
    assert np.array_equal(r1.keys, r0.keys)
    assert r1.nodes == r0.nodes


def test_cli_large_output_parallel_equals_sequential(tmp_path):
    g = gcl.permute(gcl.erdos_renyi_mean_degree(50000, 8.0, 11), 12)
    f = str(tmp_path / "big.dimacs")
    gcl.write_dimacs(f, g)
    outs = {}
    for mode in ("--parallel", "--sequential"):
        d = tmp_path / mode.strip("-")
        d.mkdir()
        _log(f"cli {mode}")
        p = subprocess.run([gcl.CLI_PATH, f, mode], cwd=d, capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stderr[-500:]
        outs[mode] = (d / "output").read_bytes()
    assert outs["--parallel"] == outs["--sequential"]
    assert outs["--parallel"].startswith(b"  0 : ")
    assert outs["--parallel"].count(b";\n") == 50000

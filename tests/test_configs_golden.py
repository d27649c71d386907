"""Reference outputs at the benchmark configurations' full sizes
(tests/golden/configs.json, written by make_golden_configs.py from the
reference's sequential binary), the serving configuration's large graphs
(G(n, p) of mean degree 8 with n = 20,000, searched by the global class, and
a 3-regular graph with n = 10,000) included: the GPU engine's `output` bytes (SHA-256),
node and automorphism counts equal the reference's; pairs built from them
(relabeled copy, double-edge swap) get the reference's verdict.  The CPU part
only checks that the seeded generators still produce the same inputs."""
import hashlib
import json
import os
import sys

import pytest

import gcl
from conftest import GOLDEN


def load():
    path = os.path.join(GOLDEN, "configs.json")
    if not os.path.exists(path):
        pytest.skip("configs.json not generated")
    with open(path) as f:
        return json.load(f)["graphs"]


def build(entries):
    by = {e["name"]: e for e in entries}
    cache = {}

    def get(name):
        if name in cache:
            return cache[name]
        r = by[name]["recipe"]
        if r[0] == "random_regular":
            g = gcl.random_regular(r[1], r[2], r[3])
        elif r[0] == "er_mean_degree":
            g = gcl.erdos_renyi_mean_degree(r[1], r[2], r[3])
        elif r[0] == "torus":
            g = gcl.torus(r[1], r[2])
        elif r[0] == "cfi":
            g = gcl.cfi(r[1], bool(r[2]), r[3])
        elif r[0] == "permute":
            g = gcl.permute(get(r[1]), r[2])
        else:
            g = gcl.edge_swap(get(r[1]), r[2])
        cache[name] = g
        return g

    return get


def test_config_inputs_reproduce(tmp_path):
    entries = load()
    get = build(entries)
    for e in entries:
        f = str(tmp_path / (e["name"] + ".dimacs"))
        gcl.write_dimacs(f, get(e["name"]))
        assert hashlib.sha256(open(f, "rb").read()).hexdigest() == e["dimacs_sha256"], e["name"]


@pytest.mark.gpu
def test_gpu_config_outputs_match_reference():
    entries = load()
    get = build(entries)
    sha = {}
    for e in entries:
        print(f"config {e['name']} n={e['n']}", flush=True)
        sys.stdout.flush()
        g = get(e["name"])
        r = gcl.canon_label(g, max_aut=e["max_aut"], engine="gpu")
        text = gcl.format_adjacency(r.keys, g.n).encode()
        assert len(text) == e["output_bytes"], e["name"]
        assert hashlib.sha256(text).hexdigest() == e["output_sha256"], e["name"]
        assert (r.nodes, r.automorphisms) == (e["nodes"], e["automorphisms"]), e["name"]
        sha[e["name"]] = e["output_sha256"]
    # verdicts of the configurations' pairs follow from the reference outputs
    for base in ("rr3_n4096_s0", "torus64x64"):
        print(f"config pairs {base}", flush=True)
        assert gcl.is_isomorphic(get(base), get(base + "_perm"), engine="gpu") == (sha[base] == sha[base + "_perm"])
        assert gcl.is_isomorphic(get(base), get(base + "_swap"), engine="gpu") == (sha[base] == sha[base + "_swap"])
    cfi = sorted({k[:-len("_straight")] for k in sha if k.endswith("_straight")})
    assert cfi, "no CFI pairs in the fixture"
    for base in cfi:
        print(f"config pair {base}", flush=True)
        assert gcl.is_isomorphic(get(base + "_straight"), get(base + "_twisted"), engine="gpu") == (
            sha[base + "_straight"] == sha[base + "_twisted"])

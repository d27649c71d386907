"""Writes tests/golden/configs.json: the reference's own sequential binary
(oracle/_ref/gpu_canon_label_ref) on graphs of the benchmark configurations'
full sizes (3-regular and torus graphs with n = 4096, CFI graphs with
n = 1000, 2000 and 4000, each with a relabeled copy and a double-edge-swap
variant where the configuration names one), and two of the serving
configuration's large graphs (G(n, p) of mean degree 8 with n = 20,000 and a
3-regular graph with n = 10,000).  The outputs are large, so the
fixture keeps their SHA-256 and length with the node and automorphism
counts; the inputs are re-created from libgcl's seeded generators, pinned by
the SHA-256 of their DIMACS text.  Takes about 10 minutes of one CPU core
for the n <= 4096 graphs (the reference needs ~150 s per 3-regular graph)
and most of an hour for the 3-regular n = 10,000 graph.

    python tests/golden/make_golden_configs.py [NAME ...]   (needs the reference sources)

With names, only those entries are (re)computed and the others are kept.
"""
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "gpu-canonical-labeling_amd"))
import gcl  # noqa: E402

REF = os.path.join(REPO, "oracle", "_ref", "gpu_canon_label_ref")
RE_AUT = re.compile(r"Total automorphisms discovered \(sequential\): (\d+)")
RE_NODES = re.compile(r"Total search tree nodes created \(sequential\): (\d+)")

# name -> (recipe); recipes are re-run by the tests through build_graph()
RECIPES = {
    "rr3_n4096_s0": ["random_regular", 4096, 3, 0],
    "rr3_n4096_s0_perm": ["permute", "rr3_n4096_s0", 1],
    "rr3_n4096_s0_swap": ["edge_swap", "rr3_n4096_s0", 2],
    "torus64x64": ["torus", 64, 64],
    "torus64x64_perm": ["permute", "torus64x64", 3],
    "torus64x64_swap": ["edge_swap", "torus64x64", 4],
    "cfi100_s1_straight": ["cfi", 100, 0, 1],
    "cfi100_s1_twisted": ["cfi", 100, 1, 1],
    "cfi200_s2_straight": ["cfi", 200, 0, 2],
    "cfi200_s2_twisted": ["cfi", 200, 1, 2],
    "cfi400_s3_straight": ["cfi", 400, 0, 3],
    "cfi400_s3_twisted": ["cfi", 400, 1, 3],
    # the serving configuration's large graphs (searched by the global class)
    "er_n20000_d8_s5": ["er_mean_degree", 20000, 8.0, 5],
    "rr3_n10000_s7": ["random_regular", 10000, 3, 7],
}


def build_graph(name, cache=None):
    cache = {} if cache is None else cache
    if name in cache:
        return cache[name]
    r = RECIPES[name]
    if r[0] == "random_regular":
        g = gcl.random_regular(r[1], r[2], r[3])
    elif r[0] == "er_mean_degree":
        g = gcl.erdos_renyi_mean_degree(r[1], r[2], r[3])
    elif r[0] == "torus":
        g = gcl.torus(r[1], r[2])
    elif r[0] == "cfi":
        g = gcl.cfi(r[1], bool(r[2]), r[3])
    elif r[0] == "permute":
        g = gcl.permute(build_graph(r[1], cache), r[2])
    else:
        g = gcl.edge_swap(build_graph(r[1], cache), r[2])
    cache[name] = g
    return g


def main(names):
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run make -C oracle ref")
    unknown = [x for x in names if x not in RECIPES]
    if unknown:
        sys.exit(f"unknown names: {unknown}")
    path = os.path.join(HERE, "configs.json")
    kept = {}
    if names and os.path.exists(path):
        with open(path) as fh:
            kept = {e["name"]: e for e in json.load(fh)["graphs"]}
    cache, out = {}, []
    with tempfile.TemporaryDirectory() as tmp:
        for name in RECIPES:
            if names and name not in names:
                if name in kept:
                    out.append(kept[name])
                continue
            g = build_graph(name, cache)
            f = os.path.join(tmp, name + ".dimacs")
            gcl.write_dimacs(f, g)
            dimacs_sha = hashlib.sha256(open(f, "rb").read()).hexdigest()
            p = subprocess.run([REF, f, "--sequential", "--max-aut=10000"], cwd=tmp, capture_output=True,
                               text=True, timeout=7200)
            if p.returncode != 0:
                sys.exit(f"reference failed on {name}: {p.stderr}")
            data = open(os.path.join(tmp, "output"), "rb").read()
            out.append(dict(name=name, recipe=RECIPES[name], dimacs_sha256=dimacs_sha, max_aut=10000, n=g.n, m=g.m,
                            nodes=int(RE_NODES.search(p.stdout).group(1)),
                            automorphisms=int(RE_AUT.search(p.stdout).group(1)),
                            output_sha256=hashlib.sha256(data).hexdigest(), output_bytes=len(data)))
            print(name, out[-1]["nodes"], out[-1]["automorphisms"], flush=True)
    with open(path, "w") as fh:
        json.dump(dict(source="reference sequential binary (oracle/_ref), --sequential --max-aut=10000",
                       graphs=out), fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])

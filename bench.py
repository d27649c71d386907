"""Benchmark: canonical labelings/s of the batched mix on one MI355X.

Workload (one "step"): canonical labeling of a fixed batch of 65,536 graphs,
n log-uniform in [16, 1024]; 50% G(n,p) with mean degree U[3,12], 30% random
d-regular (d in {3,4,5}), 20% C_a x C_b (a, b in U[4,32]); each under a seeded
random relabeling.  The batch is uploaded once (inputs resident in HBM); a
step is one run of the persistent search launch (one kernel per size class,
each on its own stream, bracketed by HIP events on that stream) and ends with
a stream synchronize; three replays of the same launches captured as a
hipGraph are timed beside it (--hipgraph times the replays instead, with two
eager steps after them for the per-class kernel times).

Printed as one JSON line with the throughput; the plain one-graph-per-launch
rate on a sample of the same batch (kernel and call latency); the roofline
figure of the dominant kernel from its algorithmic refinement bytes; the
single-graph GPU path on the graphs of the first three configurations beside
the reference's sequential binary on the same graphs (output bytes compared
in the same run); the served rates and latencies; and the CPU baseline (the
reference's sequential binary, or the oracle where that binary is absent, on
a bounded sample of the batch, one core and all cores), whose `output` files
are compared byte for byte with the batched GPU labels.
"""
import argparse
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gpu-canonical-labeling_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import numpy as np  # noqa: E402

import gcl  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "canonical labelings/sec + p99 latency, batched/served vs plain GPU, bit-exact"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_name(max_n):
    return f"search_kernel_lds<{max_n}>" if max_n in (64, 128, 256, 512, 1024, 2048, 4096, 8192) else "search_kernel_global"


def traffic_from_profile(kernel, graphs):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE, then WRITE_SIZE, each scaled by the copy-kernel calibration),
    committed under profiles/ for the same workload."""
    path = os.path.join(REPO, "profiles", "search_batch_mix_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("graphs") != graphs:
        return None
    # the profile names kernels with their parameter list, "name(KParams)"
    by_name = {n.split("(")[0]: v for n, v in t.get("kernels", {}).items()}
    k = by_name.get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")


RE_TIME = re.compile(r"Algorithm ran in: (\d+)min, (\d+)s, (\d+)ms, (\d+)\S+s, (\d+)ns")
RE_NODES = re.compile(r"Total search tree nodes created \(sequential\): (\d+)")


def ref_seconds(stdout):
    m = RE_TIME.search(stdout)
    if not m:
        return None
    mi, s, ms, us, ns = (int(x) for x in m.groups())
    return mi * 60 + s + ms * 1e-3 + us * 1e-6 + ns * 1e-9


def cpu_baseline(batch, budget_s, stride):
    """Times the reference's sequential binary (or the oracle) on graphs
    0, stride, 2*stride, ... of the batch until budget_s of CPU time is spent.
    Keeps each reference run's `output` bytes (graph index -> bytes) for the
    bit-exact check of the batched labels."""
    import oracle
    ref = oracle.REF_BIN
    sample = list(range(0, batch.num_graphs, stride))
    done, spent, outputs = 0, 0.0, {}
    kind = "reference" if os.path.exists(ref) else "port"
    with tempfile.TemporaryDirectory() as tmp:
        for i in sample:
            g = batch.graph(i)
            if kind == "reference":
                f = os.path.join(tmp, "g.dimacs")
                gcl.write_dimacs(f, g)
                p = subprocess.run([ref, f, "--sequential", "--max-aut=10000"], cwd=tmp, capture_output=True,
                                   text=True, timeout=max(60, int(budget_s * 4)))
                t = ref_seconds(p.stdout)
                if p.returncode != 0 or t is None:
                    raise RuntimeError("reference run failed: " + p.stderr[-300:])
                spent += t
                with open(os.path.join(tmp, "output"), "rb") as fh:
                    outputs[i] = fh.read()
            else:
                t0 = time.perf_counter()
                o = oracle.canon_graph(g)
                spent += time.perf_counter() - t0
                outputs[i] = oracle.format_adjacency(o["dense"]).encode()
            done += 1
            if spent >= budget_s:
                break
    return dict(value=done / spent, unit="labelings/s", cores=1, kind=kind,
                sample=f"graphs 0,{stride},{2 * stride},... of the same batch: {done} graphs, {spent:.1f} s "
                       f"of single-thread algorithm time ({'reference sequential binary' if kind == 'reference' else 'oracle'})"), \
        sample[:done], outputs


def cpu_baseline_all_cores(batch, indices):
    """The same sample through the reference's sequential binary on every
    host core: independent processes over the graphs (the path is
    single-threaded), at most 16 at a time; wall clock from the first start
    to the last exit, process start and DIMACS parse included."""
    import concurrent.futures
    import oracle
    ref = oracle.REF_BIN
    if not os.path.exists(ref) or not indices:
        return None
    workers = max(1, min(16, os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for i in indices:
            d = os.path.join(tmp, str(i))
            os.makedirs(d)
            gcl.write_dimacs(os.path.join(d, "g.dimacs"), batch.graph(i))
            dirs.append(d)

        def one(d):
            p = subprocess.run([ref, "g.dimacs", "--sequential", "--max-aut=10000"], cwd=d, capture_output=True,
                               text=True, timeout=600)
            if p.returncode != 0:
                raise RuntimeError("reference run failed: " + p.stderr[-300:])

        t0 = time.perf_counter()
        with concurrent.futures.ThreadPoolExecutor(workers) as ex:
            list(ex.map(one, dirs))
        wall = time.perf_counter() - t0
    return dict(value=len(indices) / wall, unit="labelings/s", cores=workers, kind="reference",
                sample=f"the {len(indices)} graphs of cpu_baseline, {workers} reference processes at a time, "
                       f"{wall:.1f} s wall (process start and DIMACS parse included)")


def verify_batch_vs_reference(batch, labels, outputs):
    """Each sampled graph's `output` text from the batched GPU labels against
    the reference run's own `output` file, byte for byte."""
    ok = 0
    for i, want in outputs.items():
        g = batch.graph(i)
        lab = labels[batch.vtx_off[i]:batch.vtx_off[i + 1]]
        text = gcl.format_adjacency(gcl.canonical_keys(g, lab), g.n).encode()
        if text != want:
            raise SystemExit(f"bench.py: batched output of graph {i} differs from the reference's")
        ok += 1
    return f"{ok}/{len(outputs)}"


# Single graphs of the first three benchmark configurations, built with the
# seeded generators (the recipes of tests/golden/make_golden_configs.py and
# tests/golden/make_golden.py): G(128, 0.08) seed 0 and a relabeled copy;
# 3-regular and 64x64 torus graphs with n = 4096, each with a relabeled copy
# and a double-edge-swap variant; CFI graphs over 100, 200 and 400 base
# vertices (n = 1000, 2000, 4000), straight and twisted.
CONFIG_GRAPHS = [
    ("er128_p008_s0", "C0", ("erdos_renyi", 128, 0.08, 0)),
    ("er128_p008_s0_perm", "C0", ("permute", "er128_p008_s0", 1)),
    ("rr3_n4096_s0", "C1", ("random_regular", 4096, 3, 0)),
    ("rr3_n4096_s0_perm", "C1", ("permute", "rr3_n4096_s0", 1)),
    ("rr3_n4096_s0_swap", "C1", ("edge_swap", "rr3_n4096_s0", 2)),
    ("torus64x64", "C1", ("torus", 64, 64)),
    ("torus64x64_perm", "C1", ("permute", "torus64x64", 3)),
    ("torus64x64_swap", "C1", ("edge_swap", "torus64x64", 4)),
    ("cfi100_s1_straight", "C2", ("cfi", 100, 0, 1)),
    ("cfi100_s1_twisted", "C2", ("cfi", 100, 1, 1)),
    ("cfi200_s2_straight", "C2", ("cfi", 200, 0, 2)),
    ("cfi200_s2_twisted", "C2", ("cfi", 200, 1, 2)),
    ("cfi400_s3_straight", "C2", ("cfi", 400, 0, 3)),
    ("cfi400_s3_twisted", "C2", ("cfi", 400, 1, 3)),
]
CONFIG_PAIRS = [("er128_p008_s0", "er128_p008_s0_perm"), ("rr3_n4096_s0", "rr3_n4096_s0_perm"),
                ("rr3_n4096_s0", "rr3_n4096_s0_swap"), ("torus64x64", "torus64x64_perm"),
                ("torus64x64", "torus64x64_swap"), ("cfi100_s1_straight", "cfi100_s1_twisted"),
                ("cfi200_s2_straight", "cfi200_s2_twisted"), ("cfi400_s3_straight", "cfi400_s3_twisted")]


def build_config_graphs():
    graphs = {}
    for name, _, r in CONFIG_GRAPHS:
        if r[0] == "erdos_renyi":
            graphs[name] = gcl.erdos_renyi(r[1], r[2], r[3])
        elif r[0] == "random_regular":
            graphs[name] = gcl.random_regular(r[1], r[2], r[3])
        elif r[0] == "torus":
            graphs[name] = gcl.torus(r[1], r[2])
        elif r[0] == "cfi":
            graphs[name] = gcl.cfi(r[1], bool(r[2]), r[3])
        elif r[0] == "permute":
            graphs[name] = gcl.permute(graphs[r[1]], r[2])
        else:
            graphs[name] = gcl.edge_swap(graphs[r[1]], r[2])
    return graphs


def start_reference_legs(graphs, tmp):
    """The reference's sequential binary on every configuration graph, one
    process each (the path is single-threaded), started in the background at
    a lower priority while the GPU legs run."""
    import oracle
    ref = oracle.REF_BIN
    if not os.path.exists(ref):
        return None
    legs = {}
    for name, _, _ in CONFIG_GRAPHS:
        d = os.path.join(tmp, name)
        os.makedirs(d)
        gcl.write_dimacs(os.path.join(d, "g.dimacs"), graphs[name])
        p = subprocess.Popen([ref, "g.dimacs", "--sequential", "--max-aut=10000"], cwd=d,
                             stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             preexec_fn=lambda: os.nice(10))
        legs[name] = (p, d)
    return legs


def collect_reference_legs(legs, limit_s):
    out = {}
    t_end = time.time() + limit_s
    for name, (p, d) in legs.items():
        try:
            stdout, _ = p.communicate(timeout=max(1.0, t_end - time.time()))
        except subprocess.TimeoutExpired:
            p.kill()
            p.communicate()
            log(f"reference leg {name}: not done within the bench's limit")
            out[name] = None
            continue
        with open(os.path.join(d, "output"), "rb") as fh:
            data = fh.read()
        m = RE_NODES.search(stdout)
        out[name] = {"seconds": ref_seconds(stdout), "nodes": int(m.group(1)) if m else None,
                     "sha256": hashlib.sha256(data).hexdigest()}
        log(f"reference leg {name}: {out[name]['seconds']:.2f} s")
    return out


def plain_configs(graphs):
    """The single-graph GPU path (one launch per graph, full automorphism
    ring) on the configuration graphs: kernel and wall time, nodes/s, and the
    SHA-256 of the `output` text; pair verdicts through gcl_is_isomorphic."""
    rows = {}
    for name, cfg, _ in CONFIG_GRAPHS:
        g = graphs[name]
        r = gcl.canon_label(g, max_aut=10000, engine="gpu")
        text = gcl.format_adjacency(r.keys, g.n).encode()
        rows[name] = {"config": cfg, "n": g.n, "m": g.m, "gpu_kernel_ms": round(r.kernel_ms, 3),
                      "gpu_wall_ms": round(r.total_ms, 3), "nodes": int(r.nodes),
                      "nodes_per_s": round(r.nodes / max(r.kernel_ms, 1e-6) * 1e3, 1),
                      "sha256": hashlib.sha256(text).hexdigest()}
        log(f"plain config {name}: kernel {r.kernel_ms:.1f} ms, {r.nodes} nodes")
    verdicts = {f"{a}~{b}": bool(gcl.is_isomorphic(graphs[a], graphs[b], engine="gpu")) for a, b in CONFIG_PAIRS}
    return rows, verdicts


def configs_table(rows, verdicts, ref):
    """Joins the GPU rows with the reference runs of the same graphs and the
    committed reference fixtures; exits on any difference in output bytes
    or verdicts."""
    with open(os.path.join(REPO, "tests", "golden", "configs.json")) as fh:
        fixture = {e["name"]: e["output_sha256"] for e in json.load(fh)["graphs"]}
    checked = 0
    for name, row in rows.items():
        r = (ref or {}).get(name)
        row["ref_cpu_s"] = None if r is None else round(r["seconds"], 3)
        row["ref_nodes"] = None if r is None else r["nodes"]
        if r is not None:
            if r["sha256"] != row["sha256"]:
                raise SystemExit(f"bench.py: GPU output of {name} differs from the reference run's")
            checked += 1
        if name in fixture and fixture[name] != row["sha256"]:
            raise SystemExit(f"bench.py: GPU output of {name} differs from the committed reference fixture")
        row["speedup_vs_ref_1core"] = (None if r is None or not r["seconds"] else
                                       round(r["seconds"] * 1e3 / max(row["gpu_kernel_ms"], 1e-6), 1))
    for key, v in verdicts.items():
        a, b = key.split("~")
        if (rows[a]["sha256"] == rows[b]["sha256"]) != v:
            raise SystemExit(f"bench.py: verdict of {key} does not follow the outputs")
    return {"graphs": rows, "verdicts": {k: ("ISOMORPHIC" if v else "NOT ISOMORPHIC") for k, v in verdicts.items()},
            "verified_vs_reference_run": f"{checked}/{len(rows)}",
            "cpu": "reference sequential binary, one process per graph at nice 10, concurrent with the GPU legs"}


POISSON_CONFIGURED = 1_000_000  # requests of the serving configuration's Poisson stream


def served_workload(R, seed):
    """R requests over the batch's size mix: graph i, and for about half of
    them a partner (a permuted copy or a one-swap variant, 50/50) that makes
    the request a pair test.  Returns (batch, pair_with, pairs)."""
    first = gcl.mix_batch(R, seed)
    rs = np.random.RandomState(seed + 1)
    graphs = [first.graph(i) for i in range(R)]
    pair_with = np.full(R, -1, np.int32)
    partners = []
    for i in range(R):
        if rs.rand() < 0.5:
            g = graphs[i]
            partners.append(gcl.permute(g, 1000003 + i) if rs.rand() < 0.5 else gcl.edge_swap(g, 2000003 + i))
            pair_with[i] = R + len(partners) - 1
    return gcl.Batch(graphs + partners), pair_with, len(partners)


def served(args):
    """Served measurement: requests over the same mix of sizes, half
    canonicalize and half isomorphism tests (partner: a permuted copy or a
    one-swap variant, 50/50), through the request server.  The Poisson phase
    runs --serve-poisson requests (default: the capacity phase's count): a
    stream of chunks of --serve-requests requests each, every chunk freshly
    generated (the next one is built while the current one is served)."""
    R = args.serve_requests
    t0 = time.time()
    batch, pair_with, npairs = served_workload(R, args.seed + 1)
    log(f"served workload: {R} requests ({npairs} pairs) built in {time.time() - t0:.1f}s")
    total = max(args.serve_poisson, 0) or R
    out = serve_run(args, batch, R, pair_with, npairs, "served", dict(max_n=1024), poisson_total=total,
                    chunk=lambda k: served_workload(R, args.seed + 1 + 1000 * k)[:2])
    out["workload"] = (f"same size mix (n log-uniform [16,1024]); half canonicalize, half pair tests "
                       f"(permuted copy or one-swap variant, 50/50); persistent serving kernels, one per size "
                       f"class; capacity phase {R} requests, Poisson phase {total} requests in chunks of {R} "
                       f"distinct generated requests (the configuration's stream has {POISSON_CONFIGURED}); "
                       f"the 2% of n 10k-50k graphs are in served_large")
    return out


def served_large(args):
    """The serving configuration's 2% of large graphs: requests over the same
    size mix where each is, with probability 0.02, a large graph with n
    log-uniform in [10000, 50000], half of them sparse G(n, p) of mean degree
    8 and half 3-regular; half canonicalize, half pair tests as above.  The
    server runs a second persistent kernel for the large class.  3-regular
    graphs above the server's admission cap on estimated search cost (n^2
    splitter steps for a regular graph) are refused with GCL_E_CAPACITY and
    counted."""
    R = args.serve_large_requests
    t0 = time.time()
    first = gcl.mix_batch(R, args.seed + 11)
    rs = np.random.RandomState(args.seed + 12)
    graphs, nlarge, nreg = [], 0, 0
    for i in range(R):
        if rs.rand() < 0.02:
            n = int(round(np.exp(rs.uniform(np.log(10000), np.log(50000)))))
            if rs.rand() < 0.5:
                graphs.append(gcl.erdos_renyi_mean_degree(n, 8.0, 3000017 + i))
            else:
                graphs.append(gcl.random_regular(n + (n % 2), 3, 3500017 + i))
                nreg += 1
            nlarge += 1
        else:
            graphs.append(first.graph(i))
    pair_with = np.full(R, -1, np.int32)
    partners = []
    for i in range(R):
        if rs.rand() < 0.5:
            g = graphs[i]
            partners.append(gcl.permute(g, 4000037 + i) if rs.rand() < 0.5 else gcl.edge_swap(g, 5000011 + i))
            pair_with[i] = R + len(partners) - 1
    batch = gcl.Batch(graphs + partners)
    max_nnz = max(int(g.nnz) for g in graphs + partners)
    log(f"served-large workload: {R} requests ({nlarge} large, {nreg} of them 3-regular, {len(partners)} pairs) "
        f"built in {time.time() - t0:.1f}s")
    cfg = dict(max_n=50000, max_graph_nnz=max_nnz)
    out = serve_run(args, batch, R, pair_with, len(partners), "served-large", cfg)
    out["large_requests"] = nlarge
    out["large_regular_requests"] = nreg
    out["admission_cap_splitter_steps"] = int(gcl.server_config().max_search_cost)
    out["workload"] = (f"same size mix with 2% large graphs, n log-uniform [10000,50000], half sparse G(n,p) of "
                       f"mean degree 8 and half 3-regular; half canonicalize, half pair tests (permuted copy or "
                       f"one-swap variant, 50/50); {R} requests in each phase; requests whose estimated search "
                       f"cost is above the admission cap are refused and counted in refused_capacity")
    return out


def serve_run(args, batch, R, pair_with, npairs, tag, cfg, poisson_total=None, chunk=None):
    """Phase 1 submits the R requests at once (capacity); phase 2 replays a
    Poisson stream at half that rate (latency from arrival to completion):
    poisson_total requests (default R), the first R being phase 1's and each
    further R a chunk from chunk(k) = (batch, pair_with), served back to back
    by the same server.  Requests refused at admission (GCL_E_CAPACITY) are
    counted and left out of the rates and latencies."""
    import concurrent.futures
    total = poisson_total or R
    srv = gcl.Server(max_aut=10000, lifetime_ms=600000, **cfg)
    try:
        warm = min(R, 512)
        srv.replay(batch, np.zeros(warm), pair_with[:warm])
        s0 = srv.stats()
        t1 = time.perf_counter()
        r1 = srv.replay(batch, np.zeros(R), pair_with)
        cap_s = time.perf_counter() - t1
        s1 = srv.stats()
        refused1 = r1["status"] == -6
        if np.any((r1["status"] != 0) & ~refused1):
            raise SystemExit("bench.py: served requests failed")
        served_ok = ~refused1
        labelings = int(served_ok.sum()) + int(((pair_with >= 0) & served_ok).sum())
        cap = labelings / cap_s
        rate = 0.5 * int(served_ok.sum()) / cap_s
        arr_rs = np.random.RandomState(args.seed + 3)
        parts = []
        t2 = time.perf_counter()
        with concurrent.futures.ThreadPoolExecutor(1) as ex:
            done, k, cur, nxt = 0, 0, (batch, pair_with), None
            while done < total:
                cnt = min(R, total - done)
                if done + cnt < total:
                    nxt = ex.submit(chunk, k + 1)  # built while this chunk is served
                arrival = np.cumsum(arr_rs.exponential(1.0 / rate, size=cnt))
                parts.append(srv.replay(cur[0], arrival, cur[1][:cnt]))
                done += cnt
                k += 1
                if done < total:
                    cur = nxt.result()
                    log(f"{tag}: poisson {done}/{total} requests")
        lat_s = time.perf_counter() - t2
        r2 = {f: np.concatenate([p[f] for p in parts]) for f in parts[0]}
        refused2 = r2["status"] == -6
        if np.any((r2["status"] != 0) & ~refused2):
            raise SystemExit("bench.py: served requests failed")
        st = srv.stats()
    finally:
        srv.close()
    ok2 = ~refused2
    lat = r2["latency_ms"][ok2]
    dev = r2["device_ms"][ok2]
    log(f"{tag}: capacity {cap:.0f} labelings/s ({cap_s:.2f}s), poisson {rate:.0f} req/s: "
        f"p50 {np.percentile(lat, 50):.2f} ms p99 {np.percentile(lat, 99):.2f} ms; refused {int(refused2.sum())}")
    return {"value": round(cap, 2), "unit": "labelings/s",
            "requests": R, "pair_requests": npairs, "refused_capacity": int(refused2.sum()),
            "poisson_requests": int(total), "poisson_chunk": R,
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                           "p99": round(float(np.percentile(lat, 99)), 3), "max": round(float(lat.max()), 3)},
            "poisson_rate_req_s": round(rate, 1), "poisson_wall_s": round(lat_s, 3),
            "queue_ms_p50": round(float(np.percentile(r2["queue_ms"][ok2], 50)), 3),
            "device_ms": {"p50": round(float(np.percentile(dev, 50)), 3),
                          "p99": round(float(np.percentile(dev, 99)), 3)},
            "capacity_scheduler_ms": {k: round(s1["sched_" + k + "_ms"] - s0["sched_" + k + "_ms"], 1)
                                      for k in ("publish", "collect")},
            "serving_kernels": int(st["launches"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--graphs", type=int, default=65536)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--hipgraph", action="store_true",
                    help="time the steps as replays of the captured hipGraph instead of eager launches "
                         "(per-class HIP events on each class's stream)")
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--skip-baselines", action="store_true")
    ap.add_argument("--skip-configs", action="store_true",
                    help="leave out the single-graph legs of the first three configurations")
    ap.add_argument("--ref-limit", type=float, default=600.0,
                    help="seconds the bench waits for the reference's runs of the configuration graphs")
    ap.add_argument("--autotune", type=int, default=0,
                    help="pick the launch tuning with gcl_batch_autotune on a sample of this many graphs "
                         "(untimed, before the warmup); 0 uses the defaults")
    ap.add_argument("--serve-requests", type=int, default=0,
                    help="requests of the served measurement, e.g. 20000 (0 skips it; off by default: one "
                         "served run of the current kernels faulted the GPU and its cause is not found)")
    ap.add_argument("--serve-poisson", type=int, default=0,
                    help="requests of the served Poisson phase (0: as many as --serve-requests); "
                         f"{POISSON_CONFIGURED} is the serving configuration's stream")
    ap.add_argument("--serve-large-requests", type=int, default=0,
                    help="requests of the served measurement with 2%% large graphs, e.g. 2000 (0 skips it; "
                         "off by default, as --serve-requests)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "graphs are independent, so a batch shards across GPUs without exchange.", file=sys.stderr)
        sys.exit(2)
    if gcl.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")

    t0 = time.time()
    batch = gcl.mix_batch(args.graphs, args.seed)
    G = batch.num_graphs
    log(f"generated {G} graphs, n total {int(batch.vtx_off[-1])}, nnz {int(batch.row_ptr[-1])} "
        f"in {time.time() - t0:.1f}s")
    tmp = tempfile.TemporaryDirectory()
    cpu_out = None
    ref_legs = cfg_graphs = None
    if not args.skip_baselines:
        # CPU legs first, on an idle host: the reference on a bounded sample
        # of the batch, one core, then all cores
        cpu_out = cpu_baseline(batch, args.cpu_budget, 128)
        cpu_all = cpu_baseline_all_cores(batch, cpu_out[1])
        log(f"cpu baseline: {cpu_out[0]['value']:.1f} labelings/s on 1 core")
    if not args.skip_configs:
        cfg_graphs = build_config_graphs()
        if not args.skip_baselines:
            ref_legs = start_reference_legs(cfg_graphs, tmp.name)
    tuning, tune_info = gcl.batch_tuning_default(), None
    if args.autotune > 0:
        t1 = time.time()
        tuning, tms, tried = gcl.autotune(batch, max_aut=10000, sample=args.autotune)
        tune_info = {"sample_graphs": args.autotune, "candidates": tried, "sample_kernel_ms": round(tms, 3),
                     "seconds": round(time.time() - t1, 1)}
        log(f"autotune: {tuning} ({tms:.1f} ms on the sample, {tried} candidates, {time.time() - t1:.1f}s)")
    db = gcl.DeviceBatch(batch, max_aut=10000, tuning=tuning)
    use_graph = args.hipgraph
    for w in range(args.warmup):
        k, t = db.run(use_graph=use_graph)
        log(f"warmup {w}: kernel {k:.1f} ms, wall {t:.1f} ms")
    walls, kernels, class_ms = [], [], []

    def eager_step(tag):
        k, t = db.run(use_graph=False)
        cs = db.class_stats()  # after the run: events were recorded on each class stream
        class_ms.append([c["kernel_ms"] for c in cs])
        log(f"{tag}: kernel {k:.1f} ms, wall {t:.1f} ms; per class " +
            ", ".join(f"n<={c['max_n']}: {c['kernel_ms']:.1f}" for c in cs))
        return k, t, cs

    for s in range(args.steps):
        if use_graph:
            k, t = db.run(use_graph=True)
            log(f"step {s}: graph replay {k:.1f} ms, wall {t:.1f} ms")
        else:
            k, t, cs = eager_step(f"step {s}")
        walls.append(t)
        kernels.append(k)
    # the other launch shape, untimed for `value`: replays of the captured
    # hipGraph after eager steps, or eager steps (per-class HIP events on
    # each class's stream, which a replayed graph does not carry) after replays
    other = []
    if use_graph:
        for s in range(2):
            k, t, cs = eager_step(f"eager step {s}")
            other.append(t)
    else:
        db.run(use_graph=True)  # capture, instantiation and a first replay, untimed
        for s in range(3):
            k, t = db.run(use_graph=True)
            log(f"graph replay {s}: {k:.1f} ms, wall {t:.1f} ms")
            other.append(t)
    res = db.results()
    total_s = sum(walls) / 1e3
    value = args.steps * G / total_s
    ms_per_step = sum(walls) / args.steps
    kernel_ms = sum(kernels) / args.steps
    avg_class_ms = np.mean(np.array(class_ms), axis=0)
    dom = int(np.argmax(avg_class_ms))
    dom_name, dom_ms = kernel_name(cs[dom]["max_n"]), float(avg_class_ms[dom])
    dom_alg, dom_csr = int(cs[dom]["alg_bytes"]), int(cs[dom]["csr_bytes"])
    achieved = dom_alg / (dom_ms / 1e3) / 1e9
    traffic = traffic_from_profile(dom_name, G)
    classes = []
    for c, ms in zip(cs, avg_class_ms):
        name = kernel_name(c["max_n"])
        tr = traffic_from_profile(name, G)
        classes.append({"kernel": name, "graphs": c["graphs"], "workers": c["workers"], "kernel_ms": round(float(ms), 3),
                        "alg_bytes": int(c["alg_bytes"]), "csr_bytes": int(c["csr_bytes"]),
                        "achieved_gbs": round(c["alg_bytes"] / (ms / 1e3) / 1e9, 3),
                        "hbm_traffic": tr,
                        "traffic_over_alg": None if tr is None else round(tr / max(c["alg_bytes"], 1), 3)})
    _, _, _, dev_ns = db.graph_stats()
    search_ms = dev_ns.astype(np.float64) / 1e6
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "labelings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generators)",
        "config": {"workload": f"batched mix: {G} graphs, n log-uniform [16,1024], 50% G(n,p) deg U[3,12], "
                               "30% random d-regular d in {3,4,5}, 20% C_a x C_b a,b in U[4,32], seeded relabeling",
                   "graphs": G, "seed": args.seed, "hipgraph": use_graph, "workers": res.workers,
                   "tuning": tuning, "autotune": tune_info},
        ("eager_ms_per_step" if use_graph else "hipgraph_ms_per_step"): round(float(np.mean(other)), 3),
        "search_nodes_per_s": round(res.total_nodes / (kernel_ms / 1e3), 1),
        "graph_search_ms": {"p50": round(float(np.percentile(search_ms, 50)), 3),
                            "p99": round(float(np.percentile(search_ms, 99)), 3),
                            "max": round(float(search_ms.max()), 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "kernel": dom_name, "kernel_ms": round(dom_ms, 3), "algorithmic_bytes": dom_alg,
                     "csr_bytes": dom_csr,
                     "whole_step_gbs": round(res.alg_bytes / (ms_per_step / 1e3) / 1e9, 3),
                     "classes": classes},
    }
    if not args.skip_baselines:
        # plain path: one launch per graph (allocation, copies, launch, sync)
        ps = min(args.plain_sample, G)
        idx = np.linspace(0, G - 1, ps).astype(np.int64)
        sub = gcl.Batch([batch.graph(int(i)) for i in idx])
        t1 = time.perf_counter()
        plain = gcl.batch_canon(sub, plain=True)
        t_plain = time.perf_counter() - t1
        for j, i in enumerate(idx):
            a = sub.labels_of(plain.labels, j)
            b = res.labels[batch.vtx_off[i]:batch.vtx_off[i + 1]]
            if not np.array_equal(a, b):
                raise SystemExit(f"bench.py: plain and batched labelings differ on graph {i}")
        plain_rate = ps / t_plain
        # per-graph latency of the plain path, split into the search
        # kernel's time and the rest of the call (setup, copies, teardown)
        pk, pw = [], []
        for j in range(min(ps, 128)):
            r = gcl.canon_label(sub.graph(j), engine="gpu", max_aut=64)
            pk.append(r.kernel_ms)
            pw.append(r.total_ms)
        out["plain"] = {"value": round(plain_rate, 2), "unit": "labelings/s",
                        "latency_ms": {"p50": round(float(np.percentile(pw, 50)), 3),
                                       "p99": round(float(np.percentile(pw, 99)), 3)},
                        "kernel_ms": {"p50": round(float(np.percentile(pk, 50)), 3),
                                      "p99": round(float(np.percentile(pk, 99)), 3)},
                        "sample": f"{ps} graphs evenly spaced over the batch, one launch each"}
        out["batched_vs_plain"] = round(value / plain_rate, 2)
    if cfg_graphs is not None:
        rows, verdicts = plain_configs(cfg_graphs)
    if not args.skip_baselines:
        if args.serve_requests > 0:
            out["served"] = served(args)
            out["served_vs_plain"] = round(out["served"]["value"] / plain_rate, 2)
            out["p99_latency_ms"] = out["served"]["latency_ms"]["p99"]
        if args.serve_large_requests > 0:
            out["served_large"] = served_large(args)
        out["cpu_baseline"] = cpu_out[0]
        out["cpu_baseline_all_cores"] = cpu_all
        out["verified_vs_reference"] = verify_batch_vs_reference(batch, res.labels, cpu_out[2])
        log(f"batched labels vs reference output: {out['verified_vs_reference']}")
    if cfg_graphs is not None:
        ref = collect_reference_legs(ref_legs, args.ref_limit) if ref_legs else None
        out["configs_single_graph"] = configs_table(rows, verdicts, ref)
    tmp.cleanup()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: canonical labelings/s of the batched mix on one MI355X.

Workload (one "step"): canonical labeling of a fixed batch of 65,536 graphs,
n log-uniform in [16, 1024]; 50% G(n,p) with mean degree U[3,12], 30% random
d-regular (d in {3,4,5}), 20% C_a x C_b (a, b in U[4,32]); each under a seeded
random relabeling.  The batch is uploaded once (inputs resident in HBM); a
step is one run of the persistent search launch (one kernel per size class,
each on its own stream, bracketed by HIP events on that stream; --hipgraph
replays the same launches from a captured hipGraph instead) and ends with a
stream synchronize.

Printed as one JSON line with the throughput, the plain one-graph-per-launch
rate on a sample of the same batch, the HBM roofline figure of the search
launch and the CPU baseline (the reference's sequential binary, or the oracle
where that binary is absent, on a bounded sample, one core).
"""
import argparse
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


def cpu_baseline(batch, budget_s, stride):
    """Times the reference's sequential binary (or the oracle) on graphs
    0, stride, 2*stride, ... of the batch until budget_s of CPU time is spent."""
    import oracle
    ref = oracle.REF_BIN
    sample = list(range(0, batch.num_graphs, stride))
    done, spent = 0, 0.0
    kind = "reference" if os.path.exists(ref) else "port"
    with tempfile.TemporaryDirectory() as tmp:
        for i in sample:
            g = batch.graph(i)
            if kind == "reference":
                f = os.path.join(tmp, "g.dimacs")
                gcl.write_dimacs(f, g)
                p = subprocess.run([ref, f, "--sequential", "--max-aut=10000"], cwd=tmp, capture_output=True,
                                   text=True, timeout=max(60, int(budget_s * 4)))
                m = RE_TIME.search(p.stdout)
                if p.returncode != 0 or not m:
                    raise RuntimeError("reference run failed: " + p.stderr[-300:])
                mi, s, ms, us, ns = (int(x) for x in m.groups())
                spent += mi * 60 + s + ms * 1e-3 + us * 1e-6 + ns * 1e-9
            else:
                t0 = time.perf_counter()
                oracle.canon_graph(g, want_dense=False)
                spent += time.perf_counter() - t0
            done += 1
            if spent >= budget_s:
                break
    return dict(value=done / spent, unit="labelings/s", cores=1, kind=kind,
                sample=f"graphs 0,{stride},{2 * stride},... of the same batch: {done} graphs, {spent:.1f} s "
                       f"of single-thread algorithm time ({'reference sequential binary' if kind == 'reference' else 'oracle'})"), \
        sample[:done]


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


def served(args):
    """Served measurement: requests over the same mix of sizes, half
    canonicalize and half isomorphism tests (partner: a permuted copy or a
    one-swap variant, 50/50), through the request server."""
    R = args.serve_requests
    t0 = time.time()
    first = gcl.mix_batch(R, args.seed + 1)
    rs = np.random.RandomState(args.seed + 2)
    graphs = [first.graph(i) for i in range(R)]
    pair_with = np.full(R, -1, np.int32)
    partners = []
    for i in range(R):
        if rs.rand() < 0.5:
            g = graphs[i]
            partners.append(gcl.permute(g, 1000003 + i) if rs.rand() < 0.5 else gcl.edge_swap(g, 2000003 + i))
            pair_with[i] = R + len(partners) - 1
    batch = gcl.Batch(graphs + partners)
    log(f"served workload: {R} requests ({len(partners)} pairs) built in {time.time() - t0:.1f}s")
    out = serve_run(args, batch, R, pair_with, len(partners), "served", dict(max_n=1024))
    out["workload"] = ("same size mix (n log-uniform [16,1024]); half canonicalize, half pair tests "
                       "(permuted copy or one-swap variant, 50/50); persistent serving kernels, one per size "
                       "class; the 2% of n 10k-50k graphs of the serving configuration are in served_large")
    return out


def served_large(args):
    """The serving configuration's 2% of large graphs: requests over the same
    size mix where each is, with probability 0.02, a sparse G(n, p) of mean
    degree 8 with n log-uniform in [10000, 50000]; half canonicalize, half
    pair tests as above.  The server runs a second persistent kernel for the
    large class (global-memory partition state)."""
    R = args.serve_large_requests
    t0 = time.time()
    first = gcl.mix_batch(R, args.seed + 11)
    rs = np.random.RandomState(args.seed + 12)
    graphs, nlarge = [], 0
    for i in range(R):
        if rs.rand() < 0.02:
            n = int(round(np.exp(rs.uniform(np.log(10000), np.log(50000)))))
            graphs.append(gcl.erdos_renyi_mean_degree(n, 8.0, 3000017 + i))
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
    log(f"served-large workload: {R} requests ({nlarge} large, {len(partners)} pairs) built in "
        f"{time.time() - t0:.1f}s")
    out = serve_run(args, batch, R, pair_with, len(partners), "served-large",
                    dict(max_n=50000, max_graph_nnz=max_nnz))
    out["large_requests"] = nlarge
    out["workload"] = ("same size mix with 2% sparse G(n,p) of mean degree 8, n log-uniform [10000,50000]; "
                       "half canonicalize, half pair tests (permuted copy or one-swap variant, 50/50); the "
                       "large 3-regular graphs of the serving configuration are not in this bounded run "
                       "(n + 1 search nodes and about n^2 splitter steps each)")
    return out


def serve_run(args, batch, R, pair_with, npairs, tag, cfg):
    """Phase 1 submits the R requests at once (capacity); phase 2 replays a
    Poisson stream at half that rate (latency from arrival to completion)."""
    labelings = R + npairs
    srv = gcl.Server(max_aut=10000, lifetime_ms=600000, **cfg)
    try:
        srv.replay(batch, np.zeros(min(R, 512)), pair_with[:min(R, 512)])  # warm-up
        s0 = srv.stats()
        t1 = time.perf_counter()
        r1 = srv.replay(batch, np.zeros(R), pair_with)
        cap_s = time.perf_counter() - t1
        s1 = srv.stats()
        if np.any(r1["status"] != 0):
            raise SystemExit("bench.py: served requests failed")
        cap = labelings / cap_s
        rate = 0.5 * R / cap_s
        arrival = np.cumsum(np.random.RandomState(args.seed + 3).exponential(1.0 / rate, size=R))
        t2 = time.perf_counter()
        r2 = srv.replay(batch, arrival, pair_with)
        lat_s = time.perf_counter() - t2
        if np.any(r2["status"] != 0):
            raise SystemExit("bench.py: served requests failed")
        st = srv.stats()
    finally:
        srv.close()
    lat = r2["latency_ms"]
    log(f"{tag}: capacity {cap:.0f} labelings/s ({cap_s:.2f}s), poisson {rate:.0f} req/s: "
        f"p50 {np.percentile(lat, 50):.2f} ms p99 {np.percentile(lat, 99):.2f} ms")
    return {"value": round(cap, 2), "unit": "labelings/s",
            "requests": R, "pair_requests": npairs,
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                           "p99": round(float(np.percentile(lat, 99)), 3), "max": round(float(lat.max()), 3)},
            "poisson_rate_req_s": round(rate, 1), "poisson_wall_s": round(lat_s, 3),
            "queue_ms_p50": round(float(np.percentile(r2["queue_ms"], 50)), 3),
            "device_ms": {"p50": round(float(np.percentile(r2["device_ms"], 50)), 3),
                          "p99": round(float(np.percentile(r2["device_ms"], 99)), 3)},
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
                    help="replay the run as a captured hipGraph (no per-kernel events, so no roofline)")
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--skip-baselines", action="store_true")
    ap.add_argument("--autotune", type=int, default=0,
                    help="pick the launch tuning with gcl_batch_autotune on a sample of this many graphs "
                         "(untimed, before the warmup); 0 uses the defaults")
    ap.add_argument("--serve-requests", type=int, default=20000,
                    help="requests of the served measurement (0 skips it)")
    ap.add_argument("--serve-large-requests", type=int, default=2000,
                    help="requests of the served measurement with 2%% large graphs (0 skips it)")
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
    for s in range(args.steps):
        k, t = db.run(use_graph=use_graph)
        walls.append(t)
        kernels.append(k)
        cs = db.class_stats()  # after the run: events were recorded on each class stream
        class_ms.append([c["kernel_ms"] for c in cs])
        log(f"step {s}: kernel {k:.1f} ms, wall {t:.1f} ms; per class " +
            ", ".join(f"n<={c['max_n']}: {c['kernel_ms']:.1f}" for c in cs))
    res = db.results()
    total_s = sum(walls) / 1e3
    value = args.steps * G / total_s
    ms_per_step = sum(walls) / args.steps
    kernel_ms = sum(kernels) / args.steps
    # roofline of the dominant kernel: its algorithmic CSR bytes per launch
    # over its average launch duration (HIP events on its own stream)
    if use_graph:
        # a replayed graph carries no events: whole launch instead
        dom_name, dom_ms, dom_bytes = "all size classes (hipGraph replay)", kernel_ms, res.csr_bytes
    else:
        avg_class_ms = np.mean(np.array(class_ms), axis=0)
        dom = int(np.argmax(avg_class_ms))
        dom_name, dom_ms, dom_bytes = kernel_name(cs[dom]["max_n"]), float(avg_class_ms[dom]), cs[dom]["csr_bytes"]
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
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
        "search_nodes_per_s": round(res.total_nodes / (kernel_ms / 1e3), 1),
        "graph_search_ms": {"p50": round(float(np.percentile(search_ms, 50)), 3),
                            "p99": round(float(np.percentile(search_ms, 99)), 3),
                            "max": round(float(search_ms.max()), 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6),
                     "traffic": traffic_from_profile(dom_name, G),
                     "kernel": dom_name, "kernel_ms": round(dom_ms, 3), "algorithmic_bytes": int(dom_bytes)},
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
        # per-graph latency of the plain path (one launch per graph)
        plat = []
        for j in range(min(ps, 64)):
            g = sub.graph(j)
            t2 = time.perf_counter()
            gcl.canon_label(g, engine="gpu", max_aut=64)
            plat.append((time.perf_counter() - t2) * 1e3)
        out["plain"] = {"value": round(plain_rate, 2), "unit": "labelings/s",
                        "latency_ms": {"p50": round(float(np.percentile(plat, 50)), 3),
                                       "p99": round(float(np.percentile(plat, 99)), 3)},
                        "sample": f"{ps} graphs evenly spaced over the batch, one launch each"}
        out["batched_vs_plain"] = round(value / plain_rate, 2)
        if args.serve_requests > 0:
            out["served"] = served(args)
            out["served_vs_plain"] = round(out["served"]["value"] / plain_rate, 2)
            out["p99_latency_ms"] = out["served"]["latency_ms"]["p99"]
        if args.serve_large_requests > 0:
            out["served_large"] = served_large(args)
        out["cpu_baseline"], done = cpu_baseline(batch, args.cpu_budget, 128)
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores(batch, done)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

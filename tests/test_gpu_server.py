"""Request server (continuous batching) through the C ABI: results are the
same alone and in the mix, isomorphism verdicts match the oracle, admission
refuses what does not fit, and a Poisson replay completes every request."""
import sys

import numpy as np
import pytest

import gcl
import oracle

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mix(count, seed):
    rs = np.random.RandomState(seed)
    out = []
    for i in range(count):
        n = int(np.exp(rs.uniform(np.log(16), np.log(600))))
        k = rs.randint(3)
        if k == 0:
            g = gcl.erdos_renyi_mean_degree(n, float(rs.uniform(3, 12)), seed * 1000 + i)
        elif k == 1:
            d = int(rs.choice([3, 4, 5]))
            g = gcl.random_regular(n + (n * d) % 2, d, seed * 1000 + i)
        else:
            g = gcl.torus(int(rs.randint(4, 20)), int(rs.randint(4, 20)))
        out.append(gcl.permute(g, seed * 7919 + i))
    return out


def test_server_same_alone_and_in_mix():
    graphs = _mix(90, 5)
    srv = gcl.Server(max_n=1024, max_aut=64, lifetime_ms=60000)
    try:
        _log("alone")
        alone = {}
        for i in (0, 7, 33, 60, 89):
            t = srv.submit(graphs[i])
            alone[i] = srv.wait_all([t], timeout_s=60)[t]["labels"].copy()
        _log("mix")
        tickets = [srv.submit(g) for g in graphs]
        res = srv.wait_all(tickets, timeout_s=60)
        for i, t in enumerate(tickets):
            r = res[t]
            assert r["status"] == 0, (i, r)
            keys = gcl.canonical_keys(graphs[i], r["labels"])
            o = oracle.canon_graph(graphs[i], max_aut=64)
            assert np.array_equal(keys, o["keys"]), i
            if i in alone:
                assert np.array_equal(r["labels"], alone[i]), i
        st = srv.stats()
        assert st["completed"] >= 95 and st["pending"] == 0 and st["in_flight"] == 0
        assert st["launches"] == 1  # one serving kernel for n <= 1024
    finally:
        srv.close()


def test_server_pair_verdicts_match_oracle():
    base = _mix(24, 9)
    srv = gcl.Server(max_n=1024, lifetime_ms=60000)
    try:
        pairs = []
        for i, g in enumerate(base):
            other = gcl.permute(g, 500 + i) if i % 2 == 0 else gcl.edge_swap(g, 700 + i)
            pairs.append((g, other))
        tickets = [srv.submit(a, b) for a, b in pairs]
        res = srv.wait_all(tickets, timeout_s=60)
        for (a, b), t in zip(pairs, tickets):
            want = np.array_equal(oracle.canon_graph(a)["keys"], oracle.canon_graph(b)["keys"]) if a.m == b.m else False
            assert res[t]["status"] == 0
            assert res[t]["verdict"] == (1 if want else 0), (a.n, res[t])
    finally:
        srv.close()


def test_server_admission():
    srv = gcl.Server(max_n=256, max_graph_nnz=4096, queue_bytes=64 * 1024, lifetime_ms=60000)
    try:
        big = gcl.random_regular(300, 3, 1)
        with pytest.raises(gcl.GclError) as e:
            srv.submit(big)
        assert e.value.status == -6  # capacity
        g = gcl.random_regular(200, 3, 2)
        # fill the queue budget; every accepted request still completes
        tickets, busy = [], 0
        for i in range(64):
            try:
                tickets.append(srv.submit(gcl.permute(g, i)))
            except gcl.GclError as e:
                assert e.status == -8  # busy
                busy += 1
        assert busy > 0 and tickets
        res = srv.wait_all(tickets, timeout_s=60)
        assert all(r["status"] == 0 for r in res.values())
        assert srv.stats()["refused_busy"] == busy
    finally:
        srv.close()


def test_server_replay_poisson():
    graphs = _mix(400, 13)
    batch = gcl.Batch(graphs)
    rs = np.random.RandomState(1)
    arrival = np.cumsum(rs.exponential(1.0 / 4000.0, size=len(graphs)))
    pair_with = np.full(len(graphs), -1, np.int32)
    pair_with[::4] = np.arange(1, len(graphs), 4)[: len(pair_with[::4])]
    srv = gcl.Server(max_n=1024, lifetime_ms=60000)
    try:
        _log("replay")
        r = srv.replay(batch, arrival, pair_with)
        assert np.all(r["status"] == 0)
        assert np.all(r["latency_ms"] > 0)
        assert np.all((r["verdict"] == -1) == (pair_with < 0))
        st = srv.stats()
        assert st["completed"] == len(graphs)
    finally:
        srv.close()


def test_server_golden_graphs_and_pairs(golden):
    """Every golden graph (one vertex, no edges, coloured, ...) through the
    server gives the reference's `output` bytes; golden pairs its verdicts."""
    from conftest import golden_graph, golden_output
    srv = gcl.Server(max_n=1024, max_aut=10000, lifetime_ms=60000)
    try:
        singles = [e for e in golden["singles"] if e["n"] <= 1024]
        _log(f"golden singles {len(singles)}")
        gs = [golden_graph(e["name"]) for e in singles]
        tickets = [srv.submit(g) for g in gs]
        res = srv.wait_all(tickets, timeout_s=60)
        for e, g, t in zip(singles, gs, tickets):
            r = res[t]
            assert r["status"] == 0, (e["name"], r)
            keys = gcl.canonical_keys(g, r["labels"])
            assert gcl.format_adjacency(keys, g.n).encode() == golden_output(e), e["name"]
        _log(f"golden pairs {len(golden['pairs'])}")
        pairs = [(golden_graph(p["a"]), golden_graph(p["b"])) for p in golden["pairs"]]
        tickets = [srv.submit(a, b) for a, b in pairs]
        res = srv.wait_all(tickets, timeout_s=60)
        for p, t in zip(golden["pairs"], tickets):
            r = res[t]
            assert r["status"] == 0, (p, r)
            assert ("ISOMORPHIC" if r["verdict"] == 1 else "NOT ISOMORPHIC") == p["verdict"], p
    finally:
        srv.close()


def test_server_large_class_same_as_batch_path():
    # graphs of n 10k-50k go to the server's second (global-memory) serving
    # kernel; alone and amid small requests they get the batch path's labels,
    # node counts and pair verdicts
    big = [gcl.permute(gcl.erdos_renyi_mean_degree(n, 8.0, 40 + n), n) for n in (12000, 30000)]
    ref = [gcl.canon_label(g, engine="gpu") for g in big]
    swap = gcl.edge_swap(big[0], 9)
    # batch-path results first: a batch launched while serving kernels hold
    # the GPU waits for them to exit
    want_verdict = int(gcl.is_isomorphic(big[0], swap, engine="gpu"))
    srv = gcl.Server(max_n=50000, max_graph_nnz=max(g.nnz for g in big + [swap]), lifetime_ms=60000)
    try:
        for g, r in zip(big, ref):
            _log(f"large alone n={g.n}")
            t = srv.submit(g)
            res = srv.wait_all([t], timeout_s=60)[t]
            assert res["status"] == 0
            assert np.array_equal(res["labels"], r.label) and res["nodes"] == r.nodes
        _log("large in a mix")
        small = _mix(300, 41)
        tickets, want = [], {}
        for i, g in enumerate(small):
            tickets.append(srv.submit(g))
            if i == 100:
                t = srv.submit(big[1])
                tickets.append(t)
                want[t] = ref[1]
            if i == 200:
                tp = srv.submit(big[0], swap)
                tickets.append(tp)
        got = srv.wait_all(tickets, timeout_s=90)
        assert all(got[t]["status"] == 0 for t in tickets)
        for t, r in want.items():
            assert np.array_equal(got[t]["labels"], r.label) and got[t]["nodes"] == r.nodes
        assert got[tp]["verdict"] == want_verdict
    finally:
        srv.close()


def test_server_large_regular_graph_and_admission_cap():
    """The serving configuration's large 3-regular graphs: n = 10,000 (an
    estimated 10^8 splitter steps, under the default cap of 1.5e8) is served
    within 60 s and gets the batch path's labels; n = 13,000 (1.69e8) is
    refused at admission with GCL_E_CAPACITY and counted."""
    import time
    g = gcl.permute(gcl.random_regular(10000, 3, 7), 3)
    ref = gcl.canon_label(g, engine="gpu")  # before the server holds the GPU
    too_big = gcl.random_regular(13000, 3, 8)
    srv = gcl.Server(max_n=50000, max_graph_nnz=400000, lifetime_ms=180000)
    try:
        with pytest.raises(gcl.GclError) as ei:
            srv.submit(too_big)
        assert ei.value.status == -6
        assert srv.stats()["refused_capacity"] == 1
        _log("large regular n=10000 served")
        t0 = time.time()
        t = srv.submit(g)
        res = srv.wait_all([t], timeout_s=60)[t]
        elapsed = time.time() - t0
        _log(f"served in {elapsed:.2f} s (batch path kernel {ref.kernel_ms:.0f} ms)")
        assert res["status"] == 0
        assert np.array_equal(res["labels"], ref.label) and res["nodes"] == ref.nodes
        assert elapsed < 60
    finally:
        srv.close()


def test_server_stopped_by_its_lifetime_fails_requests():
    """Serving kernels that returned because their lifetime passed take no
    more requests: a request submitted afterwards completes with
    GCL_E_TIMEOUT, and later submissions are refused with it."""
    import time
    srv = gcl.Server(max_n=256, lifetime_ms=200)
    try:
        time.sleep(1.0)
        t = srv.submit(gcl.torus(5, 6))
        r = srv.wait_all([t], timeout_s=10)[t]
        assert r["status"] == -9, r
        with pytest.raises(gcl.GclError) as ei:
            srv.submit(gcl.torus(5, 6))
        assert ei.value.status == -9
    finally:
        srv.close()

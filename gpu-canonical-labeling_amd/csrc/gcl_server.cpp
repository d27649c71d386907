// gcl_server.cpp -- request server of libgcl: admission by footprint,
// continuous batching over fixed launch slots, completion and load replay.
// The device side of a slot (buffers, kernels) is in gcl_device.hip.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <stdexcept>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gcl.h"
#include "gcl_device.h"
#include "gcl_internal.h"

namespace gcl {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    int ngraphs = 1;  // 1: canonicalize, 2: isomorphism test
    Graph g[2];
    int32_t* out_label = nullptr;
    Clock::time_point t0;
    uint64_t footprint = 0;
};

// Device and staging bytes a graph takes in a launch: row offsets, column
// ids, label and result record.
uint64_t graph_footprint(int64_t n, int64_t nnz) {
    return 8ull * (uint64_t)(n + 1) + 4ull * (uint64_t)nnz + 4ull * (uint64_t)n + sizeof(DeviceGraphResult);
}

// Estimated search cost in splitter steps.  When every vertex has the same
// degree and the same colour, the initial partition is already equitable:
// refinement at the root splits nothing, and a graph without automorphisms
// (a random regular graph) has n root children that each refine to about n
// splitter steps, n^2 in all.  Otherwise degree refinement splits the root
// and the search is short; it is costed as n.
uint64_t search_cost_estimate(const gcl_csr& g) {
    const uint64_t n = (uint64_t)g.n;
    if (g.n <= 1) return n;
    const int32_t d0 = g.row_ptr[1] - g.row_ptr[0];
    for (int32_t v = 1; v < g.n; ++v)
        if (g.row_ptr[v + 1] - g.row_ptr[v] != d0) return n;
    if (g.color)
        for (int32_t v = 1; v < g.n; ++v)
            if (g.color[v] != g.color[0]) return n;
    return n * n;
}

Graph copy_graph(const gcl_csr& c) {
    Graph g;
    g.n = c.n;
    g.row_ptr.assign(c.row_ptr, c.row_ptr + c.n + 1);
    g.col.assign(c.col, c.col + c.nnz);
    if (c.color) g.color.assign(c.color, c.color + c.n);
    return g;
}

}  // namespace

struct Server {
    gcl_server_config cfg;
    ServeEngine* eng = nullptr;
    struct Slot {
        Request* req = nullptr;
        int part = 0;  // graph 0 or 1 of the request
    };
    struct ClassState {
        std::vector<uint32_t> free_slots;
        std::vector<Slot> slots;
        size_t busy = 0;  // slots in flight
    };
    std::vector<uint32_t> cbuf = std::vector<uint32_t>(1024);  // completed slot ids
    std::vector<ClassState> cls;
    struct Live {
        int remaining = 0;
        int32_t status = 0;
        uint64_t nodes = 0, dev_ticks = 0;
        const int32_t* label[2] = {nullptr, nullptr};
        std::vector<int32_t> lab[2];
        Clock::time_point t_pub;
    };
    std::unordered_map<Request*, Live> live;
    uint64_t ticks_per_ms = 100000;
    std::atomic<uint64_t> publish_ns{0}, collect_ns{0};  // scheduler thread time
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::deque<gcl_server_result> done;
    uint64_t next_ticket = 1;
    uint64_t queued_bytes = 0;
    gcl_server_stats st{};
    bool stop = false;
    gcl_status fatal = GCL_OK;
    std::thread thr;

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (thr.joinable()) thr.join();
        if (eng) serve_destroy(eng);
        for (auto* r : pending) delete r;
        for (auto& kv : live) delete kv.first;
    }

    gcl_status create(const gcl_server_config& c) {
        cfg = c;
        if (cfg.max_n < 1 || cfg.max_n > kMaxVertices || cfg.max_graph_nnz < 0 || cfg.max_aut == 0 ||
            cfg.queue_bytes == 0 || cfg.slots < 0 || cfg.lifetime_ms < 0 || cfg.large_share_pct < 0 ||
            cfg.large_share_pct > 80 || (cfg.large_share_pct > 0 && cfg.large_share_pct < 5))
            return fail(GCL_E_INVALID, "invalid server configuration");
        ServeConfig sc;
        sc.large_share = (cfg.large_share_pct > 0 ? cfg.large_share_pct : 40) / 100.0;
        sc.max_n = cfg.max_n;
        sc.max_graph_nnz = cfg.max_graph_nnz;
        sc.max_aut = cfg.max_aut;
        sc.slots = cfg.slots;
        if (cfg.lifetime_ms > 0) sc.lifetime_ticks = device_wall_ticks_per_ms() * (uint64_t)cfg.lifetime_ms;
        ticks_per_ms = device_wall_ticks_per_ms();
        gcl_status r = serve_create(sc, &eng);
        if (r != GCL_OK) return r;
        cls.resize((size_t)serve_classes(eng));
        for (size_t k = 0; k < cls.size(); ++k) {
            const ServeClassInfo ci = serve_class_info(eng, (int)k);
            cls[k].slots.resize(ci.slots);
            for (uint32_t s = ci.slots; s-- > 0;) cls[k].free_slots.push_back(s);
        }
        st.launches = cls.size();
        thr = std::thread([this] { loop(); });
        return GCL_OK;
    }

    gcl_status submit(const gcl_csr* a, const gcl_csr* b, int32_t* out_label, uint64_t* ticket,
                      Clock::time_point t0) {
        gcl_status r;
        if ((r = gcl_csr_validate(a)) != GCL_OK) return r;
        if (b && (r = gcl_csr_validate(b)) != GCL_OK) return r;
        const gcl_csr* gs[2] = {a, b};
        const int ng = b ? 2 : 1;
        uint64_t fp = 0;
        auto refuse = [&](const char* why, int i, unsigned long long x, unsigned long long cap) {
            {
                std::lock_guard<std::mutex> lk(mu);
                ++st.refused_capacity;
            }
            return fail(GCL_E_CAPACITY, "graph (n=%d, nnz=%lld): %s %llu > %llu", gs[i]->n, (long long)gs[i]->nnz,
                        why, x, cap);
        };
        for (int i = 0; i < ng; ++i) {
            if (gs[i]->n > cfg.max_n)
                return refuse("n above the server's max_n", i, (unsigned long long)gs[i]->n,
                              (unsigned long long)cfg.max_n);
            if (serve_class_of(eng, gs[i]->n, gs[i]->nnz) < 0)
                return refuse("edges above the largest class's", i, (unsigned long long)gs[i]->nnz,
                              (unsigned long long)cfg.max_graph_nnz);
            const uint64_t cost = search_cost_estimate(*gs[i]);
            if (cfg.max_search_cost > 0 && cost > cfg.max_search_cost)
                return refuse("estimated search cost (splitter steps) above max_search_cost", i,
                              (unsigned long long)cost, (unsigned long long)cfg.max_search_cost);
            fp += graph_footprint(gs[i]->n, gs[i]->nnz);
        }
        if (fp > cfg.queue_bytes) return fail(GCL_E_CAPACITY, "request footprint exceeds queue_bytes");
        {
            // the footprint is reserved with the check, so concurrent
            // submitters cannot together pass queue_bytes
            std::lock_guard<std::mutex> lk(mu);
            if (fatal != GCL_OK) return fail(fatal, "server stopped: %s", gcl_status_string(fatal));
            if (queued_bytes + fp > cfg.queue_bytes) {
                ++st.refused_busy;
                return fail(GCL_E_BUSY, "server queue full");
            }
            queued_bytes += fp;
        }
        Request* q = new (std::nothrow) Request();
        bool copied = q != nullptr;
        if (q) {
            try {
                q->ngraphs = ng;
                for (int i = 0; i < ng; ++i) q->g[i] = copy_graph(*gs[i]);
            } catch (const std::bad_alloc&) {
                copied = false;
            }
        }
        if (!copied) {
            delete q;
            std::lock_guard<std::mutex> lk(mu);
            queued_bytes -= fp;
            return fail(GCL_E_NOMEM, "out of host memory");
        }
        q->out_label = out_label;
        q->t0 = t0;
        q->footprint = fp;
        {
            std::lock_guard<std::mutex> lk(mu);
            q->ticket = next_ticket++;
            ++st.submitted;
            pending.push_back(q);
            *ticket = q->ticket;
        }
        cv_work.notify_one();
        return GCL_OK;
    }

    // Writes graph `part` of q into slot s of class k (host-coherent memory).
    void fill(int k, uint32_t s, Request* q, int part) {
        const Graph& g = q->g[part];
        ServeSlotView v = serve_slot(eng, k, s);
        std::copy(g.row_ptr.begin(), g.row_ptr.end(), v.row);
        std::copy(g.col.begin(), g.col.end(), v.col);
        if (!g.color.empty()) {
            std::copy(g.color.begin(), g.color.end(), v.color);
            for (int32_t i = 0; i < g.n; ++i) v.init[i] = i;
            std::stable_sort(v.init, v.init + g.n, [&](int32_t x, int32_t y) { return g.color[x] < g.color[y]; });
            *v.colored = 1;
        } else {
            *v.colored = 0;
        }
        __atomic_store_n(v.n, g.n, __ATOMIC_RELAXED);
        cls[(size_t)k].slots[s] = Slot{q, part};
        ++cls[(size_t)k].busy;
        serve_publish(eng, k, s);
    }

    // Publishes pending requests (arrival order) while their classes have free slots.
    void dispatch() {
        for (;;) {
            Request* q;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (pending.empty()) return;
                q = pending.front();
            }
            int k[2] = {-1, -1};
            for (int i = 0; i < q->ngraphs; ++i) k[i] = serve_class_of(eng, q->g[i].n, q->g[i].row_ptr[q->g[i].n]);
            size_t need0 = 1, need1 = q->ngraphs == 2 ? 1 : 0;
            if (q->ngraphs == 2 && k[0] == k[1]) {
                need0 = 2;
                need1 = 0;
            }
            if (cls[(size_t)k[0]].free_slots.size() < need0) return;
            if (need1 && cls[(size_t)k[1]].free_slots.size() < need1) return;
            {
                std::lock_guard<std::mutex> lk(mu);
                pending.pop_front();
            }
            Live& L = live[q];
            L.remaining = q->ngraphs;
            L.t_pub = Clock::now();
            for (int i = 0; i < q->ngraphs; ++i) {
                auto& fs = cls[(size_t)k[i]].free_slots;
                const uint32_t s = fs.back();
                fs.pop_back();
                fill(k[i], s, q, i);
            }
            std::lock_guard<std::mutex> lk(mu);
            st.graphs_launched += (uint64_t)q->ngraphs;
        }
    }

    void finish(Request* q, Live& L) {
        gcl_server_result r{};
        r.ticket = q->ticket;
        r.n = q->g[0].n;
        r.verdict = -1;
        r.status = L.status;
        r.nodes = L.nodes;
        r.device_ms = (double)L.dev_ticks / (double)ticks_per_ms;
        if (L.status == 0 && q->out_label) std::copy(L.lab[0].begin(), L.lab[0].end(), q->out_label);
        if (L.status == 0 && q->ngraphs == 2) {
            const gcl_csr A = q->g[0].view(), B = q->g[1].view();
            if (A.n != B.n || A.nnz != B.nnz) {
                r.verdict = 0;
            } else {
                std::vector<uint32_t> ka((size_t)A.nnz), kb((size_t)B.nnz);
                canonical_keys(A, L.lab[0].data(), ka.data());
                canonical_keys(B, L.lab[1].data(), kb.data());
                r.verdict = ka == kb ? 1 : 0;
            }
        }
        const auto now = Clock::now();
        r.latency_ms = std::chrono::duration<double, std::milli>(now - q->t0).count();
        r.queue_ms = std::chrono::duration<double, std::milli>(L.t_pub - q->t0).count();
        std::lock_guard<std::mutex> lk(mu);
        queued_bytes -= q->footprint;
        done.push_back(r);
        ++st.completed;
    }

    // Collects the slots the device reported done; returns how many.
    size_t collect() {
        size_t got = 0;
        std::vector<Request*> finished;
        for (size_t k = 0; k < cls.size(); ++k) {
            auto& C = cls[k];
            for (;;) {
                const int m = C.busy > 0 ? serve_completions(eng, (int)k, cbuf.data(), (int)cbuf.size()) : 0;
                if (m == 0) break;
                for (int i = 0; i < m; ++i) {
                    const uint32_t s = cbuf[(size_t)i];
                    if (s >= C.slots.size() || C.slots[s].req == nullptr) continue;  // not a slot in flight
                    ServeSlotView v = serve_slot(eng, (int)k, s);
                    Slot sl = C.slots[s];
                    C.slots[s] = Slot{};
                    Live& L = live[sl.req];
                    const DeviceGraphResult& d = *v.out;
                    if (d.status != 0 && L.status == 0) L.status = d.status;
                    L.nodes += d.nodes;
                    L.dev_ticks += d.device_ns;
                    L.lab[sl.part].assign(v.label, v.label + sl.req->g[sl.part].n);
                    if (--L.remaining == 0) finished.push_back(sl.req);
                    --C.busy;
                    C.free_slots.push_back(s);
                    ++got;
                }
            }
        }
        for (Request* q : finished) {
            finish(q, live[q]);
            live.erase(q);
            delete q;
        }
        if (!finished.empty()) cv_done.notify_all();
        return got;
    }

    void fail_all(gcl_status e) {
        std::lock_guard<std::mutex> lk(mu);
        fatal = e;
        auto drop = [&](Request* q) {
            gcl_server_result r{};
            r.ticket = q->ticket;
            r.status = e;
            r.verdict = -1;
            r.n = q->g[0].n;
            done.push_back(r);
            queued_bytes -= q->footprint;
            ++st.completed;
            delete q;
        };
        for (auto& kv : live) drop(kv.first);
        live.clear();
        for (auto& C : cls) {
            C.busy = 0;
            for (auto& sl : C.slots) sl = Slot{};
        }
        while (!pending.empty()) {
            drop(pending.front());
            pending.pop_front();
        }
        cv_done.notify_all();
    }

    void loop() {
        auto last_check = Clock::now();
        Clock::time_point stop_at{};
        for (;;) {
            const auto c0 = Clock::now();
            const size_t got = collect();
            const auto c1 = Clock::now();
            dispatch();
            const auto c2 = Clock::now();
            collect_ns += (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(c1 - c0).count();
            publish_ns += (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(c2 - c1).count();
            bool inflight = false;
            for (auto& C : cls) inflight = inflight || C.busy > 0;
            bool waiting;
            {
                std::lock_guard<std::mutex> lk(mu);
                waiting = !pending.empty();
            }
            if ((inflight || waiting) && Clock::now() - last_check > std::chrono::milliseconds(20)) {
                last_check = Clock::now();
                // a serving kernel that failed, or that returned because its
                // lifetime passed, completes nothing more: the requests in
                // flight and queued fail, and later submissions are refused
                const gcl_status cs = serve_check(eng);
                if (cs != GCL_OK) {
                    fail_all(cs);
                    return;
                }
            }
            std::unique_lock<std::mutex> lk(mu);
            if (stop && stop_at == Clock::time_point{}) stop_at = Clock::now();
            if (stop && Clock::now() - stop_at > std::chrono::seconds(30)) {
                // the queue did not drain in time: report and leave
                lk.unlock();
                fail_all(GCL_E_INTERNAL);
                return;
            }
            if (!inflight && pending.empty()) {
                if (stop) return;
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
            } else if (got == 0) {
                cv_work.wait_for(lk, std::chrono::microseconds(20));
            }
        }
    }

    int32_t poll(gcl_server_result* out, int32_t cap, int32_t timeout_ms) {
        std::unique_lock<std::mutex> lk(mu);
        if (done.empty() && timeout_ms > 0)
            cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !done.empty(); });
        int32_t k = 0;
        while (k < cap && !done.empty()) {
            out[k++] = done.front();
            done.pop_front();
        }
        return k;
    }

    gcl_status stats(gcl_server_stats* o) {
        std::lock_guard<std::mutex> lk(mu);
        *o = st;
        o->pending = pending.size();
        o->in_flight = st.submitted - st.completed - pending.size();
        o->queued_bytes = queued_bytes;
        o->sched_publish_ms = (double)publish_ns.load() / 1e6;
        o->sched_collect_ms = (double)collect_ns.load() / 1e6;
        return GCL_OK;
    }

    gcl_status replay(const gcl_batch* b, const int32_t* pair_with, const double* arrival, int32_t count,
                      gcl_server_result* out) {
        std::unordered_map<uint64_t, int32_t> idx;
        const auto start = Clock::now();
        int32_t next = 0, got = 0;
        std::vector<gcl_server_result> buf(256);
        auto graph_of = [&](int32_t g, std::vector<int32_t>& rp) {
            const int64_t v0 = b->vtx_off[g], v1 = b->vtx_off[g + 1];
            const int64_t e0 = b->row_ptr[v0];
            rp.resize((size_t)(v1 - v0 + 1));
            for (int64_t v = v0; v <= v1; ++v) rp[(size_t)(v - v0)] = (int32_t)(b->row_ptr[v] - e0);
            gcl_csr c;
            c.n = (int32_t)(v1 - v0);
            c.nnz = b->row_ptr[v1] - e0;
            c.row_ptr = rp.data();
            c.col = b->col + e0;
            c.color = b->color ? b->color + v0 : nullptr;
            return c;
        };
        std::vector<int32_t> rpa, rpb;
        while (got < count) {
            const double now_s = std::chrono::duration<double>(Clock::now() - start).count();
            bool blocked = false;
            while (next < count && arrival[next] <= now_s) {
                const gcl_csr a = graph_of(next, rpa);
                gcl_csr bb;
                const bool pair = pair_with && pair_with[next] >= 0;
                if (pair) bb = graph_of(pair_with[next], rpb);
                const auto t0 = start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival[next]));
                uint64_t ticket = 0;
                const gcl_status r = submit(&a, pair ? &bb : nullptr, nullptr, &ticket, t0);
                if (r == GCL_E_BUSY) {  // retry this request once results drain
                    blocked = true;
                    break;
                }
                if (r == GCL_E_CAPACITY) {  // refused at admission: that is its result
                    gcl_server_result& o = out[next];
                    o = gcl_server_result{};
                    o.status = GCL_E_CAPACITY;
                    o.verdict = -1;
                    o.n = a.n;
                    ++next;
                    ++got;
                    continue;
                }
                if (r != GCL_OK) return r;
                idx[ticket] = next++;
            }
            const int32_t k = poll(buf.data(), (int32_t)buf.size(), blocked ? 1 : (next < count ? 0 : 100));
            if (k < 0) return (gcl_status)k;
            for (int32_t j = 0; j < k; ++j) {
                auto it = idx.find(buf[(size_t)j].ticket);
                if (it == idx.end()) continue;
                out[it->second] = buf[(size_t)j];
                ++got;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                if (fatal != GCL_OK && st.completed == st.submitted && done.empty() && got < count)
                    return fail(fatal, "server stopped: %s", gcl_status_string(fatal));
            }
            if (k == 0 && next < count && !blocked) {
                // sleep until the next arrival (busy-wait the last 200 us)
                const double wait = arrival[next] - std::chrono::duration<double>(Clock::now() - start).count();
                if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 2e-4));
            }
        }
        return GCL_OK;
    }
};

}  // namespace gcl

using gcl::fail;

namespace {
// Runs f, turning a C++ exception (thread creation, container growth) into a
// status: nothing is thrown across the C ABI.
template <class F>
gcl_status guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(GCL_E_NOMEM, "out of host memory");
    } catch (const std::exception& e) {
        return fail(GCL_E_INTERNAL, "server: %s", e.what());
    }
}
}  // namespace

extern "C" {

void gcl_server_config_default(gcl_server_config* c) {
    if (!c) return;
    c->max_n = 4096;
    c->max_graph_nnz = 1ll << 20;
    c->max_aut = 10000;
    c->slots = 0;
    c->queue_bytes = 1ull << 30;
    c->lifetime_ms = 0;
    c->large_share_pct = 40;
    c->max_search_cost = 150000000ull;
}

gcl_status gcl_server_create(const gcl_server_config* cfg, gcl_server** out) {
    if (!cfg || !out) return fail(GCL_E_INVALID, "null argument");
    if (gcl::device_count() <= 0) return fail(GCL_E_NODEVICE, "no HIP device visible");
    gcl::Server* s = new (std::nothrow) gcl::Server();
    if (!s) return fail(GCL_E_NOMEM, "out of host memory");
    gcl_status r = guarded([&] { return s->create(*cfg); });
    if (r != GCL_OK) {
        delete s;
        return r;
    }
    *out = reinterpret_cast<gcl_server*>(s);
    return GCL_OK;
}

gcl_status gcl_server_submit(gcl_server* s, const gcl_csr* a, const gcl_csr* b, int32_t* out_label,
                             uint64_t* ticket) {
    if (!s || !a || !ticket) return fail(GCL_E_INVALID, "null argument");
    return reinterpret_cast<gcl::Server*>(s)->submit(a, b, out_label, ticket, gcl::Clock::now());
}

int32_t gcl_server_poll(gcl_server* s, gcl_server_result* out, int32_t cap, int32_t timeout_ms) {
    if (!s || (!out && cap > 0)) return (int32_t)fail(GCL_E_INVALID, "null argument");
    return reinterpret_cast<gcl::Server*>(s)->poll(out, cap, timeout_ms);
}

gcl_status gcl_server_stats_get(gcl_server* s, gcl_server_stats* st) {
    if (!s || !st) return fail(GCL_E_INVALID, "null argument");
    return reinterpret_cast<gcl::Server*>(s)->stats(st);
}

gcl_status gcl_server_replay(gcl_server* s, const gcl_batch* graphs, const int32_t* pair_with,
                             const double* arrival_s, int32_t count, gcl_server_result* out) {
    if (!s || !graphs || !arrival_s || (!out && count > 0) || count < 0 || count > graphs->num_graphs)
        return fail(GCL_E_INVALID, "invalid replay arguments");
    for (int32_t i = 1; i < count; ++i)
        if (arrival_s[i] < arrival_s[i - 1]) return fail(GCL_E_INVALID, "arrival times must be non-decreasing");
    if (pair_with)
        for (int32_t i = 0; i < count; ++i)
            if (pair_with[i] >= graphs->num_graphs) return fail(GCL_E_INVALID, "pair index out of range");
    return guarded([&] { return reinterpret_cast<gcl::Server*>(s)->replay(graphs, pair_with, arrival_s, count, out); });
}

void gcl_server_destroy(gcl_server* s) { delete reinterpret_cast<gcl::Server*>(s); }

}  // extern "C"

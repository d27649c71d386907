// gcl_server.cpp -- request server of libgcl: admission by footprint,
// continuous batching over fixed launch slots, completion and load replay.
// The device side of a slot (buffers, kernels) is in gcl_device.hip.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
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
    std::vector<DeviceSlot*> slots;
    std::vector<std::vector<Request*>> running;  // requests of each busy slot
    std::vector<int32_t> running_graphs;
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
        for (auto* s : slots) slot_destroy(s);
        for (auto* r : pending) delete r;
        for (auto& v : running)
            for (auto* r : v) delete r;
    }

    gcl_status create(const gcl_server_config& c) {
        cfg = c;
        if (cfg.slots < 1 || cfg.max_graphs < 2 || cfg.max_vertices < 1 || cfg.max_nnz < 0 || cfg.max_n < 1 ||
            cfg.max_n > kMaxVertices || cfg.max_graph_nnz < 0 || cfg.max_aut == 0 || cfg.queue_bytes == 0)
            return fail(GCL_E_INVALID, "invalid server configuration");
        if ((int64_t)cfg.max_n * 2 > cfg.max_vertices)
            return fail(GCL_E_INVALID, "max_vertices must hold a pair of the largest graphs");
        SlotConfig sc;
        sc.max_graphs = cfg.max_graphs;
        sc.max_vertices = cfg.max_vertices;
        sc.max_nnz = cfg.max_nnz;
        sc.max_n = cfg.max_n;
        sc.max_graph_nnz = cfg.max_graph_nnz;
        sc.max_aut = cfg.max_aut;
        sc.share = std::min(cfg.slots, 2);
        for (int i = 0; i < cfg.slots; ++i) {
            DeviceSlot* s = nullptr;
            gcl_status r = slot_create(sc, &s);
            if (r != GCL_OK) return r;
            slots.push_back(s);
        }
        running.resize(slots.size());
        running_graphs.assign(slots.size(), 0);
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
        int64_t nsum = 0, msum = 0;
        for (int i = 0; i < ng; ++i) {
            if (gs[i]->n > cfg.max_n || gs[i]->nnz > cfg.max_graph_nnz)
                return fail(GCL_E_CAPACITY, "graph (n=%d, nnz=%lld) exceeds the server's max_n=%d / max_graph_nnz=%lld",
                            gs[i]->n, (long long)gs[i]->nnz, cfg.max_n, (long long)cfg.max_graph_nnz);
            fp += graph_footprint(gs[i]->n, gs[i]->nnz);
            nsum += gs[i]->n;
            msum += gs[i]->nnz;
        }
        if (nsum > cfg.max_vertices || msum > cfg.max_nnz)
            return fail(GCL_E_CAPACITY, "request exceeds the per-launch capacity");
        if (fp > cfg.queue_bytes) return fail(GCL_E_CAPACITY, "request footprint exceeds queue_bytes");
        {
            std::lock_guard<std::mutex> lk(mu);
            if (fatal != GCL_OK) return fail(fatal, "server stopped after a device error");
            if (queued_bytes + fp > cfg.queue_bytes) {
                ++st.refused_busy;
                return fail(GCL_E_BUSY, "server queue full");
            }
        }
        Request* q = new (std::nothrow) Request();
        if (!q) return fail(GCL_E_NOMEM, "out of host memory");
        q->ngraphs = ng;
        for (int i = 0; i < ng; ++i) q->g[i] = copy_graph(*gs[i]);
        q->out_label = out_label;
        q->t0 = t0;
        q->footprint = fp;
        {
            std::lock_guard<std::mutex> lk(mu);
            q->ticket = next_ticket++;
            queued_bytes += fp;
            ++st.submitted;
            pending.push_back(q);
            *ticket = q->ticket;
        }
        cv_work.notify_one();
        return GCL_OK;
    }

    // Packs pending requests (arrival order) into slot i; returns the graphs staged.
    int32_t stage(size_t i, std::vector<Request*>& taken, bool& colored) {
        SlotStaging t = slot_staging(slots[i]);
        int32_t G = 0;
        int64_t V = 0, E = 0;
        t.vtx[0] = 0;
        t.row[0] = 0;
        colored = false;
        std::lock_guard<std::mutex> lk(mu);
        while (!pending.empty()) {
            Request* q = pending.front();
            int64_t qn = 0, qm = 0;
            for (int k = 0; k < q->ngraphs; ++k) {
                qn += q->g[k].n;
                qm += q->g[k].row_ptr[q->g[k].n];
            }
            if (G + q->ngraphs > cfg.max_graphs || V + qn > cfg.max_vertices || E + qm > cfg.max_nnz) break;
            pending.pop_front();
            for (int k = 0; k < q->ngraphs; ++k) {
                const Graph& g = q->g[k];
                for (int32_t v = 0; v < g.n; ++v) t.row[V + v + 1] = E + g.row_ptr[v + 1];
                std::copy(g.col.begin(), g.col.end(), t.col + E);
                if (!g.color.empty()) colored = true;
                for (int32_t v = 0; v < g.n; ++v) t.color[V + v] = g.color.empty() ? 0 : g.color[v];
                V += g.n;
                E += g.row_ptr[g.n];
                t.vtx[++G] = V;
            }
            taken.push_back(q);
        }
        return G;
    }

    // Results of slot i, whose launch has finished.
    void complete(size_t i) {
        SlotStaging t = slot_staging(slots[i]);
        std::vector<gcl_server_result> out;
        int32_t g = 0;
        const auto now = Clock::now();
        for (Request* q : running[i]) {
            gcl_server_result r{};
            r.ticket = q->ticket;
            r.n = q->g[0].n;
            r.launch_graphs = running_graphs[i];
            r.verdict = -1;
            int32_t status = 0;
            const int32_t* lab[2] = {nullptr, nullptr};
            for (int k = 0; k < q->ngraphs; ++k, ++g) {
                const DeviceGraphResult& d = t.out[g];
                if (d.status != 0 && status == 0) status = d.status;
                r.nodes += d.nodes;
                lab[k] = t.label + t.vtx[g];
            }
            if (status == 0 && q->out_label) std::copy(lab[0], lab[0] + q->g[0].n, q->out_label);
            if (status == 0 && q->ngraphs == 2) {
                const gcl_csr A = q->g[0].view(), B = q->g[1].view();
                if (A.n != B.n || A.nnz != B.nnz) {
                    r.verdict = 0;
                } else {
                    std::vector<uint32_t> ka((size_t)A.nnz), kb((size_t)B.nnz);
                    canonical_keys(A, lab[0], ka.data());
                    canonical_keys(B, lab[1], kb.data());
                    r.verdict = ka == kb ? 1 : 0;
                }
            }
            r.status = status;
            r.latency_ms = std::chrono::duration<double, std::milli>(now - q->t0).count();
            out.push_back(r);
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            for (Request* q : running[i]) {
                queued_bytes -= q->footprint;
                delete q;
            }
            running[i].clear();
            for (auto& r : out) done.push_back(r);
            st.completed += out.size();
        }
        cv_done.notify_all();
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
        for (auto& v : running) {
            for (Request* q : v) drop(q);
            v.clear();
        }
        while (!pending.empty()) {
            drop(pending.front());
            pending.pop_front();
        }
        cv_done.notify_all();
    }

    void loop() {
        std::vector<bool> busy(slots.size(), false);
        for (;;) {
            bool any_busy = false;
            for (size_t i = 0; i < slots.size(); ++i) {
                if (!busy[i]) continue;
                bool fin = false;
                if (slot_poll(slots[i], &fin) != GCL_OK) {
                    fail_all(GCL_E_DEVICE);
                    return;
                }
                if (fin) {
                    complete(i);
                    busy[i] = false;
                } else {
                    any_busy = true;
                }
            }
            for (size_t i = 0; i < slots.size(); ++i) {
                if (busy[i]) continue;
                std::vector<Request*> taken;
                bool colored = false;
                const int32_t G = stage(i, taken, colored);
                if (G == 0) break;
                running[i] = std::move(taken);
                running_graphs[i] = G;
                if (slot_launch(slots[i], G, cfg.max_aut, colored) != GCL_OK) {
                    fail_all(GCL_E_DEVICE);
                    return;
                }
                busy[i] = any_busy = true;
                std::lock_guard<std::mutex> lk(mu);
                ++st.launches;
                st.graphs_launched += (uint64_t)G;
            }
            std::unique_lock<std::mutex> lk(mu);
            if (!any_busy) {
                if (stop && pending.empty()) return;
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
            } else {
                // a launch is running: poll again shortly (or at once when
                // a slot is free and work arrives)
                cv_work.wait_for(lk, std::chrono::microseconds(50));
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
        uint64_t inf = 0;
        for (auto& v : running) inf += v.size();
        o->in_flight = inf;
        o->queued_bytes = queued_bytes;
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
                if (fatal != GCL_OK && pending.empty()) {
                    bool idle = true;
                    for (auto& v : running) idle = idle && v.empty();
                    if (idle && done.empty() && got < count) return fail(fatal, "server stopped after a device error");
                }
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

extern "C" {

void gcl_server_config_default(gcl_server_config* c) {
    if (!c) return;
    c->slots = 3;
    c->max_graphs = 4096;
    c->max_vertices = 1ll << 21;
    c->max_nnz = 1ll << 24;
    c->max_n = 4096;
    c->max_graph_nnz = 1ll << 20;
    c->max_aut = 10000;
    c->queue_bytes = 1ull << 30;
}

gcl_status gcl_server_create(const gcl_server_config* cfg, gcl_server** out) {
    if (!cfg || !out) return fail(GCL_E_INVALID, "null argument");
    if (gcl::device_count() <= 0) return fail(GCL_E_NODEVICE, "no HIP device visible");
    gcl::Server* s = new (std::nothrow) gcl::Server();
    if (!s) return fail(GCL_E_NOMEM, "out of host memory");
    gcl_status r = s->create(*cfg);
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
    return reinterpret_cast<gcl::Server*>(s)->replay(graphs, pair_with, arrival_s, count, out);
}

void gcl_server_destroy(gcl_server* s) { delete reinterpret_cast<gcl::Server*>(s); }

}  // extern "C"

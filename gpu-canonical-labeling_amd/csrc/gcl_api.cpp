// gcl_api.cpp -- extern "C" entry points of libgcl (declared in include/gcl.h).
//
// The GPU engine never falls back to the host engine: a missing device or a
// device error is reported as a status.  GCL_ENGINE_CPU is the explicit
// --sequential mode of the reference CLI.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <vector>

#include "gcl_cpu.h"
#include "gcl_device.h"
#include "gcl_internal.h"

namespace gcl {
namespace {

// Automorphism slots per worker for batches: the canonical form does not
// depend on the ring size (every stored automorphism is genuine, so pruning
// with fewer of them is still sound); only node counts do.
constexpr uint32_t kBatchRing = 64;

gcl_status check_max_aut(uint64_t max_aut) {
    if (max_aut == 0) return fail(GCL_E_INVALID, "max_aut must be at least 1");
    return GCL_OK;
}

gcl_status validate_nonempty(const gcl_csr* g) {
    gcl_status s = gcl_csr_validate(g);
    if (s != GCL_OK) return s;
    if (g->n == 0) return fail(GCL_E_INVALID, "graph has no vertices");
    return GCL_OK;
}

gcl_status validate_batch(const gcl_batch* b) {
    if (!b) return fail(GCL_E_INVALID, "null batch");
    if (b->num_graphs < 0) return fail(GCL_E_INVALID, "negative graph count");
    if (b->num_graphs == 0) return GCL_OK;
    if (!b->vtx_off || !b->row_ptr) return fail(GCL_E_INVALID, "null batch arrays");
    if (b->vtx_off[0] != 0) return fail(GCL_E_INVALID, "vtx_off[0] must be 0");
    if (b->row_ptr[0] != 0) return fail(GCL_E_INVALID, "row_ptr[0] must be 0");
    if (b->row_ptr[b->vtx_off[b->num_graphs]] > 0 && !b->col) return fail(GCL_E_INVALID, "null col");
    for (int32_t g = 0; g < b->num_graphs; ++g) {
        const int64_t v0 = b->vtx_off[g], v1 = b->vtx_off[g + 1];
        if (v1 <= v0) return fail(GCL_E_INVALID, "graph %d has no vertices", g);
        if (v1 - v0 > kMaxVertices) return fail(GCL_E_CAPACITY, "graph %d exceeds %d vertices", g, kMaxVertices);
        const int32_t n = (int32_t)(v1 - v0);
        for (int64_t v = v0; v < v1; ++v) {
            const int64_t a = b->row_ptr[v], e = b->row_ptr[v + 1];
            if (e < a) return fail(GCL_E_INVALID, "graph %d: row_ptr not monotone", g);
            for (int64_t x = a; x < e; ++x) {
                const int32_t u = b->col[x];
                const int32_t self = (int32_t)(v - v0);
                if (u < 0 || u >= n) return fail(GCL_E_INVALID, "graph %d: neighbour out of range", g);
                if (u == self) return fail(GCL_E_INVALID, "graph %d: self loop", g);
                if (x > a && b->col[x - 1] >= u)
                    return fail(GCL_E_INVALID, "graph %d: row %d not strictly increasing", g, self);
                const int32_t* lo = b->col + b->row_ptr[v0 + u];
                const int32_t* hi = b->col + b->row_ptr[v0 + u + 1];
                if (!std::binary_search(lo, hi, self))
                    return fail(GCL_E_INVALID, "graph %d: edge %d-%d not symmetric", g, self, u);
            }
        }
    }
    return GCL_OK;
}

// Single graph as a one-element batch (row offsets widened to 64 bits).
struct OneBatch {
    int64_t vtx[2];
    std::vector<int64_t> rows;
    gcl_batch b;
    explicit OneBatch(const gcl_csr* g) {
        vtx[0] = 0;
        vtx[1] = g->n;
        rows.assign(g->row_ptr, g->row_ptr + g->n + 1);
        b.num_graphs = 1;
        b.vtx_off = vtx;
        b.row_ptr = rows.data();
        b.col = g->col;
        b.color = g->color;
    }
};

gcl_status gpu_label_one(const gcl_csr* g, uint32_t max_aut, int32_t* label, gcl_stats* st) {
    OneBatch ob(g);
    DeviceGraphResult res;
    gcl_batch_stats bs;
    gcl_status s = gpu_canon_batch(ob.b, max_aut, max_aut, label, &res, &bs);
    if (s != GCL_OK) return s;
    if (st) {
        st->nodes = res.nodes;
        st->automorphisms = res.automorphisms;
        st->refinements = res.nodes;
        st->splitter_steps = res.splitter_steps;
        st->csr_bytes = res.csr_bytes;
        st->alg_bytes = res.csr_bytes + res.part_bytes;
        st->kernel_ms = bs.kernel_ms;
        st->total_ms = bs.total_ms;
    }
    return GCL_OK;
}

}  // namespace
}  // namespace gcl

using namespace gcl;

extern "C" {

int32_t gcl_device_count(void) { return gcl::device_count(); }

gcl_status gcl_canon_label(const gcl_csr* g, uint32_t max_aut, gcl_engine engine, int32_t* out_label,
                           uint32_t* out_keys, uint8_t* out_dense, gcl_stats* stats) {
    gcl_status s = validate_nonempty(g);
    if (s != GCL_OK) return s;
    if ((s = check_max_aut(max_aut)) != GCL_OK) return s;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int32_t> label((size_t)g->n);
    std::vector<uint32_t> keys((size_t)g->nnz);
    gcl_stats st;
    memset(&st, 0, sizeof(st));
    try {
        if (engine == GCL_ENGINE_CPU) {
            s = cpu_canon_label(*g, max_aut, label.data(), keys.data(), &st);
            if (s != GCL_OK) return s;
        } else if (engine == GCL_ENGINE_GPU) {
            s = gpu_label_one(g, max_aut, label.data(), &st);
            if (s != GCL_OK) return s;
            canonical_keys(*g, label.data(), keys.data());
        } else {
            return fail(GCL_E_INVALID, "unknown engine %d", (int)engine);
        }
    } catch (const std::bad_alloc&) {
        return fail(GCL_E_NOMEM, "host allocation failed");
    }
    if (out_label) memcpy(out_label, label.data(), sizeof(int32_t) * (size_t)g->n);
    if (out_keys && g->nnz) memcpy(out_keys, keys.data(), sizeof(uint32_t) * (size_t)g->nnz);
    if (out_dense) keys_to_dense(keys.data(), g->nnz, g->n, out_dense);
    st.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return GCL_OK;
}

gcl_status gcl_canon_label_noalloc(const gcl_csr* g, uint32_t max_aut, int32_t* out_label, uint32_t* out_keys,
                                   gcl_stats* stats) {
    gcl_status s = validate_nonempty(g);
    if (s != GCL_OK) return s;
    if ((s = check_max_aut(max_aut)) != GCL_OK) return s;
    if (!out_label) return fail(GCL_E_INVALID, "null output labels");
    auto t0 = std::chrono::steady_clock::now();
    DeviceGraphResult res;
    if ((s = noalloc_canon(*g, max_aut, out_label, &res)) != GCL_OK) return s;
    try {
        if (out_keys) canonical_keys(*g, out_label, out_keys);
    } catch (const std::bad_alloc&) {
        return fail(GCL_E_NOMEM, "host allocation failed");
    }
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->nodes = res.nodes;
        stats->automorphisms = res.automorphisms;
        stats->refinements = res.nodes;
        stats->splitter_steps = res.splitter_steps;
        stats->csr_bytes = res.csr_bytes;
        stats->alg_bytes = res.csr_bytes + res.part_bytes;
        stats->kernel_ms = (double)res.device_ns / 1e6;  // device wall clock of the search
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return GCL_OK;
}

gcl_status gcl_canon_automorphisms(const gcl_csr* g, uint32_t max_aut, int32_t* out_label, int32_t* out_auts,
                                   int64_t cap_rows, int64_t* rows, gcl_stats* stats) {
    gcl_status s = validate_nonempty(g);
    if (s != GCL_OK) return s;
    if ((s = check_max_aut(max_aut)) != GCL_OK) return s;
    if (cap_rows < 0 || (cap_rows > 0 && !out_auts)) return fail(GCL_E_INVALID, "bad automorphism buffer");
    OneBatch ob(g);
    BatchTuning t = default_batch_tuning();
    t.keep_auts = 1;
    DeviceBatch* db = nullptr;
    // the single-graph ring: max_aut slots, as gcl_canon_label uses
    s = device_batch_create(ob.b, max_aut, max_aut, &db, &t);
    if (s != GCL_OK) return s;
    float ms = 0.f;
    std::vector<DeviceGraphResult> res(1);
    gcl_batch_stats bs;
    memset(&bs, 0, sizeof(bs));
    s = device_batch_run(db, false, &ms);
    if (s == GCL_OK) s = device_batch_results(db, out_label, res.data(), &bs);
    if (s == GCL_OK) s = device_batch_automorphisms(db, 0, out_auts, cap_rows, rows);
    device_batch_destroy(db);
    if (s != GCL_OK) return s;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->nodes = res[0].nodes;
        stats->automorphisms = res[0].automorphisms;
        stats->refinements = res[0].nodes;
        stats->splitter_steps = res[0].splitter_steps;
        stats->csr_bytes = res[0].csr_bytes;
        stats->alg_bytes = res[0].csr_bytes + res[0].part_bytes;
        stats->kernel_ms = ms;
    }
    return GCL_OK;
}

gcl_status gcl_set_side_files(const char* prefix) {
    gcl::set_side_file_prefix(prefix);
    return GCL_OK;
}

gcl_status gcl_nauty_parallel(const gcl_csr* g, int32_t max_aut, uint8_t* out_canon, uint32_t* node_count) {
    if (max_aut < 1) return fail(GCL_E_INVALID, "max_aut must be at least 1");
    gcl_stats st;
    gcl_status s = gcl_canon_label(g, (uint32_t)max_aut, GCL_ENGINE_GPU, nullptr, nullptr, out_canon, &st);
    if (s == GCL_OK && node_count) *node_count = (uint32_t)st.nodes;
    return s;
}

gcl_status gcl_nauty_sequential(const gcl_csr* g, uint32_t max_aut, uint8_t* out_canon, uint32_t* node_count) {
    gcl_stats st;
    gcl_status s = gcl_canon_label(g, max_aut, GCL_ENGINE_CPU, nullptr, nullptr, out_canon, &st);
    if (s == GCL_OK && node_count) *node_count = (uint32_t)st.nodes;
    return s;
}

gcl_status gcl_is_isomorphic(const gcl_csr* a, const gcl_csr* b, uint32_t max_aut, gcl_engine engine,
                             int* verdict) {
    if (!verdict) return fail(GCL_E_INVALID, "null verdict");
    gcl_status s;
    if ((s = validate_nonempty(a)) != GCL_OK) return s;
    if ((s = validate_nonempty(b)) != GCL_OK) return s;
    if ((s = check_max_aut(max_aut)) != GCL_OK) return s;
    if (a->n != b->n || a->nnz != b->nnz) {
        *verdict = 0;
        return GCL_OK;
    }
    std::vector<uint32_t> ka((size_t)a->nnz), kb((size_t)b->nnz);
    if (engine == GCL_ENGINE_GPU) {
        // both graphs in one launch
        const int32_t n = a->n;
        std::vector<int64_t> vtx = {0, n, 2 * (int64_t)n};
        std::vector<int64_t> rows((size_t)2 * n + 1);
        for (int32_t i = 0; i <= n; ++i) rows[(size_t)i] = a->row_ptr[i];
        for (int32_t i = 1; i <= n; ++i) rows[(size_t)n + (size_t)i] = a->nnz + b->row_ptr[i];
        std::vector<int32_t> col((size_t)(a->nnz + b->nnz));
        std::copy(a->col, a->col + a->nnz, col.begin());
        std::copy(b->col, b->col + b->nnz, col.begin() + a->nnz);
        std::vector<int32_t> color;
        if (a->color || b->color) {
            color.assign((size_t)2 * n, 0);
            if (a->color) std::copy(a->color, a->color + n, color.begin());
            if (b->color) std::copy(b->color, b->color + n, color.begin() + n);
        }
        gcl_batch bt;
        bt.num_graphs = 2;
        bt.vtx_off = vtx.data();
        bt.row_ptr = rows.data();
        bt.col = col.data();
        bt.color = color.empty() ? nullptr : color.data();
        std::vector<int32_t> label((size_t)2 * n);
        s = gpu_canon_batch(bt, max_aut, max_aut, label.data(), nullptr, nullptr);
        if (s != GCL_OK) return s;
        canonical_keys(*a, label.data(), ka.data());
        canonical_keys(*b, label.data() + n, kb.data());
    } else {
        if ((s = gcl_canon_label(a, max_aut, engine, nullptr, ka.data(), nullptr, nullptr)) != GCL_OK) return s;
        if ((s = gcl_canon_label(b, max_aut, engine, nullptr, kb.data(), nullptr, nullptr)) != GCL_OK) return s;
    }
    *verdict = ka == kb ? 1 : 0;
    return GCL_OK;
}

gcl_status gcl_batch_canon(const gcl_batch* b, uint32_t max_aut, gcl_engine engine, int32_t* out_label,
                           uint64_t* out_nodes, gcl_batch_stats* stats) {
    gcl_status s = validate_batch(b);
    if (s != GCL_OK) return s;
    if ((s = check_max_aut(max_aut)) != GCL_OK) return s;
    if (!out_label) return fail(GCL_E_INVALID, "null output labels");
    const int32_t G = b->num_graphs;
    if (engine == GCL_ENGINE_GPU) {
        std::vector<DeviceGraphResult> res((size_t)std::max(G, 1));
        s = gpu_canon_batch(*b, max_aut, kBatchRing, out_label, res.data(), stats);
        if (s != GCL_OK) return s;
        if (out_nodes)
            for (int32_t g = 0; g < G; ++g) out_nodes[g] = res[(size_t)g].nodes;
        return GCL_OK;
    }
    if (engine != GCL_ENGINE_CPU) return fail(GCL_E_INVALID, "unknown engine %d", (int)engine);
    auto t0 = std::chrono::steady_clock::now();
    gcl_batch_stats acc;
    memset(&acc, 0, sizeof(acc));
    const uint32_t ring = std::min(max_aut, kBatchRing);
    for (int32_t g = 0; g < G; ++g) {
        const int64_t v0 = b->vtx_off[g];
        const int32_t n = (int32_t)(b->vtx_off[g + 1] - v0);
        std::vector<int32_t> rp((size_t)n + 1);
        for (int32_t i = 0; i <= n; ++i) rp[(size_t)i] = (int32_t)(b->row_ptr[v0 + i] - b->row_ptr[v0]);
        gcl_csr c;
        c.n = n;
        c.nnz = rp[(size_t)n];
        c.row_ptr = rp.data();
        c.col = b->col + b->row_ptr[v0];
        c.color = b->color ? b->color + v0 : nullptr;
        gcl_stats st;
        s = cpu_canon_label(c, ring, out_label + v0, nullptr, &st);
        if (s != GCL_OK) return s;
        acc.nodes += st.nodes;
        acc.automorphisms += st.automorphisms;
        acc.splitter_steps += st.splitter_steps;
        acc.csr_bytes += st.csr_bytes;
        acc.alg_bytes += st.alg_bytes;
        if (out_nodes) out_nodes[g] = st.nodes;
    }
    acc.workers = 1;
    acc.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = acc;
    return GCL_OK;
}

gcl_status gcl_device_batch_create(const gcl_batch* b, uint32_t max_aut, gcl_device_batch** out) {
    if (!out) return fail(GCL_E_INVALID, "null output handle");
    gcl_status s = validate_batch(b);
    if (s != GCL_OK) return s;
    if ((s = check_max_aut(max_aut)) != GCL_OK) return s;
    DeviceBatch* db = nullptr;
    s = device_batch_create(*b, max_aut, kBatchRing, &db);
    if (s != GCL_OK) return s;
    *out = reinterpret_cast<gcl_device_batch*>(db);
    return GCL_OK;
}

gcl_status gcl_batch_tuning_default(gcl_batch_tuning* t) {
    if (!t) return fail(GCL_E_INVALID, "null tuning");
    const BatchTuning d = default_batch_tuning();
    t->spec_min_n = (int32_t)d.spec_min_n;
    t->min_class_n = (int32_t)d.min_class_n;
    return GCL_OK;
}

gcl_status gcl_device_batch_create_tuned(const gcl_batch* b, uint32_t max_aut, const gcl_batch_tuning* t,
                                         gcl_device_batch** out) {
    if (!out) return fail(GCL_E_INVALID, "null output handle");
    gcl_status s = validate_batch(b);
    if (s != GCL_OK) return s;
    if ((s = check_max_aut(max_aut)) != GCL_OK) return s;
    BatchTuning bt = default_batch_tuning();
    if (t) {
        if (t->spec_min_n < 0 || t->min_class_n < 1) return fail(GCL_E_INVALID, "negative tuning field");
        bt.spec_min_n = (uint32_t)t->spec_min_n;
        bt.min_class_n = (uint32_t)t->min_class_n;
    }
    DeviceBatch* db = nullptr;
    s = device_batch_create(*b, max_aut, kBatchRing, &db, &bt);
    if (s != GCL_OK) return s;
    *out = reinterpret_cast<gcl_device_batch*>(db);
    return GCL_OK;
}

gcl_status gcl_batch_autotune(const gcl_batch* b, uint32_t max_aut, int32_t sample, gcl_batch_tuning* best,
                              double* best_ms, int32_t* tried) {
    if (!best) return fail(GCL_E_INVALID, "null tuning");
    gcl_status s = validate_batch(b);
    if (s != GCL_OK) return s;
    if ((s = check_max_aut(max_aut)) != GCL_OK) return s;
    if (b->num_graphs < 1) return fail(GCL_E_INVALID, "empty batch");
    BatchTuning bt{};
    s = batch_autotune(*b, max_aut, kBatchRing, sample, &bt, best_ms, tried);
    if (s != GCL_OK) return s;
    best->spec_min_n = (int32_t)bt.spec_min_n;
    best->min_class_n = (int32_t)bt.min_class_n;
    return GCL_OK;
}

gcl_status gcl_device_batch_run(gcl_device_batch* db, int32_t use_graph, gcl_batch_stats* stats) {
    if (!db) return fail(GCL_E_INVALID, "null handle");
    auto t0 = std::chrono::steady_clock::now();
    float ms = 0.f;
    gcl_status s = device_batch_run(reinterpret_cast<DeviceBatch*>(db), use_graph != 0, &ms);
    if (s != GCL_OK) return s;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->kernel_ms = ms;
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return GCL_OK;
}

gcl_status gcl_device_batch_results(gcl_device_batch* db, int32_t* out_label, uint64_t* out_nodes,
                                    gcl_batch_stats* stats) {
    if (!db) return fail(GCL_E_INVALID, "null handle");
    DeviceBatch* d = reinterpret_cast<DeviceBatch*>(db);
    gcl_batch_stats bs;
    memset(&bs, 0, sizeof(bs));
    std::vector<DeviceGraphResult> res;
    if (out_nodes) res.resize((size_t)std::max(device_batch_size(d), 1));
    gcl_status s = device_batch_results(d, out_label, out_nodes ? res.data() : nullptr, &bs);
    if (out_nodes)
        for (int32_t g = 0; g < device_batch_size(d); ++g) out_nodes[g] = res[(size_t)g].nodes;
    if (stats) {
        stats->nodes = bs.nodes;
        stats->automorphisms = bs.automorphisms;
        stats->splitter_steps = bs.splitter_steps;
        stats->csr_bytes = bs.csr_bytes;
        stats->alg_bytes = bs.alg_bytes;
        stats->workers = bs.workers;
    }
    return s;
}

gcl_status gcl_device_batch_graph_stats(gcl_device_batch* db, uint64_t* nodes, uint64_t* automorphisms,
                                        uint64_t* splitter_steps, uint64_t* device_ns) {
    if (!db) return fail(GCL_E_INVALID, "null handle");
    DeviceBatch* d = reinterpret_cast<DeviceBatch*>(db);
    const int32_t G = device_batch_size(d);
    std::vector<DeviceGraphResult> res((size_t)std::max(G, 1));
    gcl_status s = device_batch_results(d, nullptr, res.data(), nullptr);
    for (int32_t g = 0; g < G; ++g) {
        if (nodes) nodes[g] = res[(size_t)g].nodes;
        if (automorphisms) automorphisms[g] = res[(size_t)g].automorphisms;
        if (splitter_steps) splitter_steps[g] = res[(size_t)g].splitter_steps;
        if (device_ns) device_ns[g] = res[(size_t)g].device_ns;
    }
    return s;
}

gcl_status gcl_device_batch_class_stats(gcl_device_batch* db, gcl_class_stats* out, int32_t cap,
                                        int32_t* count) {
    if (!db) return fail(GCL_E_INVALID, "null handle");
    if (cap > 0 && !out) return fail(GCL_E_INVALID, "null output array");
    return device_batch_class_stats(reinterpret_cast<DeviceBatch*>(db), out, cap, count);
}

void gcl_device_batch_destroy(gcl_device_batch* db) { device_batch_destroy(reinterpret_cast<DeviceBatch*>(db)); }

gcl_status gcl_batch_canon_plain(const gcl_batch* b, uint32_t max_aut, int32_t* out_label, uint64_t* out_nodes,
                                 gcl_batch_stats* stats) {
    // One launch per graph, each with its own allocations, copies and
    // synchronisation: the shape of the reference's Nauty_Parallel call
    // (reference src/nauty_parallel.cu:640-1031) repeated over the batch.
    gcl_status s = validate_batch(b);
    if (s != GCL_OK) return s;
    if ((s = check_max_aut(max_aut)) != GCL_OK) return s;
    if (!out_label) return fail(GCL_E_INVALID, "null output labels");
    auto t0 = std::chrono::steady_clock::now();
    gcl_batch_stats acc;
    memset(&acc, 0, sizeof(acc));
    for (int32_t g = 0; g < b->num_graphs; ++g) {
        const int64_t v0 = b->vtx_off[g];
        const int64_t n = b->vtx_off[g + 1] - v0;
        const int64_t vtx[2] = {0, n};
        std::vector<int64_t> rows((size_t)n + 1);
        for (int64_t i = 0; i <= n; ++i) rows[(size_t)i] = b->row_ptr[v0 + i] - b->row_ptr[v0];
        gcl_batch one;
        one.num_graphs = 1;
        one.vtx_off = vtx;
        one.row_ptr = rows.data();
        one.col = b->col + b->row_ptr[v0];
        one.color = b->color ? b->color + v0 : nullptr;
        DeviceGraphResult res;
        gcl_batch_stats bs;
        s = gpu_canon_batch(one, max_aut, kBatchRing, out_label + v0, &res, &bs);
        if (s != GCL_OK) return s;
        acc.nodes += res.nodes;
        acc.automorphisms += res.automorphisms;
        acc.splitter_steps += res.splitter_steps;
        acc.csr_bytes += res.csr_bytes;
        acc.alg_bytes += res.csr_bytes + res.part_bytes;
        acc.kernel_ms += bs.kernel_ms;
        acc.workers = std::max(acc.workers, bs.workers);
        if (out_nodes) out_nodes[g] = res.nodes;
    }
    acc.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = acc;
    return GCL_OK;
}

}  // extern "C"

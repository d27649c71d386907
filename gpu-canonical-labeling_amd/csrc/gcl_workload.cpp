// gcl_workload.cpp -- seeded benchmark workloads of libgcl: the batched mix
// of graphs of uneven size (gcl_gen_mix_batch).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <thread>
#include <vector>

#include "gcl_internal.h"

namespace gcl {
namespace {

uint64_t mix_seed(uint64_t seed, uint64_t i) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ULL + (i + 1) * 0xD1B54A32D192ED03ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Item {
    Graph g;
    int32_t kind = 0;
    gcl_status st = GCL_OK;
};

void make_item(uint64_t seed, int32_t idx, int32_t n_min, int32_t n_max, Item& out) {
    Rng rng(mix_seed(seed, (uint64_t)idx));
    const double u = rng.uniform();
    const int32_t kind = u < 0.5 ? 0 : (u < 0.8 ? 1 : 2);
    int32_t n = (int32_t)std::floor(std::exp(std::log((double)n_min) +
                                             rng.uniform() * (std::log((double)n_max + 1.0) - std::log((double)n_min))));
    n = std::min(std::max(n, n_min), n_max);
    std::vector<std::pair<int32_t, int32_t>> pairs;
    const uint64_t sub_seed = rng.next();
    if (kind == 0) {
        const double mean = 3.0 + 9.0 * rng.uniform();
        const double p = n > 1 ? std::min(1.0, mean / (double)(n - 1)) : 0.0;
        Rng r2(sub_seed);
        // G(n, p) by geometric skipping (Batagelj & Brandes 2005)
        if (p > 0.0 && p < 1.0) {
            const double lq = std::log(1.0 - p);
            int64_t v = 1, w = -1;
            while (v < n) {
                w = w + 1 + (int64_t)std::floor(std::log(1.0 - r2.uniform()) / lq);
                while (w >= v && v < n) {
                    w -= v;
                    ++v;
                }
                if (v < n) pairs.push_back({(int32_t)w, (int32_t)v});
            }
        } else if (p >= 1.0) {
            for (int32_t i = 0; i < n; ++i)
                for (int32_t j = i + 1; j < n; ++j) pairs.push_back({i, j});
        }
    } else if (kind == 1) {
        const int32_t d = 3 + (int32_t)rng.below(3);
        if (((int64_t)n * d) % 2 != 0) n = n + 1 <= n_max ? n + 1 : n - 1;
        out.st = random_regular_pairs(n, d, sub_seed, pairs);
        if (out.st != GCL_OK) return;
    } else {
        const int32_t a = 4 + (int32_t)rng.below(29), b = 4 + (int32_t)rng.below(29);
        n = a * b;
        for (int32_t i = 0; i < a; ++i)
            for (int32_t j = 0; j < b; ++j) {
                const int32_t v = i * b + j, r = ((i + 1) % a) * b + j, c = i * b + (j + 1) % b;
                pairs.push_back({std::min(v, r), std::max(v, r)});
                pairs.push_back({std::min(v, c), std::max(v, c)});
            }
    }
    // seeded random relabeling (Fisher-Yates)
    std::vector<int32_t> sigma((size_t)n);
    std::iota(sigma.begin(), sigma.end(), 0);
    for (int32_t i = n - 1; i > 0; --i) std::swap(sigma[(size_t)i], sigma[(size_t)rng.below((uint64_t)i + 1)]);
    for (auto& e : pairs) {
        const int32_t x = sigma[(size_t)e.first], y = sigma[(size_t)e.second];
        e = {std::min(x, y), std::max(x, y)};
    }
    out.g = graph_from_pairs(n, pairs);
    out.kind = kind;
}

}  // namespace
}  // namespace gcl

using namespace gcl;

extern "C" {

gcl_status gcl_gen_mix_batch(int32_t count, uint64_t seed, int32_t n_min, int32_t n_max, int32_t threads,
                             gcl_owned_batch** out) {
    if (!out || count < 0 || n_min < 2 || n_max < n_min || n_max > kMaxVertices)
        return fail(GCL_E_INVALID, "bad mix batch arguments");
    std::vector<Item> items((size_t)count);
    int32_t nt = threads > 0 ? threads : (int32_t)std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, std::max(count, 1));
    std::vector<std::thread> pool;
    for (int32_t t = 0; t < nt; ++t)
        pool.emplace_back([&, t]() {
            for (int32_t i = t; i < count; i += nt) make_item(seed, i, n_min, n_max, items[(size_t)i]);
        });
    for (auto& th : pool) th.join();
    int64_t total_n = 0, total_nnz = 0;
    for (auto& it : items) {
        if (it.st != GCL_OK) return it.st;
        total_n += it.g.n;
        total_nnz += (int64_t)it.g.col.size();
    }
    gcl_owned_batch* b = new (std::nothrow) gcl_owned_batch();
    if (!b) return fail(GCL_E_NOMEM, "out of host memory");
    b->vtx_buf = (int64_t*)malloc(sizeof(int64_t) * ((size_t)count + 1));
    b->row_buf = (int64_t*)malloc(sizeof(int64_t) * ((size_t)total_n + 1));
    b->col_buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(total_nnz, 1));
    b->kind = (int32_t*)malloc(sizeof(int32_t) * (size_t)std::max(count, 1));
    if (!b->vtx_buf || !b->row_buf || !b->col_buf || !b->kind) {
        gcl_owned_batch_free(b);
        return fail(GCL_E_NOMEM, "out of host memory");
    }
    int64_t v = 0, e = 0;
    b->vtx_buf[0] = 0;
    for (int32_t i = 0; i < count; ++i) {
        const Graph& g = items[(size_t)i].g;
        for (int32_t r = 0; r < g.n; ++r) b->row_buf[v + r] = e + g.row_ptr[(size_t)r];
        if (!g.col.empty()) memcpy(b->col_buf + e, g.col.data(), sizeof(int32_t) * g.col.size());
        v += g.n;
        e += (int64_t)g.col.size();
        b->vtx_buf[i + 1] = v;
        b->kind[i] = items[(size_t)i].kind;
    }
    b->row_buf[total_n] = total_nnz;
    b->batch.num_graphs = count;
    b->batch.vtx_off = b->vtx_buf;
    b->batch.row_ptr = b->row_buf;
    b->batch.col = b->col_buf;
    b->batch.color = nullptr;
    *out = b;
    return GCL_OK;
}

void gcl_owned_batch_free(gcl_owned_batch* b) {
    if (!b) return;
    free(b->vtx_buf);
    free(b->row_buf);
    free(b->col_buf);
    free(b->kind);
    delete b;
}

}  // extern "C"

// gcl_graph.cpp -- graph model, validation, DIMACS I/O, text output and the
// seeded generators of libgcl.
//
// DIMACS acceptance follows CSRGraph::fromDIMACS (reference
// src/CSRGraph.cpp:40-119): 'c' comments, "p <tag> n m", "n v colour" and
// "e u v" with 1-based ids, loops dropped, both directions stored, rows
// sorted.  Two documented differences: out-of-range ids and repeated edges
// are rejected with GCL_E_INVALID (the reference writes out of bounds or
// keeps the duplicate, which corrupts its refinement).
#include "gcl_internal.h"

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <numeric>
#include <sstream>

namespace gcl {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

gcl_status fail(gcl_status s, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return s;
}

Graph graph_from_pairs(int32_t n, const std::vector<std::pair<int32_t, int32_t>>& pairs) {
    Graph g;
    g.n = n;
    g.row_ptr.assign((size_t)n + 1, 0);
    for (auto& e : pairs) {
        g.row_ptr[e.first + 1]++;
        g.row_ptr[e.second + 1]++;
    }
    for (int32_t i = 0; i < n; ++i) g.row_ptr[i + 1] += g.row_ptr[i];
    g.col.assign((size_t)g.row_ptr[n], 0);
    std::vector<int32_t> fill(g.row_ptr.begin(), g.row_ptr.end() - 1);
    for (auto& e : pairs) {
        g.col[fill[e.first]++] = e.second;
        g.col[fill[e.second]++] = e.first;
    }
    for (int32_t i = 0; i < n; ++i) std::sort(g.col.begin() + g.row_ptr[i], g.col.begin() + g.row_ptr[i + 1]);
    return g;
}

gcl_status graph_to_owned(Graph&& g, gcl_graph** out) {
    if (!out) return fail(GCL_E_INVALID, "null output pointer");
    gcl_graph* r = new (std::nothrow) gcl_graph();
    if (!r) return fail(GCL_E_NOMEM, "out of host memory");
    size_t nn = (size_t)g.n;
    r->row_ptr_buf = (int32_t*)malloc(sizeof(int32_t) * (nn + 1));
    r->col_buf = (int32_t*)malloc(sizeof(int32_t) * (g.col.size() ? g.col.size() : 1));
    r->color_buf = g.color.empty() ? nullptr : (int32_t*)malloc(sizeof(int32_t) * (nn ? nn : 1));
    if (!r->row_ptr_buf || !r->col_buf || (!g.color.empty() && !r->color_buf)) {
        free(r->row_ptr_buf); free(r->col_buf); free(r->color_buf);
        delete r;
        return fail(GCL_E_NOMEM, "out of host memory");
    }
    memcpy(r->row_ptr_buf, g.row_ptr.data(), sizeof(int32_t) * (nn + 1));
    if (!g.col.empty()) memcpy(r->col_buf, g.col.data(), sizeof(int32_t) * g.col.size());
    if (r->color_buf) memcpy(r->color_buf, g.color.data(), sizeof(int32_t) * nn);
    r->csr.n = g.n;
    r->csr.nnz = (int64_t)g.col.size();
    r->csr.row_ptr = r->row_ptr_buf;
    r->csr.col = r->col_buf;
    r->csr.color = r->color_buf;
    *out = r;
    return GCL_OK;
}

static uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

Rng::Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Rng::next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}

double Rng::uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }

uint64_t Rng::below(uint64_t bound) {
    // Lemire's multiply-shift with rejection (unbiased)
    if (bound <= 1) return 0;
    __uint128_t m = (__uint128_t)next() * bound;
    uint64_t l = (uint64_t)m;
    if (l < bound) {
        uint64_t t = (0 - bound) % bound;
        while (l < t) {
            m = (__uint128_t)next() * bound;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

void canonical_keys(const gcl_csr& g, const int32_t* label, uint32_t* keys) {
    const int32_t n = g.n;
    std::vector<int32_t> pos((size_t)n);
    for (int32_t i = 0; i < n; ++i) pos[label[i]] = i;
    int64_t k = 0;
    for (int32_t i = 0; i < n; ++i) {
        int32_t v = label[i];
        int64_t b = k;
        for (int32_t e = g.row_ptr[v]; e < g.row_ptr[v + 1]; ++e)
            keys[k++] = (uint32_t)i * (uint32_t)n + (uint32_t)pos[g.col[e]];
        std::sort(keys + b, keys + k);
    }
}

void keys_to_dense(const uint32_t* keys, int64_t nnz, int32_t n, uint8_t* dense) {
    memset(dense, 0, (size_t)n * (size_t)n);
    for (int64_t k = 0; k < nnz; ++k) dense[keys[k]] = 1;
}

gcl_status random_regular_pairs(int32_t n, int32_t d, uint64_t seed,
                                std::vector<std::pair<int32_t, int32_t>>& pairs) {
    // Configuration model: a uniform perfect matching of the n*d stubs,
    // drawn pair by pair and restarted at the first loop or repeated edge,
    // which gives the uniform distribution over simple d-regular graphs.
    if (n < 1 || n > kMaxVertices || d < 0 || d >= n || ((int64_t)n * d) % 2 != 0)
        return fail(GCL_E_INVALID, "no simple %d-regular graph on %d vertices", d, n);
    Rng rng(seed);
    const int64_t stubs = (int64_t)n * d;
    std::vector<int32_t> pts((size_t)stubs);
    std::vector<int32_t> nbr((size_t)stubs);
    std::vector<int32_t> fill((size_t)n);
    for (int attempt = 0; attempt < 10000000; ++attempt) {
        for (int64_t i = 0; i < stubs; ++i) pts[(size_t)i] = (int32_t)(i / d);
        std::fill(fill.begin(), fill.end(), 0);
        pairs.clear();
        bool ok = true;
        for (int64_t i = 0; i < stubs && ok; i += 2) {
            const int64_t j = i + 1 + (int64_t)rng.below((uint64_t)(stubs - i - 1));
            std::swap(pts[(size_t)i + 1], pts[(size_t)j]);
            const int32_t a = pts[(size_t)i], b = pts[(size_t)i + 1];
            if (a == b) {
                ok = false;
                break;
            }
            for (int32_t k = 0; k < fill[(size_t)a]; ++k)
                if (nbr[(size_t)a * d + k] == b) ok = false;
            if (!ok) break;
            nbr[(size_t)a * d + fill[(size_t)a]++] = b;
            nbr[(size_t)b * d + fill[(size_t)b]++] = a;
            pairs.push_back({std::min(a, b), std::max(a, b)});
        }
        if (ok) return GCL_OK;
    }
    return fail(GCL_E_INTERNAL, "configuration model did not produce a simple graph");
}

}  // namespace gcl

using namespace gcl;

extern "C" {

const char* gcl_last_error(void) { return g_error.c_str(); }

const char* gcl_status_string(gcl_status s) {
    switch (s) {
        case GCL_OK: return "ok";
        case GCL_E_INVALID: return "invalid argument";
        case GCL_E_NOMEM: return "out of memory";
        case GCL_E_DEVICE: return "device error";
        case GCL_E_NODEVICE: return "no device";
        case GCL_E_IO: return "i/o error";
        case GCL_E_CAPACITY: return "capacity exceeded";
        case GCL_E_INTERNAL: return "internal error";
        case GCL_E_BUSY: return "busy";
        case GCL_E_TIMEOUT: return "time limit exceeded";
    }
    return "unknown status";
}

gcl_status gcl_csr_validate(const gcl_csr* g) {
    if (!g) return fail(GCL_E_INVALID, "null graph");
    if (g->n < 0) return fail(GCL_E_INVALID, "negative vertex count");
    if (g->n > kMaxVertices)
        return fail(GCL_E_CAPACITY, "n=%d exceeds the supported maximum %d", g->n, kMaxVertices);
    if (!g->row_ptr) return fail(GCL_E_INVALID, "null row_ptr");
    const int32_t n = g->n;
    if (g->row_ptr[0] != 0) return fail(GCL_E_INVALID, "row_ptr[0] must be 0");
    for (int32_t i = 0; i < n; ++i)
        if (g->row_ptr[i + 1] < g->row_ptr[i]) return fail(GCL_E_INVALID, "row_ptr not monotone at %d", i);
    if ((int64_t)g->row_ptr[n] != g->nnz)
        return fail(GCL_E_INVALID, "nnz=%" PRId64 " does not match row_ptr[n]=%d", g->nnz, g->row_ptr[n]);
    if (g->nnz > 0 && !g->col) return fail(GCL_E_INVALID, "null col");
    for (int32_t u = 0; u < n; ++u) {
        for (int32_t e = g->row_ptr[u]; e < g->row_ptr[u + 1]; ++e) {
            int32_t v = g->col[e];
            if (v < 0 || v >= n) return fail(GCL_E_INVALID, "neighbour %d of vertex %d out of range", v, u);
            if (v == u) return fail(GCL_E_INVALID, "self loop at vertex %d", u);
            if (e > g->row_ptr[u] && g->col[e - 1] >= v)
                return fail(GCL_E_INVALID, "row %d not strictly increasing (repeated or unsorted neighbour)", u);
            // symmetry: u must appear in row v (binary search)
            const int32_t* b = g->col + g->row_ptr[v];
            const int32_t* en = g->col + g->row_ptr[v + 1];
            if (!std::binary_search(b, en, u)) return fail(GCL_E_INVALID, "edge %d-%d is not symmetric", u, v);
        }
    }
    return GCL_OK;
}

gcl_status gcl_canonical_keys(const gcl_csr* g, const int32_t* label, uint32_t* out_keys) {
    if (!g || !label || (!out_keys && g->nnz > 0)) return fail(GCL_E_INVALID, "null argument");
    canonical_keys(*g, label, out_keys);
    return GCL_OK;
}

void gcl_graph_free(gcl_graph* g) {
    if (!g) return;
    free(g->row_ptr_buf);
    free(g->col_buf);
    free(g->color_buf);
    delete g;
}

gcl_status gcl_read_dimacs(const char* path, gcl_graph** out) {
    if (!path || !out) return fail(GCL_E_INVALID, "null argument");
    std::ifstream in(path);
    if (!in.is_open()) return fail(GCL_E_IO, "Failed to open DIMACS file: %s", path);
    long long n = 0;
    bool have_p = false;
    std::vector<std::pair<int64_t, int64_t>> raw;
    std::vector<std::pair<int64_t, int32_t>> colors;
    std::string line;
    long long lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        if (line.empty() || line[0] == 'c') continue;
        std::istringstream iss(line);
        char type = 0;
        iss >> type;
        if (type == 'p') {
            std::string tag;
            long long nn = -1, mm = -1;
            iss >> tag >> nn >> mm;
            if (nn < 0) return fail(GCL_E_INVALID, "%s:%lld: bad problem line", path, lineno);
            n = nn;
            have_p = true;
        } else if (type == 'n') {
            long long vid = 0, col = 0;
            iss >> vid >> col;
            // the reference reads the id as unsigned (CSRGraph.cpp:62-68): 0
            // is skipped and a negative id wraps past n, so its colour is
            // ignored by the range check below (CSRGraph.cpp:86-90)
            if (vid < 1) continue;
            colors.push_back({vid - 1, (int32_t)col});
        } else if (type == 'e') {
            long long u = 0, v = 0;
            iss >> u >> v;
            if (u == v) continue;
            raw.push_back({u - 1, v - 1});
        }
    }
    (void)have_p;
    if (n > kMaxVertices) return fail(GCL_E_CAPACITY, "%s: n=%lld exceeds %d", path, n, kMaxVertices);
    std::vector<std::pair<int32_t, int32_t>> pairs;
    pairs.reserve(raw.size());
    for (auto& e : raw) {
        if (e.first < 0 || e.first >= n || e.second < 0 || e.second >= n)
            return fail(GCL_E_INVALID, "%s: edge %lld-%lld has a vertex outside 1..%lld", path,
                        (long long)e.first + 1, (long long)e.second + 1, n);
        pairs.push_back({(int32_t)std::min(e.first, e.second), (int32_t)std::max(e.first, e.second)});
    }
    std::sort(pairs.begin(), pairs.end());
    for (size_t i = 1; i < pairs.size(); ++i)
        if (pairs[i] == pairs[i - 1])
            return fail(GCL_E_INVALID, "%s: repeated edge %d-%d (multigraphs are not supported)", path,
                        pairs[i].first + 1, pairs[i].second + 1);
    Graph g = graph_from_pairs((int32_t)n, pairs);
    if (!colors.empty()) {
        g.color.assign((size_t)n, 0);
        for (auto& c : colors)
            if (c.first >= 0 && c.first < n) g.color[(size_t)c.first] = c.second;
    }
    return graph_to_owned(std::move(g), out);
}

gcl_status gcl_write_dimacs(const char* path, const gcl_csr* g) {
    if (!path || !g) return fail(GCL_E_INVALID, "null argument");
    FILE* f = fopen(path, "w");
    if (!f) return fail(GCL_E_IO, "could not open %s for writing", path);
    fprintf(f, "p edge %d %lld\n", g->n, (long long)(g->nnz / 2));
    if (g->color)
        for (int32_t v = 0; v < g->n; ++v) fprintf(f, "n %d %d\n", v + 1, g->color[v]);
    for (int32_t u = 0; u < g->n; ++u)
        for (int32_t e = g->row_ptr[u]; e < g->row_ptr[u + 1]; ++e)
            if (g->col[e] > u) fprintf(f, "e %d %d\n", u + 1, g->col[e] + 1);
    if (fclose(f) != 0) return fail(GCL_E_IO, "write to %s failed", path);
    return GCL_OK;
}

int64_t gcl_format_adjacency_keys(const uint32_t* keys, int64_t nnz, int32_t n, char* buf, int64_t cap) {
    int64_t len = 0;
    char tmp[48];
    auto emit = [&](const char* s, int k) {
        if (buf && len + k < cap) memcpy(buf + len, s, (size_t)k);
        len += k;
    };
    int64_t k = 0;
    for (int32_t i = 0; i < n; ++i) {
        int w = snprintf(tmp, sizeof(tmp), "%3d : ", i);
        emit(tmp, w);
        int width = w;
        bool first = true;
        const uint64_t lo = (uint64_t)i * (uint64_t)n, hi = lo + (uint64_t)n;
        while (k < nnz && keys[k] < hi) {
            if (keys[k] >= lo) {
                int t = snprintf(tmp, sizeof(tmp), " %u", (unsigned)(keys[k] - lo));
                if (!first && width + t + 1 > 80) {
                    emit("\n    ", 5);
                    width = 4;
                }
                emit(tmp, t);
                width += t;
                first = false;
            }
            ++k;
        }
        emit(";\n", 2);
    }
    if (buf && len < cap) buf[len] = 0;
    return len;
}

// ------------------------------------------------------------------ generators

gcl_status gcl_gen_erdos_renyi(int32_t n, double p, uint64_t seed, gcl_graph** out) {
    if (n < 0 || n > kMaxVertices || !(p >= 0.0 && p <= 1.0)) return fail(GCL_E_INVALID, "bad G(n,p) arguments");
    Rng rng(seed);
    std::vector<std::pair<int32_t, int32_t>> pairs;
    if (p > 0.0) {
        if (p >= 1.0) {
            for (int32_t i = 0; i < n; ++i)
                for (int32_t j = i + 1; j < n; ++j) pairs.push_back({i, j});
        } else {
            // geometric skipping over the pairs (i<j) in row-major order
            // (Batagelj & Brandes, "Efficient generation of large random networks", 2005)
            const double lq = std::log(1.0 - p);
            int64_t v = 1, w = -1;
            while (v < n) {
                double r = rng.uniform();
                w = w + 1 + (int64_t)std::floor(std::log(1.0 - r) / lq);
                while (w >= v && v < n) {
                    w -= v;
                    ++v;
                }
                if (v < n) pairs.push_back({(int32_t)w, (int32_t)v});
            }
        }
    }
    return graph_to_owned(graph_from_pairs(n, pairs), out);
}

gcl_status gcl_gen_erdos_renyi_mean_degree(int32_t n, double mean_degree, uint64_t seed, gcl_graph** out) {
    if (n < 2) return gcl_gen_erdos_renyi(n, 0.0, seed, out);
    double p = mean_degree / (double)(n - 1);
    if (p > 1.0) p = 1.0;
    return gcl_gen_erdos_renyi(n, p, seed, out);
}

gcl_status gcl_gen_random_regular(int32_t n, int32_t d, uint64_t seed, gcl_graph** out) {
    std::vector<std::pair<int32_t, int32_t>> pairs;
    gcl_status st = random_regular_pairs(n, d, seed, pairs);
    if (st != GCL_OK) return st;
    return graph_to_owned(graph_from_pairs(n, pairs), out);
}

gcl_status gcl_gen_torus(int32_t a, int32_t b, gcl_graph** out) {
    if (a < 3 || b < 3 || (int64_t)a * b > kMaxVertices) return fail(GCL_E_INVALID, "torus sides must be >= 3");
    std::vector<std::pair<int32_t, int32_t>> pairs;
    for (int32_t i = 0; i < a; ++i)
        for (int32_t j = 0; j < b; ++j) {
            int32_t v = i * b + j;
            int32_t r = ((i + 1) % a) * b + j;
            int32_t c = i * b + (j + 1) % b;
            pairs.push_back({std::min(v, r), std::max(v, r)});
            pairs.push_back({std::min(v, c), std::max(v, c)});
        }
    return graph_to_owned(graph_from_pairs(a * b, pairs), out);
}

gcl_status gcl_gen_cfi(int32_t base_n, int32_t twisted, uint64_t seed, gcl_graph** out) {
    // Cai-Fuerer-Immerman graph over a random 3-regular base graph: each base
    // vertex becomes 4 middle vertices (even subsets of its 3 edge slots) and
    // 2 end vertices per slot; end vertices of matching slots are joined
    // straight, or crossed on the first base edge when twisted.
    if (base_n < 4 || base_n % 2 != 0 || (int64_t)base_n * 10 > kMaxVertices)
        return fail(GCL_E_INVALID, "CFI base size must be even and >= 4");
    gcl_graph* base = nullptr;
    gcl_status st = gcl_gen_random_regular(base_n, 3, seed, &base);
    if (st != GCL_OK) return st;
    const int32_t* rp = base->csr.row_ptr;
    const int32_t* cl = base->csr.col;
    auto mid = [](int32_t v, int32_t s) { return v * 10 + s; };                   // s in 0..3
    auto end = [](int32_t v, int32_t slot, int32_t bit) { return v * 10 + 4 + slot * 2 + bit; };
    static const int even_sets[4] = {0x0, 0x3, 0x5, 0x6};  // {}, {0,1}, {0,2}, {1,2}
    std::vector<std::pair<int32_t, int32_t>> pairs;
    for (int32_t v = 0; v < base_n; ++v)
        for (int32_t s = 0; s < 4; ++s)
            for (int32_t slot = 0; slot < 3; ++slot) {
                int32_t bit = (even_sets[s] >> slot) & 1;
                pairs.push_back({mid(v, s), end(v, slot, bit)});
            }
    bool twist_done = false;
    for (int32_t u = 0; u < base_n; ++u)
        for (int32_t e = rp[u]; e < rp[u + 1]; ++e) {
            int32_t v = cl[e];
            if (v <= u) continue;
            int32_t su = e - rp[u];
            int32_t sv = (int32_t)(std::lower_bound(cl + rp[v], cl + rp[v + 1], u) - (cl + rp[v]));
            bool tw = twisted && !twist_done;
            twist_done = twist_done || tw;
            for (int32_t bit = 0; bit < 2; ++bit) {
                int32_t a = end(u, su, bit), b = end(v, sv, tw ? 1 - bit : bit);
                pairs.push_back({std::min(a, b), std::max(a, b)});
            }
        }
    gcl_graph_free(base);
    return graph_to_owned(graph_from_pairs(base_n * 10, pairs), out);
}

gcl_status gcl_gen_permute(const gcl_csr* g, uint64_t seed, gcl_graph** out) {
    if (!g) return fail(GCL_E_INVALID, "null graph");
    Rng rng(seed);
    std::vector<int32_t> sigma((size_t)g->n);
    std::iota(sigma.begin(), sigma.end(), 0);
    for (int32_t i = g->n - 1; i > 0; --i) std::swap(sigma[(size_t)i], sigma[(size_t)rng.below((uint64_t)i + 1)]);
    std::vector<std::pair<int32_t, int32_t>> pairs;
    for (int32_t u = 0; u < g->n; ++u)
        for (int32_t e = g->row_ptr[u]; e < g->row_ptr[u + 1]; ++e)
            if (g->col[e] > u) {
                int32_t a = sigma[(size_t)u], b = sigma[(size_t)g->col[e]];
                pairs.push_back({std::min(a, b), std::max(a, b)});
            }
    Graph h = graph_from_pairs(g->n, pairs);
    if (g->color) {
        h.color.assign((size_t)g->n, 0);
        for (int32_t v = 0; v < g->n; ++v) h.color[(size_t)sigma[(size_t)v]] = g->color[v];
    }
    return graph_to_owned(std::move(h), out);
}

gcl_status gcl_gen_edge_swap(const gcl_csr* g, uint64_t seed, gcl_graph** out) {
    if (!g) return fail(GCL_E_INVALID, "null graph");
    std::vector<std::pair<int32_t, int32_t>> edges;
    for (int32_t u = 0; u < g->n; ++u)
        for (int32_t e = g->row_ptr[u]; e < g->row_ptr[u + 1]; ++e)
            if (g->col[e] > u) edges.push_back({u, g->col[e]});
    if (edges.size() < 2) return fail(GCL_E_INVALID, "edge swap needs at least two edges");
    auto has = [&](int32_t a, int32_t b) {
        const int32_t* lo = g->col + g->row_ptr[a];
        const int32_t* hi = g->col + g->row_ptr[a + 1];
        return std::binary_search(lo, hi, b);
    };
    Rng rng(seed);
    for (int attempt = 0; attempt < 1000000; ++attempt) {
        size_t i = (size_t)rng.below(edges.size()), j = (size_t)rng.below(edges.size());
        if (i == j) continue;
        int32_t a = edges[i].first, b = edges[i].second, c = edges[j].first, d = edges[j].second;
        if (rng.next() & 1) std::swap(c, d);
        // (a,b),(c,d) -> (a,d),(c,b): degree preserving, must stay simple
        if (a == c || a == d || b == c || b == d) continue;
        if (has(a, d) || has(c, b)) continue;
        edges[i] = {std::min(a, d), std::max(a, d)};
        edges[j] = {std::min(c, b), std::max(c, b)};
        Graph h = graph_from_pairs(g->n, edges);
        if (g->color) h.color.assign(g->color, g->color + g->n);
        return graph_to_owned(std::move(h), out);
    }
    return fail(GCL_E_INVALID, "no valid double edge swap found");
}

}  // extern "C"

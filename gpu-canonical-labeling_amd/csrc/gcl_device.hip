// gcl_device.hip -- device engine of libgcl: individualization-refinement
// canonical labeling on CDNA4 (gfx950), one wave64 per search worker.
//
// A worker runs the exact traversal of the reference's sequential engine
// (reference src/sequential.cpp:216-670) for one graph at a time, pulling
// graph ids from a launch-wide counter (persistent kernel).  All 64 lanes
// execute the search's control flow in lockstep; the data-parallel inner
// work of every step runs across the lanes:
//   * refinement (reference src/partition_refinement.cpp:150-485): neighbour
//     counts with LDS atomics, in-cell regrouping by per-cell ranks, splitter
//     choice by ballot over an active-position bitmap;
//   * backtracking: rebuild of the cell structure from a bitmap of cell
//     boundaries minus the undone ones (replaces reconstruct_at_level);
//   * target cells: member list sorted by id through a vertex bitmap;
//   * orbit pruning: lock-free union-find (union by minimum) over the
//     automorphisms fixing the current sequence;
//   * leaves: sort-free comparison of the current relabeled graph with the
//     stored first / best certificates (sorted canonical rows).
// Hot partition arrays live in LDS for n <= 2048 (one template instance per
// size class) and in per-worker global scratch above that.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "gcl_device.h"
#include "gcl_search.h"

namespace gcl {
namespace {

#define GCL_HIP(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            return fail(GCL_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));     \
        }                                                                                 \
    } while (0)

typedef uint16_t ix_t;
constexpr int kWave = 64;

struct LevelRec {
    uint32_t tc_first, tc_len, cursor, arena_off;
    int64_t last_num_aut;
};

// ------------------------------------------------------------------ layout
// Per-worker global scratch.  Offsets are 256-byte aligned.
struct Layout {
    uint64_t undo_id, undo_tag, seq, first_seq, best_seq, lv, arena_mem, arena_rep, firstcode, canon_code,
        first_label, best_label, first_keys, best_keys, first_row, best_row, subgroup, auts, hot16, hot32, total;
};

__host__ __device__ inline uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

// Number of 16-bit and 32-bit hot arrays of a worker.
constexpr int kHot16 = 9;   // elem pos cend cfirst tmp sa sb tv tcl
constexpr int kHot32 = 5;   // cnt hits c1 c2 sv

__host__ __device__ inline Layout make_layout(uint32_t maxn, uint64_t max_nnz, uint32_t ring, uint64_t arena,
                                              bool hot_global) {
    Layout L;
    uint64_t o = 0;
    const uint64_t n1 = (uint64_t)maxn + 2;
    L.undo_id = o; o = align256(o + 2 * n1);
    L.undo_tag = o; o = align256(o + 2 * n1);
    L.seq = o; o = align256(o + 2 * n1);
    L.first_seq = o; o = align256(o + 2 * n1);
    L.best_seq = o; o = align256(o + 2 * n1);
    L.lv = o; o = align256(o + sizeof(LevelRec) * n1);
    L.arena_mem = o; o = align256(o + 2 * arena);
    L.arena_rep = o; o = align256(o + 2 * arena);
    L.firstcode = o; o = align256(o + 4 * n1);
    L.canon_code = o; o = align256(o + 4 * n1);
    L.first_label = o; o = align256(o + 2 * n1);
    L.best_label = o; o = align256(o + 2 * n1);
    L.first_keys = o; o = align256(o + 4 * (max_nnz + 1));
    L.best_keys = o; o = align256(o + 4 * (max_nnz + 1));
    L.first_row = o; o = align256(o + 4 * n1);
    L.best_row = o; o = align256(o + 4 * n1);
    L.subgroup = o; o = align256(o + 4 * (uint64_t)ring);
    L.auts = o; o = align256(o + 2 * (uint64_t)ring * maxn);
    L.hot16 = o; if (hot_global) o = align256(o + 2ull * kHot16 * n1);
    L.hot32 = o; if (hot_global) o = align256(o + 4ull * kHot32 * n1);
    L.total = o;
    return L;
}

struct KParams {
    int32_t num_items;
    const int32_t* items;       // graph ids processed by this launch
    const int64_t* vtx_off;     // G+1
    const int64_t* row_ptr;     // total_n+1 (global edge offsets)
    const int32_t* col;         // local ids
    const int32_t* init_order;  // per vertex: initial order for coloured graphs, or NULL
    const int32_t* color;       // per vertex or NULL
    int32_t* out_label;
    DeviceGraphResult* out;
    unsigned int* next_item;
    uint8_t* scratch;
    uint64_t scratch_stride;
    uint32_t maxn;
    uint64_t max_nnz;
    uint32_t ring;
    uint32_t max_aut;
    uint64_t arena_cap;
    uint64_t step_budget;       // splitter steps + search iterations per graph
};

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return (int)threadIdx.x; }
__device__ __forceinline__ void wsync() { __syncthreads(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}
__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor(x, o, kWave));
    return x;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x = max(x, (uint32_t)__shfl_xor(x, o, kWave));
    return x;
}
// inclusive prefix sum across the wave
__device__ __forceinline__ uint32_t wave_scan(uint32_t x) {
    const int l = lane_id();
    for (int o = 1; o < kWave; o <<= 1) {
        uint32_t y = __shfl_up(x, o, kWave);
        if (l >= o) x += y;
    }
    return x;
}
// ascending bitonic sort of one key per lane
__device__ __forceinline__ uint32_t wave_sort(uint32_t key) {
    const int l = lane_id();
#pragma unroll
    for (int k = 2; k <= kWave; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            uint32_t other = __shfl_xor(key, j, kWave);
            bool up = (l & k) == 0;
            bool lower = (l & j) == 0;
            uint32_t mn = min(key, other), mx = max(key, other);
            key = (lower == up) ? mn : mx;
        }
    }
    return key;
}

// ------------------------------------------------------------------ worker
struct Worker {
    // graph
    uint32_t n, nw;
    uint64_t nnz;
    const int64_t* rowp;
    const int32_t* col;
    // hot partition state (LDS or global)
    ix_t *elem, *pos, *cend, *cfirst, *tmp, *sa, *sb, *tv, *tcl;
    uint32_t *cnt, *hits, *c1, *c2, *sv;
    uint64_t *active, *nsmask, *vmask;  // LDS bitmaps
    uint32_t* ctr;                      // LDS counters
    uint32_t ncells, nactive;
    // per-worker global scratch
    ix_t *undo_id, *undo_tag, *seq, *first_seq, *best_seq, *arena_mem, *arena_rep, *first_label, *best_label,
        *auts;
    LevelRec* lv;
    int32_t *firstcode, *canon_code;
    uint32_t *first_keys, *best_keys, *first_row, *best_row, *subgroup;
    uint32_t nundo;
    uint64_t arena_cap;
    uint32_t ring, max_aut;
    uint64_t step_budget;
    // search state (reference src/sequential.h:24-58)
    uint32_t level, seq_size, tc_level_reached, eqlevfirst, first_seq_size, best_seq_size;
    int eqlevcanon, comp_canon, canon_level, code;
    bool first_discovered;
    int64_t num_aut;
    uint64_t nodes, steps, csr_bytes, budget;
    int32_t status;

    // --------------------------------------------------------- bit helpers
    __device__ bool test_bit(const uint64_t* m, uint32_t p) const { return (m[p >> 6] >> (p & 63)) & 1ull; }
    __device__ void add_active(uint32_t p) {
        if (p < n && !test_bit(active, p)) {
            if (lane_id() == 0) active[p >> 6] |= 1ull << (p & 63);
            ++nactive;
        }
    }
    __device__ void remove_active(uint32_t p) {
        if (p < n && test_bit(active, p)) {
            if (lane_id() == 0) active[p >> 6] &= ~(1ull << (p & 63));
            --nactive;
        }
    }
    __device__ void set_ns(uint32_t p, bool on) {
        if (lane_id() == 0) {
            if (on) nsmask[p >> 6] |= 1ull << (p & 63);
            else nsmask[p >> 6] &= ~(1ull << (p & 63));
        }
    }
    // first set bit at position >= start, or -1
    __device__ int32_t first_set_from(const uint64_t* m, uint32_t start) const {
        const int l = lane_id();
        for (uint32_t base = start >> 6; base < nw; base += kWave) {
            uint32_t w = base + (uint32_t)l;
            uint64_t bits = (w < nw) ? m[w] : 0ull;
            if (w == (start >> 6)) bits &= ~0ull << (start & 63);
            uint64_t b = ballot(bits != 0ull);
            if (b) {
                int fl = __builtin_ctzll(b);
                uint32_t lo = __shfl((uint32_t)bits, fl, kWave);
                uint32_t hi = __shfl((uint32_t)(bits >> 32), fl, kWave);
                uint64_t word = ((uint64_t)hi << 32) | lo;
                return (int32_t)((base + (uint32_t)fl) * 64u + (uint32_t)__builtin_ctzll(word));
            }
        }
        return -1;
    }

    // ------------------------------------------------------- cell rebuild
    // Rebuilds cend / cfirst / nsmask / ncells from vmask, a bitmap of cell
    // last positions.  O(n/64) wave iterations.
    __device__ void rebuild_from_lasts() {
        const int l = lane_id();
        // next nonzero word at or after each word: scan from the end
        uint32_t carry = 0xFFFFu;
        uint32_t total = 0;
        const uint32_t nchunks = (nw + kWave - 1) / kWave;
        for (int32_t ch = (int32_t)nchunks - 1; ch >= 0; --ch) {
            uint32_t w = (uint32_t)ch * kWave + (uint32_t)l;
            uint64_t bits = (w < nw) ? vmask[w] : 0ull;
            total += wave_sum((uint32_t)__popcll(bits));
            // suffix min of (bits ? w : inf) within the chunk, then with carry
            uint32_t v = bits ? w : 0xFFFFu;
            for (int o = 1; o < kWave; o <<= 1) {
                uint32_t y = __shfl_down(v, o, kWave);
                if (l + o < kWave) v = min(v, y);
            }
            v = min(v, carry);
            if (w < nw) tmp[w] = (ix_t)v;  // next nonzero word at or after w
            carry = __shfl(v, 0, kWave);
        }
        wsync();
        for (uint32_t base = 0; base < n; base += kWave) {
            uint32_t p = base + (uint32_t)l;
            if (p < n) {
                uint32_t w = p >> 6;
                uint64_t bits = vmask[w] >> (p & 63);
                uint32_t last;
                if (bits) {
                    last = p + (uint32_t)__builtin_ctzll(bits);
                } else {
                    uint32_t w2 = tmp[w + 1];  // the last position is always a cell last
                    last = w2 * 64u + (uint32_t)__builtin_ctzll(vmask[w2]);
                }
                cend[p] = (ix_t)last;
                bool is_first = (p == 0) || test_bit(vmask, p - 1);
                if (is_first) cfirst[last] = (ix_t)p;
            }
        }
        // nsmask: first positions of cells longer than one
        for (uint32_t base = 0; base < nw; base += kWave) {
            uint32_t w = base + (uint32_t)l;
            if (w < nw) {
                uint64_t lasts = vmask[w];
                uint64_t prev_last_bit = (w == 0) ? 1ull : (vmask[w - 1] >> 63);
                uint64_t firsts = (lasts << 1) | prev_last_bit;  // p first <=> p-1 last (or p == 0)
                uint64_t singles = firsts & lasts;
                uint64_t valid = (w * 64u + 64u <= n) ? ~0ull : ((1ull << (n - w * 64u)) - 1ull);
                nsmask[w] = (firsts & ~singles) & valid;
            }
        }
        ncells = total;
        wsync();
    }

    __device__ void lasts_from_cend() {
        const int l = lane_id();
        for (uint32_t base = 0; base < n; base += kWave) {
            uint32_t p = base + (uint32_t)l;
            bool is_last = p < n && cend[p] == p;
            uint64_t b = ballot(is_last);
            if (l == 0) vmask[base >> 6] = b;
        }
        wsync();
    }

    // --------------------------------------------------------- partitions
    __device__ void init_partition(const int32_t* init_order, const int32_t* color, int64_t vbase) {
        // initializePartition_s (reference src/partition_refinement.cpp:28-68):
        // vertices in (colour, id) order, one cell per colour run.  The host
        // provides the order for coloured graphs.
        const int l = lane_id();
        for (uint32_t base = 0; base < n; base += kWave) {
            uint32_t p = base + (uint32_t)l;
            bool is_last = false;
            if (p < n) {
                uint32_t v = init_order ? (uint32_t)init_order[vbase + p] : p;
                elem[p] = (ix_t)v;
                pos[v] = (ix_t)p;
                if (p + 1 == n) is_last = true;
                else if (color) {
                    uint32_t v2 = init_order ? (uint32_t)init_order[vbase + p + 1] : p + 1;
                    is_last = color[vbase + v] != color[vbase + v2];
                }
            }
            uint64_t b = ballot(is_last);
            if (l == 0) vmask[base >> 6] = b;
        }
        for (uint32_t base = 0; base < n; base += kWave) {
            uint32_t p = base + (uint32_t)l;
            if (p < n) {
                cnt[p] = 0; hits[p] = 0; c1[p] = 0; c2[p] = 0; sv[p] = 0;
                sa[p] = 0;
            }
        }
        for (uint32_t w = (uint32_t)l; w < nw; w += kWave) active[w] = 0;
        nactive = 0;
        wsync();
        rebuild_from_lasts();
        for (uint32_t w = (uint32_t)l; w < nw; w += kWave) vmask[w] = 0;
        wsync();
    }

    __device__ void push_undo_uniform(uint32_t id, uint32_t tag) {
        undo_id[nundo] = (ix_t)id;   // every lane stores the same value
        undo_tag[nundo] = (ix_t)tag;
        ++nundo;
    }

    // Refinement step for a splitter of size 1 whose vertex has at most 64
    // neighbours (reference src/partition_refinement.cpp:253-468 with
    // splitter.length == 1): cells holding neighbours split into
    // (neighbours, others); uniform cells add nothing to the code.
    __device__ void singleton_fast(uint32_t w, uint32_t tag, uint32_t& longcode, int& hint) {
        const int l = lane_id();
        const int64_t e0 = rowp[w];
        const uint32_t deg = (uint32_t)(rowp[w + 1] - e0);
        csr_bytes += 12 + 4ull * deg;
        const bool valid = (uint32_t)l < deg;
        uint32_t u = 0, c = 0, f = 0, r = 0;
        bool ns = false;
        if (valid) {
            u = (uint32_t)col[e0 + l];
            c = cend[pos[u]];
            f = cfirst[c];
            ns = f != c;
        }
        if (ns) {
            cnt[u] = 1;
            r = atomicAdd(&hits[c], 1u);
        }
        wsync();
        uint32_t k = 0, L = 0, q = 0;
        bool split = false, out = false;
        if (ns) {
            k = hits[c];
            L = c - f + 1;
            split = k < L;
            q = pos[u];
            out = split && q >= f + k;
        }
        if (split) {
            uint32_t x = elem[f + r];
            if (cnt[x] == 0) {
                uint32_t h = atomicAdd(&c1[c], 1u);
                tmp[f + h] = (ix_t)(f + r);
            }
        }
        uint32_t o = 0;
        if (out) o = atomicAdd(&c2[c], 1u);
        wsync();
        if (out) {
            uint32_t hp = tmp[f + o];
            uint32_t x = elem[hp];
            elem[hp] = (ix_t)u;
            pos[u] = (ix_t)hp;
            elem[q] = (ix_t)x;
            pos[x] = (ix_t)q;
        }
        wsync();
        const bool owner = ns && r == 0;
        if (split) cend[f + r] = (ix_t)(f + k - 1);
        if (owner) {
            hits[c] = 0;
            c1[c] = 0;
            c2[c] = 0;
            if (split) {
                cfirst[f + k - 1] = (ix_t)f;
                cfirst[c] = (ix_t)(f + k);
            }
        }
        if (ns) cnt[u] = 0;
        const uint64_t sm = ballot(owner && split);
        const uint32_t nsplit = (uint32_t)__popcll(sm);
        if (owner && split) {
            uint32_t idx = nundo + lanes_below(sm);
            undo_id[idx] = (ix_t)(f + k - 1);
            undo_tag[idx] = (ix_t)tag;
        }
        nundo += nsplit;
        ncells += nsplit;
        wsync();
        if (nsplit == 0) return;
        // serial part in position order (reference 399-435)
        uint32_t key = (owner && split) ? ((f << 6) | (uint32_t)l) : 0xFFFFFFFFu;
        key = wave_sort(key);
        const uint32_t src = key & 63u;
        const uint32_t sf = key >> 6;
        const uint32_t sk = __shfl(k, (int)src, kWave);
        const uint32_t sL = __shfl(L, (int)src, kWave);
        for (uint32_t t = 0; t < nsplit; ++t) {
            const uint32_t F = __builtin_amdgcn_readlane(sf, t);
            const uint32_t K = __builtin_amdgcn_readlane(sk, t);
            const uint32_t LL = __builtin_amdgcn_readlane(sL, t);
            longcode = GCL_MASH(longcode, F + K - 1);
            if (test_bit(active, F) || K >= LL - K) {
                add_active(F + K);
                if (LL - K == 1) hint = (int)(F + K);
            } else {
                add_active(F);
                if (K == 1) hint = (int)F;
            }
            set_ns(F, K > 1);
            set_ns(F + K, LL - K > 1);
        }
        wsync();
    }

    // General refinement step: splitter cell [p, e] of any size (reference
    // src/partition_refinement.cpp:197-473).  Counts with LDS atomics, then
    // every touched cell is regrouped by a counting sort on its key
    // (count for |W| > 1; neighbours first for |W| == 1), then one pass over
    // the pre-step non-singleton cells in position order does the hash chain
    // and the active-set rules.
    __device__ void general_step(uint32_t p, uint32_t e, uint32_t tag, uint32_t& longcode, int& hint) {
        const int l = lane_id();
        const uint32_t s = e - p + 1;
        if (l == 0) {
            ctr[0] = 0;
            ctr[1] = 0;
        }
        wsync();
        uint32_t degsum = 0;
        for (uint32_t base = 0; base < s; base += kWave) {
            uint32_t i = base + (uint32_t)l;
            if (i < s) {
                uint32_t w = elem[p + i];
                int64_t a = rowp[w], b = rowp[w + 1];
                degsum += (uint32_t)(b - a);
                for (int64_t x = a; x < b; ++x) {
                    uint32_t u = (uint32_t)col[x];
                    if (atomicAdd(&cnt[u], 1u) == 0u) {
                        uint32_t idx = atomicAdd(&ctr[0], 1u);
                        tv[idx] = (ix_t)u;
                    }
                }
            }
        }
        csr_bytes += 12ull * s + 4ull * wave_sum(degsum);
        wsync();
        const uint32_t ntv = ctr[0];
        for (uint32_t base = 0; base < ntv; base += kWave) {
            uint32_t i = base + (uint32_t)l;
            if (i < ntv) {
                uint32_t u = tv[i];
                uint32_t c = cend[pos[u]];
                if (cfirst[c] != c) {
                    if (atomicAdd(&hits[c], 1u) == 0u) {
                        uint32_t idx = atomicAdd(&ctr[1], 1u);
                        tcl[idx] = (ix_t)c;
                    }
                }
            }
        }
        wsync();
        const uint32_t ntc = ctr[1];
        // regroup each touched cell
        for (uint32_t j = 0; j < ntc; ++j) {
            const uint32_t c = tcl[j];
            const uint32_t f = cfirst[c];
            const uint32_t L = c - f + 1;
            uint32_t kmin = 0xFFFFFFFFu, kmax = 0;
            for (uint32_t base = 0; base < L; base += kWave) {
                uint32_t i = base + (uint32_t)l;
                if (i < L) {
                    uint32_t v = elem[f + i];
                    uint32_t kv = (s == 1) ? 1u - cnt[v] : cnt[v];
                    kmin = min(kmin, kv);
                    kmax = max(kmax, kv);
                }
            }
            kmin = wave_min(kmin);
            kmax = wave_max(kmax);
            if (l == 0) hits[c] = 0;
            if (kmin == kmax) {
                if (l == 0) {
                    sa[f] = 1;
                    sv[f] = (s == 1) ? 1u - kmin : kmin;
                    tcl[j] = (ix_t)f;
                }
                wsync();
                continue;
            }
            const uint32_t R = kmax - kmin + 1;
            for (uint32_t b = (uint32_t)l; b < R; b += kWave) c1[b] = 0;
            wsync();
            for (uint32_t base = 0; base < L; base += kWave) {
                uint32_t i = base + (uint32_t)l;
                if (i < L) {
                    uint32_t v = elem[f + i];
                    uint32_t kv = (s == 1) ? 1u - cnt[v] : cnt[v];
                    sb[v] = (ix_t)atomicAdd(&c1[kv - kmin], 1u);
                }
            }
            wsync();
            // exclusive scan of bucket sizes into c2, fragments counted
            uint32_t run = 0, nfrag = 0;
            for (uint32_t base = 0; base < R; base += kWave) {
                uint32_t b = base + (uint32_t)l;
                uint32_t sz = b < R ? c1[b] : 0u;
                uint32_t inc = wave_scan(sz);
                if (b < R) c2[b] = run + inc - sz;
                run += __shfl(inc, kWave - 1, kWave);
                nfrag += (uint32_t)__popcll(ballot(sz != 0));
            }
            wsync();
            for (uint32_t base = 0; base < L; base += kWave) {
                uint32_t i = base + (uint32_t)l;
                if (i < L) {
                    uint32_t v = elem[f + i];
                    uint32_t kv = (s == 1) ? 1u - cnt[v] : cnt[v];
                    tmp[f + c2[kv - kmin] + sb[v]] = (ix_t)v;
                }
            }
            wsync();
            for (uint32_t base = 0; base < L; base += kWave) {
                uint32_t i = base + (uint32_t)l;
                bool push = false;
                uint32_t last = 0;
                if (i < L) {
                    uint32_t P = f + i;
                    uint32_t v = tmp[P];
                    elem[P] = (ix_t)v;
                    pos[v] = (ix_t)P;
                    uint32_t kv = (s == 1) ? 1u - cnt[v] : cnt[v];
                    uint32_t b = kv - kmin;
                    uint32_t first = f + c2[b];
                    last = first + c1[b] - 1;
                    cend[P] = (ix_t)last;
                    if (P == first) {
                        cfirst[last] = (ix_t)first;
                        push = last != c;
                    }
                }
                uint64_t pm = ballot(push);
                if (push) {
                    uint32_t idx = nundo + lanes_below(pm);
                    undo_id[idx] = (ix_t)last;
                    undo_tag[idx] = (ix_t)tag;
                }
                nundo += (uint32_t)__popcll(pm);
            }
            ncells += nfrag - 1;
            wsync();
            for (uint32_t b = (uint32_t)l; b < R; b += kWave) {
                c1[b] = 0;
                c2[b] = 0;
            }
            if (l == 0) {
                sa[f] = 2;
                sv[f] = c;
                tcl[j] = (ix_t)f;
            }
            wsync();
        }
        // serial pass over the pre-step non-singleton cells (reference 253-468)
        for (uint32_t wi = 0; wi < nw; ++wi) {
            uint64_t bits = nsmask[wi];
            while (bits) {
                const uint32_t f = wi * 64u + (uint32_t)__builtin_ctzll(bits);
                bits &= bits - 1;
                const uint32_t st = sa[f];
                if (st == 0) {
                    if (s > 1) longcode = GCL_MASH(longcode, f);
                    continue;
                }
                if (st == 1) {
                    if (s > 1) longcode = GCL_MASH(longcode, sv[f] + f);
                    continue;
                }
                const uint32_t cold = sv[f];
                if (s == 1) {
                    const uint32_t K = (uint32_t)cend[f] - f + 1;
                    const uint32_t LL = cold - f + 1;
                    longcode = GCL_MASH(longcode, f + K - 1);
                    if (test_bit(active, f) || K >= LL - K) {
                        add_active(f + K);
                        if (LL - K == 1) hint = (int)(f + K);
                    } else {
                        add_active(f);
                        if (K == 1) hint = (int)f;
                    }
                    continue;
                }
                uint32_t max_size = 0, max_first = f;
                for (uint32_t q = f; q <= cold;) {
                    const uint32_t last = cend[q];
                    const uint32_t size = last - q + 1;
                    longcode = GCL_MASH(longcode, cnt[elem[q]] + q);
                    if (size > max_size) {
                        max_size = size;
                        max_first = q;
                    }
                    q = last + 1;
                }
                const bool was_active = test_bit(active, f);
                for (uint32_t q = (uint32_t)cend[f] + 1; q <= cold;) {
                    const uint32_t last = cend[q];
                    add_active(q);
                    if (last == q) hint = (int)q;
                    q = last + 1;
                }
                if (!was_active) {
                    add_active(f);
                    remove_active(max_first);
                }
            }
        }
        wsync();
        // nsmask of the new fragments; reset the per-cell state (tcl holds old firsts)
        for (uint32_t base = 0; base < ntc; base += kWave) {
            uint32_t i = base + (uint32_t)l;
            if (i < ntc) {
                const uint32_t f = tcl[i];
                if (sa[f] == 2) {
                    const uint32_t cold = sv[f];
                    for (uint32_t q = f; q <= cold;) {
                        const uint32_t last = cend[q];
                        const uint64_t bit = 1ull << (q & 63);
                        if (last > q) atomicOr((unsigned long long*)&nsmask[q >> 6], (unsigned long long)bit);
                        else atomicAnd((unsigned long long*)&nsmask[q >> 6], ~(unsigned long long)bit);
                        q = last + 1;
                    }
                }
                sa[f] = 0;
            }
        }
        wsync();
        for (uint32_t base = 0; base < ntv; base += kWave) {
            uint32_t i = base + (uint32_t)l;
            if (i < ntv) cnt[tv[i]] = 0;
        }
        wsync();
    }

    // Partition_s_refinement (reference src/partition_refinement.cpp:150-485)
    __device__ int refine(int32_t init_active, uint32_t tag) {
        const int l = lane_id();
        if (init_active < 0) {
            uint32_t total = 0;
            for (uint32_t base = 0; base < n; base += kWave) {
                uint32_t p = base + (uint32_t)l;
                bool first = p < n && cfirst[cend[p]] == p;
                uint64_t b = ballot(first);
                if (l == 0) active[base >> 6] = b;
                total += (uint32_t)__popcll(b);
            }
            nactive = total;
            wsync();
        } else {
            add_active((uint32_t)init_active);
            wsync();
        }
        uint32_t longcode = ncells;
        int hint = 0;
        while (ncells < n && nactive > 0) {
            if (budget == 0) {
                status = GCL_E_INTERNAL;  // step budget exhausted
                break;
            }
            --budget;
            int32_t sp = first_set_from(active, (uint32_t)hint);
            if (sp < 0) sp = first_set_from(active, 0);
            const uint32_t p = (uint32_t)sp;
            remove_active(p);
            wsync();
            ++steps;
            const uint32_t e = cend[p];
            longcode = GCL_MASH(longcode, p + e);
            if (e > p) longcode = GCL_MASH(longcode, e - p + 1);
            if (e == p) {
                const uint32_t w = elem[p];
                if (rowp[w + 1] - rowp[w] <= kWave) singleton_fast(w, tag, longcode, hint);
                else general_step(p, e, tag, longcode, hint);
            } else {
                general_step(p, e, tag, longcode, hint);
            }
        }
        if (nactive) {
            for (uint32_t w = (uint32_t)l; w < nw; w += kWave) active[w] = 0;
            nactive = 0;
            wsync();
        }
        longcode = GCL_MASH(longcode, ncells);
        return GCL_CLEANUP(longcode);
    }

    // split_by_and_refine (reference src/partition_refinement.cpp:486-539)
    __device__ void individualize_and_refine(uint32_t v, uint32_t f, uint32_t tag) {
        const uint32_t c = cend[f];
        const uint32_t q = pos[v];
        const uint32_t x = elem[f];
        wsync();
        if (lane_id() == 0) {
            elem[f] = (ix_t)v;
            pos[v] = (ix_t)f;
            elem[q] = (ix_t)x;
            pos[x] = (ix_t)q;
            cend[f] = (ix_t)f;
            cfirst[f] = (ix_t)f;
            cfirst[c] = (ix_t)(f + 1);
        }
        set_ns(f, false);
        set_ns(f + 1, c > f + 1);
        push_undo_uniform(f, tag);
        ++ncells;
        wsync();
        code = refine((int32_t)f, tag);
    }

    // Undo every cell created above `lev`: drop those cell boundaries and
    // rebuild the cell structure (reconstruct_at_level, reference
    // src/partition_refinement.cpp:541-604, up to in-cell order).
    __device__ void undo_to(uint32_t lev) {
        // tags are non-decreasing along the log: first entry with tag > lev
        uint32_t lo = 0, hi = nundo;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((uint32_t)undo_tag[mid] > lev) hi = mid;
            else lo = mid + 1;
        }
        const uint32_t k0 = lo;
        if (k0 == nundo) return;
        lasts_from_cend();
        const int l = lane_id();
        for (uint32_t i = k0 + (uint32_t)l; i < nundo; i += kWave) {
            uint32_t x = undo_id[i];
            atomicAnd((unsigned long long*)&vmask[x >> 6], ~(1ull << (x & 63)));
        }
        wsync();
        nundo = k0;
        rebuild_from_lasts();
        for (uint32_t w = (uint32_t)l; w < nw; w += kWave) vmask[w] = 0;
        wsync();
    }

    // --------------------------------------------------------- search ops
    // target_cell_selector with hint -1 (reference
    // src/partition_refinement.cpp:607-655): first non-singleton cell, members
    // listed in increasing vertex id.
    __device__ bool select_target() {
        const int32_t f = first_set_from(nsmask, 0);
        if (f < 0) return true;
        const int l = lane_id();
        const uint32_t c = cend[(uint32_t)f];
        const uint32_t t = c - (uint32_t)f + 1;
        const uint32_t idx = tc_level_reached;
        const uint64_t off = idx == 0 ? 0 : (uint64_t)lv[idx - 1].arena_off + lv[idx - 1].tc_len;
        if (off + t > arena_cap) {
            status = GCL_E_CAPACITY;
            return false;
        }
        for (uint32_t i = (uint32_t)l; i < t; i += kWave) {
            uint32_t v = elem[(uint32_t)f + i];
            atomicOr((unsigned long long*)&vmask[v >> 6], 1ull << (v & 63));
        }
        wsync();
        uint32_t run = 0;
        for (uint32_t base = 0; base < nw; base += kWave) {
            uint32_t w = base + (uint32_t)l;
            uint64_t bits = w < nw ? vmask[w] : 0ull;
            uint32_t c0 = (uint32_t)__popcll(bits);
            uint32_t inc = wave_scan(c0);
            uint32_t j = run + inc - c0;
            while (bits) {
                uint32_t b = (uint32_t)__builtin_ctzll(bits);
                bits &= bits - 1;
                arena_mem[off + j] = (ix_t)(w * 64u + b);
                arena_rep[off + j] = (ix_t)(w * 64u + b);
                ++j;
            }
            if (w < nw) vmask[w] = 0;
            run += __shfl(inc, kWave - 1, kWave);
        }
        LevelRec r;
        r.tc_first = (uint32_t)f;
        r.tc_len = t;
        r.cursor = 0;
        r.arena_off = (uint32_t)off;
        r.last_num_aut = 0;
        lv[idx] = r;  // uniform store by every lane
        ++tc_level_reached;
        wsync();
        return true;
    }

    // subgroup_fixing_sequence + compute_orbits_fast (reference
    // src/perm_group.cpp:5-116), lock-free union by minimum.
    __device__ void compute_orbits(uint32_t li) {
        const int l = lane_id();
        const LevelRec r = lv[li];
        const int64_t upto = min(num_aut, (int64_t)ring);
        const int64_t from = r.last_num_aut < 0 ? 0 : r.last_num_aut;
        uint32_t nsub = 0;
        for (int64_t i = from; i < upto; ++i) {
            const ix_t* a = auts + (uint64_t)i * n;
            bool bad = false;
            for (uint32_t base = 0; base < seq_size; base += kWave) {
                uint32_t j = base + (uint32_t)l;
                bool b = j < seq_size && a[seq[j]] != seq[j];
                if (ballot(b)) {
                    bad = true;
                    break;
                }
            }
            if (!bad) {
                subgroup[nsub] = (uint32_t)i;  // uniform store
                ++nsub;
            }
        }
        if (nsub == 0) return;
        // cnt is zero between refinements; use it as the parent array
        uint32_t* parent = cnt;
        for (uint32_t v = (uint32_t)l; v < n; v += kWave) parent[v] = v;
        wsync();
        for (uint32_t i = (uint32_t)l; i < r.tc_len; i += kWave)
            parent[arena_mem[r.arena_off + i]] = arena_rep[r.arena_off + i];
        wsync();
        for (uint32_t s = 0; s < nsub; ++s) {
            const ix_t* a = auts + (uint64_t)subgroup[s] * n;
            for (uint32_t v = (uint32_t)l; v < n; v += kWave) {
                uint32_t img = a[v];
                if (img == v) continue;
                // parents only ever point to smaller ids, so every walk is
                // shorter than n; the caps only bound a corrupted structure
                uint32_t x = v, y = img;
                for (uint32_t attempt = 0; attempt < n; ++attempt) {
                    for (uint32_t h = 0; h < n; ++h) {
                        uint32_t px = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (px == x) break;
                        x = px;
                    }
                    for (uint32_t h = 0; h < n; ++h) {
                        uint32_t py = __hip_atomic_load(&parent[y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (py == y) break;
                        y = py;
                    }
                    if (x == y) break;
                    uint32_t lo = min(x, y), hi = max(x, y);
                    if (atomicCAS(&parent[hi], hi, lo) == hi) break;
                    x = lo;
                    y = hi;
                }
            }
            wsync();
        }
        for (uint32_t i = (uint32_t)l; i < r.tc_len; i += kWave) {
            uint32_t x = arena_mem[r.arena_off + i];
            for (uint32_t h = 0; h < n && parent[x] != x; ++h) x = parent[x];
            arena_rep[r.arena_off + i] = (ix_t)x;
        }
        wsync();
        for (uint32_t v = (uint32_t)l; v < n; v += kWave) parent[v] = 0;
        wsync();
    }

    // Certificate of the current leaf: sorted canonical rows (keys i*n+j)
    // and row offsets (the sparse form of the reference's n x n hash,
    // src/CSRGraph.cpp:332-351).
    __device__ void store_cert(uint32_t* keys, uint32_t* row, ix_t* label) {
        const int l = lane_id();
        uint32_t run = 0;
        for (uint32_t base = 0; base < n; base += kWave) {
            uint32_t i = base + (uint32_t)l;
            uint32_t v = 0, d = 0;
            if (i < n) {
                v = elem[i];
                d = (uint32_t)(rowp[v + 1] - rowp[v]);
            }
            uint32_t inc = wave_scan(d);
            uint32_t o = run + inc - d;
            if (i < n) {
                row[i] = o;
                label[i] = (ix_t)v;
                const int64_t a = rowp[v];
                for (uint32_t t = 0; t < d; ++t) {
                    uint32_t key = i * n + pos[(uint32_t)col[a + t]];
                    uint32_t z = t;
                    while (z > 0 && keys[o + z - 1] > key) {
                        keys[o + z] = keys[o + z - 1];
                        --z;
                    }
                    keys[o + z] = key;
                }
            }
            run += __shfl(inc, kWave - 1, kWave);
        }
        if (l == 0) row[n] = run;
        wsync();
    }

    // compare_adj_upper (reference src/sequential.cpp:114-124) of the current
    // leaf against a stored certificate, without sorting the current rows:
    // at the first row whose neighbour sets differ, the graph holding the
    // smallest element of the symmetric difference is the greater one.
    __device__ int compare_cert(const uint32_t* keys, const uint32_t* row) {
        const int l = lane_id();
        for (uint32_t base = 0; base < n; base += kWave) {
            uint32_t i = base + (uint32_t)l;
            bool differs = false;
            if (i < n) {
                const uint32_t v = elem[i];
                const int64_t a = rowp[v];
                const uint32_t d = (uint32_t)(rowp[v + 1] - a);
                const uint32_t r0 = row[i], r1 = row[i + 1];
                if (r1 - r0 != d) {
                    differs = true;
                } else {
                    for (uint32_t t = 0; t < d && !differs; ++t) {
                        const uint32_t key = i * n + pos[(uint32_t)col[a + t]];
                        uint32_t lo = r0, hi = r1;
                        while (lo < hi) {
                            uint32_t mid = (lo + hi) >> 1;
                            if (keys[mid] < key) lo = mid + 1;
                            else hi = mid;
                        }
                        if (lo == r1 || keys[lo] != key) differs = true;
                    }
                }
            }
            const uint64_t dm = ballot(differs);
            if (dm) {
                const uint32_t fl = (uint32_t)__builtin_ctzll(dm);
                const uint32_t i0 = base + fl;
                // smallest element of each side of the symmetric difference
                const uint32_t v = elem[i0];
                const int64_t a = rowp[v];
                const uint32_t d = (uint32_t)(rowp[v + 1] - a);
                const uint32_t r0 = row[i0], r1 = row[i0 + 1];
                uint32_t min_cur = 0xFFFFFFFFu, min_ref = 0xFFFFFFFFu;
                for (uint32_t t = (uint32_t)l; t < d; t += kWave) {
                    const uint32_t key = i0 * n + pos[(uint32_t)col[a + t]];
                    bool found = false;
                    for (uint32_t z = r0; z < r1; ++z)
                        if (keys[z] == key) {
                            found = true;
                            break;
                        }
                    if (!found) min_cur = min(min_cur, key);
                }
                for (uint32_t z = r0 + (uint32_t)l; z < r1; z += kWave) {
                    const uint32_t key = keys[z];
                    bool found = false;
                    for (uint32_t t = 0; t < d; ++t)
                        if (i0 * n + pos[(uint32_t)col[a + t]] == key) {
                            found = true;
                            break;
                        }
                    if (!found) min_ref = min(min_ref, key);
                }
                min_cur = wave_min(min_cur);
                min_ref = wave_min(min_ref);
                return min_cur < min_ref ? 1 : -1;
            }
        }
        return 0;
    }

    __device__ void store_aut(const ix_t* ref_label) {
        const size_t slot = (size_t)(num_aut % (int64_t)ring);
        ix_t* dst = auts + slot * n;
        for (uint32_t j = (uint32_t)lane_id(); j < n; j += kWave) dst[ref_label[j]] = elem[j];
        ++num_aut;
        wsync();
    }

    __device__ int gca_level(const ix_t* a, uint32_t na) {
        // get_gca_level (reference src/sequential.cpp:183-213)
        const uint32_t m = min(na, seq_size);
        const int l = lane_id();
        for (uint32_t base = 0; base < m; base += kWave) {
            uint32_t k = base + (uint32_t)l;
            uint64_t dm = ballot(k < m && a[k] != seq[k]);
            if (dm) return (int)(base + (uint32_t)__builtin_ctzll(dm)) + 1;
        }
        return -1;
    }

    __device__ void backtrack(uint32_t lev) {
        // backtrack_to (reference src/sequential.cpp:650-670)
        if (lev == 0) {
            level = 0;
            return;
        }
        undo_to(lev);
        if (lev < eqlevfirst) eqlevfirst = lev;
        if ((int)lev <= eqlevcanon) {
            eqlevcanon = (int)lev;
            comp_canon = 0;
        }
        seq_size = lev - 1;
        tc_level_reached = lev;
        level = lev;
    }

    __device__ void first_terminal() {
        // first_terminal (reference src/sequential.cpp:259-303)
        eqlevfirst = level;
        firstcode[level] = kSentinel;
        store_cert(first_keys, first_row, first_label);
        store_cert(best_keys, best_row, best_label);
        for (uint32_t i = (uint32_t)lane_id(); i < seq_size; i += kWave) {
            first_seq[i] = seq[i];
            best_seq[i] = seq[i];
        }
        first_seq_size = best_seq_size = seq_size;
        first_discovered = true;
        canon_level = eqlevcanon = (int)level;
        comp_canon = 0;
        for (uint32_t i = (uint32_t)lane_id(); i < level; i += kWave) canon_code[i] = firstcode[i];
        wsync();
        canon_code[level] = kSentinel;
        wsync();
    }

    __device__ bool first_path_node() {
        // first_path_node (reference src/sequential.cpp:216-257)
        if (ncells == n) {
            first_terminal();
            backtrack(level - 1);
            return true;
        }
        if (!select_target()) return false;
        LevelRec r = lv[tc_level_reached - 1];
        const uint32_t v = arena_mem[r.arena_off + r.cursor];
        r.cursor++;
        lv[tc_level_reached - 1].cursor = r.cursor;
        ++nodes;
        seq[seq_size] = (ix_t)v;
        ++seq_size;
        individualize_and_refine(v, r.tc_first, level + 1);
        ++level;
        firstcode[level - 1] = code;
        return true;
    }

    __device__ void classify() {
        // classify_node (reference src/sequential.cpp:396-624)
        const uint32_t L = level;
        if (eqlevfirst == L - 1 && code == firstcode[L - 1]) eqlevfirst = L;
        if (eqlevcanon == (int)L - 1) {
            const int cc = canon_code[L - 1];
            if (code < cc) comp_canon = -1;
            else if (code > cc) comp_canon = 1;
            else {
                comp_canon = 0;
                eqlevcanon = (int)L;
            }
        }
        if (comp_canon > 0) canon_code[L - 1] = code;
        int decision = kNormal;
        if (eqlevfirst != L && comp_canon < 0) {
            decision = kBad;
        } else if (ncells == n) {
            if (eqlevfirst == L && compare_cert(first_keys, first_row) == 0) decision = kAutFirst;
            if (decision == kNormal) {
                if (comp_canon == 0) {
                    if ((int)L < canon_level) comp_canon = 1;
                    else comp_canon = compare_cert(best_keys, best_row);
                }
                if (comp_canon == 0) decision = kAutBest;
                else if (comp_canon > 0) decision = kNewBest;
                else decision = kBad;
            }
        }
        wsync();
        switch (decision) {
            case kNormal:
                break;
            case kAutFirst: {
                store_aut(first_label);
                const int gl = gca_level(first_seq, first_seq_size);
                backtrack(gl >= 0 ? (uint32_t)gl : L - 1);
                break;
            }
            case kAutBest: {
                store_aut(best_label);
                const int gl = gca_level(best_seq, best_seq_size);
                backtrack(gl >= 0 ? (uint32_t)gl : L - 1);
                break;
            }
            case kNewBest: {
                canon_level = eqlevcanon = (int)L;
                comp_canon = 0;
                canon_code[L] = kSentinel;
                store_cert(best_keys, best_row, best_label);
                for (uint32_t i = (uint32_t)lane_id(); i < seq_size; i += kWave) best_seq[i] = seq[i];
                best_seq_size = seq_size;
                wsync();
                backtrack(L - 1);
                break;
            }
            case kBad:
                backtrack(L - 1);
                break;
        }
    }

    __device__ bool other_path_node() {
        // other_path_node (reference src/sequential.cpp:304-392)
        if (ncells < n && (eqlevfirst == level || comp_canon >= 0)) {
            if (tc_level_reached < level) {
                if (!select_target()) return false;
            }
            const uint32_t li = level - 1;
            LevelRec r = lv[li];
            if (r.cursor >= r.tc_len) {
                backtrack(level - 1);
                return true;
            }
            if (num_aut > 0 && num_aut > r.last_num_aut) {
                compute_orbits(li);
                lv[li].last_num_aut = num_aut;
            }
            // next member that is its own orbit representative
            const int l = lane_id();
            uint32_t cur = r.cursor;
            while (cur < r.tc_len) {
                uint32_t i = cur + (uint32_t)l;
                bool ok = i < r.tc_len && arena_rep[r.arena_off + i] == arena_mem[r.arena_off + i];
                uint64_t m = ballot(ok);
                if (m) {
                    cur += (uint32_t)__builtin_ctzll(m);
                    break;
                }
                cur += kWave;
            }
            if (cur >= r.tc_len) {
                lv[li].cursor = r.tc_len;
                backtrack(level - 1);
                return true;
            }
            const uint32_t v = arena_mem[r.arena_off + cur];
            lv[li].cursor = cur + 1;
            ++nodes;
            seq[seq_size] = (ix_t)v;
            ++seq_size;
            wsync();
            individualize_and_refine(v, r.tc_first, level + 1);
            ++level;
        }
        classify();
        return true;
    }

    __device__ void run(const int32_t* init_order, const int32_t* color, int64_t vbase) {
        nodes = 0;
        steps = 0;
        csr_bytes = 0;
        status = 0;
        num_aut = 0;
        nundo = 0;
        level = 0;
        seq_size = 0;
        tc_level_reached = 0;
        eqlevfirst = 0;
        eqlevcanon = 0;
        comp_canon = 0;
        canon_level = 0;
        code = 0;
        first_seq_size = best_seq_size = 0;
        first_discovered = false;
        budget = step_budget;
        init_partition(init_order, color, vbase);
        // search_tree_traversal (reference src/sequential.cpp:626-648)
        code = refine(-1, 1);
        level = 1;
        firstcode[0] = code;
        ++nodes;
        wsync();
        while (level >= 1 && status == 0) {
            if (budget == 0) {
                status = GCL_E_INTERNAL;
                break;
            }
            --budget;
            bool ok = first_discovered ? other_path_node() : first_path_node();
            if (!ok) break;
        }
    }
};

template <int MAXN>
struct LdsHot {
    ix_t a16[kHot16][MAXN];
    uint32_t a32[kHot32][MAXN];
    uint64_t bm[3][MAXN / 64 + 1];
    uint32_t ctr[4];
};

struct LdsGlobalHot {
    uint64_t bm[3][1025];
    uint32_t ctr[4];
};

__device__ void bind_scratch(Worker& W, const KParams& P, uint8_t* base, const Layout& L) {
    W.undo_id = (ix_t*)(base + L.undo_id);
    W.undo_tag = (ix_t*)(base + L.undo_tag);
    W.seq = (ix_t*)(base + L.seq);
    W.first_seq = (ix_t*)(base + L.first_seq);
    W.best_seq = (ix_t*)(base + L.best_seq);
    W.lv = (LevelRec*)(base + L.lv);
    W.arena_mem = (ix_t*)(base + L.arena_mem);
    W.arena_rep = (ix_t*)(base + L.arena_rep);
    W.firstcode = (int32_t*)(base + L.firstcode);
    W.canon_code = (int32_t*)(base + L.canon_code);
    W.first_label = (ix_t*)(base + L.first_label);
    W.best_label = (ix_t*)(base + L.best_label);
    W.first_keys = (uint32_t*)(base + L.first_keys);
    W.best_keys = (uint32_t*)(base + L.best_keys);
    W.first_row = (uint32_t*)(base + L.first_row);
    W.best_row = (uint32_t*)(base + L.best_row);
    W.subgroup = (uint32_t*)(base + L.subgroup);
    W.auts = (ix_t*)(base + L.auts);
    W.arena_cap = P.arena_cap;
    W.ring = P.ring;
    W.max_aut = P.max_aut;
    W.step_budget = P.step_budget;
}

__device__ void worker_loop(Worker& W, const KParams& P) {
    const int l = lane_id();
    while (true) {
        uint32_t item = 0;
        if (l == 0) item = atomicAdd(P.next_item, 1u);
        item = __shfl(item, 0, kWave);
        if (item >= (uint32_t)P.num_items) break;
        const int32_t g = P.items[item];
        const int64_t vbase = P.vtx_off[g];
        W.n = (uint32_t)(P.vtx_off[g + 1] - vbase);
        W.nw = (W.n + 63) / 64;
        W.rowp = P.row_ptr + vbase;
        W.nnz = (uint64_t)(W.rowp[W.n] - W.rowp[0]);
        W.col = P.col;
        if (W.n == 0) {
            if (l == 0) {
                DeviceGraphResult res = {0, 0, 0, 0, (int32_t)GCL_E_INVALID, 0};
                P.out[g] = res;
            }
            continue;
        }
        W.run(P.init_order, P.color, vbase);
        for (uint32_t i = (uint32_t)l; i < W.n; i += kWave) P.out_label[vbase + i] = (int32_t)W.best_label[i];
        if (l == 0) {
            DeviceGraphResult res;
            res.nodes = W.nodes;
            res.automorphisms = (uint64_t)W.num_aut;
            res.splitter_steps = W.steps;
            res.csr_bytes = W.csr_bytes;
            res.status = W.status;
            res.pad = 0;
            P.out[g] = res;
        }
        wsync();
    }
}

template <int MAXN>
__global__ __launch_bounds__(kWave) void search_kernel_lds(KParams P) {
    __shared__ LdsHot<MAXN> S;
    Worker W;
    uint8_t* base = P.scratch + (uint64_t)blockIdx.x * P.scratch_stride;
    const Layout L = make_layout(P.maxn, P.max_nnz, P.ring, P.arena_cap, false);
    bind_scratch(W, P, base, L);
    W.elem = S.a16[0]; W.pos = S.a16[1]; W.cend = S.a16[2]; W.cfirst = S.a16[3]; W.tmp = S.a16[4];
    W.sa = S.a16[5]; W.sb = S.a16[6]; W.tv = S.a16[7]; W.tcl = S.a16[8];
    W.cnt = S.a32[0]; W.hits = S.a32[1]; W.c1 = S.a32[2]; W.c2 = S.a32[3]; W.sv = S.a32[4];
    W.active = S.bm[0]; W.nsmask = S.bm[1]; W.vmask = S.bm[2];
    W.ctr = S.ctr;
    for (int w = lane_id(); w < MAXN / 64 + 1; w += kWave) S.bm[2][w] = 0;
    wsync();
    worker_loop(W, P);
}

__global__ __launch_bounds__(kWave) void search_kernel_global(KParams P) {
    __shared__ LdsGlobalHot S;
    Worker W;
    uint8_t* base = P.scratch + (uint64_t)blockIdx.x * P.scratch_stride;
    const Layout L = make_layout(P.maxn, P.max_nnz, P.ring, P.arena_cap, true);
    bind_scratch(W, P, base, L);
    const uint64_t n1 = (uint64_t)P.maxn + 2;
    ix_t* h16 = (ix_t*)(base + L.hot16);
    uint32_t* h32 = (uint32_t*)(base + L.hot32);
    W.elem = h16 + 0 * n1; W.pos = h16 + 1 * n1; W.cend = h16 + 2 * n1; W.cfirst = h16 + 3 * n1;
    W.tmp = h16 + 4 * n1; W.sa = h16 + 5 * n1; W.sb = h16 + 6 * n1; W.tv = h16 + 7 * n1; W.tcl = h16 + 8 * n1;
    W.cnt = h32 + 0 * n1; W.hits = h32 + 1 * n1; W.c1 = h32 + 2 * n1; W.c2 = h32 + 3 * n1; W.sv = h32 + 4 * n1;
    W.active = S.bm[0]; W.nsmask = S.bm[1]; W.vmask = S.bm[2];
    W.ctr = S.ctr;
    for (int w = lane_id(); w < 1025; w += kWave) S.bm[2][w] = 0;
    wsync();
    worker_loop(W, P);
}

// ------------------------------------------------------------------- host
struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

constexpr int kClasses = 6;
constexpr uint32_t kClassN[kClasses] = {64, 128, 256, 512, 1024, 2048};

template <int MAXN>
hipError_t launch_lds(const KParams& P, int grid, hipStream_t s) {
    hipLaunchKernelGGL(search_kernel_lds<MAXN>, dim3(grid), dim3(kWave), 0, s, P);
    return hipGetLastError();
}

hipError_t launch_class(int cls, const KParams& P, int grid, hipStream_t s) {
    switch (cls) {
        case 0: return launch_lds<64>(P, grid, s);
        case 1: return launch_lds<128>(P, grid, s);
        case 2: return launch_lds<256>(P, grid, s);
        case 3: return launch_lds<512>(P, grid, s);
        case 4: return launch_lds<1024>(P, grid, s);
        case 5: return launch_lds<2048>(P, grid, s);
        default:
            hipLaunchKernelGGL(search_kernel_global, dim3(grid), dim3(kWave), 0, s, P);
            return hipGetLastError();
    }
}

const void* class_kernel(int cls) {
    switch (cls) {
        case 0: return (const void*)search_kernel_lds<64>;
        case 1: return (const void*)search_kernel_lds<128>;
        case 2: return (const void*)search_kernel_lds<256>;
        case 3: return (const void*)search_kernel_lds<512>;
        case 4: return (const void*)search_kernel_lds<1024>;
        case 5: return (const void*)search_kernel_lds<2048>;
        default: return (const void*)search_kernel_global;
    }
}

// Zeroes the per-class work counters at the start of a run (a kernel rather
// than a memset, so that the run can be captured in a graph).
__global__ void reset_counters(unsigned int* ctr, int count) {
    int i = (int)threadIdx.x;
    if (i < count) ctr[i] = 0u;
}

}  // namespace

// Per-graph bound on splitter steps plus search iterations; a graph that
// exhausts it reports GCL_E_INTERNAL instead of running on.  GCL_STEP_BUDGET
// overrides the default.
uint64_t step_budget() {
    const char* e = getenv("GCL_STEP_BUDGET");
    if (e && *e) return strtoull(e, nullptr, 10);
    return 1ull << 40;
}

int32_t device_count() {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

// A batch resident on device 0: CSR arrays, per-class worker scratch, output
// buffers, streams and (optionally) a captured graph of one run.
struct DeviceBatch {
    struct Cls {
        int cls = 0;
        int32_t count = 0;
        int64_t grid = 0;
        size_t item_off = 0;
        KParams P;
        DevBuf scratch;
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
    };
    int32_t G = 0;
    int64_t total_n = 0, total_nnz = 0;
    uint32_t max_aut = 1, ring = 1;
    DevBuf vtx, row, col, color, init, label, out, items, ctr;
    std::vector<Cls> classes;
    hipStream_t main = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, start = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t workers = 0;

    ~DeviceBatch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& c : classes) {
            if (c.done) (void)hipEventDestroy(c.done);
            if (c.stream) (void)hipStreamDestroy(c.stream);
        }
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (start) (void)hipEventDestroy(start);
        if (main) (void)hipStreamDestroy(main);
    }

    gcl_status create(const gcl_batch& b, uint32_t max_aut_, uint32_t ring_cap) {
        if (max_aut_ == 0) return fail(GCL_E_INVALID, "max_aut must be >= 1");
        G = b.num_graphs;
        if (G < 0) return fail(GCL_E_INVALID, "negative graph count");
        if (device_count() <= 0) return fail(GCL_E_NODEVICE, "no HIP device visible");
        GCL_HIP(hipSetDevice(0));
        max_aut = max_aut_;
        ring = std::max(1u, std::min(max_aut, ring_cap));
        total_n = G > 0 ? b.vtx_off[G] : 0;
        total_nnz = G > 0 ? b.row_ptr[total_n] : 0;

        // size classes, larger graphs first within a class
        std::vector<std::vector<int32_t>> items_h(kClasses + 1);
        std::vector<uint32_t> cls_maxn(kClasses + 1, 0);
        std::vector<uint64_t> cls_maxnnz(kClasses + 1, 0);
        for (int32_t g = 0; g < G; ++g) {
            const int64_t n = b.vtx_off[g + 1] - b.vtx_off[g];
            if (n > kMaxVertices)
                return fail(GCL_E_CAPACITY, "graph %d has n=%lld > %d", g, (long long)n, kMaxVertices);
            int cls = kClasses;
            for (int c = 0; c < kClasses; ++c)
                if ((uint32_t)n <= kClassN[c]) {
                    cls = c;
                    break;
                }
            items_h[cls].push_back(g);
            cls_maxn[cls] = std::max(cls_maxn[cls], (uint32_t)n);
            const uint64_t nnz = (uint64_t)(b.row_ptr[b.vtx_off[g + 1]] - b.row_ptr[b.vtx_off[g]]);
            cls_maxnnz[cls] = std::max(cls_maxnnz[cls], nnz);
        }
        for (auto& v : items_h)
            std::stable_sort(v.begin(), v.end(), [&](int32_t x, int32_t y) {
                return b.vtx_off[x + 1] - b.vtx_off[x] > b.vtx_off[y + 1] - b.vtx_off[y];
            });
        const bool any_color = b.color != nullptr;
        std::vector<int32_t> init_order;
        if (any_color) {
            init_order.resize((size_t)total_n);
            for (int32_t g = 0; g < G; ++g) {
                const int64_t v0 = b.vtx_off[g], v1 = b.vtx_off[g + 1];
                for (int64_t v = v0; v < v1; ++v) init_order[(size_t)v] = (int32_t)(v - v0);
                std::stable_sort(init_order.begin() + v0, init_order.begin() + v1,
                                 [&](int32_t x, int32_t y) { return b.color[v0 + x] < b.color[v0 + y]; });
            }
        }
        hipDeviceProp_t prop;
        GCL_HIP(hipGetDeviceProperties(&prop, 0));
        GCL_HIP(hipMalloc(&vtx.p, sizeof(int64_t) * (size_t)(G + 1)));
        GCL_HIP(hipMalloc(&row.p, sizeof(int64_t) * (size_t)(total_n + 1)));
        GCL_HIP(hipMalloc(&col.p, sizeof(int32_t) * (size_t)std::max<int64_t>(total_nnz, 1)));
        GCL_HIP(hipMalloc(&label.p, sizeof(int32_t) * (size_t)std::max<int64_t>(total_n, 1)));
        GCL_HIP(hipMalloc(&out.p, sizeof(DeviceGraphResult) * (size_t)std::max(G, 1)));
        GCL_HIP(hipMalloc(&items.p, sizeof(int32_t) * (size_t)std::max(G, 1)));
        GCL_HIP(hipMalloc(&ctr.p, sizeof(unsigned int) * 64));
        GCL_HIP(hipMemcpy(vtx.p, b.vtx_off, sizeof(int64_t) * (size_t)(G + 1), hipMemcpyHostToDevice));
        GCL_HIP(hipMemcpy(row.p, b.row_ptr, sizeof(int64_t) * (size_t)(total_n + 1), hipMemcpyHostToDevice));
        if (total_nnz > 0)
            GCL_HIP(hipMemcpy(col.p, b.col, sizeof(int32_t) * (size_t)total_nnz, hipMemcpyHostToDevice));
        if (any_color) {
            GCL_HIP(hipMalloc(&color.p, sizeof(int32_t) * (size_t)std::max<int64_t>(total_n, 1)));
            GCL_HIP(hipMalloc(&init.p, sizeof(int32_t) * (size_t)std::max<int64_t>(total_n, 1)));
            GCL_HIP(hipMemcpy(color.p, b.color, sizeof(int32_t) * (size_t)total_n, hipMemcpyHostToDevice));
            GCL_HIP(hipMemcpy(init.p, init_order.data(), sizeof(int32_t) * (size_t)total_n, hipMemcpyHostToDevice));
        }
        std::vector<int32_t> flat;
        for (auto& v : items_h) flat.insert(flat.end(), v.begin(), v.end());
        if (!flat.empty())
            GCL_HIP(hipMemcpy(items.p, flat.data(), sizeof(int32_t) * flat.size(), hipMemcpyHostToDevice));
        GCL_HIP(hipStreamCreateWithFlags(&main, hipStreamNonBlocking));
        GCL_HIP(hipEventCreate(&ev0));
        GCL_HIP(hipEventCreate(&ev1));
        GCL_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
        size_t item_off = 0;
        for (int cls = 0; cls <= kClasses; ++cls) {
            const size_t cnt = items_h[cls].size();
            if (cnt == 0) continue;
            Cls C;
            C.cls = cls;
            C.count = (int32_t)cnt;
            C.item_off = item_off;
            const uint32_t maxn = cls < kClasses ? kClassN[cls] : cls_maxn[cls];
            const uint64_t max_nnz = cls_maxnnz[cls];
            // target-cell members over all levels of a path: n(n+1)/2 at worst,
            // capped for the large classes (GCL_E_CAPACITY past the cap)
            const uint64_t arena = std::min<uint64_t>((uint64_t)maxn * (maxn + 1) / 2, 64ull * maxn) + 64;
            const bool hot_global = cls == kClasses;
            const Layout L = make_layout(maxn, max_nnz, ring, arena, hot_global);
            int per_cu = 0;
            GCL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, class_kernel(cls), kWave, 0));
            if (per_cu <= 0) per_cu = 1;
            const int64_t resident = (int64_t)per_cu * prop.multiProcessorCount;
            size_t free_b = 0, total_b = 0;
            GCL_HIP(hipMemGetInfo(&free_b, &total_b));
            const int64_t mem_workers = (int64_t)((double)free_b * 0.5 / (double)L.total);
            int64_t grid = std::min<int64_t>({(int64_t)cnt, resident, std::max<int64_t>(mem_workers, 1)});
            C.grid = std::max<int64_t>(grid, 1);
            GCL_HIP(hipMalloc(&C.scratch.p, (size_t)(L.total * (uint64_t)C.grid)));
            GCL_HIP(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
            GCL_HIP(hipEventCreateWithFlags(&C.done, hipEventDisableTiming));
            KParams& P = C.P;
            P.num_items = (int32_t)cnt;
            P.items = (const int32_t*)items.p + item_off;
            P.vtx_off = (const int64_t*)vtx.p;
            P.row_ptr = (const int64_t*)row.p;
            P.col = (const int32_t*)col.p;
            P.init_order = any_color ? (const int32_t*)init.p : nullptr;
            P.color = any_color ? (const int32_t*)color.p : nullptr;
            P.out_label = (int32_t*)label.p;
            P.out = (DeviceGraphResult*)out.p;
            P.next_item = (unsigned int*)ctr.p + cls;
            P.scratch = (uint8_t*)C.scratch.p;
            P.scratch_stride = L.total;
            P.maxn = maxn;
            P.max_nnz = max_nnz;
            P.ring = ring;
            P.max_aut = max_aut;
            P.arena_cap = arena;
            P.step_budget = step_budget();
            workers += (int32_t)C.grid;
            item_off += cnt;
            classes.push_back(std::move(C));
        }
        return GCL_OK;
    }

    // reset counters, fork one stream per size class, join
    hipError_t enqueue() {
        hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, main, (unsigned int*)ctr.p, 64);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if ((e = hipEventRecord(start, main)) != hipSuccess) return e;
        for (auto& c : classes) {
            if ((e = hipStreamWaitEvent(c.stream, start, 0)) != hipSuccess) return e;
            if ((e = launch_class(c.cls, c.P, (int)c.grid, c.stream)) != hipSuccess) return e;
            if ((e = hipEventRecord(c.done, c.stream)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(main, c.done, 0)) != hipSuccess) return e;
        }
        return hipSuccess;
    }

    gcl_status run(bool use_graph, float* ms) {
        GCL_HIP(hipSetDevice(0));
        if (use_graph && !exec) {
            GCL_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeGlobal));
            hipError_t e = enqueue();
            hipGraph_t gr = nullptr;
            hipError_t e2 = hipStreamEndCapture(main, &gr);
            if (e != hipSuccess) return fail(GCL_E_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
            if (e2 != hipSuccess) return fail(GCL_E_DEVICE, "graph capture failed: %s", hipGetErrorString(e2));
            graph = gr;
            GCL_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        }
        GCL_HIP(hipEventRecord(ev0, main));
        if (use_graph) {
            GCL_HIP(hipGraphLaunch(exec, main));
        } else {
            hipError_t e = enqueue();
            if (e != hipSuccess) return fail(GCL_E_DEVICE, "kernel launch failed: %s", hipGetErrorString(e));
        }
        GCL_HIP(hipEventRecord(ev1, main));
        hipError_t se = hipStreamSynchronize(main);
        if (se != hipSuccess) return fail(GCL_E_DEVICE, "search kernel failed: %s", hipGetErrorString(se));
        float t = 0.f;
        GCL_HIP(hipEventElapsedTime(&t, ev0, ev1));
        if (ms) *ms = t;
        return GCL_OK;
    }

    gcl_status results(int32_t* out_label, DeviceGraphResult* res_out, gcl_batch_stats* st) {
        GCL_HIP(hipSetDevice(0));
        std::vector<DeviceGraphResult> res((size_t)std::max(G, 1));
        if (G > 0)
            GCL_HIP(hipMemcpy(res.data(), out.p, sizeof(DeviceGraphResult) * (size_t)G, hipMemcpyDeviceToHost));
        if (out_label && total_n > 0)
            GCL_HIP(hipMemcpy(out_label, label.p, sizeof(int32_t) * (size_t)total_n, hipMemcpyDeviceToHost));
        uint64_t nodes = 0, auts = 0, steps = 0, bytes = 0;
        gcl_status worst = GCL_OK;
        for (int32_t g = 0; g < G; ++g) {
            nodes += res[(size_t)g].nodes;
            auts += res[(size_t)g].automorphisms;
            steps += res[(size_t)g].splitter_steps;
            bytes += res[(size_t)g].csr_bytes;
            if (res[(size_t)g].status != 0 && worst == GCL_OK) {
                worst = (gcl_status)res[(size_t)g].status;
                set_error("graph %d: device search failed (%s)", g, gcl_status_string(worst));
            }
            if (res_out) res_out[g] = res[(size_t)g];
        }
        if (st) {
            st->nodes = nodes;
            st->automorphisms = auts;
            st->splitter_steps = steps;
            st->csr_bytes = bytes;
            st->workers = workers;
        }
        return worst;
    }
};

gcl_status device_batch_create(const gcl_batch& b, uint32_t max_aut, uint32_t ring_cap, DeviceBatch** out) {
    DeviceBatch* db = new (std::nothrow) DeviceBatch();
    if (!db) return fail(GCL_E_NOMEM, "out of host memory");
    gcl_status s = db->create(b, max_aut, ring_cap);
    if (s != GCL_OK) {
        delete db;
        return s;
    }
    *out = db;
    return GCL_OK;
}

gcl_status device_batch_run(DeviceBatch* db, bool use_graph, float* ms) { return db->run(use_graph, ms); }

gcl_status device_batch_results(DeviceBatch* db, int32_t* out_label, DeviceGraphResult* res, gcl_batch_stats* st) {
    return db->results(out_label, res, st);
}

void device_batch_destroy(DeviceBatch* db) { delete db; }

int32_t device_batch_size(const DeviceBatch* db) { return db->G; }

gcl_status gpu_canon_batch(const gcl_batch& b, uint32_t max_aut, uint32_t ring_cap, int32_t* out_label,
                           DeviceGraphResult* results, gcl_batch_stats* st) {
    auto t0 = std::chrono::steady_clock::now();
    DeviceBatch* db = nullptr;
    gcl_status s = device_batch_create(b, max_aut, ring_cap, &db);
    if (s != GCL_OK) return s;
    float ms = 0.f;
    s = db->run(false, &ms);
    gcl_batch_stats bs;
    memset(&bs, 0, sizeof(bs));
    if (s == GCL_OK) s = db->results(out_label, results, &bs);
    delete db;
    if (s != GCL_OK) return s;
    bs.kernel_ms = ms;
    bs.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (st) *st = bs;
    return GCL_OK;
}

}  // namespace gcl

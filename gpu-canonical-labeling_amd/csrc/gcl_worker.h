// gcl_worker.h -- device search worker of libgcl (included by gcl_device.hip).
//
// One wave64 runs the exact traversal of the reference's sequential engine
// (reference src/sequential.cpp:216-670) for one graph at a time, pulling
// graph ids from a launch-wide counter (persistent kernel).  All 64 lanes
// execute the search's control flow in lockstep; the data-parallel inner
// work of every step runs across the lanes:
//   * refinement (reference src/partition_refinement.cpp:150-485): neighbour
//     counts and per-cell ranks with LDS atomics, in-cell regrouping by hole /
//     outsider pairing, active-set and cell-mask updates by the owner lane of
//     each split cell, and the order-dependent hash chain over split cells by
//     lane ranks;
//   * backtracking: restore of a per-level snapshot of the cell-boundary
//     bitmap and a rebuild of the cell arrays (replaces reconstruct_at_level;
//     the order of vertices inside a cell is never observed);
//   * target cells: member list sorted by id through a vertex bitmap;
//   * orbit pruning: lock-free union-find (union by minimum);
//   * leaves: sort-free comparison of the current relabeled graph with the
//     stored first / best certificates (sorted canonical rows).
// Root speculation: a worker whose graph is large publishes the remaining
// root children on a board; idle workers refine them ahead of time and the
// owner adopts the results in the exact sequential order, so the traversal,
// node counts and canonical form are unchanged.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gcl_device.h"
#include "gcl_search.h"

namespace gcl {
namespace dev {

typedef uint16_t ix_t;
constexpr int kWave = 64;
constexpr uint32_t kMaxRecords = 4096;  // speculation records per launch (64 mask words)
constexpr uint32_t kNoKey = 0u;
constexpr uint32_t kFailedCode = 0xFFFFFFFFu;  // speculation item whose refinement failed
// Set in an item's code word when the helper left its partition out: the
// code alone rejects the child (see help_item).
constexpr uint32_t kCodeOnly = 0x40000000u;

struct LevelRec {
    uint32_t tc_first, tc_len, cursor, arena_off;
    int64_t last_num_aut;
};

__host__ __device__ inline uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

// ------------------------------------------------------------------ layout
// Per-worker global scratch.
struct Layout {
    uint64_t snap, seq, first_seq, best_seq, lv, arena_mem, arena_rep, firstcode, canon_code, first_label,
        best_label, first_keys, best_keys, first_row, best_row, subgroup, auts, hot16, hot32, hot64, lists, total;
    uint32_t levels;  // snapshot levels available
};

constexpr int kHot16 = 6;  // elem pos ce tmp(=sb) tv tcl
constexpr int kHot32 = 2;  // cnt hits
constexpr int kHot64 = 2;  // split-cell bitmap, old-end bitmap
// LDS classes from this size keep the touched-vertex and touched-cell lists
// of a general step in global scratch (the step keeps them in its lanes when
// it has at most a wave of (member, neighbour) pairs), which leaves LDS room
// for more workers per CU.
constexpr uint32_t kListsGlobalN = 512;
// the touched-vertex and touched-cell lists of one worker (maxn + 2 entries each)
__host__ __device__ inline uint64_t lists_bytes(uint32_t maxn) { return align256(2ull * 2ull * ((uint64_t)maxn + 2)); }
// element stride of the global hot arrays: even, so that the 16-bit arrays
// that double as packed counters start on a 32-bit word
__host__ __device__ inline uint64_t hot_stride(uint32_t maxn) { return ((uint64_t)maxn + 2 + 63) & ~uint64_t(63); }

// Automorphism ring of one graph: per slot a fix length (below), then the
// slots' images (ring x maxn entries).
__host__ __device__ inline uint64_t aut_fix_bytes(uint32_t ring) { return align256(4 * (uint64_t)ring); }
__host__ __device__ inline uint64_t aut_region_bytes(uint32_t ring, uint32_t maxn) {
    return aut_fix_bytes(ring) + align256(2 * (uint64_t)ring * maxn);
}

// pooled_auts: the automorphism rings live in a pool, one per graph of the
// launch, instead of in each worker's scratch
__host__ __device__ inline Layout make_layout(uint32_t maxn, uint64_t max_nnz, uint32_t ring, uint64_t arena,
                                              uint32_t levels, bool hot_global, bool pooled_auts = false) {
    Layout L;
    uint64_t o = 0;
    const uint64_t n1 = (uint64_t)maxn + 2;
    const uint64_t nw1 = (uint64_t)maxn / 64 + 2;
    L.levels = levels;
    L.snap = o; o = align256(o + 8 * nw1 * (uint64_t)levels);
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
    L.auts = o; if (!pooled_auts) o += aut_region_bytes(ring, maxn);
    L.hot16 = o; if (hot_global) o = align256(o + 2ull * kHot16 * hot_stride(maxn));
    L.hot32 = o; if (hot_global) o = align256(o + 4ull * kHot32 * (hot_stride(maxn) / 2 + 1));
    L.hot64 = o; if (hot_global) o = align256(o + 8ull * kHot64 * nw1);
    L.lists = o; if (!hot_global && maxn >= kListsGlobalN) o = align256(o + lists_bytes(maxn));
    L.total = o;
    return L;
}

// Speculation record: one per board slot.
struct SpecHeader {
    uint32_t epoch;     // publication number (item states carry it)
    uint32_t graph;     // graph id of the publication
    uint32_t count;     // items published
    uint32_t next;      // helper claim counter
    uint32_t tc_first;  // first position of the root target cell
    uint32_t n;
    int32_t first1;     // the first path's code at level 2
    int32_t canon1;     // the best path's code at level 2 (only grows; the owner refreshes it)
};

// A record has a fixed part, sized by the launch's largest n and its item
// capacity `slots` (header, root partition, per-item vertex, state, code and
// counters), and a payload of refined partitions, laid out by the n of the
// graph that publishes (its element vector, then its cell-boundary bitmap, per
// item).  A record of a given size thus holds the root children of a graph in
// proportion to 1/n: every root child of a 10,000-vertex graph fits in the
// records of a class whose largest graph has 50,000 vertices.
struct SpecLayout {
    uint64_t header, root_elem, root_lasts, vert, state, code, ncells, steps, payload;
    uint32_t slots;
};

__host__ __device__ inline SpecLayout make_spec_layout(uint32_t maxn, uint32_t slots) {
    SpecLayout L;
    const uint64_t n1 = (uint64_t)maxn + 2;
    const uint64_t nw1 = (uint64_t)maxn / 64 + 2;
    uint64_t o = 0;
    L.slots = slots;
    L.header = o; o = align256(o + sizeof(SpecHeader));
    L.root_elem = o; o = align256(o + 2 * n1);
    L.root_lasts = o; o = align256(o + 8 * nw1);
    L.vert = o; o = align256(o + 2 * (uint64_t)slots);
    L.state = o; o = align256(o + 4 * (uint64_t)slots);
    L.code = o; o = align256(o + 4 * (uint64_t)slots);
    L.ncells = o; o = align256(o + 4 * (uint64_t)slots);
    // per item: splitter steps, CSR bytes, partition bytes (and one spare word)
    L.steps = o; o = align256(o + 32 * (uint64_t)slots);
    L.payload = o;
    return L;
}
__host__ __device__ inline uint64_t align128(uint64_t x) { return (x + 127) & ~uint64_t(127); }
// payload bytes of one item of a graph with n vertices, and the offset of its
// bitmap inside the item
__host__ __device__ inline uint64_t spec_item_lasts(uint32_t n) { return align128(2 * ((uint64_t)n + 2)); }
__host__ __device__ inline uint64_t spec_item_bytes(uint32_t n) {
    return spec_item_lasts(n) + align128(8 * ((uint64_t)n / 64 + 2));
}
// items of a graph with n vertices that a record of `stride` bytes holds
__host__ __device__ inline uint32_t spec_item_cap(const SpecLayout& L, uint64_t stride, uint32_t n) {
    const uint64_t room = stride > L.payload ? (stride - L.payload) / spec_item_bytes(n) : 0;
    return (uint32_t)(room < L.slots ? room : L.slots);
}
// the record size with metadata for `slots` items that holds the partitions
// of `items` of them for a graph with maxn vertices
__host__ __device__ inline uint64_t spec_record_bytes(uint32_t maxn, uint32_t slots, uint32_t items) {
    return align256(make_spec_layout(maxn, slots).payload + spec_item_bytes(maxn) * (uint64_t)items);
}

// Descent speculation (single large graphs).  After its first leaf, a worker
// publishes a descent record: for each level L >= 2 of the first path, the
// members of the level's target cell after the first-path vertex (up to
// kDescPerLevel of them), deepest level first.  A helper that claims item
// (L, v) rebuilds the first path's partition at level L from the record,
// individualizes v and follows the leftmost path below it (the smallest
// member of each level's target cell, which is always its orbit's
// representative), recording per level what the owner's search would
// compute there; the owner replays such a descent when its exact sequential
// traversal reaches (L, v) on the first path.
constexpr uint32_t kDescPerLevel = 8;    // items per first-path level
constexpr uint32_t kDescMembers = 16;    // target cells larger than this end a recorded descent
struct DescHeader {
    uint32_t epoch, graph, count, next, n, levels;  // levels: first-path snapshots copied
    uint32_t pad[2];
};
// One recorded level of a descent.
struct DescLevel {
    uint32_t vert, code, ncells;  // vertex individualized, refinement code, cells after it
    uint32_t tc_first, tc_len;    // this level's target cell (tc_len 0: none recorded)
    uint32_t pad;
    uint64_t steps, csr, part;    // splitter steps and algorithmic bytes of the refinement
    uint16_t mem[kDescMembers];   // target cell members in increasing id
};
struct DescLayout {
    uint64_t header, elem, snaps, item_level, item_vert, state, nrec, level_first, level_count, payload;
    uint64_t item_bytes, lev_bytes, lasts_off, elem_off;
    uint32_t items, depth, path;
};
// maxn: the launch's largest n; items, depth: item capacity and levels
// recorded per item; path: first-path levels whose snapshots the record holds
__host__ __device__ inline DescLayout make_desc_layout(uint32_t maxn, uint32_t items, uint32_t depth, uint32_t path) {
    DescLayout L;
    const uint64_t n1 = (uint64_t)maxn + 2;
    const uint64_t nw1 = (uint64_t)maxn / 64 + 2;
    uint64_t o = 0;
    L.items = items;
    L.depth = depth;
    L.path = path;
    L.header = o; o = align256(o + sizeof(DescHeader));
    L.elem = o; o = align256(o + 2 * n1);
    L.snaps = o; o = align256(o + 8 * nw1 * ((uint64_t)path + 1));
    L.item_level = o; o = align256(o + 2 * (uint64_t)items);
    L.item_vert = o; o = align256(o + 2 * (uint64_t)items);
    L.state = o; o = align256(o + 4 * (uint64_t)items);
    L.nrec = o; o = align256(o + 4 * (uint64_t)items);
    L.level_first = o; o = align256(o + 4 * ((uint64_t)path + 1));
    L.level_count = o; o = align256(o + 4 * ((uint64_t)path + 1));
    L.payload = o;
    // per item: `depth` DescLevel records, then `depth` bitmaps, then the
    // element vector of the last recorded level
    L.lev_bytes = sizeof(DescLevel);
    L.lasts_off = align256(L.lev_bytes * depth);
    L.elem_off = L.lasts_off + align256(8 * nw1 * depth);
    L.item_bytes = L.elem_off + align256(2 * n1);
    return L;
}
__host__ __device__ inline uint64_t desc_record_bytes(const DescLayout& L) {
    return L.payload + L.item_bytes * (uint64_t)L.items;
}
// nrec word of an item: levels recorded, and these flags
constexpr uint32_t kDescLeaf = 1u << 30;    // the last recorded level is a leaf
constexpr uint32_t kDescFailed = 1u << 31;  // the helper's refinement failed

// Serving mode: requests published by the host while the kernel runs.  The
// host side (fine-grained host memory, read with system-scope loads) holds
// the queue and each request slot's graph; the claiming worker copies the
// graph into the device arrays of its slot (vtx_off / row_ptr / col / color /
// init_order of KParams, where graph 2s is slot s) and writes the labels, the
// result record and the slot's done flag back to host memory.
struct ServeParams {
    const uint32_t* tail;    // requests published (host)
    const uint32_t* stop;    // 1: no more requests will be published (host)
    const uint32_t* ring;    // slot ids in publication order, ring_mask + 1 entries (host)
    uint32_t ring_mask;
    uint32_t slots;
    const int32_t* h_n;      // per slot (host)
    const int32_t* h_row;    // per slot: maxn + 1 local row offsets (host)
    const int32_t* h_col;    // per slot: max_nnz neighbour ids (host)
    const int32_t* h_color;  // per slot: maxn colours (host)
    const int32_t* h_init;   // per slot: vertices in (colour, id) order (host)
    const int32_t* h_colored;  // per slot: 1 when the graph has colours (host)
    // completion queue (host): entry c = ((c + 1) << 32) | slot, written once
    // the slot's labels and result record are in host memory
    unsigned long long* cq;
    uint32_t cq_mask;        // entries - 1 (at least the slot count)
    unsigned int* cq_count;  // completions so far (device memory)
    // device-memory mirror of the host queue: idle workers poll these; one
    // worker at a time (the one that moves poll_at forward) reads the host
    uint32_t* tail_seen;     // largest host tail read so far
    uint32_t* stop_seen;     // 1 once the host's stop flag was read
    unsigned long long* poll_at;  // wall-clock tick before which nobody reads the host
    uint32_t poll_ticks;     // wall-clock ticks between host reads
};

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
    // per-worker touched lists (kListsGlobalN classes) outside the workers'
    // scratch, lists_stride bytes each; NULL: in each worker's scratch
    uint8_t* lists;
    uint64_t lists_stride;
    uint32_t maxn;
    uint64_t max_nnz;
    uint32_t ring;
    uint32_t max_aut;
    uint64_t arena_cap;
    uint32_t levels;
    uint64_t step_budget;
    // root speculation
    uint8_t* spec;              // spec_records records
    uint64_t spec_stride;
    uint32_t spec_slots;        // root children per record
    uint32_t spec_records;      // <= kMaxRecords
    uint32_t spec_min_n;        // 0 disables speculation
    unsigned long long* slot_free;   // bit r set while record r is free
    unsigned long long* board_mask;  // bit r set while record r is published
    unsigned int* pending;      // running graphs that may still publish
    uint64_t time_limit;        // wall-clock ticks a worker runs at most (its loop ends)
    // a search (and every wait inside it) gives up with GCL_E_TIMEOUT after
    // graph_ticks_base + graph_ticks_per_n2 * n^2 ticks, or at the worker's
    // end, whichever comes first
    uint64_t graph_ticks_base, graph_ticks_per_n2;
    DiagRecord* diag;           // per-worker records (grid entries), or NULL
    uint8_t* aut_pool;          // automorphism rings, one per item (aut_stride bytes), or NULL
    uint64_t aut_stride;
    // descent speculation: records, per-record layout sizes, free set and
    // board (as for the root records); dspec == nullptr disables it
    uint8_t* dspec;
    uint64_t dspec_stride;
    uint32_t dspec_records, dspec_items, dspec_depth, dspec_path;
    unsigned long long* dslot_free;
    unsigned long long* dboard;
    ServeParams sv;             // serving mode when sv.tail != nullptr
};

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return (int)threadIdx.x; }
// ordering of LDS traffic inside the one wave of a workgroup
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// ordering of global-memory traffic between lanes
__device__ __forceinline__ void group_sync() { __syncthreads(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t readlane(uint32_t v, int i) { return (uint32_t)__builtin_amdgcn_readlane(v, i); }
// Wave-wide scans with DPP lane moves (no LDS round trips): shifts by 1, 2,
// 4 and 8 inside each row of 16 lanes, then row sums carried into the later
// rows with row_bcast:15 (rows 1, 3) and row_bcast:31 (rows 2, 3).  Every
// lane of the wave must be active.
#define GCL_DPP(x, ctrl, rows) ((uint32_t)__builtin_amdgcn_update_dpp(0, (int)(x), ctrl, rows, 0xF, false))
__device__ __forceinline__ uint32_t wave_scan(uint32_t x) {  // inclusive sum
    x += GCL_DPP(x, 0x111, 0xF);
    x += GCL_DPP(x, 0x112, 0xF);
    x += GCL_DPP(x, 0x114, 0xF);
    x += GCL_DPP(x, 0x118, 0xF);
    x += GCL_DPP(x, 0x142, 0xA);
    x += GCL_DPP(x, 0x143, 0xC);
    return x;
}
__device__ __forceinline__ uint32_t wave_scan_max(uint32_t x) {  // inclusive max
    x = max(x, GCL_DPP(x, 0x111, 0xF));
    x = max(x, GCL_DPP(x, 0x112, 0xF));
    x = max(x, GCL_DPP(x, 0x114, 0xF));
    x = max(x, GCL_DPP(x, 0x118, 0xF));
    x = max(x, GCL_DPP(x, 0x142, 0xA));
    x = max(x, GCL_DPP(x, 0x143, 0xC));
    return x;
}
// reductions: the result is wave-uniform (a scalar register)
// A copy of x in a vector register: arithmetic on it stays on the vector
// unit even though x is the same in every lane.
__device__ __forceinline__ uint32_t to_vgpr(uint32_t x) {
    uint32_t v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
    return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) { return readlane(wave_scan(x), kWave - 1); }
__device__ __forceinline__ uint32_t wave_max(uint32_t x) { return readlane(wave_scan_max(x), kWave - 1); }
__device__ __forceinline__ uint32_t wave_min(uint32_t x) { return ~wave_max(~x); }
// Values that are the same in every lane (loaded from LDS, from the worker
// object, or reduced over the wave) are moved to scalar registers, so that
// the arithmetic and branches on them are scalar instructions.
// (the builtin returns a signed int: each half is taken as uint32_t before
// it is widened, so that the low half is not sign-extended)
__device__ __forceinline__ uint32_t uni(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int32_t uni(int32_t x) { return (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)x); }
__device__ __forceinline__ uint64_t uni(uint64_t x) {
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)x);
    return ((uint64_t)hi << 32) | (uint64_t)lo;
}
__device__ __forceinline__ int64_t uni(int64_t x) { return (int64_t)uni((uint64_t)x); }
__device__ __forceinline__ bool uni(bool x) { return __builtin_amdgcn_readfirstlane((uint32_t)x) != 0u; }
template <class T>
__device__ __forceinline__ T* uni(T* ptr) {
    return (T*)uni((uint64_t)ptr);
}
// polling load: no cache invalidation per poll; pair with agent_acquire()
// once the awaited value has been seen
__device__ __forceinline__ uint32_t agent_peek(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void agent_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void agent_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool agent_cas(uint32_t* p, uint32_t expect, uint32_t desired) {
    return __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                                __HIP_MEMORY_SCOPE_AGENT);
}

// ----------------------------------------------------------- hot storage

// Load through a global-memory pointer (the pointers kept in the LDS worker
// object lose their address space, which would make every load a flat one).
template <class T>
__device__ __forceinline__ T gload(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}

// 16-bit counter arrays stored as pairs in 32-bit words: loads and stores are
// 16-bit, increments and compare-and-swap go through the containing word
// (values stay below 2^16 since every count is at most n <= 65535).  Loads
// and stores are plain: wave_sync orders them against the other lanes.
#define GCL_PACKED16(NAME, BASE)                                                                  \
    __device__ __forceinline__ uint32_t NAME##_ld(uint32_t i) {                                   \
        return ((const uint16_t*)(BASE))[i];                                                      \
    }                                                                                             \
    __device__ __forceinline__ void NAME##_st(uint32_t i, uint32_t v) {                           \
        ((uint16_t*)(BASE))[i] = (uint16_t)v;                                                     \
    }                                                                                             \
    __device__ __forceinline__ uint32_t NAME##_add(uint32_t i) {                                  \
        const uint32_t sh = (i & 1u) * 16u;                                                       \
        return (atomicAdd(&(BASE)[i >> 1], 1u << sh) >> sh) & 0xFFFFu;                            \
    }                                                                                             \
    __device__ __forceinline__ bool NAME##_cas(uint32_t i, uint32_t expect, uint32_t desired) {   \
        const uint32_t sh = (i & 1u) * 16u;                                                       \
        uint32_t* w = &(BASE)[i >> 1];                                                            \
        uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);      \
        for (int k = 0; k < 64; ++k) {                                                            \
            if (((old >> sh) & 0xFFFFu) != expect) return false;                                  \
            const uint32_t nw_ = (old & ~(0xFFFFu << sh)) | (desired << sh);                      \
            const uint32_t got = atomicCAS(w, old, nw_);                                          \
            if (got == old) return true;                                                          \
            old = got;                                                                            \
        }                                                                                         \
        return false;                                                                             \
    }

// Row offsets are kept in LDS below 1024 vertices; the larger classes read
// them from global memory, which leaves room for more workers per CU.
template <int MAXN>
struct LdsHot {
    static constexpr int NW = MAXN / 64 + 2;
    static constexpr int NROW = MAXN < 1024 ? MAXN + 1 : 1;
    static constexpr bool kListG = (uint32_t)MAXN >= kListsGlobalN;
    static constexpr int NLW = kListG ? 1 : MAXN / 2 + 1;
    // tmp is also sb: the staging area of a large-cell regroup and the list
    // of non-singleton cells are used one after the other; in a singleton
    // step it holds each touched cell's hole list and, at the cell's last
    // position, its hole / outsider counters (Worker::singleton_step)
    // ce: the cell array (cell_end / cell_first merged, see Worker::cend)
    ix_t elem[MAXN], pos[MAXN], ce[MAXN], tmp[MAXN];
    // 16-bit counters packed in pairs (atomics on the containing word)
    uint32_t cnt[MAXN / 2 + 1], hits[MAXN / 2 + 1];
    // touched-vertex / touched-cell lists of a general step (in global
    // scratch from kListsGlobalN vertices: kListG)
    uint32_t tvw[NLW], tclw[NLW];
    uint32_t rowl[NROW];
    // split / oldend: per general step, the split cells (by first position)
    // and their old last positions, all zero between steps; split is also
    // the vertex bitmap of a target cell's members (vmask), used between
    // refinements only and cleared after each use
    uint64_t active[NW], nsmask[NW], lasts[NW];
    uint64_t split[NW], oldend[NW];
    uint32_t ctr[4];
    const int64_t* rowp;
    int64_t rbase;  // rowp[0]
    ix_t* tvg;      // the lists in global scratch (kListG)
    ix_t* tclg;
};

// Hot state in LDS: addresses are compile-time constants.
template <int MAXN>
struct HotL {
    static constexpr bool kLds = true;
    static constexpr bool kDesc = false;
    static constexpr bool kRowLds = MAXN < 1024;
    __device__ __forceinline__ static LdsHot<MAXN>& S() {
        __shared__ LdsHot<MAXN> s;
        return s;
    }
    __device__ __forceinline__ ix_t& elem(uint32_t i) { return S().elem[i]; }
    __device__ __forceinline__ ix_t& pos(uint32_t i) { return S().pos[i]; }
    __device__ __forceinline__ ix_t& ce(uint32_t i) { return S().ce[i]; }
    __device__ __forceinline__ ix_t& tmp(uint32_t i) { return S().tmp[i]; }
    __device__ __forceinline__ ix_t& sb(uint32_t i) { return S().tmp[i]; }
    static constexpr bool kListG = LdsHot<MAXN>::kListG;
    __device__ __forceinline__ ix_t& tv(uint32_t i) {
        if constexpr (kListG) return S().tvg[i];
        else return ((ix_t*)S().tvw)[i];
    }
    __device__ __forceinline__ ix_t& tcl(uint32_t i) {
        if constexpr (kListG) return S().tclg[i];
        else return ((ix_t*)S().tclw)[i];
    }
    GCL_PACKED16(cnt, S().cnt)
    GCL_PACKED16(hits, S().hits)
    // adds inc to the 16-bit entry tmp[i] through its 32-bit word (the
    // other half is untouched: the entry never carries); returns the old entry
    __device__ __forceinline__ uint32_t tmp_add(uint32_t i, uint32_t inc) {
        const uint32_t sh = (i & 1u) * 16u;
        return (atomicAdd((uint32_t*)S().tmp + (i >> 1), inc << sh) >> sh) & 0xFFFFu;
    }
    __device__ __forceinline__ uint64_t* split() { return S().split; }
    __device__ __forceinline__ uint64_t* oldend() { return S().oldend; }
    __device__ __forceinline__ uint64_t* active() { return S().active; }
    __device__ __forceinline__ uint64_t* nsmask() { return S().nsmask; }
    __device__ __forceinline__ uint64_t* vmask() { return S().split; }
    __device__ __forceinline__ uint64_t* lasts() { return S().lasts; }
    __device__ __forceinline__ uint32_t* ctr() { return S().ctr; }
    __device__ __forceinline__ uint32_t row(uint32_t v) {
        if (kRowLds) return S().rowl[v];
        return (uint32_t)(gload(S().rowp + v) - S().rbase);
    }
    __device__ __forceinline__ const int64_t* rows() { return S().rowp; }
    __device__ __forceinline__ void sync() { wave_sync(); }
    // order list writes before other lanes' reads (global lists: the
    // stores must have completed)
    __device__ __forceinline__ void lsync() {
        if constexpr (kListG) group_sync();
        else wave_sync();
    }
    __device__ __forceinline__ void bind_lists(uint8_t* p, uint32_t maxn) {
        if constexpr (kListG) {
            S().tvg = (ix_t*)p;
            S().tclg = (ix_t*)p + maxn + 2;
        }
    }
    __device__ __forceinline__ void bind(const int64_t* rowp, uint32_t n, uint8_t*, uint32_t) {
        if (kRowLds) {
            for (uint32_t i = (uint32_t)lane_id(); i <= n; i += kWave) S().rowl[i] = (uint32_t)(rowp[i] - rowp[0]);
        } else {
            S().rowp = rowp;
            S().rbase = rowp[0];
        }
    }
};

struct LdsBits {
    uint64_t active[1026], nsmask[1026], vmask[1026], lasts[1026];
    uint32_t ctr[4];
};

// Hot state in per-worker global scratch (graphs above the LDS classes);
// the bitmaps stay in LDS.
struct HotG {
    static constexpr bool kLds = false;
    static constexpr bool kDesc = false;
    ix_t* h16;
    uint32_t* h32;
    uint64_t* h64;
    uint64_t n1, nw1;  // array stride (hot_stride), bitmap words
    const int64_t* rowp;
    __device__ __forceinline__ static LdsBits& B() {
        __shared__ LdsBits b;
        return b;
    }
    __device__ __forceinline__ ix_t& elem(uint32_t i) { return h16[i]; }
    __device__ __forceinline__ ix_t& pos(uint32_t i) { return h16[n1 + i]; }
    __device__ __forceinline__ ix_t& ce(uint32_t i) { return h16[2 * n1 + i]; }
    __device__ __forceinline__ ix_t& tmp(uint32_t i) { return h16[3 * n1 + i]; }
    __device__ __forceinline__ ix_t& sb(uint32_t i) { return h16[3 * n1 + i]; }
    __device__ __forceinline__ ix_t& tv(uint32_t i) { return h16[4 * n1 + i]; }
    __device__ __forceinline__ ix_t& tcl(uint32_t i) { return h16[5 * n1 + i]; }
    GCL_PACKED16(cnt, h32)
    GCL_PACKED16(hits, h32 + (n1 / 2 + 1))
    // tmp_add as in HotL (n1 is a multiple of 64, so tmp is word aligned)
    __device__ __forceinline__ uint32_t tmp_add(uint32_t i, uint32_t inc) {
        const uint32_t sh = (i & 1u) * 16u;
        return (atomicAdd((uint32_t*)(h16 + 3 * n1) + (i >> 1), inc << sh) >> sh) & 0xFFFFu;
    }
    __device__ __forceinline__ void lsync() { group_sync(); }
    __device__ __forceinline__ void bind_lists(uint8_t*, uint32_t) {}
    __device__ __forceinline__ uint64_t* split() { return h64; }
    __device__ __forceinline__ uint64_t* oldend() { return h64 + nw1; }
    __device__ __forceinline__ uint64_t* active() { return B().active; }
    __device__ __forceinline__ uint64_t* nsmask() { return B().nsmask; }
    __device__ __forceinline__ uint64_t* vmask() { return B().vmask; }
    __device__ __forceinline__ uint64_t* lasts() { return B().lasts; }
    __device__ __forceinline__ uint32_t* ctr() { return B().ctr; }
    static constexpr bool kRowLds = false;
    __device__ __forceinline__ uint32_t row(uint32_t v) { return (uint32_t)(gload(rowp + v) - gload(rowp)); }
    __device__ __forceinline__ const int64_t* rows() { return rowp; }
    __device__ __forceinline__ void sync() { group_sync(); }
    __device__ __forceinline__ void bind(const int64_t* rp, uint32_t, uint8_t* base, uint32_t maxn) {
        rowp = rp;
        n1 = hot_stride(maxn);
        nw1 = (uint64_t)maxn / 64 + 2;
        (void)base;
    }
};

// A hot-state type with descent speculation compiled in.  Only the kernels
// of the classes where single large graphs run carry it, so that the
// kernels of the batch's many small graphs keep their register budget.
template <class Base>
struct WithDesc : Base {
    static constexpr bool kDesc = true;
};

// Register copies of the worker fields that every refinement step touches
// (the worker object itself lives in LDS).
struct Ctx {
    uint32_t n, nw, ncells, nactive;
    uint64_t steps, csr_bytes;
    // this refinement's CSR words read by singleton steps, 3 + deg(w) per
    // step (the splitter's id and row offsets, its neighbour ids): below
    // 2n * 67 < 2^32, folded into csr_bytes at its end
    uint32_t swords;
    // algorithmic bytes beyond the CSR reads: 8 per non-singleton cell hashed
    // under a non-singleton splitter, 8 per vertex of a cell that splits, 8n
    // per refinement (the element vector loaded and stored once)
    uint64_t part_bytes;
    // per lane: sizes of the split cells this lane regrouped or owned since
    // the last flush into part_bytes (one vector add per split instead of a
    // lane read; at most 2^14 steps between flushes keeps it below 2^32)
    uint32_t lane_split;
    const int32_t* col;
    const int64_t* rp;  // global row offsets and rp[0] (rows outside LDS)
    int64_t rb;
};

// ------------------------------------------------------------------ worker
template <class Hot>
struct Worker {
    Hot H;
    // graph
    uint32_t n, nw;
    uint64_t nnz;
    const int64_t* rowp;  // global row offsets of the graph (n+1)
    const int32_t* col;   // the graph's neighbour ids (local)
    uint32_t ncells, nactive;
    // per-worker global scratch
    uint64_t* snap;
    ix_t *seq, *first_seq, *best_seq, *arena_mem, *arena_rep, *first_label, *best_label, *auts;
    uint32_t* aut_fix;  // per ring slot: how much of seq its automorphism fixes
    uint32_t fix_low;   // lowest seq_size since aut_fix was last truncated
    LevelRec* lv;
    int32_t *firstcode, *canon_code;
    uint32_t *first_keys, *best_keys, *first_row, *best_row, *subgroup;
    // the launch parameters (constant over the launch): the ring size,
    // snapshot levels, arena capacity, step budget and the speculation
    // records' layouts are read from them where used rather than kept here
    const KParams* kp;
    __device__ __forceinline__ uint32_t ring() const { return gload(&kp->ring); }
    __device__ __forceinline__ uint32_t levels() const { return gload(&kp->levels); }
    __device__ __forceinline__ uint64_t arena_cap() const { return gload(&kp->arena_cap); }
    __device__ __forceinline__ uint64_t step_budget() const { return gload(&kp->step_budget); }
    __device__ __forceinline__ SpecLayout sl() const {
        return make_spec_layout(gload(&kp->maxn), gload(&kp->spec_slots));
    }
    __device__ __forceinline__ DescLayout dly() const {
        return make_desc_layout(gload(&kp->maxn), gload(&kp->dspec_items), gload(&kp->dspec_depth),
                                gload(&kp->dspec_path));
    }
    uint64_t life;      // wall-clock tick at which this worker's loop ends
    uint64_t deadline;  // tick at which the current search gives up (<= life)
    // search state (reference src/sequential.h:24-58)
    uint32_t level, seq_size, tc_level_reached, eqlevfirst, first_seq_size, best_seq_size;
    int eqlevcanon, comp_canon, canon_level, code;
    int64_t num_aut;
    uint64_t nodes, steps, csr_bytes, part_bytes, budget;
    int32_t status;
    // diagnostics: totals of this worker over the launch (DiagRecord)
    uint64_t dg_nodes, dg_steps, dg_helped, dg_search, dg_wait, dg_help, dg_idle;
    // phase times of the own search (refine, classify, orbits, backtracking
    // past an exhausted cell) and root children adopted from helpers or
    // rejected by their code; kept only with diagnostics on
    uint64_t dg_refine, dg_classify, dg_orbits, dg_backtrack;
    uint32_t dg_adopted, dg_rejected;
    __device__ __forceinline__ uint64_t dg_t0() const { return dg_on ? wall_clock64() : 0ull; }
    __device__ __forceinline__ void dg_add(uint64_t& acc, uint64_t t0) {
        if (dg_on) acc += wall_clock64() - t0;
    }
    uint32_t dg_graphs;
    int32_t spec_slot;       // board slot published by this worker, or -1
    // descent record published by this worker for its current best path
    // (-1: none), and the record it retired when a new best leaf was found
    // (-1: none), whose claimed items may still be in progress
    int32_t d_slot, d_old;
    uint32_t d_epoch, d_old_epoch;
    uint8_t* d_rec;
    uint32_t spec_epoch, spec_cursor, spec_count;
    uint8_t* spec_rec;
    // flags (together, to keep the object small: it lives in LDS)
    bool first_discovered;
    bool root_changed;   // the partition held at level 1 is not the root's (own helping)
    bool dg_on;
    bool d_renew;        // a new best leaf: publish its path's descents
    bool spec_pending;   // counted in *pending until it publishes or ends

    // ---------------------------------------------------------- cell array
    // One 16-bit entry per position, ce[p]: the last position of p's cell
    // when p is not that last position, and the cell's first position at its
    // last position (a cell is named by its last position).  Every entry of
    // a cell other than its last is above p, the last one at most p, so
    //   cell_end(p)   = max(ce[p], p)
    //   cell_first(c) = ce[c]          (c a last position)
    // A fragment [a, b] is written as ce[p] = (p == b ? a : b).
    __device__ __forceinline__ uint32_t cend(uint32_t p) { return max((uint32_t)H.ce(p), p); }
    __device__ __forceinline__ uint32_t cfirst(uint32_t c) { return H.ce(c); }
    __device__ __forceinline__ static uint32_t ce_val(uint32_t p, uint32_t a, uint32_t b) { return p == b ? a : b; }

    // --------------------------------------------------------- bit helpers
    __device__ __forceinline__ static bool test_bit(const uint64_t* m, uint32_t p) {
        return (m[p >> 6] >> (p & 63)) & 1ull;
    }
    // 32-bit views of a bitmap (bit p in word p >> 5): one bit tested or
    // updated without 64-bit shifts.  or_bits32 ors in b << (p & 31),
    // and_bits32 keeps bit p only when b is 1, xor_bits32 toggles it when b
    // is 1 (b is 0 or 1).
    __device__ __forceinline__ static bool test_bit32(const uint64_t* m, uint32_t p) {
        return (((const uint32_t*)m)[p >> 5] >> (p & 31)) & 1u;
    }
    __device__ __forceinline__ static void or_bits32(uint64_t* m, uint32_t p, uint32_t b) {
        atomicOr((uint32_t*)m + (p >> 5), b << (p & 31));
    }
    __device__ __forceinline__ static void and_bits32(uint64_t* m, uint32_t p, uint32_t b) {
        atomicAnd((uint32_t*)m + (p >> 5), ~((1u - b) << (p & 31)));
    }
    __device__ __forceinline__ static void xor_bits32(uint64_t* m, uint32_t p, uint32_t b) {
        atomicXor((uint32_t*)m + (p >> 5), b << (p & 31));
    }
    __device__ __forceinline__ static void set_bit_atomic(uint64_t* m, uint32_t p) {
        atomicOr((unsigned long long*)&m[p >> 6], (unsigned long long)(1ull << (p & 63)));
    }
    __device__ __forceinline__ static void clear_bit_atomic(uint64_t* m, uint32_t p) {
        atomicAnd((unsigned long long*)&m[p >> 6], ~(unsigned long long)(1ull << (p & 63)));
    }
    // uniform single-bit update by lane 0
    __device__ __forceinline__ void set_bit_u(uint64_t* m, uint32_t p, bool on) {
        if (lane_id() == 0) {
            if (on) m[p >> 6] |= 1ull << (p & 63);
            else m[p >> 6] &= ~(1ull << (p & 63));
        }
    }
    // row offset of v (local to the graph's col) inside a refinement
    __device__ __forceinline__ uint32_t xrow(const Ctx& X, uint32_t v) {
        if constexpr (Hot::kRowLds) return H.row(v);
        else return (uint32_t)(gload(X.rp + v) - X.rb);
    }
    // p is wave-uniform
    __device__ __forceinline__ void add_active(Ctx& X, uint32_t p) {
        if (p < X.n && !uni(test_bit(H.active(), p))) {
            set_bit_u(H.active(), p, true);
            ++X.nactive;
        }
    }
    __device__ __forceinline__ void remove_active(Ctx& X, uint32_t p) {
        if (p < X.n && uni(test_bit(H.active(), p))) {
            set_bit_u(H.active(), p, false);
            --X.nactive;
        }
    }
    // first set bit at position >= start, or -1
    // First set bit at or after `start`, or -1; *wordp (if given) receives
    // the whole bitmap word that holds it.
    __device__ __forceinline__ static int32_t first_set_from(const uint64_t* m, uint32_t start, uint32_t nw,
                                                             uint64_t* wordp = nullptr) {
        const int l = lane_id();
        for (uint32_t base = start >> 6; base < nw; base += kWave) {
            const uint32_t w = base + (uint32_t)l;
            const uint64_t raw = (w < nw) ? m[w] : 0ull;
            uint64_t bits = raw;
            if (w == (start >> 6)) bits &= ~0ull << (start & 63);
            const uint64_t b = ballot(bits != 0ull);
            if (b) {
                const int fl = __builtin_ctzll(b);
                const uint32_t lo = readlane((uint32_t)bits, fl);
                const uint32_t hi = readlane((uint32_t)(bits >> 32), fl);
                const uint64_t word = ((uint64_t)hi << 32) | lo;
                if (wordp)
                    *wordp = ((uint64_t)readlane((uint32_t)(raw >> 32), fl) << 32) | readlane((uint32_t)raw, fl);
                return uni((int32_t)((base + (uint32_t)fl) * 64u + (uint32_t)__builtin_ctzll(word)));
            }
        }
        return -1;
    }

    // ------------------------------------------------------- cell rebuild
    // Rebuilds cend / cfirst / nsmask / ncells from the cell-boundary bitmap
    // `lasts` (bit p: p is the last position of a cell).
    __device__ void rebuild_from_lasts() {
        const int l = lane_id();
        uint64_t* lasts = H.lasts();
        uint32_t carry = 0xFFFFu;
        uint32_t total = 0;
        const uint32_t nchunks = (nw + kWave - 1) / kWave;
        for (int32_t ch = (int32_t)nchunks - 1; ch >= 0; --ch) {
            const uint32_t w = (uint32_t)ch * kWave + (uint32_t)l;
            const uint64_t bits = (w < nw) ? lasts[w] : 0ull;
            total += wave_sum((uint32_t)__popcll(bits));
            uint32_t v = bits ? w : 0xFFFFu;
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t y = __shfl_down(v, o, kWave);
                if (l + o < kWave) v = min(v, y);
            }
            v = min(v, carry);
            if (w < nw) H.tmp(w) = (ix_t)v;  // next nonzero word at or after w
            carry = __shfl(v, 0, kWave);
        }
        H.sync();
        for (uint32_t base = 0; base < n; base += kWave) {
            const uint32_t p = base + (uint32_t)l;
            if (p < n) {
                const uint32_t w = p >> 6;
                const uint64_t bits = lasts[w] >> (p & 63);
                uint32_t last;
                if (bits) {
                    last = p + (uint32_t)__builtin_ctzll(bits);
                } else {
                    const uint32_t w2 = H.tmp(w + 1);  // the last position is always a cell last
                    last = w2 * 64u + (uint32_t)__builtin_ctzll(lasts[w2]);
                }
                // the first position of a cell writes its last entry
                if (p != last) H.ce(p) = (ix_t)last;
                if (p == 0 || test_bit(lasts, p - 1)) H.ce(last) = (ix_t)p;
            }
        }
        for (uint32_t base = 0; base < nw; base += kWave) {
            const uint32_t w = base + (uint32_t)l;
            if (w < nw) {
                const uint64_t ls = lasts[w];
                const uint64_t prev = (w == 0) ? 1ull : (lasts[w - 1] >> 63);
                const uint64_t firsts = (ls << 1) | prev;  // p first <=> p-1 last, or p == 0
                const uint64_t valid = (w * 64u + 64u <= n) ? ~0ull : ((1ull << (n - w * 64u)) - 1ull);
                H.nsmask()[w] = (firsts & ~ls) & valid;
            }
        }
        ncells = total;
        H.sync();
    }

    __device__ void positions_from_elem() {
        for (uint32_t p = (uint32_t)lane_id(); p < n; p += kWave) H.pos(H.elem(p)) = (ix_t)p;
        H.sync();
    }

    __device__ void snapshot(uint32_t lev) {
        // cell boundaries of the level-`lev` node, for backtracking
        if (lev >= levels()) {
            status = GCL_E_CAPACITY;
            return;
        }
        uint64_t* dst = snap + (uint64_t)lev * (nw + 1);
        for (uint32_t w = (uint32_t)lane_id(); w < nw; w += kWave) dst[w] = H.lasts()[w];
    }

    // --------------------------------------------------------- partitions
    __device__ void init_partition(const int32_t* init_order, const int32_t* color, int64_t vbase) {
        // initializePartition_s (reference src/partition_refinement.cpp:28-68):
        // vertices in (colour, id) order, one cell per colour run.
        const int l = lane_id();
        for (uint32_t base = 0; base < n; base += kWave) {
            const uint32_t p = base + (uint32_t)l;
            bool is_last = false;
            if (p < n) {
                const uint32_t v = init_order ? (uint32_t)init_order[vbase + p] : p;
                H.elem(p) = (ix_t)v;
                if (p + 1 == n) is_last = true;
                else if (color) {
                    const uint32_t v2 = init_order ? (uint32_t)init_order[vbase + p + 1] : p + 1;
                    is_last = color[vbase + v] != color[vbase + v2];
                }
                H.cnt_st(p, 0);
                H.hits_st(p, 0);
            }
            const uint64_t b = ballot(is_last);
            if (l == 0) H.lasts()[base >> 6] = b;
        }
        for (uint32_t w = (uint32_t)l; w < nw; w += kWave) {
            H.active()[w] = 0;
            H.vmask()[w] = 0;
            H.split()[w] = 0;
            H.oldend()[w] = 0;
        }
        nactive = 0;
        H.sync();
        positions_from_elem();
        rebuild_from_lasts();
    }

    // Refinement step, splitter {w} with deg(w) <= 64 (reference
    // src/partition_refinement.cpp:253-468, splitter.length == 1): cells
    // holding neighbours split into (neighbours, others).
    // u: this lane's neighbour of the splitter (lanes below deg), already
    // loaded by the caller
    __device__ __forceinline__ void singleton_step(Ctx& X, uint32_t u, uint32_t deg, uint32_t& longcode, int& hint) {
        const int l = lane_id();
        X.swords += 3 + deg;
        uint32_t c = 0, f = 0, r = 0, q = 0;
        bool ns = false;
        if ((uint32_t)l < deg) {
            // the neighbour mark (read only at positions of non-singleton
            // cells, so a neighbour in a singleton cell may carry it too)
            H.cnt_st(u, 1);
            q = H.pos(u);
            // ce[q] == q exactly when q's cell is the singleton [q, q]
            const uint32_t v = H.ce(q);
            ns = v != q;
            c = max(v, q);
            if (ns) {
                r = H.hits_add(c);
                f = cfirst(c);
                // the first toucher clears the cell's hole / outsider
                // counters: tmp[c], outside the cell's hole list [f, f+k)
                // once it splits
                if (r == 0) H.tmp(c) = 0;
            }
        }
        H.sync();
        uint32_t k = 0, L = 0;
        bool split = false, out = false, hole = false;
        if (ns) {
            k = H.hits_ld(c);
            L = c - f + 1;
            split = k < L;
            out = split && q >= f + k;
            if (split) hole = H.cnt_ld(H.elem(f + r)) == 0;
        }
        // one atomic per lane numbers its hole (low byte of tmp[c]) and its
        // outsider (high byte); a cell has at most 64 of each
        uint32_t o = 0;
        if (hole || out) {
            const uint32_t old = H.tmp_add(c, (hole ? 1u : 0u) | (out ? 0x100u : 0u));
            if (hole) H.tmp(f + (old & 0xFFu)) = (ix_t)(f + r);
            o = old >> 8;
        }
        H.sync();
        if (out) {  // move u into the front block over a non-neighbour
            const uint32_t hp = H.tmp(f + o);
            const uint32_t x = H.elem(hp);
            H.elem(hp) = (ix_t)u;
            H.pos(u) = (ix_t)hp;
            H.elem(q) = (ix_t)x;
            H.pos(x) = (ix_t)q;
        }
        H.sync();
        const bool owner = ns && r == 0;
        const bool own = owner && split;
        if (split) H.ce(f + r) = (ix_t)ce_val(f + r, f, f + k - 1);  // the front fragment [f, f+k-1]
        uint32_t hkey = kNoKey;
        if (owner) H.hits_st(c, 0);
        if (own) {
            H.ce(c) = (ix_t)(f + k);  // the back fragment [f+k, c]
            // active-set rule (414-435): one new active cell per split, the
            // back fragment when the cell was active or the front one is not
            // smaller, else the front one; a singleton one is the hint
            // candidate
            const bool back = test_bit32(H.active(), f) || k >= L - k;
            const uint32_t a = back ? f + k : f;
            or_bits32(H.active(), a, 1u);
            if ((back ? L - k : k) == 1) hkey = ((f + 1) << 16) | a;
            // non-singleton marks: f's stays unless the front fragment is a
            // singleton; f+k's is set unless the back one is
            and_bits32(H.nsmask(), f, k == 1 ? 0u : 1u);
            or_bits32(H.nsmask(), f + k, L - k > 1 ? 1u : 0u);
            or_bits32(H.lasts(), f + k - 1, 1u);
        }
        // counts back to zero (lanes without a neighbour hold u = 0, which
        // is cleared as well: every count ends the step at zero)
        H.cnt_st(u, 0);
        const uint64_t om = ballot(own);
        const uint32_t nsplit = (uint32_t)__popcll(om);
        if (nsplit == 0) {
            H.sync();
            return;
        }
        X.ncells += nsplit;
        X.nactive += nsplit;
        // boundary hash chain in position order (399-409): x = f+k-1 grows
        // with f; the hint is the candidate of the last split that has one
        const uint32_t x = f + k - 1;
        if (own) X.lane_split += L;
        if (nsplit == 1) {
            const int j = __builtin_ctzll(om);
            longcode = GCL_MASH(longcode, readlane(x, j));
            const uint32_t hj = readlane(hkey, j);
            if (hj != kNoKey) hint = (int)(hj & 0xFFFFu);
            H.sync();
            return;
        }
        uint32_t rank = 0;
        for (uint64_t m = om; m; m &= m - 1) {
            const uint32_t xj = readlane(x, __builtin_ctzll(m));
            if (own && xj < x) ++rank;
        }
        for (uint32_t t = 0; t < nsplit; ++t) {
            const int j = __builtin_ctzll(ballot(own && rank == t));
            longcode = GCL_MASH(longcode, readlane(x, j));
            const uint32_t hj = readlane(hkey, j);
            if (hj != kNoKey) hint = (int)(hj & 0xFFFFu);
        }
        H.sync();
    }

    // Regroups one touched cell [f, c] of at most kSmallCell members by its
    // key, entirely in this lane (sorting network in registers).  Returns the
    // number of fragments; a cell that splits is marked for the hash pass
    // (bit f of split, bit c of oldend).
    static constexpr uint32_t kSmallCell = 8;
    template <uint32_t N>  // network size: a power of two >= the cell size
    __device__ __forceinline__ uint32_t regroup_small(uint32_t f, uint32_t c, uint32_t s) {
        const uint32_t L = c - f + 1;
        // every slot loads (slots past the cell repeat its last member and
        // take the sentinel key): no divergent branches, and the loads of
        // all slots are in flight together
        uint32_t kk[N];
#pragma unroll
        for (uint32_t t = 0; t < N; ++t) kk[t] = H.elem(f + min(t, L - 1));
#pragma unroll
        for (uint32_t t = 0; t < N; ++t) {
            const uint32_t cv = H.cnt_ld(kk[t]);
            const uint32_t kv = (s == 1) ? 1u - cv : cv;
            kk[t] = t < L ? ((kv << 16) | kk[t]) : 0xFFFFFFFFu;
        }
        // bitonic network; each compare-exchange is a min and a max
#pragma unroll
        for (uint32_t k = 2; k <= N; k <<= 1) {
#pragma unroll
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
                for (uint32_t i = 0; i < N; ++i) {
                    const uint32_t ixj = i ^ j;
                    if (ixj > i) {
                        const uint32_t lo = min(kk[i], kk[ixj]), hi = max(kk[i], kk[ixj]);
                        const bool up = (i & k) == 0;
                        kk[i] = up ? lo : hi;
                        kk[ixj] = up ? hi : lo;
                    }
                }
            }
        }
        // each slot's fragment: its first (fs) and last (ls) slot (a key
        // never equals the sentinel's, so fragments end at the cell's end)
        uint32_t fs[N], ls[N];
        fs[0] = 0;
#pragma unroll
        for (uint32_t t = 1; t < N; ++t) fs[t] = (kk[t] >> 16) == (kk[t - 1] >> 16) ? fs[t - 1] : t;
        ls[N - 1] = N - 1;
#pragma unroll
        for (int t = (int)N - 2; t >= 0; --t) ls[t] = (kk[t] >> 16) == (kk[t + 1] >> 16) ? ls[t + 1] : (uint32_t)t;
        uint32_t nfrag = 0;
#pragma unroll
        for (uint32_t t = 0; t < N; ++t) nfrag += (t < L && fs[t] == t) ? 1u : 0u;
        if (nfrag == 1) return 1;  // one key: nothing moves
#pragma unroll
        for (uint32_t t = 0; t < N; ++t) {
            if (t < L) {
                const uint32_t v = kk[t] & 0xFFFFu, p = f + t;
                H.elem(p) = (ix_t)v;
                H.pos(v) = (ix_t)p;
                H.ce(p) = (ix_t)(f + (ls[t] == t ? fs[t] : ls[t]));
                or_bits32(H.lasts(), p, (ls[t] == t && p != c) ? 1u : 0u);
            }
        }
        or_bits32(H.split(), f, 1u);
        or_bits32(H.oldend(), c, 1u);
        return nfrag;
    }

    // Split of a large touched cell [f, c] under a splitter of more than one
    // vertex when few of its members are touched (at most a wave, at most a
    // third of the cell: the outsiders' positions and ids are staged in
    // tmp[f + tc, f + 3tc)).  Fragments go in increasing count, so the
    // untouched members (count 0) stay a block at the front and the touched
    // ones form the tail [B, c], grouped by count: touched members outside
    // the tail swap with untouched ones inside it (ranks from ballots), the
    // tail is ordered by count, and the front block only has its cell end
    // rewritten.  The fragments, their order and the cell arrays end as the
    // whole-cell counting sort leaves them (the order inside a fragment is
    // unobservable).  Returns false, having changed nothing but scratch in
    // tmp[f, c], when the cell does not qualify.
    // This front end finds the cell's touched members in the touched-vertex
    // list of the step (ntv entries).
    __device__ bool sparse_split_list(Ctx& X, uint32_t f, uint32_t c, uint32_t L, uint32_t ntv) {
        const int l = lane_id();
        // touched members of this cell, listed in tmp[f, f + tc)
        uint32_t tc = 0;
        for (uint32_t base = 0; base < ntv; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            bool in = false;
            uint32_t u = 0;
            if (i < ntv) {
                u = H.tv(i);
                in = cend(H.pos(u)) == c;
            }
            const uint64_t m = ballot(in);
            const uint32_t at = tc + lanes_below(m);
            if (in && at < L) H.tmp(f + at) = (ix_t)u;
            tc += (uint32_t)__popcll(m);
        }
        tc = uni(tc);
        if (tc == 0 || tc > (uint32_t)kWave || 3 * tc > L) return false;
        H.sync();
        const bool act = (uint32_t)l < tc;
        const uint32_t u = act ? (uint32_t)H.tmp(f + (uint32_t)l) : 0u;
        sparse_split_core(X, f, c, L, tc, act, u);
        return true;
    }
    // The same when every touched vertex of the step is held by a lane: the
    // lanes with `in` hold the cell's touched members u.
    __device__ bool sparse_split_lanes(Ctx& X, uint32_t f, uint32_t c, uint32_t L, bool in, uint32_t u) {
        const uint32_t tc = (uint32_t)__popcll(ballot(in));
        if (tc == 0 || 3 * tc > L) return false;
        sparse_split_core(X, f, c, L, tc, in, u);
        return true;
    }
    // tc touched members, one in each lane with `member` (u: the member)
    __device__ void sparse_split_core(Ctx& X, uint32_t f, uint32_t c, uint32_t L, uint32_t tc, bool member, uint32_t u) {
        const int l = lane_id();
        const uint32_t B = c - tc + 1;  // first position of the touched block
        const bool act = (uint32_t)l < tc;
        uint32_t q = 0, hv = 0;
        bool out = false, hole = false;
        if (member) {
            q = H.pos(u);
            out = q < B;
        }
        if (act) {
            hv = H.elem(B + (uint32_t)l);
            hole = H.cnt_ld(hv) == 0u;
        }
        const uint64_t om = ballot(out), hm = ballot(hole);
        // outsider of rank r swaps with the hole of rank r (as many of each)
        if (out) {
            const uint32_t r = lanes_below(om);
            H.tmp(f + tc + r) = (ix_t)q;
            H.tmp(f + 2 * tc + r) = (ix_t)u;
        }
        H.sync();
        if (hole) {
            const uint32_t r = lanes_below(hm);
            const uint32_t qo = H.tmp(f + tc + r), uo = H.tmp(f + 2 * tc + r);
            H.elem(B + (uint32_t)l) = (ix_t)uo;
            H.pos(uo) = (ix_t)(B + (uint32_t)l);
            H.elem(qo) = (ix_t)hv;
            H.pos(hv) = (ix_t)qo;
        }
        H.sync();
        // order the tail by count: one group per count value present
        uint32_t v = 0, kv = 0;
        if (act) {
            v = H.elem(B + (uint32_t)l);
            kv = H.cnt_ld(v);
        }
        const uint32_t kmin = uni(wave_min(act ? kv : 0xFFFFFFFFu));
        const uint32_t kmax = uni(wave_max(act ? kv : 0u));
        uint32_t newpos = 0, gstart = 0, gend = 0, run = 0, ngroups = 0;
        for (uint32_t val = kmin; val <= kmax; ++val) {
            const uint64_t m = ballot(act && kv == val);
            if (!m) continue;
            const uint32_t g = (uint32_t)__popcll(m);
            if (act && kv == val) {
                newpos = B + run + lanes_below(m);
                gstart = B + run;
                gend = B + run + g - 1;
            }
            run += g;
            ++ngroups;
        }
        H.sync();
        if (act) {
            H.elem(newpos) = (ix_t)v;
            H.pos(v) = (ix_t)newpos;
            H.ce(newpos) = (ix_t)ce_val(newpos, gstart, gend);
            if (newpos == gstart && gend != c) set_bit_atomic(H.lasts(), gend);
        }
        // the untouched block [f, B - 1]
        for (uint32_t pp = f + (uint32_t)l; pp < B; pp += kWave) H.ce(pp) = (ix_t)ce_val(pp, f, B - 1);
        if (l == 0) {
            set_bit_atomic(H.lasts(), B - 1);
            set_bit_atomic(H.split(), f);
            set_bit_atomic(H.oldend(), c);
        }
        X.ncells += ngroups;
        H.sync();
    }

    // Hash chain and active-set rules of a step with a splitter of more than
    // one vertex (reference src/partition_refinement.cpp:253-468), with the
    // pre-step non-singleton cells listed in sb[0, nns) in position order.
    // The chain's values are, in position order, one per cell that did not
    // split (its common count plus its first position) and one per fragment
    // of a cell that did (the fragment's count plus its first position).
    // Each lane takes one cell: it walks its own fragments, writes its values
    // at their place in the chain (hits, all zero here, as 16-bit entries)
    // and applies the cell's active-set rules, which touch only the cell's
    // positions, and updates the fragments' non-singleton marks and clears
    // the cell's split marks; the chain itself then runs over the gathered
    // values.  The hint ends as the last singleton fragment after the first,
    // in position order.  LDS classes only (values below 2^16).
    __device__ void hash_pass_parallel(Ctx& X, uint32_t nns, uint32_t& longcode, int& hint) {
        const int l = lane_id();
        uint32_t nval = 0;
        int32_t dact = 0;
        for (uint32_t base = 0; base < nns; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            const bool valid = i < nns;
            uint32_t f = 0, nv = 0;
            bool sp = false;
            if (valid) {
                f = H.sb(i);
                sp = test_bit32(H.split(), f);
                nv = 1;
                if (sp) {
                    // its fragments: up to the old last position (oldend)
                    nv = 0;
                    for (uint32_t q = f;;) {
                        const uint32_t last = cend(q);
                        ++nv;
                        if (test_bit32(H.oldend(), last) || last + 1 >= X.n) break;
                        q = last + 1;
                    }
                }
            }
            const uint32_t inc = wave_scan(nv);
            uint32_t k = nval + inc - nv;
            int32_t cand = -1;
            if (valid && !sp) H.hits_st(k, H.cnt_ld(H.elem(f)) + f);
            if (sp) {
                // Every fragment after the first becomes active: its first
                // position was inside the cell, so never active before, and
                // each adds one active cell.  A cell that was not active
                // becomes active at f and loses its largest fragment, which
                // leaves the count as it is (reference 414-460).
                const bool was_active = test_bit32(H.active(), f);
                uint32_t max_size = 0, max_first = f, q = f, last = f;
                for (uint32_t t = 0; t < nv; ++t) {
                    last = cend(q);
                    const uint32_t size = last - q + 1;
                    H.hits_st(k++, H.cnt_ld(H.elem(q)) + q);
                    if (size > max_size) {
                        max_size = size;
                        max_first = q;
                    }
                    const bool later = q != f;
                    or_bits32(H.active(), q, later ? 1u : 0u);
                    if (later && last == q) cand = (int32_t)q;
                    // the fragment's non-singleton mark (set before at f
                    // only): toggled where it changes
                    xor_bits32(H.nsmask(), q, (last > q) != !later ? 1u : 0u);
                    q = last + 1;
                }
                and_bits32(H.oldend(), last, 0u);
                and_bits32(H.split(), f, 0u);
                dact += (int32_t)nv - 1;
                if (!was_active) {
                    or_bits32(H.active(), f, 1u);
                    and_bits32(H.active(), max_first, 0u);
                }
            }
            // the last candidate in position order wins
            const uint64_t cm = ballot(cand >= 0);
            if (cm) hint = (int)readlane((uint32_t)cand, 63 - __builtin_clzll(cm));
            nval += readlane(inc, kWave - 1);
        }
        X.nactive += (uint32_t)uni((int32_t)wave_sum((uint32_t)dact));
        H.sync();
        // The MASH chain over hits[0, nval) in order.  Each MASH's mask is
        // deferred to the end (the low 15 bits of (a ^ C) + x depend only on
        // the low 15 bits of a and x), and the chain runs in a vector
        // register, equal in all lanes: one lane read and one xor-add per
        // value, with no round trip through the scalar unit.
        uint32_t lc = to_vgpr(longcode);
        const uint32_t kx = to_vgpr(065435u);
        for (uint32_t base = 0; base < nval; base += kWave) {
            const uint32_t x = base + (uint32_t)l < nval ? H.hits_ld(base + (uint32_t)l) : 0u;
            const uint32_t here = min((uint32_t)kWave, nval - base);
            // whole groups of 8 lanes with constant lane indices (one branch
            // per group), then the last here % 8 values
            const uint32_t full = here >> 3;
#pragma unroll
            for (uint32_t g = 0; g < 8; ++g) {
                if (g < full) {
#pragma unroll
                    for (uint32_t u = 0; u < 8; ++u) lc = (lc ^ kx) + readlane(x, (int)(8 * g + u));
                }
            }
            for (uint32_t t = full * 8; t < here; ++t) lc = (lc ^ kx) + readlane(x, (int)t);
        }
        longcode = uni(lc) & 077777u;
        H.sync();
        for (uint32_t b = (uint32_t)l; b < nval; b += kWave) H.hits_st(b, 0);
    }

    // Counting-sort split of a large touched cell [f, c] by its members'
    // keys (count for |W| > 1; neighbours first for |W| == 1); a cell whose
    // keys are all equal stays as it is.  hits is all zero on entry and on
    // exit: it holds the buckets (indices 0..R-1).
    __device__ void split_big_cell(Ctx& X, uint32_t f, uint32_t c, uint32_t s) {
        const int l = lane_id();
        const uint32_t L = c - f + 1;
        uint32_t kmin = 0xFFFFFFFFu, kmax = 0;
        for (uint32_t base = 0; base < L; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            if (i < L) {
                const uint32_t v = H.elem(f + i);
                const uint32_t kv = (s == 1) ? 1u - H.cnt_ld(v) : H.cnt_ld(v);
                kmin = min(kmin, kv);
                kmax = max(kmax, kv);
            }
        }
        kmin = uni(wave_min(kmin));
        kmax = uni(wave_max(kmax));
        if (kmin == kmax) return;  // uniform: nothing moves
        X.part_bytes += 8ull * L;
        const uint32_t R = kmax - kmin + 1;
        for (uint32_t base = 0; base < L; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            if (i < L) {
                const uint32_t v = H.elem(f + i);
                const uint32_t kv = (s == 1) ? 1u - H.cnt_ld(v) : H.cnt_ld(v);
                H.hits_add(kv - kmin);
            }
        }
        H.sync();
        // bucket sizes -> exclusive offsets, in place
        uint32_t run = 0, nfrag = 0;
        for (uint32_t base = 0; base < R; base += kWave) {
            const uint32_t b = base + (uint32_t)l;
            const uint32_t sz = b < R ? H.hits_ld(b) : 0u;
            const uint32_t inc = wave_scan(sz);
            if (b < R) H.hits_st(b, run + inc - sz);
            run += readlane(inc, kWave - 1);
            nfrag += (uint32_t)__popcll(ballot(sz != 0));
        }
        H.sync();
        // each member takes the next slot of its bucket (the order inside
        // a fragment is unobservable); offsets end as bucket ends
        for (uint32_t base = 0; base < L; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            if (i < L) {
                const uint32_t v = H.elem(f + i);
                const uint32_t kv = (s == 1) ? 1u - H.cnt_ld(v) : H.cnt_ld(v);
                H.tmp(f + H.hits_add(kv - kmin)) = (ix_t)v;
            }
        }
        H.sync();
        for (uint32_t base = 0; base < L; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            if (i < L) {
                const uint32_t P = f + i;
                const uint32_t v = H.tmp(P);
                H.elem(P) = (ix_t)v;
                H.pos(v) = (ix_t)P;
                const uint32_t kv = (s == 1) ? 1u - H.cnt_ld(v) : H.cnt_ld(v);
                const uint32_t b = kv - kmin;
                const uint32_t first = f + (b == 0 ? 0u : H.hits_ld(b - 1));
                const uint32_t last = f + H.hits_ld(b) - 1;
                H.ce(P) = (ix_t)ce_val(P, first, last);
                if (P == first && last != c) set_bit_atomic(H.lasts(), last);
            }
        }
        X.ncells += nfrag - 1;
        H.sync();
        for (uint32_t b = (uint32_t)l; b < R; b += kWave) H.hits_st(b, 0);
        if (l == 0) {
            set_bit_atomic(H.split(), f);
            set_bit_atomic(H.oldend(), c);
        }
        H.sync();
    }

    // The end of a general step, once every touched cell is regrouped: the
    // hash chain and active-set rules over the pre-step non-singleton cells
    // in position order, and the cell masks of the new fragments.
    __device__ void general_tail(Ctx& X, uint32_t s, uint32_t& longcode, int& hint) {
        const int l = lane_id();
        // compact list of the pre-step non-singleton cells (position order),
        // a lane per 32-bit half of an nsmask word (fewer bits per lane than
        // whole words, and 32-bit bit scans)
        uint32_t nns = 0;
        const uint32_t nh = 2 * X.nw;
        const uint32_t* half = (const uint32_t*)H.nsmask();
        for (uint32_t base = 0; base < nh; base += kWave) {
            const uint32_t w = base + (uint32_t)l;
            uint32_t bits = w < nh ? half[w] : 0u;
            const uint32_t c0 = (uint32_t)__popc(bits);
            const uint32_t inc = wave_scan(c0);
            uint32_t j = nns + inc - c0;
            while (bits) {
                H.sb(j++) = (ix_t)(w * 32u + (uint32_t)__builtin_ctz(bits));
                bits &= bits - 1;
            }
            nns += readlane(inc, kWave - 1);
        }
        H.sync();
        // hash chain / active rules in position order (reference 253-468)
        if (s > 1) {
            X.part_bytes += 8ull * nns;  // the non-singleton cell headers hashed
            hash_pass_parallel(X, nns, longcode, hint);
        } else {
            // single-vertex splitter: cells that did not split add nothing to
            // the chain; each split cell, in position order, split in two
            // (neighbours first)
            for (uint32_t base = 0; base < nns; base += kWave) {
                const uint32_t i = base + (uint32_t)l;
                uint32_t f = 0;
                bool sp = false;
                if (i < nns) {
                    f = H.sb(i);
                    sp = test_bit(H.split(), f);
                }
                for (uint64_t m = ballot(sp); m; m &= m - 1) {
                    const uint32_t fc = readlane(f, __builtin_ctzll(m));
                    const uint32_t K = uni(cend(fc)) - fc + 1;
                    const uint32_t LL = uni(cend(fc + K)) - fc + 1;
                    longcode = GCL_MASH(longcode, fc + K - 1);
                    if (uni(test_bit(H.active(), fc)) || K >= LL - K) {
                        add_active(X, fc + K);
                        if (LL - K == 1) hint = (int)(fc + K);
                    } else {
                        add_active(X, fc);
                        if (K == 1) hint = (int)fc;
                    }
                }
            }
        }
        H.sync();
        // cell masks of the new fragments; clear the split marks.  The
        // marks are keyed by old first positions: walk the ns list again
        // (done by the parallel hash pass when it ran)
        if (s == 1)
        for (uint32_t base = 0; base < nns; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            if (i < nns) {
                const uint32_t f = H.sb(i);
                if (test_bit(H.split(), f)) {
                    for (uint32_t q = f;;) {
                        const uint32_t last = cend(q);
                        if (last > q) set_bit_atomic(H.nsmask(), q);
                        else clear_bit_atomic(H.nsmask(), q);
                        if (test_bit(H.oldend(), last)) {
                            clear_bit_atomic(H.oldend(), last);
                            break;
                        }
                        if (last + 1 >= X.n) break;
                        q = last + 1;
                    }
                    clear_bit_atomic(H.split(), f);
                }
            }
        }
        H.sync();
    }

    // General refinement step: splitter cell [p, e] of any size (reference
    // src/partition_refinement.cpp:197-473): counts with LDS atomics; touched
    // cells regrouped by key (count for |W| > 1; neighbours first for
    // |W| == 1), small cells one per lane, large ones by the whole wave; then
    // the hash chain and active-set rules over the pre-step non-singleton
    // cells in position order.
    __device__ void general_step(Ctx& X, uint32_t p, uint32_t e, uint32_t& longcode, int& hint) {
        const int l = lane_id();
        const uint32_t s = e - p + 1;
        if (s <= (uint32_t)kWave) {
            uint32_t a = 0, d = 0;
            if ((uint32_t)l < s) {
                const uint32_t w = H.elem(p + (uint32_t)l);
                a = xrow(X, w);
                d = xrow(X, w + 1) - a;
            }
            const uint32_t incl = wave_scan(d);
            const uint32_t total = readlane(incl, kWave - 1);
            if (total <= (uint32_t)kWave) {
                general_step_lanes(X, s, a, d, incl, total, longcode, hint);
                return;
            }
        }
        general_step_list(X, p, s, longcode, hint);
    }

    // General step with at most a wave of splitter members and of (member,
    // neighbour) pairs (a, d, incl: each member lane's row offset, degree and
    // inclusive degree sum; total: the pairs).  A touched vertex stays in
    // the lane that counted it first, and the first of those lanes in each
    // touched non-singleton cell owns the cell: it regroups a small cell
    // itself and names a large one to the wave.  No touched lists are kept.
    __device__ void general_step_lanes(Ctx& X, uint32_t s, uint32_t a, uint32_t d, uint32_t incl, uint32_t total,
                                       uint32_t& longcode, int& hint) {
        const int l = lane_id();
        // one (member, neighbour) pair per lane: lane t takes the first member
        // whose degree run [incl - d, incl) ends after t
        const uint32_t off = a + d - incl;  // col index of a pair = t + off of its member
        uint32_t x = 0;
        bool found = false;
        for (uint32_t j = 0; j < s; ++j) {
            const uint32_t ij = readlane(incl, (int)j);
            const uint32_t oj = readlane(off, (int)j);
            if (!found && (uint32_t)l < ij) {
                found = true;
                x = (uint32_t)l + oj;
            }
        }
        bool first = false;
        uint32_t u = 0;
        if ((uint32_t)l < total) {
            u = (uint32_t)gload(X.col + x);
            first = H.cnt_add(u) == 0u;
        }
        X.csr_bytes += 12ull * s + 4ull * total;
        H.sync();
        // the counts are final: the cell of each touched vertex, and its owner
        uint32_t c = 0, f = 0;
        bool own = false;
        if (first) {
            c = cend(H.pos(u));
            f = cfirst(c);
            if (f != c) own = H.hits_add(c) == 0u;
        }
        H.sync();
        bool small = false, big = false;
        if (own) {
            H.hits_st(c, 0);
            small = c - f + 1 <= kSmallCell;
            big = !small;
        }
        // the 4-wide network when no small cell has more than 4 members
        uint32_t nf = 1;
        if (ballot(small && c - f + 1 > 4)) {
            if (small) nf = regroup_small<8>(f, c, s);
        } else {
            if (small) nf = regroup_small<4>(f, c, s);
        }
        if (nf > 1) X.lane_split += c - f + 1;
        X.ncells += uni(wave_sum(nf - 1));
        H.sync();
        // hits is all zero from here on (the buckets of split_big_cell)
        for (uint64_t bm = ballot(big); bm; bm &= bm - 1) {
            const int j = __builtin_ctzll(bm);
            const uint32_t cb = uni(readlane(c, j)), fb = uni(readlane(f, j));
            const uint32_t L = cb - fb + 1;
            if (s > 1 && sparse_split_lanes(X, fb, cb, L, first && c == cb, u)) {
                X.part_bytes += 8ull * L;
                continue;
            }
            split_big_cell(X, fb, cb, s);
        }
        general_tail(X, s, longcode, hint);
        if (first) H.cnt_st(u, 0);  // counts back to zero
        H.sync();
    }

    // General step for any splitter: the touched vertices and touched cells
    // are listed in tv and tcl.
    __device__ void general_step_list(Ctx& X, uint32_t p, uint32_t s, uint32_t& longcode, int& hint) {
        const int l = lane_id();
        // counts: cnt[u] = neighbours of u in the splitter; the first touch of
        // a vertex appends it to the tv list (ranks from ballots)
        uint32_t ntv = 0;
        uint64_t degsum = 0;
        for (uint32_t base = 0; base < s; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            uint32_t a = 0, d = 0;
            if (i < s) {
                const uint32_t w = H.elem(p + i);
                a = xrow(X, w);
                d = xrow(X, w + 1) - a;
            }
            const uint32_t incl = wave_scan(d);
            const uint32_t total = readlane(incl, kWave - 1);
            degsum += total;
            const uint32_t nv = min(s - base, (uint32_t)kWave);
            if (total <= (uint32_t)kWave) {
                // one (vertex, neighbour) pair per lane: lane t takes the
                // first vertex whose degree run [incl - d, incl) ends after t
                const uint32_t off = a + d - incl;  // col index of a pair = t + off of its vertex
                uint32_t x = 0;
                bool found = false;
                for (uint32_t j = 0; j < nv; ++j) {
                    const uint32_t ij = readlane(incl, (int)j);
                    const uint32_t oj = readlane(off, (int)j);
                    if (!found && (uint32_t)l < ij) {
                        found = true;
                        x = (uint32_t)l + oj;
                    }
                }
                bool first = false;
                uint32_t u = 0;
                if ((uint32_t)l < total) {
                    u = (uint32_t)gload(X.col + x);
                    first = H.cnt_add(u) == 0u;
                }
                const uint64_t fm = ballot(first);
                if (first) H.tv(ntv + lanes_below(fm)) = (ix_t)u;
                ntv += (uint32_t)__popcll(fm);
            } else {
                const uint32_t dmax = wave_max(d);
                for (uint32_t t = 0; t < dmax; ++t) {
                    bool first = false;
                    uint32_t u = 0;
                    if (t < d) {
                        u = (uint32_t)gload(X.col + a + t);
                        first = H.cnt_add(u) == 0u;
                    }
                    const uint64_t fm = ballot(first);
                    if (first) H.tv(ntv + lanes_below(fm)) = (ix_t)u;
                    ntv += (uint32_t)__popcll(fm);
                }
            }
        }
        X.csr_bytes += 12ull * s + 4ull * degsum;
        H.lsync();
        // touched non-singleton cells, each once (hits[c] counts its touches)
        uint32_t ntc = 0;
        for (uint32_t base = 0; base < ntv; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            bool firstc = false;
            uint32_t c = 0;
            if (i < ntv) {
                const uint32_t u = H.tv(i);
                c = cend(H.pos(u));
                if (cfirst(c) != c) firstc = H.hits_add(c) == 0u;
            }
            const uint64_t cm = ballot(firstc);
            if (firstc) H.tcl(ntc + lanes_below(cm)) = (ix_t)c;
            ntc += (uint32_t)__popcll(cm);
        }
        H.lsync();
        // small cells: one lane each; large ones are compacted to the front
        // of the touched list (entry nbig + rank <= i was already read)
        uint32_t nbig = 0, nfrag_extra = 0;
        for (uint32_t base = 0; base < ntc; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            bool big = false, small = false;
            uint32_t c = 0, f = 0;
            if (i < ntc) {
                c = H.tcl(i);
                f = cfirst(c);
                H.hits_st(c, 0);
                small = c - f + 1 <= kSmallCell;
                big = !small;
            }
            // the 4-wide network when no small cell here has more than 4 members
            uint32_t nf = 1;
            if (ballot(small && c - f + 1 > 4)) {
                if (small) nf = regroup_small<8>(f, c, s);
            } else {
                if (small) nf = regroup_small<4>(f, c, s);
            }
            nfrag_extra += nf - 1;
            if (nf > 1) X.lane_split += c - f + 1;
            const uint64_t bm = ballot(big);
            if (big) H.tcl(nbig + lanes_below(bm)) = (ix_t)c;
            nbig += (uint32_t)__popcll(bm);
        }
        X.ncells += uni(wave_sum(nfrag_extra));
        H.lsync();
        // hits is all zero from here on (the buckets of split_big_cell)
        for (uint32_t j = 0; j < nbig; ++j) {
            const uint32_t c = uni((uint32_t)H.tcl(j));
            const uint32_t f = uni(cfirst(c));
            const uint32_t L = c - f + 1;
            if (s > 1 && sparse_split_list(X, f, c, L, ntv)) {
                X.part_bytes += 8ull * L;
                continue;
            }
            split_big_cell(X, f, c, s);
        }
        general_tail(X, s, longcode, hint);
        // counts back to zero
        for (uint32_t base = 0; base < ntv; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            if (i < ntv) H.cnt_st(H.tv(i), 0);
        }
        H.lsync();
    }

    // 8 bytes per vertex of the split cells counted in the lanes (every lane
    // active; each lane's count is below 2^32, and the halves' sums fit too)
    __device__ __forceinline__ static void flush_split_bytes(Ctx& X) {
        const uint32_t lo = uni(wave_sum(X.lane_split & 0xFFFFu));
        const uint32_t hi = uni(wave_sum(X.lane_split >> 16));
        X.part_bytes += 8ull * ((uint64_t)lo + ((uint64_t)hi << 16));
        X.lane_split = 0;
    }

    // Partition_s_refinement (reference src/partition_refinement.cpp:150-485)
    __device__ int refine(int32_t init_active) {
        const int l = lane_id();
        Ctx X;
        X.n = uni(n);
        X.nw = uni(nw);
        X.ncells = uni(ncells);
        X.nactive = uni(nactive);
        X.steps = uni(steps);
        X.csr_bytes = uni(csr_bytes);
        X.swords = 0;
        X.part_bytes = uni(part_bytes) + 8ull * X.n;
        X.lane_split = 0;
        X.col = uni(col);
        X.rp = uni(H.rows());
        X.rb = Hot::kRowLds ? 0 : uni(gload(X.rp));
        init_active = uni(init_active);
        if (init_active < 0) {
            uint32_t total = 0;
            for (uint32_t base = 0; base < X.n; base += kWave) {
                const uint32_t p = base + (uint32_t)l;
                const bool first = p < X.n && cfirst(cend(p)) == p;
                const uint64_t b = ballot(first);
                if (l == 0) H.active()[base >> 6] = b;
                total += (uint32_t)__popcll(b);
            }
            X.nactive = total;
        } else {
            add_active(X, (uint32_t)init_active);
        }
        H.sync();
        uint32_t longcode = X.ncells;
        int hint = 0;
        // the step budget, counted down in 32 bits (a refinement takes at
        // most 2n steps)
        const uint64_t bud0 = uni(budget);
        const uint32_t bud32 = bud0 > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)bud0;
        uint32_t nst = 0;  // steps of this refinement
        // Singleton splitters of degree <= 64 run back to back in the inner
        // loop, and a larger splitter leaves it for one general step: the
        // values that the singleton steps carry then stay in the same
        // registers from one step to the next.
        bool more = true;
        while (more) {
            uint32_t p = 0, e = 0;
            for (;;) {
                if (!(X.ncells < X.n && X.nactive > 0)) {
                    more = false;
                    break;
                }
                if (nst == bud32) {
                    status = GCL_E_INTERNAL;  // step budget exhausted
                    more = false;
                    break;
                }
                // splitter: the first active position at or after the hint, else
                // the first one; the hint's own word is tested first (one scalar load)
                const uint32_t hw = (uint32_t)hint >> 6;
                const uint64_t hword = uni(H.active()[hw]);
                const uint64_t near = hword & (~0ull << ((uint32_t)hint & 63u));
                int32_t sp;
                uint64_t word = hword;  // the active word holding the splitter
                if (near) {
                    sp = (int32_t)(hw * 64u + (uint32_t)__builtin_ctzll(near));
                } else {
                    sp = first_set_from(H.active(), (hw + 1) * 64u, X.nw, &word);
                    if (sp < 0) sp = first_set_from(H.active(), 0, X.nw, &word);
                }
                p = uni((uint32_t)sp);
                // clear its bit from the word in hand (no other lane writes the
                // bitmap between steps)
                if (l == 0) H.active()[p >> 6] = uni(word) & ~(1ull << (p & 63u));
                --X.nactive;
                ++nst;
                // the cell end and the splitter's vertex (used when the cell is a
                // singleton) are independent loads
                e = uni(cend(p));
                const uint32_t wp = uni((uint32_t)H.elem(p));
                // a singleton splitter's neighbour ids are requested before the
                // sync, so that the load overlaps it and the bookkeeping
                uint32_t w = 0, e0 = 0, deg = 0, u = 0;
                if (e == p) {
                    w = wp;
                    e0 = uni(xrow(X, w));
                    deg = uni(xrow(X, w + 1)) - e0;
                    if (deg <= (uint32_t)kWave && (uint32_t)l < deg) u = (uint32_t)gload(X.col + e0 + (uint32_t)l);
                }
                longcode = GCL_MASH(longcode, p + e);
                if (e > p) longcode = GCL_MASH(longcode, e - p + 1);
                H.sync();
                if (e != p || deg > (uint32_t)kWave) break;
                singleton_step(X, u, deg, longcode, hint);
                // (an LDS class has n <= 8192, so a lane's count stays below
                // n^2 < 2^32 over a whole refinement)
                if constexpr (!Hot::kLds)
                    if ((nst & 0x3FFFu) == 0) flush_split_bytes(X);
            }
            if (!more) break;
            general_step(X, p, e, longcode, hint);
            if constexpr (!Hot::kLds)
                if ((nst & 0x3FFFu) == 0) flush_split_bytes(X);
        }
        flush_split_bytes(X);
        budget = bud0 - nst;
        X.steps += nst;
        X.csr_bytes += 4ull * X.swords;
        if (X.nactive) {
            for (uint32_t w = (uint32_t)l; w < X.nw; w += kWave) H.active()[w] = 0;
            X.nactive = 0;
            H.sync();
        }
        ncells = X.ncells;
        nactive = 0;
        steps = X.steps;
        csr_bytes = X.csr_bytes;
        part_bytes = X.part_bytes;
        longcode = GCL_MASH(longcode, X.ncells);
        return GCL_CLEANUP(longcode);
    }

    // split_by_and_refine (reference src/partition_refinement.cpp:486-539)
    __device__ void individualize_and_refine(uint32_t v, uint32_t f) {
        const uint32_t c = cend(f);
        const uint32_t q = H.pos(v);
        const uint32_t x = H.elem(f);
        H.sync();
        if (lane_id() == 0) {
            H.elem(f) = (ix_t)v;
            H.pos(v) = (ix_t)f;
            H.elem(q) = (ix_t)x;
            H.pos(x) = (ix_t)q;
            H.ce(f) = (ix_t)f;          // {v}
            H.ce(c) = (ix_t)(f + 1);    // the rest [f+1, c]
        }
        set_bit_u(H.nsmask(), f, false);
        set_bit_u(H.nsmask(), f + 1, c > f + 1);
        set_bit_u(H.lasts(), f, true);
        ++ncells;
        H.sync();
        code = refine((int32_t)f);
    }

    // Backtrack the partition to its state at `lev` (reconstruct_at_level,
    // reference src/partition_refinement.cpp:541-604, up to in-cell order).
    __device__ void restore_level(uint32_t lev) {
        const uint64_t* src = snap + (uint64_t)lev * (nw + 1);
        for (uint32_t w = (uint32_t)lane_id(); w < nw; w += kWave) H.lasts()[w] = src[w];
        H.sync();
        rebuild_from_lasts();
        if (lev == 1) root_changed = false;
    }

    // --------------------------------------------------------- search ops
    // target_cell_selector with hint -1 (reference
    // src/partition_refinement.cpp:607-655): first non-singleton cell, members
    // in increasing vertex id.
    __device__ bool select_target() {
        const int32_t f = first_set_from(H.nsmask(), 0, nw);
        if (f < 0) return true;
        const int l = lane_id();
        const uint32_t c = cend((uint32_t)f);
        const uint32_t t = c - (uint32_t)f + 1;
        const uint32_t idx = tc_level_reached;
        const uint64_t off = idx == 0 ? 0 : (uint64_t)lv[idx - 1].arena_off + lv[idx - 1].tc_len;
        if (off + t > arena_cap()) {
            status = GCL_E_CAPACITY;
            return false;
        }
        for (uint32_t i = (uint32_t)l; i < t; i += kWave) set_bit_atomic(H.vmask(), H.elem((uint32_t)f + i));
        H.sync();
        uint32_t run = 0;
        for (uint32_t base = 0; base < nw; base += kWave) {
            const uint32_t w = base + (uint32_t)l;
            uint64_t bits = w < nw ? H.vmask()[w] : 0ull;
            const uint32_t c0 = (uint32_t)__popcll(bits);
            const uint32_t inc = wave_scan(c0);
            uint32_t j = run + inc - c0;
            while (bits) {
                const uint32_t b = (uint32_t)__builtin_ctzll(bits);
                bits &= bits - 1;
                arena_mem[off + j] = (ix_t)(w * 64u + b);
                arena_rep[off + j] = (ix_t)(w * 64u + b);
                ++j;
            }
            if (w < nw) H.vmask()[w] = 0;
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
        group_sync();
        return true;
    }

    // subgroup_fixing_sequence + compute_orbits_fast (reference
    // src/perm_group.cpp:5-116), union by minimum without locks.
    __device__ void compute_orbits(uint32_t li) {
        const int l = lane_id();
        const LevelRec r = lv[li];
        const int64_t upto = min(num_aut, (int64_t)ring());
        const int64_t from = r.last_num_aut < 0 ? 0 : r.last_num_aut;
        // ring slots whose automorphism fixes the sequence: one slot per lane;
        // aut_fix[i] = 2k+1 when seq[k] is the first entry the slot moves,
        // 2k when it fixes seq[0, k) and the rest is unchecked, so each seq
        // entry is checked once per slot until backtracking rewrites it
        truncate_fix();
        const uint32_t ss = seq_size;
        uint32_t nsub = 0;
        // the fix words of up to 8 * 64 slots are loaded together (one
        // memory latency per chunk instead of one per 64 slots); most slots
        // are settled by their word alone
        constexpr int kChunk = 8;
        for (int64_t base = from; base < upto; base += kWave * kChunk) {
            uint32_t xs[kChunk];
#pragma unroll
            for (int c = 0; c < kChunk; ++c) {
                const int64_t i = base + c * kWave + l;
                xs[c] = i < upto ? aut_fix[i] : 0u;
            }
#pragma unroll
            for (int c = 0; c < kChunk; ++c) {
                const int64_t i = base + c * kWave + l;
                bool ok = false;
                if (i < upto) {
                    uint32_t x = xs[c];
                    if ((x >> 1) >= ss) {
                        ok = true;
                    } else if (!(x & 1u)) {
                        const ix_t* a = auts + (uint64_t)i * n;
                        ok = true;
                        for (uint32_t j = x >> 1; j < ss; ++j) {
                            const uint32_t sj = seq[j];
                            if (a[sj] != sj) {
                                ok = false;
                                x = (j << 1) | 1u;
                                break;
                            }
                        }
                        if (ok) x = ss << 1;
                        aut_fix[i] = x;
                    }
                }
                const uint64_t m = ballot(ok);
                if (ok) subgroup[nsub + lanes_below(m)] = (uint32_t)i;
                nsub += (uint32_t)__popcll(m);
            }
        }
        group_sync();
        if (nsub == 0) return;
        // An automorphism that fixes the sequence maps the node's partition
        // to itself (refinement is label-invariant), so it permutes the
        // target cell: the orbits of the cell's members are the components of
        // v -> a(v) over the members alone.  The union-find lives in cnt
        // (zero between steps) at the members' entries, seeded with the
        // representatives of the automorphisms already applied at this level.
        const uint32_t t = r.tc_len, off = r.arena_off, tf = r.tc_first;
        for (uint32_t i = (uint32_t)l; i < t; i += kWave) H.cnt_st(arena_mem[off + i], arena_rep[off + i]);
        H.sync();
        bool outside = false;
        // (automorphism, member) pairs, one per lane; 32-bit indices (t <= n
        // < 2^16, so a chunk of 2^15 automorphisms stays below 2^31 pairs)
        for (uint32_t s0 = 0; s0 < nsub; s0 += 32768u) {
          const uint32_t pairs = min(nsub - s0, 32768u) * t;
          for (uint32_t k = (uint32_t)l; k < pairs; k += kWave) {
            const uint32_t q = k / t;
            const uint32_t si = s0 + q, i = k - q * t;
            const uint32_t v = arena_mem[off + i];
            const uint32_t img = auts[(uint64_t)subgroup[si] * n + v];
            if (img == v) continue;
            if ((uint32_t)H.pos(img) - tf >= t) {
                outside = true;
                continue;
            }
            // parents only point to smaller ids, so walks are shorter than
            // the cell; the caps only bound a corrupted structure
            uint32_t x = v, y = img;
            for (uint32_t attempt = 0; attempt < n; ++attempt) {
                for (uint32_t h = 0; h < n; ++h) {
                    const uint32_t px = H.cnt_ld(x);
                    if (px == x) break;
                    x = px;
                }
                for (uint32_t h = 0; h < n; ++h) {
                    const uint32_t py = H.cnt_ld(y);
                    if (py == y) break;
                    y = py;
                }
                if (x == y) break;
                const uint32_t lo = min(x, y), hi = max(x, y);
                if (H.cnt_cas(hi, hi, lo)) break;
                x = lo;
                y = hi;
            }
          }
        }
        H.sync();
        for (uint32_t i = (uint32_t)l; i < t; i += kWave) {
            uint32_t x = arena_mem[off + i];
            for (uint32_t h = 0; h < n && H.cnt_ld(x) != x; ++h) x = H.cnt_ld(x);
            arena_rep[off + i] = (ix_t)x;
        }
        H.sync();
        for (uint32_t i = (uint32_t)l; i < t; i += kWave) H.cnt_st(arena_mem[off + i], 0);
        if (ballot(outside)) status = GCL_E_INTERNAL;  // not an automorphism of this node
        group_sync();
    }

    // Certificate of the current leaf: sorted canonical rows (keys i*n+j)
    // with row offsets, the sparse form of the reference's n x n leaf hash
    // (src/CSRGraph.cpp:332-351).
    __device__ void store_cert(uint32_t* keys, uint32_t* row, ix_t* label) {
        const int l = lane_id();
        const int32_t* cl = uni(col);
        uint32_t run = 0;
        for (uint32_t base = 0; base < n; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            uint32_t v = 0, d = 0;
            if (i < n) {
                v = H.elem(i);
                d = H.row(v + 1) - H.row(v);
            }
            const uint32_t inc = wave_scan(d);
            const uint32_t o = run + inc - d;
            if (i < n && d <= 8) {
                row[i] = o;
                label[i] = (ix_t)v;
                const uint32_t a = H.row(v);
                uint32_t kk[8];
#pragma unroll
                for (uint32_t t = 0; t < 8; ++t) kk[t] = t < d ? i * n + H.pos((uint32_t)gload(cl + a + t)) : 0xFFFFFFFFu;
#pragma unroll
                for (uint32_t k = 2; k <= 8; k <<= 1)
#pragma unroll
                    for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                        for (uint32_t q = 0; q < 8; ++q) {
                            const uint32_t qj = q ^ j;
                            if (qj > q) {
                                const bool up = (q & k) == 0;
                                const uint32_t x0 = kk[q], x1 = kk[qj];
                                if ((x0 > x1) == up) {
                                    kk[q] = x1;
                                    kk[qj] = x0;
                                }
                            }
                        }
#pragma unroll
                for (uint32_t t = 0; t < 8; ++t)
                    if (t < d) keys[o + t] = kk[t];
            } else if (i < n) {
                row[i] = o;
                label[i] = (ix_t)v;
                const uint32_t a = H.row(v);
                for (uint32_t t = 0; t < d; ++t) {
                    const uint32_t key = i * n + H.pos((uint32_t)gload(cl + a + t));
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
        group_sync();
    }

    // compare_adj_upper (reference src/sequential.cpp:114-124) of the current
    // leaf against a stored certificate, without sorting the current rows: at
    // the first row whose neighbour sets differ, the graph holding the
    // smallest element of the symmetric difference is the greater one.
    __device__ int compare_cert(const uint32_t* keys, const uint32_t* row) {
        const int l = lane_id();
        const int32_t* cl = uni(col);
        keys = uni(keys);
        row = uni(row);
        for (uint32_t base = 0; base < n; base += kWave) {
            const uint32_t i = base + (uint32_t)l;
            bool differs = false;
            if (i < n) {
                const uint32_t v = H.elem(i);
                const uint32_t a = H.row(v);
                const uint32_t d = H.row(v + 1) - a;
                const uint32_t r0 = gload(row + i), r1 = gload(row + i + 1);
                if (r1 - r0 != d) {
                    differs = true;
                } else if (d <= 8) {
                    uint32_t ref[8], cur[8];
#pragma unroll
                    for (uint32_t t = 0; t < 8; ++t) {
                        ref[t] = t < d ? gload(keys + r0 + t) : 0xFFFFFFFFu;
                        cur[t] = t < d ? (uint32_t)gload(cl + a + t) : 0u;
                    }
#pragma unroll
                    for (uint32_t t = 0; t < 8; ++t) {
                        if (t < d) {
                            const uint32_t key = i * n + H.pos(cur[t]);
                            bool found = false;
#pragma unroll
                            for (uint32_t z = 0; z < 8; ++z) found |= ref[z] == key;
                            differs |= !found;
                        }
                    }
                } else {
                    for (uint32_t t = 0; t < d && !differs; ++t) {
                        const uint32_t key = i * n + H.pos((uint32_t)gload(cl + a + t));
                        uint32_t lo = r0, hi = r1;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (gload(keys + mid) < key) lo = mid + 1;
                            else hi = mid;
                        }
                        if (lo == r1 || gload(keys + lo) != key) differs = true;
                    }
                }
            }
            const uint64_t dm = ballot(differs);
            if (dm) {
                const uint32_t i0 = base + (uint32_t)__builtin_ctzll(dm);
                const uint32_t v = H.elem(i0);
                const uint32_t a = H.row(v);
                const uint32_t d = H.row(v + 1) - a;
                const uint32_t r0 = gload(row + i0), r1 = gload(row + i0 + 1);
                uint32_t min_cur = 0xFFFFFFFFu, min_ref = 0xFFFFFFFFu;
                for (uint32_t t = (uint32_t)l; t < d; t += kWave) {
                    const uint32_t key = i0 * n + H.pos((uint32_t)gload(cl + a + t));
                    bool found = false;
                    for (uint32_t z = r0; z < r1; ++z)
                        if (gload(keys + z) == key) {
                            found = true;
                            break;
                        }
                    if (!found) min_cur = min(min_cur, key);
                }
                for (uint32_t z = r0 + (uint32_t)l; z < r1; z += kWave) {
                    const uint32_t key = gload(keys + z);
                    bool found = false;
                    for (uint32_t t = 0; t < d; ++t)
                        if (i0 * n + H.pos((uint32_t)gload(cl + a + t)) == key) {
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

    __device__ void bind_auts(uint8_t* region) {
        aut_fix = (uint32_t*)region;
        auts = region ? (ix_t*)(region + aut_fix_bytes(ring())) : nullptr;
    }

    __device__ void store_aut(const ix_t* ref_label) {
        const uint64_t slot = (uint64_t)(num_aut % (int64_t)ring());
        ix_t* dst = auts + slot * n;
        for (uint32_t j = (uint32_t)lane_id(); j < n; j += kWave) dst[ref_label[j]] = H.elem(j);
        if (lane_id() == 0) aut_fix[slot] = 0;  // nothing of seq checked yet
        ++num_aut;
        group_sync();
    }

    // The sequence shrank to fix_low entries at some point since the fix
    // lengths were last brought up to date (backtracking only records the
    // lowest length; truncations to s1, then s2 act as one to min(s1, s2)):
    // lengths beyond it, or a mismatch at an entry since rewritten, go back
    // to "matches the first fix_low entries, rest unchecked".
    __device__ void truncate_fix() {
        const int64_t cnt = min(num_aut, (int64_t)ring());
        const uint32_t low = fix_low;
        fix_low = 0xFFFFFFFFu;
        if (cnt == 0 || low == 0xFFFFFFFFu) return;
        const uint32_t keep = low << 1;
        for (int64_t i = lane_id(); i < cnt; i += kWave) {
            const uint32_t x = aut_fix[i];
            if (x > keep) aut_fix[i] = keep;
        }
        group_sync();
    }

    __device__ int gca_level(const ix_t* a, uint32_t na) {
        // get_gca_level (reference src/sequential.cpp:183-213)
        const uint32_t m = min(na, seq_size);
        const int l = lane_id();
        for (uint32_t base = 0; base < m; base += kWave) {
            const uint32_t k = base + (uint32_t)l;
            const uint64_t dm = ballot(k < m && a[k] != seq[k]);
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
        restore_level(lev);
        if (lev < eqlevfirst) eqlevfirst = lev;
        if ((int)lev <= eqlevcanon) {
            eqlevcanon = (int)lev;
            comp_canon = 0;
        }
        seq_size = lev - 1;
        fix_low = min(fix_low, seq_size);
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
        group_sync();
        canon_code[level] = kSentinel;
        group_sync();
    }

    // descend from the current node (level `level`) into child v
    __device__ void descend(uint32_t v, uint32_t tc_first) {
        ++nodes;
        seq[seq_size] = (ix_t)v;
        ++seq_size;
        group_sync();
        const uint64_t t0 = dg_t0();
        individualize_and_refine(v, tc_first);
        dg_add(dg_refine, t0);
        ++level;
        snapshot(level);
    }

    __device__ bool first_path_node(const KParams& P, uint32_t graph) {
        // first_path_node (reference src/sequential.cpp:216-257)
        if (ncells == n) {
            first_terminal();
            if constexpr (Hot::kDesc)
                if (P.dspec) desc_publish(P, graph, level);
            backtrack(level - 1);
            return true;
        }
        if (!select_target()) return false;
        LevelRec r = lv[tc_level_reached - 1];
        const uint32_t v = arena_mem[r.arena_off + r.cursor];
        lv[tc_level_reached - 1].cursor = r.cursor + 1;
        descend(v, r.tc_first);
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
        group_sync();
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
                if constexpr (Hot::kDesc) d_renew = true;
                group_sync();
                backtrack(L - 1);
                break;
            }
            case kBad:
                backtrack(L - 1);
                break;
        }
    }

    // ------------------------------------------------- root speculation
    __device__ SpecHeader* spec_header() { return (SpecHeader*)(spec_rec + sl().header); }
    __device__ uint32_t* spec_state() { return (uint32_t*)(spec_rec + sl().state); }

    // This graph will not publish (any more): it leaves the launch's count of
    // graphs that may, which lets helpers leave once nothing is published.
    __device__ void settle_pending(const KParams& P) {
        if (spec_pending) {
            spec_pending = false;
            if (lane_id() == 0) atomicSub(P.pending, 1u);
        }
    }
    // Publishes the remaining root children (orbit representatives at and
    // after the cursor) for other workers to refine.  Called at level 1.  The
    // graph leaves the pending count only once its record is on the board.
    __device__ void spec_publish(const KParams& P, const LevelRec& r, uint32_t graph) {
        // count candidates
        uint32_t cnt = 0;
        for (uint32_t i = r.cursor + (uint32_t)lane_id(); i < r.tc_len; i += kWave)
            cnt += arena_rep[r.arena_off + i] == arena_mem[r.arena_off + i] ? 1u : 0u;
        cnt = wave_sum(cnt);
        if (cnt < 2) {
            settle_pending(P);
            return;
        }
        // take a free record
        int32_t slot = -1;
        if (lane_id() == 0) {
            const uint32_t nwr = (P.spec_records + 63) / 64;
            for (uint32_t w = 0; w < nwr && slot < 0; ++w) {
                uint64_t m = __hip_atomic_load(&P.slot_free[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                while (m) {
                    const uint64_t bit = m & (~m + 1);
                    const uint64_t old = atomicAnd(&P.slot_free[w], ~(unsigned long long)bit);
                    if (old & bit) {
                        slot = (int32_t)(w * 64u + (uint32_t)__builtin_ctzll(bit));
                        break;
                    }
                    m = old & ~bit;
                }
            }
        }
        slot = __shfl(slot, 0, kWave);
        if (slot < 0) {
            settle_pending(P);
            return;
        }
        spec_slot = slot;
        spec_rec = P.spec + (uint64_t)slot * P.spec_stride;
        SpecHeader* hd = spec_header();
        const uint32_t epoch = hd->epoch + 1;
        spec_epoch = epoch;
        const uint32_t count = min(cnt, spec_item_cap(sl(), P.spec_stride, n));
        spec_count = count;
        spec_cursor = 0;
        // root partition and item list
        ix_t* re = (ix_t*)(spec_rec + sl().root_elem);
        uint64_t* rl = (uint64_t*)(spec_rec + sl().root_lasts);
        const uint64_t* s1 = snap + 1ull * (nw + 1);
        for (uint32_t p = (uint32_t)lane_id(); p < n; p += kWave) re[p] = H.elem(p);
        for (uint32_t w = (uint32_t)lane_id(); w < nw; w += kWave) rl[w] = s1[w];
        ix_t* vert = (ix_t*)(spec_rec + sl().vert);
        uint32_t run = 0;
        for (uint32_t base = r.cursor; base < r.tc_len && run < count; base += kWave) {
            const uint32_t i = base + (uint32_t)lane_id();
            const bool ok = i < r.tc_len && arena_rep[r.arena_off + i] == arena_mem[r.arena_off + i];
            const uint64_t m = ballot(ok);
            const uint32_t j = run + lanes_below(m);
            if (ok && j < count) vert[j] = arena_mem[r.arena_off + i];
            run += (uint32_t)__popcll(m);
        }
        if (lane_id() == 0) {
            hd->graph = graph;
            hd->count = count;
            hd->tc_first = r.tc_first;
            hd->n = n;
            hd->first1 = firstcode[1];
            __hip_atomic_store(&hd->canon1, canon_code[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&hd->next, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // data and header before the item states that make items claimable
        __threadfence();
        group_sync();
        uint32_t* state = spec_state();
        for (uint32_t j = (uint32_t)lane_id(); j < count; j += kWave) state[j] = epoch << 2;
        __threadfence();
        group_sync();
        if (lane_id() == 0) {
            agent_store(&hd->epoch, epoch);
            __threadfence();
            atomicOr(&P.board_mask[slot >> 6], (unsigned long long)(1ull << (slot & 63)));
            // the record is on the board before the graph leaves the pending
            // count: a helper that reads the count at 0 then finds the record
            __threadfence();
        }
        settle_pending(P);
        group_sync();
    }

    // Root child v: adopt a helper's refinement, or claim and refine it here.
    // Returns kTaken when v was adopted (partition at level 2, code set),
    // kRejected when its code alone decides it (see below), kNotTaken when
    // it must be refined here.
    static constexpr int kNotTaken = 0, kTaken = 1, kRejected = 2;
    __device__ int spec_take(const KParams& P, uint32_t v) {
        if (spec_slot < 0) return kNotTaken;
        const ix_t* vert = (const ix_t*)(spec_rec + sl().vert);
        uint32_t* state = spec_state();
        const uint32_t E = spec_epoch << 2;
        // skip (and withdraw) published items that became non-representatives
        while (spec_cursor < spec_count && (uint32_t)vert[spec_cursor] < v) {
            if (lane_id() == 0) agent_cas(&state[spec_cursor], E | 0u, E | 3u);
            ++spec_cursor;
        }
        if (spec_cursor >= spec_count || (uint32_t)vert[spec_cursor] != v) return kNotTaken;
        const uint32_t j = spec_cursor++;
        // the best path's code at level 2 as it is now: helpers that finish
        // later leave out the partitions of children it rejects
        if (lane_id() == 0)
            __hip_atomic_store(&spec_header()->canon1, canon_code[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t won = 0;
        if (lane_id() == 0) won = agent_cas(&state[j], E | 0u, E | 3u) ? 1u : 0u;
        won = __shfl(won, 0, kWave);
        if (won) return kNotTaken;  // withdrawn: refine it here
        // a helper has it: refine later children meanwhile, then copy its partition
        const uint64_t tw = wall_clock64();
        while (agent_peek(&state[j]) != (E | 2u)) {
            if (wall_clock64() > deadline) {
                status = GCL_E_TIMEOUT;  // a helper did not finish in time
                return kNotTaken;
            }
            if (!spec_help_own(P, E)) __builtin_amdgcn_s_sleep(2);
        }
        dg_wait += wall_clock64() - tw;
        agent_acquire();
        const uint32_t hw = uni(((const uint32_t*)(spec_rec + sl().code))[j]);
        if (hw == kFailedCode) {
            status = GCL_E_INTERNAL;  // the helper's refinement failed
            return kNotTaken;
        }
        const uint32_t hc = hw & ~kCodeOnly;
        const uint64_t* cnt_rec = (const uint64_t*)(spec_rec + sl().steps) + 4ull * j;
        // A child whose code is below the best path's at this level and
        // differs from the first path's is classified kBad before its
        // partition is looked at (classify_node, reference
        // src/sequential.cpp:396-624), and the backtrack that follows puts
        // back the root partition.  In the root state (level 1, eqlevcanon
        // 1, comp_canon 0: every root child after the first path returns to
        // it) such a child costs its counters only.  When this worker refined
        // a child of its own record meanwhile, the partition it holds is that
        // child's, and the level-1 snapshot puts the root's back.
        if (level == 1 && seq_size == 0 && eqlevfirst <= 1 && eqlevcanon == 1 && comp_canon == 0) {
            const bool eq_first = eqlevfirst == 1 && (int)hc == uni(firstcode[1]);
            if (!eq_first && (int)hc < uni(canon_code[1])) {
                steps += cnt_rec[0];
                csr_bytes += cnt_rec[1];
                part_bytes += cnt_rec[2];
                if (root_changed) restore_level(1);
                tc_level_reached = 1;  // what backtrack(1) leaves (the rest is unchanged)
                return kRejected;
            }
        }
        // a helper leaves the partition out only for a child that the code
        // rejects; should this worker not be in the state that rejects it,
        // the child is refined here
        if (hw & kCodeOnly) return kNotTaken;
        const uint8_t* item = spec_rec + sl().payload + (uint64_t)j * spec_item_bytes(n);
        const ix_t* ie = (const ix_t*)item;
        const uint64_t* il = (const uint64_t*)(item + spec_item_lasts(n));
        for (uint32_t p = (uint32_t)lane_id(); p < n; p += kWave) H.elem(p) = ie[p];
        for (uint32_t w = (uint32_t)lane_id(); w < nw; w += kWave) H.lasts()[w] = il[w];
        code = (int)hc;
        steps += cnt_rec[0];
        csr_bytes += cnt_rec[1];
        part_bytes += cnt_rec[2];
        H.sync();
        positions_from_elem();
        rebuild_from_lasts();
        return kTaken;
    }
    // While a helper refines the child this worker waits for, claim a later
    // unclaimed child of its own record and refine it as a helper would.
    // The adoption that follows replaces the partition; the counters that
    // the adoption adds from the record are restored here, so nothing is
    // counted twice.  Returns true when it refined one.
    __device__ bool spec_help_own(const KParams& P, uint32_t E) {
        SpecHeader* hd = spec_header();
        uint32_t idx = 0xFFFFFFFFu;
        if (lane_id() == 0 &&
            __hip_atomic_load(&hd->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < spec_count) {
            const uint32_t t = atomicAdd(&hd->next, 1u);
            if (t < spec_count && agent_cas(&spec_state()[t], E | 0u, E | 1u)) idx = t;
        }
        idx = uni(readlane(idx, 0));
        if (idx == 0xFFFFFFFFu) return false;
        const uint64_t st0 = steps, by0 = csr_bytes, pb0 = part_bytes, bud0 = budget;
        const int32_t s0 = status;
        root_changed = true;  // help_item works in this worker's partition
        help_item(P, spec_rec, idx, E);
        steps = st0;
        csr_bytes = by0;
        part_bytes = pb0;
        budget = bud0;
        status = s0;
        return true;
    }

    // Ends the publication: withdraw unclaimed items, wait for running ones.
    __device__ void spec_finish(const KParams& P) {
        settle_pending(P);
        if (spec_slot < 0) return;
        uint32_t* state = spec_state();
        const uint32_t E = spec_epoch << 2;
        if (lane_id() == 0)
            atomicAnd(&P.board_mask[spec_slot >> 6], ~(unsigned long long)(1ull << (spec_slot & 63)));
        group_sync();
        bool held = false;
        for (uint32_t j = (uint32_t)lane_id(); j < spec_count; j += kWave) {
            agent_cas(&state[j], E | 0u, E | 3u);
            while (agent_peek(&state[j]) == (E | 1u) && !(held = wall_clock64() > deadline))
                __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
        group_sync();
        // A helper that still holds an item after the deadline may still
        // write into the record: the record then stays out of the free set
        // for the rest of the launch, so that no later publication reuses it.
        if (ballot(held)) {
            if (status == 0) status = GCL_E_TIMEOUT;
        } else if (lane_id() == 0) {
            atomicOr(&P.slot_free[spec_slot >> 6], (unsigned long long)(1ull << (spec_slot & 63)));
        }
        group_sync();
        spec_slot = -1;
    }

    // ---------------------------------------------- descent speculation
    __device__ DescHeader* desc_header() { return (DescHeader*)(d_rec + dly().header); }
    __device__ uint32_t* desc_state() { return (uint32_t*)(d_rec + dly().state); }

    // At a new best leaf of level d (the first leaf is the first of them):
    // lv[0 .. d-2] are the best path's target cells, each cursor just after
    // the path's vertex, snap[1 .. d] the path's boundary snapshots and
    // H.elem the leaf's element order (backtracking since has only rebuilt
    // cells).  Publishes items (L, v) for L = 2 .. d-1, v the members of
    // level L's target cell after the path's vertex, deepest level first.
    __device__ void desc_publish(const KParams& P, uint32_t graph, uint32_t leaf_level) {
        d_slot = -1;
        if (P.spec_min_n == 0 || n < P.spec_min_n || leaf_level < 3) return;
        const int l = lane_id();
        const uint32_t lmax = min(leaf_level - 1, P.dspec_path);
        if (lmax < 2) return;
        // take a free record
        int32_t slot = -1;
        if (l == 0) {
            const uint32_t nwr = (P.dspec_records + 63) / 64;
            for (uint32_t w = 0; w < nwr && slot < 0; ++w) {
                uint64_t m = __hip_atomic_load(&P.dslot_free[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                while (m) {
                    const uint64_t bit = m & (~m + 1);
                    const uint64_t old = atomicAnd(&P.dslot_free[w], ~(unsigned long long)bit);
                    if (old & bit) {
                        slot = (int32_t)(w * 64u + (uint32_t)__builtin_ctzll(bit));
                        break;
                    }
                    m = old & ~bit;
                }
            }
        }
        slot = __shfl(slot, 0, kWave);
        if (slot < 0) return;
        d_rec = P.dspec + (uint64_t)slot * P.dspec_stride;
        const uint64_t nw1 = (uint64_t)P.maxn / 64 + 2;
        uint8_t* rec = d_rec;
        uint32_t* lfirst = (uint32_t*)(rec + dly().level_first);
        uint32_t* lcount = (uint32_t*)(rec + dly().level_count);
        ix_t* ilev = (ix_t*)(rec + dly().item_level);
        ix_t* ivert = (ix_t*)(rec + dly().item_vert);
        // items, deepest level first: lane q of a chunk takes level lmax - q
        uint32_t count = 0;
        for (uint32_t base = 0; base + 2 <= lmax; base += kWave) {
            const uint32_t q = base + (uint32_t)l;
            const bool valid = q + 2 <= lmax;
            const uint32_t L = valid ? lmax - q : 0;
            uint32_t c = 0;
            LevelRec r;
            if (valid) {
                r = lv[L - 1];
                c = r.cursor < r.tc_len ? min(r.tc_len - r.cursor, kDescPerLevel) : 0u;
            }
            const uint32_t inc = wave_scan(c);
            const uint32_t off = count + inc - c;
            const uint32_t take = valid ? (off >= dly().items ? 0u : min(c, dly().items - off)) : 0u;
            if (valid) {
                lfirst[L] = off;
                lcount[L] = take;
                for (uint32_t i = 0; i < take; ++i) {
                    ilev[off + i] = (ix_t)L;
                    ivert[off + i] = arena_mem[r.arena_off + r.cursor + i];
                }
            }
            count = min(count + readlane(inc, kWave - 1), dly().items);
        }
        count = uni(count);
        if (count == 0) {
            if (l == 0) atomicOr(&P.dslot_free[slot >> 6], (unsigned long long)(1ull << (slot & 63)));
            group_sync();
            return;
        }
        // levels 0 and 1 have no items
        if (l < 2) {
            lfirst[l] = 0;
            lcount[l] = 0;
        }
        // the first leaf's element vector and the first path's snapshots
        ix_t* re = (ix_t*)(rec + dly().elem);
        for (uint32_t p = (uint32_t)l; p < n; p += kWave) re[p] = H.elem(p);
        uint64_t* rs = (uint64_t*)(rec + dly().snaps);
        for (uint32_t L = 2; L <= lmax; ++L)
            for (uint32_t w = (uint32_t)l; w < nw; w += kWave) rs[(uint64_t)L * nw1 + w] = snap[(uint64_t)L * (nw + 1) + w];
        DescHeader* hd = desc_header();
        const uint32_t epoch = hd->epoch + 1;
        d_epoch = epoch;
        if (l == 0) {
            hd->graph = graph;
            hd->count = count;
            hd->n = n;
            hd->levels = lmax;
            __hip_atomic_store(&hd->next, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __threadfence();
        group_sync();
        uint32_t* state = desc_state();
        for (uint32_t j = (uint32_t)l; j < count; j += kWave) state[j] = epoch << 2;
        __threadfence();
        group_sync();
        if (l == 0) {
            agent_store(&hd->epoch, epoch);
            __threadfence();
            atomicOr(&P.dboard[slot >> 6], (unsigned long long)(1ull << (slot & 63)));
        }
        d_slot = slot;
        group_sync();
    }

    // Takes record `slot` off the board and withdraws its unclaimed items;
    // with `wait`, also waits for the claimed ones (up to the deadline) and
    // returns the record to the free set (a record a helper still holds at
    // the deadline stays out of it for the rest of the launch).
    __device__ void desc_end(const KParams& P, int32_t slot, uint32_t epoch, bool wait) {
        uint8_t* rec = P.dspec + (uint64_t)slot * P.dspec_stride;
        uint32_t* state = (uint32_t*)(rec + dly().state);
        const uint32_t E = epoch << 2;
        const uint32_t count = uni(((const DescHeader*)(rec + dly().header))->count);
        if (lane_id() == 0) atomicAnd(&P.dboard[slot >> 6], ~(unsigned long long)(1ull << (slot & 63)));
        group_sync();
        bool held = false;
        for (uint32_t j = (uint32_t)lane_id(); j < count; j += kWave) {
            agent_cas(&state[j], E | 0u, E | 3u);
            if (wait)
                while (agent_peek(&state[j]) == (E | 1u) && !(held = wall_clock64() > deadline))
                    __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
        group_sync();
        if (!wait) return;
        if (ballot(held)) {
            if (status == 0) status = GCL_E_TIMEOUT;
        } else if (lane_id() == 0) {
            atomicOr(&P.dslot_free[slot >> 6], (unsigned long long)(1ull << (slot & 63)));
        }
        group_sync();
    }
    // Ends both publications (the current path's and the retired one).
    __device__ void desc_finish(const KParams& P) {
        if (d_slot >= 0) desc_end(P, d_slot, d_epoch, true);
        if (d_old >= 0) desc_end(P, d_old, d_old_epoch, true);
        d_slot = d_old = -1;
    }
    // A new best leaf: the current record is retired without waiting for
    // its claimed descents (the record retired before, long since claimed,
    // is ended first), and the best path's descents are published.
    __device__ void desc_republish(const KParams& P, uint32_t graph) {
        if (d_old >= 0) desc_end(P, d_old, d_old_epoch, true);
        d_old = -1;
        if (d_slot >= 0) {
            desc_end(P, d_slot, d_epoch, false);
            d_old = d_slot;
            d_old_epoch = d_epoch;
            d_slot = -1;
        }
        desc_publish(P, graph, (uint32_t)canon_level);
    }

    // The owner at level L (its node is the first path's, as the caller
    // checks with the sequence) is about to descend into v.  Returns kTaken
    // when a helper's recorded descent was replayed: the levels it reached
    // were classified as the search would classify them, and the search
    // continues from the state the last classification left.
    __device__ int desc_take(const KParams& P, uint32_t v) {
        const uint32_t L = level;
        if (L < 2 || L > dly().path || L > uni(desc_header()->levels)) return kNotTaken;
        // on the best path, whose descents the record holds: seq[0 .. L-2]
        // equals best_seq[0 .. L-2]
        if (L - 1 > best_seq_size) return kNotTaken;
        for (uint32_t base = 0; base < L - 1; base += kWave) {
            const uint32_t i = base + (uint32_t)lane_id();
            if (ballot(i < L - 1 && seq[i] != best_seq[i])) return kNotTaken;
        }
        const uint32_t lf = uni(((const uint32_t*)(d_rec + dly().level_first))[L]);
        const uint32_t lc = uni(((const uint32_t*)(d_rec + dly().level_count))[L]);
        const ix_t* ivert = (const ix_t*)(d_rec + dly().item_vert);
        uint32_t j = 0xFFFFFFFFu;
        for (uint32_t i = 0; i < lc; ++i)
            if ((uint32_t)ivert[lf + i] == v) {
                j = lf + i;
                break;
            }
        j = uni(j);
        if (j == 0xFFFFFFFFu) return kNotTaken;
        uint32_t* state = desc_state();
        const uint32_t E = d_epoch << 2;
        uint32_t won = 0;
        if (lane_id() == 0) won = agent_cas(&state[j], E | 0u, E | 3u) ? 1u : 0u;
        if (uni(readlane(won, 0))) return kNotTaken;  // withdrawn: descend here
        const uint64_t tw = wall_clock64();
        while (agent_peek(&state[j]) != (E | 2u)) {
            if (wall_clock64() > deadline) {
                status = GCL_E_TIMEOUT;  // a helper did not finish in time
                return kNotTaken;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        dg_wait += wall_clock64() - tw;
        agent_acquire();
        const uint32_t nr = uni(((const uint32_t*)(d_rec + dly().nrec))[j]);
        const uint32_t nrec = nr & 0xFFFFu;
        if ((nr & kDescFailed) || nrec == 0) return kNotTaken;
        ++dg_adopted;
        desc_replay(P, d_rec + dly().payload + (uint64_t)j * dly().item_bytes, nrec);
        return kTaken;
    }

    // Replays a recorded descent from level L: per level what descend(),
    // classify() and the next other_path_node() would do, with the recorded
    // refinement results in place of refining.
    __device__ void desc_replay(const KParams& P, const uint8_t* item, uint32_t nrec) {
        const int l = lane_id();
        const DescLevel* R = (const DescLevel*)item;
        const uint64_t* LS = (const uint64_t*)(item + dly().lasts_off);
        const ix_t* FE = (const ix_t*)(item + dly().elem_off);
        const uint64_t nw1 = (uint64_t)P.maxn / 64 + 2;
        const uint32_t L0 = level;
        // the element vector of the last recorded level refines every level
        // of the descent and level L0: cells come from the snapshots
        for (uint32_t p = (uint32_t)l; p < n; p += kWave) H.elem(p) = FE[p];
        H.sync();
        positions_from_elem();
        for (uint32_t k = 0; k < nrec; ++k) {
            const uint32_t vk = uni(R[k].vert);
            const uint64_t sk = uni(R[k].steps);
            // descend(vk), the refinement taken from the record
            ++nodes;
            if (l == 0) seq[seq_size] = (ix_t)vk;
            ++seq_size;
            code = (int)uni(R[k].code);
            ncells = uni(R[k].ncells);
            steps += sk;
            csr_bytes += uni(R[k].csr);
            part_bytes += uni(R[k].part);
            if (budget <= sk) {
                status = GCL_E_INTERNAL;  // step budget exhausted
                return;
            }
            budget -= sk;
            ++level;
            if (level >= levels()) {
                status = GCL_E_CAPACITY;
                return;
            }
            uint64_t* dst = snap + (uint64_t)level * (nw + 1);
            for (uint32_t w = (uint32_t)l; w < nw; w += kWave) dst[w] = LS[(uint64_t)k * nw1 + w];
            if (k + 1 == nrec) {
                for (uint32_t w = (uint32_t)l; w < nw; w += kWave) H.lasts()[w] = LS[(uint64_t)k * nw1 + w];
                H.sync();
                rebuild_from_lasts();
            }
            group_sync();
            const uint64_t tc0 = dg_t0();
            classify();
            dg_add(dg_classify, tc0);
            // backtracked (a leaf, or pruned), failed, or the last recorded level
            if (status != 0 || level != L0 + k + 1 || k + 1 == nrec) return;
            // the run loop's next iteration, and other_path_node at this
            // level: it is a non-leaf not pruned, so it selects its target
            if (budget == 0) {
                status = GCL_E_INTERNAL;
                return;
            }
            --budget;
            if (wall_clock64() > deadline) {
                status = GCL_E_TIMEOUT;
                return;
            }
            const uint32_t tf = uni(R[k].tc_first), tl = uni(R[k].tc_len);
            const uint32_t idx = tc_level_reached;
            if (tl == 0 || idx + 1 != level || uni(R[k + 1].vert) != (uint32_t)R[k].mem[0]) {
                status = GCL_E_INTERNAL;  // not the record of a leftmost descent
                return;
            }
            const uint64_t off = idx == 0 ? 0 : (uint64_t)lv[idx - 1].arena_off + lv[idx - 1].tc_len;
            if (off + tl > arena_cap()) {
                status = GCL_E_CAPACITY;
                return;
            }
            for (uint32_t i = (uint32_t)l; i < tl; i += kWave) {
                arena_mem[off + i] = R[k].mem[i];
                arena_rep[off + i] = R[k].mem[i];
            }
            LevelRec r;
            r.tc_first = tf;
            r.tc_len = tl;
            r.cursor = 0;
            r.arena_off = (uint32_t)off;
            r.last_num_aut = 0;
            lv[idx] = r;
            ++tc_level_reached;
            group_sync();
            if (num_aut > 0) {
                const uint64_t to0 = dg_t0();
                compute_orbits(idx);
                dg_add(dg_orbits, to0);
                lv[idx].last_num_aut = num_aut;
            }
            // the smallest member is its orbit's representative
            lv[idx].cursor = 1;
            group_sync();
        }
    }

    // Records the leftmost descent of item j of a published descent record.
    __device__ void help_desc(const KParams& P, uint8_t* rec, uint32_t j, uint32_t E) {
        const int l = lane_id();
        const DescLayout L = make_desc_layout(P.maxn, P.dspec_items, P.dspec_depth, P.dspec_path);
        const DescHeader* hd = (const DescHeader*)(rec + L.header);
        const uint32_t graph = hd->graph;
        const int64_t vbase = P.vtx_off[graph];
        n = (uint32_t)(P.vtx_off[graph + 1] - vbase);
        nw = (n + 63) / 64;
        rowp = P.row_ptr + vbase;
        col = P.col + rowp[0];
        H.bind(rowp, n, nullptr, P.maxn);
        const uint64_t nw1 = (uint64_t)P.maxn / 64 + 2;
        const uint32_t lev = ((const ix_t*)(rec + L.item_level))[j];
        const uint32_t v = ((const ix_t*)(rec + L.item_vert))[j];
        const ix_t* re = (const ix_t*)(rec + L.elem);
        const uint64_t* rs = (const uint64_t*)(rec + L.snaps) + (uint64_t)lev * nw1;
        for (uint32_t p = (uint32_t)l; p < n; p += kWave) {
            H.elem(p) = re[p];
            H.cnt_st(p, 0);
            H.hits_st(p, 0);
        }
        for (uint32_t w = (uint32_t)l; w < nw; w += kWave) {
            H.lasts()[w] = rs[w];
            H.active()[w] = 0;
            H.vmask()[w] = 0;
            H.split()[w] = 0;
            H.oldend()[w] = 0;
        }
        nactive = 0;
        H.sync();
        positions_from_elem();
        rebuild_from_lasts();
        uint8_t* item = rec + L.payload + (uint64_t)j * L.item_bytes;
        DescLevel* R = (DescLevel*)item;
        uint64_t* LS = (uint64_t*)(item + L.lasts_off);
        ix_t* FE = (ix_t*)(item + L.elem_off);
        budget = step_budget();
        status = 0;
        // level lev's target cell: the first non-singleton cell
        int32_t f = first_set_from(H.nsmask(), 0, nw);
        uint32_t k = 0, flags = 0, vk = v;
        while (f >= 0) {
            const uint64_t s0 = steps, b0 = csr_bytes, p0 = part_bytes;
            individualize_and_refine(vk, (uint32_t)f);
            if (status != 0) {
                flags = kDescFailed;
                break;
            }
            if (l == 0) {
                DescLevel& d = R[k];
                d.vert = vk;
                d.code = (uint32_t)code;
                d.ncells = ncells;
                d.tc_first = 0;
                d.tc_len = 0;
                d.pad = 0;
                d.steps = steps - s0;
                d.csr = csr_bytes - b0;
                d.part = part_bytes - p0;
            }
            for (uint32_t w = (uint32_t)l; w < nw; w += kWave) LS[(uint64_t)k * nw1 + w] = H.lasts()[w];
            ++k;
            if (ncells == n) {
                flags = kDescLeaf;
                break;
            }
            if (k == L.depth) break;
            // the next level's target cell and its members in increasing id
            const int32_t f2 = first_set_from(H.nsmask(), 0, nw);
            if (f2 < 0) break;
            const uint32_t c2 = cend((uint32_t)f2);
            const uint32_t t = uni(c2) - (uint32_t)f2 + 1;
            if (t > kDescMembers) break;
            if ((uint32_t)l < t) set_bit_atomic(H.vmask(), H.elem((uint32_t)f2 + (uint32_t)l));
            H.sync();
            const int32_t m0 = first_set_from(H.vmask(), 0, nw);  // the smallest member
            uint32_t run = 0;
            for (uint32_t base = 0; base < nw; base += kWave) {
                const uint32_t w = base + (uint32_t)l;
                uint64_t bits = w < nw ? H.vmask()[w] : 0ull;
                const uint32_t c0 = (uint32_t)__popcll(bits);
                const uint32_t inc = wave_scan(c0);
                uint32_t q = run + inc - c0;
                while (bits) {
                    R[k - 1].mem[q++] = (uint16_t)(w * 64u + (uint32_t)__builtin_ctzll(bits));
                    bits &= bits - 1;
                }
                if (w < nw) H.vmask()[w] = 0;
                run += readlane(inc, kWave - 1);
            }
            if (l == 0) {
                R[k - 1].tc_first = (uint32_t)f2;
                R[k - 1].tc_len = t;
            }
            H.sync();
            vk = uni((uint32_t)m0);
            f = f2;
        }
        for (uint32_t p = (uint32_t)l; p < n; p += kWave) FE[p] = H.elem(p);
        if (l == 0) ((uint32_t*)(rec + L.nrec))[j] = k | flags;
        __threadfence();
        group_sync();
        if (l == 0) agent_store(&((uint32_t*)(rec + L.state))[j], E | 2u);
        group_sync();
    }

    __device__ bool other_path_node(const KParams& P, uint32_t graph) {
        // other_path_node (reference src/sequential.cpp:304-392)
        if (ncells < n && (eqlevfirst == level || comp_canon >= 0)) {
            if (tc_level_reached < level) {
                if (!select_target()) return false;
            }
            const uint32_t li = level - 1;
            LevelRec r = lv[li];
            if (r.cursor >= r.tc_len) {
                const uint64_t t0 = dg_t0();
                backtrack(level - 1);
                dg_add(dg_backtrack, t0);
                return true;
            }
            if (num_aut > 0 && num_aut > r.last_num_aut) {
                const uint64_t t0 = dg_t0();
                compute_orbits(li);
                dg_add(dg_orbits, t0);
                lv[li].last_num_aut = num_aut;
            }
            // next member that is its own orbit representative
            const int l = lane_id();
            uint32_t cur = r.cursor;
            while (cur < r.tc_len) {
                const uint32_t i = cur + (uint32_t)l;
                const bool ok = i < r.tc_len && arena_rep[r.arena_off + i] == arena_mem[r.arena_off + i];
                const uint64_t m = ballot(ok);
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
            r.cursor = cur;
            // the first path's descents are all behind once the root is reached again
            if constexpr (Hot::kDesc)
                if (level == 1 && (d_slot >= 0 || d_old >= 0)) desc_finish(P);
            if (level == 1 && spec_pending) spec_publish(P, r, graph);
            const uint32_t v = arena_mem[r.arena_off + cur];
            lv[li].cursor = cur + 1;
            const int taken = level == 1 ? spec_take(P, v) : kNotTaken;
            if (taken == kRejected) {
                ++dg_rejected;
                // classified and backtracked already (spec_take)
                ++nodes;
                return true;
            }
            if (taken == kTaken) {
                ++dg_adopted;
                ++nodes;
                seq[seq_size] = (ix_t)v;
                ++seq_size;
                ++level;
                group_sync();
                snapshot(level);
            } else {
                // a descent a helper recorded from this first-path node is
                // replayed (and classified level by level) instead
                if constexpr (Hot::kDesc)
                    if (level >= 2 && d_slot >= 0 && desc_take(P, v) == kTaken) return true;
                descend(v, r.tc_first);
            }
        }
        const uint64_t t0 = dg_t0();
        classify();
        dg_add(dg_classify, t0);
        return true;
    }

    __device__ void run(const KParams& P, uint32_t graph, int64_t vbase) {
        nodes = 0;
        steps = 0;
        csr_bytes = 0;
        part_bytes = 0;
        status = 0;
        {
            // this search's time limit (never past the worker's own end)
            const uint64_t now = wall_clock64();
            const uint64_t nn = (uint64_t)n * (uint64_t)n;
            const uint64_t per = P.graph_ticks_per_n2;
            uint64_t lim = P.graph_ticks_base;
            lim = (per != 0 && nn > (~0ull - lim) / per) ? ~0ull : lim + per * nn;
            deadline = lim > life - min(life, now) ? life : now + lim;
        }
        num_aut = 0;
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
        budget = step_budget();
        spec_slot = -1;
        d_slot = d_old = -1;
        d_renew = false;
        root_changed = false;
        fix_low = 0xFFFFFFFFu;
        // counted in *pending by the host before the launch
        spec_pending = P.spec_min_n > 0 && n >= P.spec_min_n;
        init_partition(P.init_order, P.color, vbase);
        // search_tree_traversal (reference src/sequential.cpp:626-648)
        code = refine(-1);
        level = 1;
        snapshot(1);
        firstcode[0] = code;
        ++nodes;
        group_sync();
        while (level >= 1 && status == 0) {
            if (budget == 0) {
                status = GCL_E_INTERNAL;  // step budget exhausted
                break;
            }
            if (wall_clock64() > deadline) {
                status = GCL_E_TIMEOUT;
                break;
            }
            --budget;
            const bool ok = first_discovered ? other_path_node(P, graph) : first_path_node(P, graph);
            if (!ok) break;
            if constexpr (Hot::kDesc) {
                if (d_renew) {
                    d_renew = false;
                    if (P.dspec && status == 0 && level > 1) desc_republish(P, graph);
                }
            }
        }
        if constexpr (Hot::kDesc) desc_finish(P);
        spec_finish(P);
    }

    // Refines one published root child of another worker's graph.
    __device__ void help_item(const KParams& P, uint8_t* rec, uint32_t j, uint32_t E) {
        const SpecLayout L = make_spec_layout(P.maxn, P.spec_slots);
        const SpecHeader* hd = (const SpecHeader*)(rec + L.header);
        const uint32_t graph = hd->graph;
        const uint32_t tcf = hd->tc_first;
        const int64_t vbase = P.vtx_off[graph];
        n = (uint32_t)(P.vtx_off[graph + 1] - vbase);
        nw = (n + 63) / 64;
        rowp = P.row_ptr + vbase;
        col = P.col + rowp[0];
        H.bind(rowp, n, nullptr, P.maxn);
        const ix_t* re = (const ix_t*)(rec + L.root_elem);
        const uint64_t* rl = (const uint64_t*)(rec + L.root_lasts);
        for (uint32_t p = (uint32_t)lane_id(); p < n; p += kWave) {
            H.elem(p) = re[p];
            H.cnt_st(p, 0);
            H.hits_st(p, 0);
        }
        for (uint32_t w = (uint32_t)lane_id(); w < nw; w += kWave) {
            H.lasts()[w] = rl[w];
            H.active()[w] = 0;
            H.vmask()[w] = 0;
            H.split()[w] = 0;
            H.oldend()[w] = 0;
        }
        nactive = 0;
        H.sync();
        positions_from_elem();
        rebuild_from_lasts();
        const uint32_t v = ((const ix_t*)(rec + L.vert))[j];
        const uint64_t s0 = steps, b0 = csr_bytes, pb0 = part_bytes;
        budget = step_budget();
        status = 0;
        individualize_and_refine(v, tcf);
        // The owner classifies a root child whose code is below the best
        // path's at level 2 and differs from the first path's as kBad from
        // the code alone (spec_take).  The best path's code only grows, so a
        // child below the value read here is rejected whenever the owner
        // takes it: its partition is not written (most root children of a
        // graph without automorphisms are such children).
        const int32_t f1 = uni(hd->first1);
        const int32_t c1 = uni((int32_t)__hip_atomic_load(&hd->canon1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        const bool code_only = status == 0 && code != f1 && code < c1;
        if (!code_only) {
            // j < count, which the publisher bounded by the record's capacity for n
            uint8_t* item = rec + L.payload + (uint64_t)j * spec_item_bytes(n);
            ix_t* oe = (ix_t*)item;
            uint64_t* ol = (uint64_t*)(item + spec_item_lasts(n));
            for (uint32_t p = (uint32_t)lane_id(); p < n; p += kWave) oe[p] = H.elem(p);
            for (uint32_t w = (uint32_t)lane_id(); w < nw; w += kWave) ol[w] = H.lasts()[w];
        }
        if (lane_id() == 0) {
            // a failed refinement is handed to the owner as a code no
            // refinement produces (codes are below 2^15)
            ((uint32_t*)(rec + L.code))[j] = status ? kFailedCode : ((uint32_t)code | (code_only ? kCodeOnly : 0u));
            ((uint32_t*)(rec + L.ncells))[j] = ncells;
            uint64_t* cr = (uint64_t*)(rec + L.steps) + 4ull * j;
            cr[0] = steps - s0;
            cr[1] = csr_bytes - b0;
            cr[2] = part_bytes - pb0;
        }
        __threadfence();
        group_sync();
        if (lane_id() == 0) agent_store(&((uint32_t*)(rec + L.state))[j], E | 2u);
        group_sync();
    }
};

// Tries to claim and record one published descent.  Returns 1 when it did,
// 0 when descents are published but none could be claimed now, -1 when none
// is published.
template <class Hot>
__device__ __forceinline__ int help_desc_once(Worker<Hot>& W, const KParams& P) {
    uint32_t claim = 0xFFFFFFFFu, E = 0, slot_sel = 0, seen = 0;
    if (lane_id() == 0) {
        const DescLayout L = make_desc_layout(P.maxn, P.dspec_items, P.dspec_depth, P.dspec_path);
        const uint32_t nwr = (P.dspec_records + 63) / 64;
        for (uint32_t w = 0; w < nwr && claim == 0xFFFFFFFFu; ++w) {
            const uint64_t bits = __hip_atomic_load(&P.dboard[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            seen |= bits != 0ull;
            for (uint64_t b = bits; b && claim == 0xFFFFFFFFu; b &= b - 1) {
                const uint32_t slot = w * 64u + (uint32_t)__builtin_ctzll(b);
                uint8_t* rec = P.dspec + (uint64_t)slot * P.dspec_stride;
                DescHeader* hd = (DescHeader*)(rec + L.header);
                // the claim below carries the epoch, which validates these reads
                const uint32_t e = agent_peek(&hd->epoch) << 2;
                const uint32_t cnt = __hip_atomic_load(&hd->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (__hip_atomic_load(&hd->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cnt) continue;
                const uint32_t idx = atomicAdd(&hd->next, 1u);
                if (idx < cnt && idx < L.items && agent_cas(&((uint32_t*)(rec + L.state))[idx], e | 0u, e | 1u)) {
                    claim = idx;
                    E = e;
                    slot_sel = slot;
                }
            }
        }
    }
    claim = uni(readlane(claim, 0));
    seen = uni(readlane(seen, 0));
    if (claim == 0xFFFFFFFFu) return seen ? 0 : -1;
    E = uni(readlane(E, 0));
    slot_sel = uni(readlane(slot_sel, 0));
    W.help_desc(P, P.dspec + (uint64_t)slot_sel * P.dspec_stride, claim, E);
    return 1;
}

// Tries to claim and refine one published descent or root child.  Returns 1
// when it did some work, 0 when there was nothing to take now, -1 when
// helping is over.
template <class Hot>
__device__ __forceinline__ int help_once(Worker<Hot>& W, const KParams& P) {
    const int l = lane_id();
    // descents first: the owner of a first path waits for them deepest first
    int dres = -1;
    if constexpr (Hot::kDesc) {
        if (P.dspec) {
            dres = help_desc_once(W, P);
            if (dres > 0) return 1;
        }
    }
    const uint32_t nwr = (P.spec_records + 63) / 64;
    const uint64_t word = (uint32_t)l < nwr
                              ? __hip_atomic_load(&P.board_mask[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : 0ull;
    const uint64_t nz = ballot(word != 0ull);
    if (!nz) {
        uint32_t pend = 0, again = 0;
        if (l == 0) {
            pend = __hip_atomic_load(P.pending, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            // a graph publishes before it leaves the count: with the count
            // at 0, a second look at the board sees every record published
            if (pend == 0)
                for (uint32_t w = 0; w < nwr && !again; ++w)
                    again = __hip_atomic_load(&P.board_mask[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull;
        }
        pend = __shfl(pend, 0, kWave);
        again = __shfl(again, 0, kWave);
        return pend == 0 && !again && dres < 0 ? -1 : 0;
    }
    const SpecLayout L = make_spec_layout(P.maxn, P.spec_slots);
    // spread helpers: start at a word and bit that depend on the worker
    const uint32_t rot = (blockIdx.x * 2654435761u) >> 16;
    uint32_t claim = 0xFFFFFFFFu, E = 0, slot_sel = 0;
    const uint32_t nnz_words = (uint32_t)__popcll(nz);
    for (uint32_t t = 0; t < nnz_words && claim == 0xFFFFFFFFu; ++t) {
        // t-th nonzero word after the rotation
        uint64_t m = nz;
        const uint32_t skip = (rot + t) % nnz_words;
        for (uint32_t k = 0; k < skip; ++k) m &= m - 1;
        const int wl = __builtin_ctzll(m);
        const uint64_t bits = ((uint64_t)readlane((uint32_t)(word >> 32), wl) << 32) | readlane((uint32_t)word, wl);
        if (l == 0) {
            const uint32_t r2 = rot % 64u;
            const uint64_t rb = (bits >> r2) | (r2 ? (bits << (64 - r2)) : 0ull);
            for (uint64_t b = rb; b && claim == 0xFFFFFFFFu; b &= b - 1) {
                const uint32_t slot = (uint32_t)wl * 64u + ((uint32_t)__builtin_ctzll(b) + r2) % 64u;
                uint8_t* rec = P.spec + (uint64_t)slot * P.spec_stride;
                SpecHeader* hd = (SpecHeader*)(rec + L.header);
                // the state word carries the epoch, so the claim below
                // (acquire) validates this relaxed read
                const uint32_t e = agent_peek(&hd->epoch) << 2;
                const uint32_t cnt = __hip_atomic_load(&hd->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (__hip_atomic_load(&hd->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cnt) continue;
                const uint32_t idx = atomicAdd(&hd->next, 1u);
                if (idx < cnt && idx < L.slots && agent_cas(&((uint32_t*)(rec + L.state))[idx], e | 0u, e | 1u)) {
                    claim = idx;
                    E = e;
                    slot_sel = slot;
                }
            }
        }
        claim = __shfl(claim, 0, kWave);
    }
    if (claim == 0xFFFFFFFFu) return 0;
    E = __shfl(E, 0, kWave);
    slot_sel = __shfl(slot_sel, 0, kWave);
    W.help_item(P, P.spec + (uint64_t)slot_sel * P.spec_stride, claim, E);
    return 1;
}

// help_once, with its time and the children refined counted for diagnostics.
template <class Hot>
__device__ __forceinline__ int help_counted(Worker<Hot>& W, const KParams& P) {
    const uint64_t t0 = wall_clock64();
    const int h = help_once(W, P);
    if (h > 0) {
        W.dg_help += wall_clock64() - t0;
        ++W.dg_helped;
    }
    return h;
}

template <class Hot>
__device__ void bind_scratch(Worker<Hot>& W, const KParams& P, uint8_t* base, const Layout& L) {
    W.kp = &P;
    W.snap = (uint64_t*)(base + L.snap);
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
    W.bind_auts(P.aut_pool ? nullptr : base + L.auts);
    W.H.bind_lists(P.lists ? P.lists + (uint64_t)blockIdx.x * P.lists_stride : base + L.lists, P.maxn);
    const uint64_t now = wall_clock64();
    W.life = P.time_limit > ~0ull - now ? ~0ull : now + P.time_limit;
    W.deadline = W.life;
    W.steps = 0;
    W.dg_nodes = W.dg_steps = W.dg_helped = W.dg_search = W.dg_wait = W.dg_help = W.dg_idle = 0;
    W.dg_refine = W.dg_classify = W.dg_orbits = W.dg_backtrack = 0;
    W.dg_adopted = W.dg_rejected = 0;
    W.dg_on = P.diag != nullptr;
    W.dg_graphs = 0;
}

// Writes this worker's diagnostics record (lane 0), when the launch has them.
template <class Hot>
__device__ void write_diag(Worker<Hot>& W, const KParams& P, uint64_t t_start) {
    if (P.diag == nullptr || lane_id() != 0) return;
    DiagRecord r;
    r.hw_id = __smid();
    r.graphs = W.dg_graphs;
    r.nodes = W.dg_nodes;
    r.steps = W.dg_steps;
    r.helped = W.dg_helped;
    r.t_total = wall_clock64() - t_start;
    r.t_search = W.dg_search;
    r.t_wait = W.dg_wait;
    r.t_help = W.dg_help;
    r.t_idle = W.dg_idle;
    r.t_refine = W.dg_refine;
    r.t_classify = W.dg_classify;
    r.t_orbits = W.dg_orbits;
    r.t_backtrack = W.dg_backtrack;
    r.adopted = W.dg_adopted;
    r.rejected = W.dg_rejected;
    P.diag[blockIdx.x] = r;
}

// Searches graph g (vertices vtx_off[g] .. vtx_off[g+1]-1) and writes its
// labels and result record.
template <class Hot>
__device__ __forceinline__ void process_graph(Worker<Hot>& W, const KParams& P, int32_t g) {
    const int l = lane_id();
    const int64_t vbase = P.vtx_off[g];
    W.n = (uint32_t)(P.vtx_off[g + 1] - vbase);
    W.nw = (W.n + 63) / 64;
    W.rowp = P.row_ptr + vbase;
    W.col = P.col + W.rowp[0];
    W.nnz = (uint64_t)(W.rowp[W.n] - W.rowp[0]);
    W.H.bind(W.rowp, W.n, nullptr, P.maxn);
    W.H.sync();
    const uint64_t t0 = wall_clock64();
    W.run(P, (uint32_t)g, vbase);
    const uint64_t t1 = wall_clock64();
    W.dg_search += t1 - t0;
    W.dg_nodes += W.nodes;
    W.dg_steps += W.steps;
    ++W.dg_graphs;
    for (uint32_t i = (uint32_t)l; i < W.n; i += kWave) P.out_label[vbase + i] = (int32_t)W.best_label[i];
    if (l == 0) {
        DeviceGraphResult res;
        res.nodes = W.nodes;
        res.automorphisms = (uint64_t)W.num_aut;
        res.splitter_steps = W.steps;
        res.csr_bytes = W.csr_bytes;
        res.part_bytes = W.part_bytes;
        res.device_ns = t1 - t0;
        res.status = W.status;
        res.pad = 0;
        P.out[g] = res;
    }
    group_sync();
}

template <class Hot>
__device__ __forceinline__ void worker_loop(Worker<Hot>& W, const KParams& P) {
    const int l = lane_id();
    const uint64_t t_start = wall_clock64();
    bool graphs_left = true;
    uint32_t idle = 1;
    while (true) {
        // help-first: root children of a published large graph are on the
        // batch's critical path; a new graph is taken when there are none
        if (P.spec_min_n && help_counted(W, P) > 0) continue;
        if (graphs_left) {
            uint32_t item = 0;
            if (l == 0) item = atomicAdd(P.next_item, 1u);
            item = __shfl(item, 0, kWave);
            if (item < (uint32_t)P.num_items) {
                if (P.aut_pool) W.bind_auts(P.aut_pool + (uint64_t)item * P.aut_stride);
                process_graph(W, P, P.items[item]);
                continue;
            }
            graphs_left = false;
        }
        if (P.spec_min_n == 0 || wall_clock64() > W.life) break;
        const int h = help_counted(W, P);
        if (h < 0) break;
        if (h == 0) {
            // idle: back off (up to ~100 us) to keep the polling traffic low
            const uint64_t ti = wall_clock64();
            for (uint32_t i = 0; i < idle; ++i) __builtin_amdgcn_s_sleep(127);
            W.dg_idle += wall_clock64() - ti;
            idle = idle < 32 ? idle * 2 : 32;
        } else {
            idle = 1;
        }
    }
    write_diag(W, P, t_start);
}

// Poll of host-coherent memory.  Relaxed: an acquire at system scope would
// invalidate the caches that the searches of the other workers depend on, and
// the request data read after a poll is itself uncached host memory, read
// only once the poll's value is known.
__device__ __forceinline__ uint32_t sys_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Lane 0: reports slot as done.  Its labels and result record were written
// and made visible to the host (__threadfence_system) before.  The host reads
// entries in order; an entry whose sequence number is not the next one it
// expects has not been written yet.
__device__ __forceinline__ void serve_complete(const ServeParams& S, uint32_t slot) {
    const uint32_t c = atomicAdd(S.cq_count, 1u);
    __hip_atomic_store(&S.cq[c & S.cq_mask], ((unsigned long long)(c + 1u) << 32) | slot, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

// Serving loop: claims published requests of this class until the host
// stops the queue (or the worker's deadline passes).  Root children of large
// graphs published by other workers are helped first, as in worker_loop.
template <class Hot>
__device__ __forceinline__ void serve_loop(Worker<Hot>& W, const KParams& P) {
    const int l = lane_id();
    const ServeParams& S = P.sv;
    uint32_t idle = 1;
    while (true) {
        if (P.spec_min_n && help_once(W, P) > 0) {
            idle = 1;
            continue;
        }
        uint32_t slot = 0xFFFFFFFFu, stop = 0, retry = 0;
        if (l == 0) {
            const uint32_t h = __hip_atomic_load(P.next_item, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t t = __hip_atomic_load(S.tail_seen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (h >= t) {
                // the queue looks empty: refresh the mirror from the host if
                // nobody did within poll_ticks.  One worker per period reads
                // host memory, and the lease's CAS is only tried once the
                // period has passed: thousands of idle workers polling over
                // PCIe, or all trying the CAS, would hold up the memory
                // requests of the workers that are searching.
                const uint64_t now = wall_clock64();
                unsigned long long due = __hip_atomic_load(S.poll_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (now >= due && __hip_atomic_compare_exchange_strong(S.poll_at, &due, now + S.poll_ticks,
                                                                       __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_AGENT)) {
                    const uint32_t hs = sys_load(S.stop);
                    // read after the stop flag: the host publishes before it
                    // stops, so a stop read here comes with the final tail
                    const uint32_t ht = sys_load(S.tail);
                    if (ht > t) {
                        __hip_atomic_fetch_max(S.tail_seen, ht, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        t = ht;
                    }
                    if (hs) __hip_atomic_store(S.stop_seen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            if (h < t) {
                uint32_t expect = h;
                if (__hip_atomic_compare_exchange_strong(P.next_item, &expect, h + 1, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    slot = sys_load(&S.ring[h & S.ring_mask]);
                else
                    retry = 1;  // another worker took it: try again at once
            } else if (__hip_atomic_load(S.stop_seen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) {
                // the tail stored before the stop flag is visible now
                stop = __hip_atomic_load(S.tail_seen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= h ? 1u : 0u;
                retry = 1u - stop;
            }
        }
        slot = uni(readlane(slot, 0));
        if (uni(readlane(retry, 0))) continue;
        if (slot != 0xFFFFFFFFu && slot < S.slots) {
            // copy the request's graph into the slot's device arrays (the
            // host memory is read with vector loads: the scalar cache is not
            // kept coherent with host writes while the kernel runs)
            const uint64_t n1 = (uint64_t)P.maxn + 1;
            const int32_t* hr = S.h_row + (uint64_t)slot * n1;
            const uint32_t n = uni((uint32_t)__hip_atomic_load(&S.h_n[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
            const uint32_t nnz = n <= P.maxn
                                     ? uni((uint32_t)__hip_atomic_load(&hr[n], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
                                     : 0xFFFFFFFFu;
            if (n > P.maxn || nnz > P.max_nnz) {  // not a request this kernel can hold
                if (l == 0) {
                    DeviceGraphResult res{};
                    res.status = GCL_E_INVALID;
                    P.out[2 * slot] = res;
                    __threadfence_system();
                    serve_complete(S, slot);
                }
                continue;
            }
            const int64_t vb = (int64_t)((uint64_t)slot * n1);
            const int64_t cb = (int64_t)((uint64_t)slot * P.max_nnz);
            int64_t* drow = const_cast<int64_t*>(P.row_ptr) + vb;
            for (uint32_t i = (uint32_t)l; i <= n; i += kWave) drow[i] = cb + hr[i];
            const int32_t* hc = S.h_col + (uint64_t)slot * P.max_nnz;
            int32_t* dcol = const_cast<int32_t*>(P.col) + cb;
            for (uint32_t e = (uint32_t)l; e < nnz; e += kWave) dcol[e] = hc[e];
            const bool colored =
                uni((uint32_t)__hip_atomic_load(&S.h_colored[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) != 0u;
            int32_t* dcolor = const_cast<int32_t*>(P.color) + vb;
            int32_t* dinit = const_cast<int32_t*>(P.init_order) + vb;
            const int32_t* hcol = S.h_color + (uint64_t)slot * P.maxn;
            const int32_t* hini = S.h_init + (uint64_t)slot * P.maxn;
            for (uint32_t i = (uint32_t)l; i < n; i += kWave) {
                dcolor[i] = colored ? hcol[i] : 0;
                dinit[i] = colored ? hini[i] : (int32_t)i;
            }
            int64_t* dvtx = const_cast<int64_t*>(P.vtx_off);
            if (l == 0) {
                dvtx[2 * slot] = vb;
                dvtx[2 * slot + 1] = vb + (int64_t)n;
            }
            __threadfence();
            group_sync();
            process_graph(W, P, (int32_t)(2 * slot));
            // labels and record reach host memory before the done flag
            __threadfence_system();
            group_sync();
            if (l == 0) serve_complete(S, slot);
            idle = 1;
            continue;
        }
        if (uni(readlane(stop, 0)) || wall_clock64() > W.life) break;
        // nothing to do: back off (up to ~0.9 ms).  With thousands of idle
        // workers a request or a published root child is still seen within
        // microseconds, and their polls (loads, and the lease's CAS on one
        // L2 line) stay too rare to delay the memory requests of the
        // workers that are searching.
        for (uint32_t i = 0; i < idle; ++i) __builtin_amdgcn_s_sleep(127);
        idle = idle < 256 ? idle * 2 : 256;
    }
}

// The worker object lives in LDS: its uniform fields are reloaded where used
// instead of occupying registers for the whole persistent loop.  The launch
// parameters are read from device memory through a pointer (scalar loads, as
// from the kernel-argument segment), so that a search function the compiler
// keeps out of line can take them by reference without a private copy.  The
// classes up to 512 vertices are compiled for 4 waves per SIMD (at most 128
// VGPRs; the allocator spills about 30 registers of long-lived values, and
// the extra waves more than pay for it: batch step 1.53 -> 1.50 s); class 512
// is then bound by LDS to 14 waves per CU, the larger classes to fewer.
template <class Hot>
__device__ __forceinline__ void search_body(const KParams& P) {
    __shared__ Worker<Hot> W;
    uint8_t* base = P.scratch + (uint64_t)blockIdx.x * P.scratch_stride;
    const Layout L = make_layout(P.maxn, P.max_nnz, P.ring, P.arena_cap, P.levels, !Hot::kLds, P.aut_pool != nullptr);
    bind_scratch(W, P, base, L);
    if constexpr (!Hot::kLds) {
        W.H.h16 = (ix_t*)(base + L.hot16);
        W.H.h32 = (uint32_t*)(base + L.hot32);
        W.H.h64 = (uint64_t*)(base + L.hot64);
        W.H.n1 = hot_stride(P.maxn);
        W.H.nw1 = (uint64_t)P.maxn / 64 + 2;
    }
    worker_loop(W, P);
}

template <int MAXN>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(MAXN <= 512 ? 4 : (MAXN <= 1024 ? 2 : 1))))
void search_kernel_lds(const KParams* __restrict__ Pp) {
    search_body<HotL<MAXN>>(*Pp);
}

// The same with descent speculation (launches of a few large graphs).
template <int MAXN>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(MAXN <= 1024 ? 2 : 1)))
void search_kernel_lds_desc(const KParams* __restrict__ Pp) {
    search_body<WithDesc<HotL<MAXN>>>(*Pp);
}

// Serving kernel of the same size class: runs until the host stops its queue.
template <int MAXN>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(MAXN <= 512 ? 4 : (MAXN <= 1024 ? 2 : 1))))
void serve_kernel_lds(const KParams* __restrict__ Pp) {
    const KParams& P = *Pp;
    __shared__ Worker<HotL<MAXN>> W;
    uint8_t* base = P.scratch + (uint64_t)blockIdx.x * P.scratch_stride;
    const Layout L = make_layout(P.maxn, P.max_nnz, P.ring, P.arena_cap, P.levels, false, P.aut_pool != nullptr);
    bind_scratch(W, P, base, L);
    serve_loop(W, P);
}

__global__ __launch_bounds__(kWave) void search_kernel_global(const KParams* __restrict__ Pp) {
    search_body<HotG>(*Pp);
}

__global__ __launch_bounds__(kWave) void search_kernel_global_desc(const KParams* __restrict__ Pp) {
    search_body<WithDesc<HotG>>(*Pp);
}

__global__ __launch_bounds__(kWave) void serve_kernel_global(const KParams* __restrict__ Pp) {
    const KParams& P = *Pp;
    __shared__ Worker<HotG> W;
    uint8_t* base = P.scratch + (uint64_t)blockIdx.x * P.scratch_stride;
    const Layout L = make_layout(P.maxn, P.max_nnz, P.ring, P.arena_cap, P.levels, true, P.aut_pool != nullptr);
    bind_scratch(W, P, base, L);
    W.H.h16 = (ix_t*)(base + L.hot16);
    W.H.h32 = (uint32_t*)(base + L.hot32);
    W.H.h64 = (uint64_t*)(base + L.hot64);
    W.H.n1 = hot_stride(P.maxn);
    W.H.nw1 = (uint64_t)P.maxn / 64 + 2;
    serve_loop(W, P);
}

}  // namespace dev
}  // namespace gcl

// gcl_device.h -- host-side interface of the device engine (internal).
#pragma once

#include <string>

#include "gcl_internal.h"

namespace gcl {

// Per-worker diagnostics of one launch, written when the launch is given a
// record array (side files): where the worker ran and how its time went.
// Times are device wall-clock ticks.
struct DiagRecord {
    uint32_t hw_id;    // __smid(): XCC in bits 9:6, shader engine in 5:4, CU in 3:0 (gfx950)
    uint32_t graphs;   // graphs this worker searched
    uint64_t nodes;    // their search-tree nodes
    uint64_t steps;    // their splitter steps (with root children refined for them)
    uint64_t helped;   // root children this worker refined for other workers
    uint64_t t_total;  // whole loop
    uint64_t t_search; // searching its own graphs (t_wait included)
    uint64_t t_wait;   // within t_search: waiting for root children from helpers
    uint64_t t_help;   // refining root children for other workers
    uint64_t t_idle;   // backing off with nothing to do
    // within t_search: individualize + refine, node classification (leaf
    // certificates, automorphisms, the backtracking it decides), orbit
    // updates, and backtracking past an exhausted target cell
    uint64_t t_refine, t_classify, t_orbits, t_backtrack;
    uint32_t adopted;   // root children taken refined from a helper
    uint32_t rejected;  // root children rejected by the helper's code alone
};

// Per-graph result of a device launch.
struct DeviceGraphResult {
    uint64_t nodes;
    uint64_t automorphisms;
    uint64_t splitter_steps;
    uint64_t csr_bytes;
    uint64_t device_ns;  // wall time of the graph's search on the device
    uint64_t part_bytes; // algorithmic bytes of refinement beyond csr_bytes
    int32_t status;      // 0 ok, otherwise a gcl_status
    int32_t pad;
};

// Canonical labeling of every graph of a flat batch on device 0.
//   out_label: vtx_off[G] ints (label[i] = vertex at canonical position i)
//   results:   G entries or NULL
//   ring_cap:  automorphism slots per worker; the search runs with
//              min(max_aut, ring_cap) slots (the canonical form does not
//              depend on it, only the node count does)
gcl_status gpu_canon_batch(const gcl_batch& b, uint32_t max_aut, uint32_t ring_cap, int32_t* out_label,
                           DeviceGraphResult* results, gcl_batch_stats* st);

// Batch resident on the device (opaque); see gcl_device_batch_* in gcl.h.
struct DeviceBatch;
// Launch choices of a device batch: graphs with n >= spec_min_n publish their
// root children (0: no speculation); graphs below min_class_n run in the
// min_class_n size class (fewer, larger classes).
struct BatchTuning {
    uint32_t spec_min_n;
    uint32_t min_class_n;
    uint32_t keep_auts = 0;  // one automorphism ring per graph, readable after the run
};
BatchTuning default_batch_tuning();
// Automorphisms stored for graph g in the last run of a batch created with
// keep_auts: min(found, ring) rows of n images (row-major, a[v] per vertex v)
// into out (cap_rows rows at most); *rows receives the rows stored.
gcl_status device_batch_automorphisms(DeviceBatch* db, int32_t g, int32_t* out, int64_t cap_rows, int64_t* rows);
gcl_status device_batch_create(const gcl_batch& b, uint32_t max_aut, uint32_t ring_cap, DeviceBatch** out,
                               const BatchTuning* tuning = nullptr);
// Times every candidate tuning on an evenly strided sample of at most
// `sample` graphs of b (one untimed run, then the best of two) and returns the
// fastest; `tried` (optional) receives the number of candidates.
gcl_status batch_autotune(const gcl_batch& b, uint32_t max_aut, uint32_t ring_cap, int32_t sample, BatchTuning* best,
                          double* best_ms, int32_t* tried);
gcl_status device_batch_run(DeviceBatch* db, bool use_graph, float* ms);
gcl_status device_batch_results(DeviceBatch* db, int32_t* out_label, DeviceGraphResult* res, gcl_batch_stats* st);
gcl_status device_batch_class_stats(DeviceBatch* db, gcl_class_stats* out, int32_t cap, int32_t* count);
void device_batch_destroy(DeviceBatch* db);
int32_t device_batch_size(const DeviceBatch* db);

// Persistent serving kernels (one per size class) and their request slots.
struct ServeConfig {
    int32_t max_n = 4096;             // largest graph
    int64_t max_graph_nnz = 1 << 20;  // directed edges of the largest graph
    uint32_t max_aut = 10000;
    uint32_t ring_cap = 64;           // automorphism slots per worker
    int32_t slots = 0;                // request slots per class (0: twice the workers)
    uint64_t lifetime_ticks = ~0ull;  // device wall-clock ticks a worker serves at most
    // share of the chip's worker slots for the class above 1024 vertices
    // (when max_n > 1024); the first class gets 0.9 minus this
    double large_share = 0.4;
    uint64_t scratch_bytes_per_class = 2ull << 30;
    // the large class's workers hold snapshots and arenas sized by max_n
    // (about 55 MB each at 50,000 vertices)
    uint64_t scratch_bytes_large = 32ull << 30;
    // the large class keeps up to 32 records (large graphs publishing at
    // once), each holding every root child of a 10,000-vertex graph
    uint64_t spec_bytes_large = 16ull << 30;
    // speculation records: one per worker, each large enough for every root
    // child of a class-sized graph (a record that holds fewer leaves the rest
    // to the owning worker alone: n = 1024 needs ~4.3 GiB over 1,843 records)
    uint64_t spec_bytes_per_class = 8ull << 30;
    // a server is mostly idle between arrivals, so graphs from n = 32 on
    // spread their root children over the idle workers (latency); a batch
    // keeps every worker busy and publishes from n = 128 on
    uint32_t spec_min_n = 32;
};
struct ServeClassInfo {
    uint32_t maxn;
    uint64_t max_nnz;
    uint32_t slots;
    int64_t workers;
};
struct ServeSlotView {  // host-coherent arrays of one request slot
    int32_t* n;
    int32_t* row;       // maxn + 1 local offsets
    int32_t* col;
    int32_t* color;
    int32_t* init;
    int32_t* colored;
    const int32_t* label;
    const DeviceGraphResult* out;
};
struct ServeEngine;
gcl_status serve_create(const ServeConfig& c, ServeEngine** out);
void serve_destroy(ServeEngine* s);  // stops the kernels (after the queues drain) and frees
int serve_classes(const ServeEngine* s);
ServeClassInfo serve_class_info(const ServeEngine* s, int k);
int serve_class_of(const ServeEngine* s, int64_t n, int64_t nnz);  // -1: too large
ServeSlotView serve_slot(ServeEngine* s, int k, uint32_t slot);
void serve_publish(ServeEngine* s, int k, uint32_t slot);  // slot data written before
// Slots of class k completed since the last call, in completion order (up to
// cap); their labels and result records are readable.  One caller at a time.
int serve_completions(ServeEngine* s, int k, uint32_t* out, int cap);
void serve_stop(ServeEngine* s);
gcl_status serve_wait(ServeEngine* s);
// GCL_E_DEVICE once a serving kernel has failed, GCL_E_TIMEOUT once one has
// returned because its lifetime passed
gcl_status serve_check(ServeEngine* s);

// Canonical labeling of one graph (n <= 1024, nnz <= 32768) with no device
// allocation: every device buffer lives in a static device array (see
// gcl_canon_label_noalloc in gcl.h).  out_label: n ints (host).
gcl_status noalloc_canon(const gcl_csr& g, uint32_t max_aut, int32_t* out_label, DeviceGraphResult* res);

int32_t device_count();
uint64_t device_wall_ticks_per_ms();
uint64_t step_budget();
// Side files written after each search launch (empty: none); see
// gcl_set_side_files in gcl.h.
void set_side_file_prefix(const char* prefix);
std::string side_file_prefix();
// Smallest graph whose root children are published to helpers
// (GCL_SPEC_MIN_N overrides; 0 disables).
uint32_t spec_min_n(uint32_t dflt = 128);
// Wall-clock limits in device ticks: a launch (or a serving worker) runs at
// most `launch` ticks, a search at most base + per_n2 * n^2.
struct TimeLimits {
    uint64_t launch = ~0ull, base = ~0ull, per_n2 = 0;
};
// GCL_TIME_LIMIT_MS, when set, bounds the launch and each search by that
// many milliseconds (0: no limit).  Otherwise a search may take
// 60 s + n^2 / 10^4 ms and a launch 60 s + the sum of its graphs' n^2 /
// 10^4 ms (sum_n2).
TimeLimits time_limits(double sum_n2);

}  // namespace gcl

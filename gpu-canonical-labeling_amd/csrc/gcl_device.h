// gcl_device.h -- host-side interface of the device engine (internal).
#pragma once

#include "gcl_internal.h"

namespace gcl {

// Per-graph result of a device launch.
struct DeviceGraphResult {
    uint64_t nodes;
    uint64_t automorphisms;
    uint64_t splitter_steps;
    uint64_t csr_bytes;
    uint64_t device_ns;  // wall time of the graph's search on the device
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
gcl_status device_batch_create(const gcl_batch& b, uint32_t max_aut, uint32_t ring_cap, DeviceBatch** out);
gcl_status device_batch_run(DeviceBatch* db, bool use_graph, float* ms);
gcl_status device_batch_results(DeviceBatch* db, int32_t* out_label, DeviceGraphResult* res, gcl_batch_stats* st);
gcl_status device_batch_class_stats(DeviceBatch* db, gcl_class_stats* out, int32_t cap, int32_t* count);
void device_batch_destroy(DeviceBatch* db);
int32_t device_batch_size(const DeviceBatch* db);

// Launch slot of the server: fixed device buffers and pinned staging.
struct SlotConfig {
    int32_t max_graphs = 0;       // graphs per launch
    int64_t max_vertices = 0;     // sum of n per launch
    int64_t max_nnz = 0;          // sum of directed edges per launch
    int32_t max_n = 0;            // largest graph
    int64_t max_graph_nnz = 0;    // directed edges of the largest graph
    uint32_t max_aut = 1;
    uint32_t ring_cap = 64;       // automorphism slots per worker
    int32_t share = 2;            // launches meant to be resident at once
    int32_t global_workers = 256; // workers of the n > 4096 class
    uint64_t scratch_bytes_per_class = 2ull << 30;
    uint64_t spec_bytes_per_class = 1ull << 30;
};
struct SlotStaging {  // pinned host arrays of a slot
    int64_t* vtx;
    int64_t* row;
    int32_t* col;
    int32_t* color;
    int32_t* label;
    DeviceGraphResult* out;
};
struct DeviceSlot;
gcl_status slot_create(const SlotConfig& c, DeviceSlot** out);
void slot_destroy(DeviceSlot* s);
SlotStaging slot_staging(DeviceSlot* s);
int slot_class_of(const DeviceSlot* s, int64_t n, uint64_t nnz);  // -1 when the graph does not fit
gcl_status slot_launch(DeviceSlot* s, int32_t graphs, uint32_t max_aut, bool colored);  // staged batch
gcl_status slot_poll(DeviceSlot* s, bool* finished);

int32_t device_count();
uint64_t step_budget();
uint32_t spec_min_n();
uint64_t time_limit_ticks();

}  // namespace gcl

// gcl_device.hip -- device engine of libgcl: launch sizing, device-resident
// batches and hipGraph capture around the search worker of gcl_worker.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "gcl_device.h"
#include "gcl_search.h"
#include "gcl_worker.h"

namespace gcl {
namespace {

using namespace dev;

#define GCL_HIP(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            return fail(GCL_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));     \
        }                                                                                 \
    } while (0)

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

constexpr int kClasses = 8;
// classes of at most this many graphs get descent records (two per graph)
constexpr uint32_t kDescGraphs = 8;
constexpr uint32_t kClassN[kClasses] = {64, 128, 256, 512, 1024, 2048, 4096, 8192};

template <int MAXN>
hipError_t launch_lds(const KParams* P, int grid, hipStream_t s) {
    hipLaunchKernelGGL(search_kernel_lds<MAXN>, dim3(grid), dim3(kWave), 0, s, P);
    return hipGetLastError();
}

template <int MAXN>
hipError_t launch_lds_desc(const KParams* P, int grid, hipStream_t s) {
    hipLaunchKernelGGL(search_kernel_lds_desc<MAXN>, dim3(grid), dim3(kWave), 0, s, P);
    return hipGetLastError();
}

// Classes from 1024 vertices up have a kernel with descent speculation.
constexpr int kDescMinClass = 4;

hipError_t launch_class(int cls, bool desc, const KParams* P, int grid, hipStream_t s) {
    if (desc && cls >= kDescMinClass) {
        switch (cls) {
            case 4: return launch_lds_desc<1024>(P, grid, s);
            case 5: return launch_lds_desc<2048>(P, grid, s);
            case 6: return launch_lds_desc<4096>(P, grid, s);
            case 7: return launch_lds_desc<8192>(P, grid, s);
            default:
                hipLaunchKernelGGL(search_kernel_global_desc, dim3(grid), dim3(kWave), 0, s, P);
                return hipGetLastError();
        }
    }
    switch (cls) {
        case 0: return launch_lds<64>(P, grid, s);
        case 1: return launch_lds<128>(P, grid, s);
        case 2: return launch_lds<256>(P, grid, s);
        case 3: return launch_lds<512>(P, grid, s);
        case 4: return launch_lds<1024>(P, grid, s);
        case 5: return launch_lds<2048>(P, grid, s);
        case 6: return launch_lds<4096>(P, grid, s);
        case 7: return launch_lds<8192>(P, grid, s);
        default:
            hipLaunchKernelGGL(search_kernel_global, dim3(grid), dim3(kWave), 0, s, P);
            return hipGetLastError();
    }
}

const void* class_kernel(int cls, bool desc = false) {
    if (desc && cls >= kDescMinClass) {
        switch (cls) {
            case 4: return (const void*)search_kernel_lds_desc<1024>;
            case 5: return (const void*)search_kernel_lds_desc<2048>;
            case 6: return (const void*)search_kernel_lds_desc<4096>;
            case 7: return (const void*)search_kernel_lds_desc<8192>;
            default: return (const void*)search_kernel_global_desc;
        }
    }
    switch (cls) {
        case 0: return (const void*)search_kernel_lds<64>;
        case 1: return (const void*)search_kernel_lds<128>;
        case 2: return (const void*)search_kernel_lds<256>;
        case 3: return (const void*)search_kernel_lds<512>;
        case 4: return (const void*)search_kernel_lds<1024>;
        case 5: return (const void*)search_kernel_lds<2048>;
        case 6: return (const void*)search_kernel_lds<4096>;
        case 7: return (const void*)search_kernel_lds<8192>;
        default: return (const void*)search_kernel_global;
    }
}

// Zeroes the per-class work counters at the start of a run (a kernel rather
// than a memset, so that the run can be captured in a graph).
__global__ void reset_counters(unsigned int* ctr, const unsigned int* init, int count) {
    int i = (int)threadIdx.x;
    if (i < count) ctr[i] = init[i];
}

}  // namespace

// Per-graph bound on splitter steps plus search iterations; a graph that
// exhausts it reports GCL_E_INTERNAL instead of running on.  GCL_STEP_BUDGET
// overrides the default.
namespace {
std::mutex g_side_mu;
std::string g_side_prefix;
}  // namespace

void set_side_file_prefix(const char* prefix) {
    std::lock_guard<std::mutex> lk(g_side_mu);
    g_side_prefix = prefix ? prefix : "";
}

std::string side_file_prefix() {
    std::lock_guard<std::mutex> lk(g_side_mu);
    return g_side_prefix;
}

uint64_t step_budget() {
    const char* e = getenv("GCL_STEP_BUDGET");
    if (e && *e) return strtoull(e, nullptr, 10);
    return 1ull << 40;
}

// Wall-clock limits of a launch and of each search in it, in device ticks.
// Past its limit a search stops with GCL_E_TIMEOUT, and so does any wait
// inside it; past the launch's limit an idle worker leaves instead of waiting
// for more work.  So no device wait is unbounded, whatever the protocol state.
// The default allows a search 60 s + n^2 / 10^4 ms: a graph without
// automorphisms and with an equitable initial partition (a random regular
// graph) costs about n^2 splitter steps, a few seconds of the whole chip at
// n = 10,000.
TimeLimits time_limits(double sum_n2) {
    const uint64_t tpm = device_wall_ticks_per_ms();
    TimeLimits t;
    const char* e = getenv("GCL_TIME_LIMIT_MS");
    if (e && *e) {
        const uint64_t ms = strtoull(e, nullptr, 10);
        if (ms == 0) return t;  // no limit
        t.launch = t.base = ms * tpm;
        t.per_n2 = 0;
        return t;
    }
    t.base = 60000ull * tpm;
    t.per_n2 = std::max<uint64_t>(1, tpm / 10000);
    const double launch = 60000.0 * (double)tpm + sum_n2 * (double)t.per_n2;
    t.launch = launch >= 1.8e19 ? ~0ull : (uint64_t)launch;
    return t;
}

// Graphs with at least this many vertices publish their root children for
// idle workers (0 disables); GCL_SPEC_MIN_N overrides the default.
uint32_t spec_min_n(uint32_t dflt) {
    const char* e = getenv("GCL_SPEC_MIN_N");
    if (e && *e) return (uint32_t)strtoul(e, nullptr, 10);
    return dflt;
}

uint64_t device_wall_ticks_per_ms() {
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess || rate_khz <= 0)
        rate_khz = 100000;
    return (uint64_t)rate_khz;
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
        bool desc = false;  // the class's kernel carries descent speculation
        KParams P;
        DevBuf scratch, spec, board, diag, autpool, dspec, dboard;
        uint64_t spec_work = 0;  // root children of the class's large graphs
        uint32_t spec_big = 0;   // graphs of the class that may publish
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        hipEvent_t t0 = nullptr, t1 = nullptr;  // timing of the class kernel on its stream
        std::vector<int32_t> graphs;            // host copy of the class's items
        uint64_t aut_stride = 0;                // bytes per graph's ring in autpool (0: not pooled)
    };
    int32_t G = 0;
    int64_t total_n = 0, total_nnz = 0;
    uint32_t max_aut = 1, ring = 1;
    DevBuf vtx, row, col, color, init, label, out, items, ctr, ctr_init;
    DevBuf params;  // the classes' launch parameters (KParams), read by the kernels
    std::vector<Cls> classes;
    hipStream_t main = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, start = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t workers = 0;
    bool timed_run = false;  // last run carried per-class events
    std::string side;        // side-file prefix taken at creation ("" = none)
    BatchTuning tune{};
    std::vector<int64_t> vtx_h;  // host copy of the graph offsets

    ~DeviceBatch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& c : classes) {
            if (c.done) (void)hipEventDestroy(c.done);
            if (c.t0) (void)hipEventDestroy(c.t0);
            if (c.t1) (void)hipEventDestroy(c.t1);
            if (c.stream) (void)hipStreamDestroy(c.stream);
        }
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (start) (void)hipEventDestroy(start);
        if (main) (void)hipStreamDestroy(main);
    }

    gcl_status create(const gcl_batch& b, uint32_t max_aut_, uint32_t ring_cap, const BatchTuning& t) {
        side = side_file_prefix();
        tune = t;
        if (max_aut_ == 0) return fail(GCL_E_INVALID, "max_aut must be >= 1");
        G = b.num_graphs;
        if (G < 0) return fail(GCL_E_INVALID, "negative graph count");
        // an empty batch may come with null arrays: nothing of it is read
        if (G > 0) vtx_h.assign(b.vtx_off, b.vtx_off + G + 1);
        else vtx_h.assign(1, 0);
        if (device_count() <= 0) return fail(GCL_E_NODEVICE, "no HIP device visible");
        GCL_HIP(hipSetDevice(0));
        max_aut = max_aut_;
        ring = std::max(1u, std::min(max_aut, ring_cap));
        total_n = G > 0 ? b.vtx_off[G] : 0;
        total_nnz = G > 0 ? b.row_ptr[total_n] : 0;
        double sum_n2 = 0.0;
        for (int32_t g = 0; g < G; ++g) {
            const double n = (double)(b.vtx_off[g + 1] - b.vtx_off[g]);
            sum_n2 += n * n;
        }
        const TimeLimits lim = time_limits(sum_n2);

        // size classes, larger graphs first within a class
        std::vector<std::vector<int32_t>> items_h(kClasses + 1);
        std::vector<uint32_t> cls_maxn(kClasses + 1, 0);
        std::vector<uint64_t> cls_maxnnz(kClasses + 1, 0);
        std::vector<uint64_t> cls_spec_work(kClasses + 1, 0);
        std::vector<uint32_t> cls_spec_big(kClasses + 1, 0);
        for (int32_t g = 0; g < G; ++g) {
            const int64_t n = b.vtx_off[g + 1] - b.vtx_off[g];
            if (n > kMaxVertices)
                return fail(GCL_E_CAPACITY, "graph %d has n=%lld > %d", g, (long long)n, kMaxVertices);
            int cls = kClasses;
            for (int c = 0; c < kClasses; ++c)
                if ((uint32_t)n <= kClassN[c] && kClassN[c] >= tune.min_class_n) {
                    cls = c;
                    break;
                }
            items_h[cls].push_back(g);
            cls_maxn[cls] = std::max(cls_maxn[cls], (uint32_t)n);
            const uint64_t nnz = (uint64_t)(b.row_ptr[b.vtx_off[g + 1]] - b.row_ptr[b.vtx_off[g]]);
            cls_maxnnz[cls] = std::max(cls_maxnnz[cls], nnz);
            if ((uint32_t)n >= tune.spec_min_n) {
                cls_spec_work[cls] += (uint64_t)n;
                cls_spec_big[cls] += 1;
            }
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
        GCL_HIP(hipMalloc(&ctr_init.p, sizeof(unsigned int) * 64));
        GCL_HIP(hipMemcpy(vtx.p, vtx_h.data(), sizeof(int64_t) * (size_t)(G + 1), hipMemcpyHostToDevice));
        if (G > 0)
            GCL_HIP(hipMemcpy(row.p, b.row_ptr, sizeof(int64_t) * (size_t)(total_n + 1), hipMemcpyHostToDevice));
        else
            GCL_HIP(hipMemset(row.p, 0, sizeof(int64_t)));
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
            // per-level snapshots of the cell-boundary bitmap (depth <= n)
            const uint32_t levels = std::min<uint32_t>(maxn + 2, 4096u);
            const uint32_t spec_min = tune.spec_min_n;
            const bool spec = spec_min > 0 && cls_maxn[cls] >= spec_min;
            C.desc = spec && cnt <= kDescGraphs && cls >= kDescMinClass;
            int per_cu = 0;
            GCL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, class_kernel(cls, C.desc), kWave, 0));
            if (per_cu <= 0) per_cu = 1;
            const int64_t resident = (int64_t)per_cu * prop.multiProcessorCount;
            // automorphism rings: only the worker that owns a graph stores
            // automorphisms, so a class with fewer graphs than workers keeps
            // one ring per graph in a pool (large rings: the reference-shaped
            // entry points use max_aut rows of n)
            const uint64_t ring_bytes = aut_region_bytes(ring, maxn);
            const bool pool_rings = tune.keep_auts || (ring_bytes > (1ull << 20) && (int64_t)cnt < resident);
            const Layout L = make_layout(maxn, max_nnz, ring, arena, levels, hot_global, pool_rings);
            size_t free_b = 0, total_b = 0;
            GCL_HIP(hipMemGetInfo(&free_b, &total_b));
            const double pool_b = pool_rings ? (double)ring_bytes * (double)cnt : 0.0;
            const int64_t mem_workers = (int64_t)(((double)free_b - pool_b) * 0.5 / (double)L.total);
            // with speculation every resident worker runs: the ones without a
            // graph refine published root children of the others
            const int64_t want = spec ? (int64_t)(cnt + cls_spec_work[cls]) : (int64_t)cnt;
            int64_t grid = std::min<int64_t>({want, resident, std::max<int64_t>(mem_workers, 1)});
            C.grid = std::max<int64_t>(grid, 1);
            uint32_t spec_slots = 0, spec_records = 0;
            uint64_t spec_stride = 0;
            if (spec) {
                // one record per graph that can publish concurrently; a record
                // holds every root child unless that exceeds the memory budget
                spec_records = (uint32_t)std::min<uint64_t>({(uint64_t)cls_spec_big[cls], (uint64_t)C.grid,
                                                             (uint64_t)kMaxRecords});
                // item metadata for every root child of a class-sized graph;
                // the partition payload as the budget allows (a record holds
                // items in proportion to 1/n of the graph that publishes)
                const uint64_t budget = 8ull << 30;
                spec_slots = maxn;
                spec_stride = std::min<uint64_t>(spec_record_bytes(maxn, spec_slots, spec_slots),
                                                 std::max<uint64_t>(spec_record_bytes(maxn, spec_slots, 64),
                                                                    budget / std::max<uint32_t>(spec_records, 1)));
                spec_stride = align256(spec_stride);
                GCL_HIP(hipMalloc(&C.spec.p, (size_t)(spec_stride * spec_records)));
                GCL_HIP(hipMemset(C.spec.p, 0, (size_t)(spec_stride * spec_records)));
                const uint32_t nwr = (spec_records + 63) / 64;
                std::vector<unsigned long long> masks(2 * (size_t)nwr, 0ull);
                for (uint32_t r = 0; r < spec_records; ++r) masks[r / 64] |= 1ull << (r % 64);  // all free
                GCL_HIP(hipMalloc(&C.board.p, sizeof(unsigned long long) * masks.size()));
                GCL_HIP(hipMemcpy(C.board.p, masks.data(), sizeof(unsigned long long) * masks.size(),
                                  hipMemcpyHostToDevice));
            }
            // descent records: a class of a few graphs (single-graph and pair
            // calls) has idle workers to spare for every graph's first path
            uint32_t d_records = 0, d_items = 0, d_depth = 0, d_path = 0;
            uint64_t d_stride = 0;
            if (C.desc) {
                // two per graph: the current best path's and the one retired
                // at the last new best leaf
                d_records = 2 * (uint32_t)cnt;
                d_depth = std::min<uint32_t>(512u, maxn);
                d_path = std::min<uint32_t>(levels - 1, 2048u);
                d_items = 1024;
                const DescLayout D = make_desc_layout(maxn, d_items, d_depth, d_path);
                const uint64_t budget = 4ull << 30;
                if (desc_record_bytes(D) > budget)
                    d_items = (uint32_t)std::max<uint64_t>(64, (budget - D.payload) / D.item_bytes);
                d_stride = align256(desc_record_bytes(make_desc_layout(maxn, d_items, d_depth, d_path)));
                GCL_HIP(hipMalloc(&C.dspec.p, (size_t)(d_stride * d_records)));
                GCL_HIP(hipMemset(C.dspec.p, 0, (size_t)(d_stride * d_records)));
                std::vector<unsigned long long> masks(2, 0ull);
                for (uint32_t r = 0; r < d_records; ++r) masks[0] |= 1ull << r;  // all free, none published
                GCL_HIP(hipMalloc(&C.dboard.p, sizeof(unsigned long long) * masks.size()));
                GCL_HIP(hipMemcpy(C.dboard.p, masks.data(), sizeof(unsigned long long) * masks.size(),
                                  hipMemcpyHostToDevice));
            }
            GCL_HIP(hipMalloc(&C.scratch.p, (size_t)(L.total * (uint64_t)C.grid)));
            if (pool_rings) GCL_HIP(hipMalloc(&C.autpool.p, (size_t)(ring_bytes * cnt)));
            if (!side.empty()) {
                GCL_HIP(hipMalloc(&C.diag.p, sizeof(DiagRecord) * (size_t)C.grid));
                GCL_HIP(hipMemset(C.diag.p, 0, sizeof(DiagRecord) * (size_t)C.grid));
            }
            GCL_HIP(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
            GCL_HIP(hipEventCreateWithFlags(&C.done, hipEventDisableTiming));
            GCL_HIP(hipEventCreate(&C.t0));
            GCL_HIP(hipEventCreate(&C.t1));
            C.graphs = items_h[cls];
            KParams& P = C.P;
            memset(&P, 0, sizeof(P));
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
            P.levels = levels;
            P.step_budget = step_budget();
            P.spec = (uint8_t*)C.spec.p;
            P.spec_stride = spec_stride;
            P.spec_slots = spec_slots;
            P.spec_min_n = spec ? spec_min : 0u;
            P.time_limit = lim.launch;
            P.graph_ticks_base = lim.base;
            P.graph_ticks_per_n2 = lim.per_n2;
            P.spec_records = spec_records;
            P.slot_free = spec ? (unsigned long long*)C.board.p : nullptr;
            P.board_mask = spec ? (unsigned long long*)C.board.p + (spec_records + 63) / 64 : nullptr;
            P.pending = (unsigned int*)ctr.p + 32 + cls;
            P.diag = side.empty() ? nullptr : (DiagRecord*)C.diag.p;
            P.aut_pool = pool_rings ? (uint8_t*)C.autpool.p : nullptr;
            P.aut_stride = pool_rings ? ring_bytes : 0;
            C.aut_stride = P.aut_stride;
            if (d_records > 0) {
                P.dspec = (uint8_t*)C.dspec.p;
                P.dspec_stride = d_stride;
                P.dspec_records = d_records;
                P.dspec_items = d_items;
                P.dspec_depth = d_depth;
                P.dspec_path = d_path;
                P.dslot_free = (unsigned long long*)C.dboard.p;
                P.dboard = (unsigned long long*)C.dboard.p + 1;
            }
            workers += (int32_t)C.grid;
            item_off += cnt;
            classes.push_back(std::move(C));
        }
        // per-class work counters start at 0; the pending counters start at
        // the number of graphs that may publish root children, so helpers
        // never leave before those graphs have published or finished
        std::vector<unsigned int> init_ctr(64, 0u);
        for (auto& c : classes)
            if (c.P.spec_min_n) init_ctr[32 + c.cls] = cls_spec_big[c.cls];
        GCL_HIP(hipMemcpy(ctr_init.p, init_ctr.data(), sizeof(unsigned int) * 64, hipMemcpyHostToDevice));
        if (!classes.empty()) {
            std::vector<KParams> ph;
            for (auto& c : classes) ph.push_back(c.P);
            GCL_HIP(hipMalloc(&params.p, sizeof(KParams) * ph.size()));
            GCL_HIP(hipMemcpy(params.p, ph.data(), sizeof(KParams) * ph.size(), hipMemcpyHostToDevice));
        }
        return GCL_OK;
    }

    // reset counters, fork one stream per size class, join; `timed` brackets
    // each class kernel with events on its stream (HIP cannot time events
    // recorded inside a captured graph, so the replayed graph has none)
    hipError_t enqueue(bool timed) {
        hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, main, (unsigned int*)ctr.p,
                           (const unsigned int*)ctr_init.p, 64);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if ((e = hipEventRecord(start, main)) != hipSuccess) return e;
        // largest class first, so that its workers take the CUs first; the
        // joins go to the main stream only after every launch, since streams
        // share hardware queues and a wait queued ahead of a launch on the
        // same queue would hold that launch back
        for (auto it = classes.rbegin(); it != classes.rend(); ++it) {
            Cls& c = *it;
            const KParams* dP = (const KParams*)params.p + (it.base() - 1 - classes.begin());
            if ((e = hipStreamWaitEvent(c.stream, start, 0)) != hipSuccess) return e;
            if (timed && (e = hipEventRecord(c.t0, c.stream)) != hipSuccess) return e;
            if ((e = launch_class(c.cls, c.desc, dP, (int)c.grid, c.stream)) != hipSuccess) return e;
            if (timed && (e = hipEventRecord(c.t1, c.stream)) != hipSuccess) return e;
            if ((e = hipEventRecord(c.done, c.stream)) != hipSuccess) return e;
        }
        for (auto& c : classes)
            if ((e = hipStreamWaitEvent(main, c.done, 0)) != hipSuccess) return e;
        return hipSuccess;
    }

    gcl_status run(bool use_graph, float* ms) {
        GCL_HIP(hipSetDevice(0));
        if (use_graph && !exec) {
            GCL_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeGlobal));
            hipError_t e = enqueue(false);
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
            hipError_t e = enqueue(true);
            if (e != hipSuccess) return fail(GCL_E_DEVICE, "kernel launch failed: %s", hipGetErrorString(e));
        }
        GCL_HIP(hipEventRecord(ev1, main));
        timed_run = !use_graph;
        hipError_t se = hipStreamSynchronize(main);
        if (se != hipSuccess) return fail(GCL_E_DEVICE, "search kernel failed: %s", hipGetErrorString(se));
        float t = 0.f;
        GCL_HIP(hipEventElapsedTime(&t, ev0, ev1));
        if (ms) *ms = t;
        if (!side.empty()) return write_side_files();
        return GCL_OK;
    }

    // Side files of the last run (counterpart of the reference's per-block
    // CSVs, src/nauty_parallel.cu:882-981): <prefix>_counters.csv, one row
    // per worker of every size-class kernel, and <prefix>_nodes_per_cu.csv.
    gcl_status write_side_files() {
        std::vector<std::pair<int, DiagRecord>> rows;
        for (auto& c : classes) {
            std::vector<DiagRecord> r((size_t)c.grid);
            GCL_HIP(hipMemcpy(r.data(), c.diag.p, sizeof(DiagRecord) * r.size(), hipMemcpyDeviceToHost));
            for (auto& x : r) rows.emplace_back((int)c.P.maxn, x);
        }
        const std::string fc = side + "_counters.csv", fn = side + "_nodes_per_cu.csv";
        FILE* f = fopen(fc.c_str(), "w");
        if (!f) return fail(GCL_E_IO, "cannot write %s", fc.c_str());
        fprintf(f, "WORKER,CLASS_MAX_N,XCC,SE,CU,GRAPHS,NODES,SPLITTER_STEPS,HELPED_CHILDREN,"
                   "SEARCH_pct,WAIT_CHILDREN_pct,HELP_pct,IDLE_pct,TOTAL_TICKS,"
                   "REFINE_pct,CLASSIFY_pct,ORBITS_pct,BACKTRACK_pct,ADOPTED_CHILDREN,REJECTED_CHILDREN\n");
        std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint64_t> per_cu;
        for (size_t i = 0; i < rows.size(); ++i) {
            const DiagRecord& d = rows[i].second;
            const uint32_t xcc = d.hw_id >> 6, se = (d.hw_id >> 4) & 3u, cu = d.hw_id & 15u;
            const double tot = d.t_total ? (double)d.t_total : 1.0;
            const uint64_t own = d.t_search > d.t_wait ? d.t_search - d.t_wait : 0;
            fprintf(f, "%zu,%d,%u,%u,%u,%u,%llu,%llu,%llu,%.2f,%.2f,%.2f,%.2f,%llu,%.2f,%.2f,%.2f,%.2f,%u,%u\n", i,
                    rows[i].first, xcc, se, cu, d.graphs, (unsigned long long)d.nodes, (unsigned long long)d.steps,
                    (unsigned long long)d.helped, 100.0 * (double)own / tot, 100.0 * (double)d.t_wait / tot,
                    100.0 * (double)d.t_help / tot, 100.0 * (double)d.t_idle / tot, (unsigned long long)d.t_total,
                    100.0 * (double)d.t_refine / tot, 100.0 * (double)d.t_classify / tot,
                    100.0 * (double)d.t_orbits / tot, 100.0 * (double)d.t_backtrack / tot, d.adopted, d.rejected);
            per_cu[std::make_tuple(xcc, se, cu)] += d.nodes;
        }
        fclose(f);
        f = fopen(fn.c_str(), "w");
        if (!f) return fail(GCL_E_IO, "cannot write %s", fn.c_str());
        fprintf(f, "xcc,se,cu,nodes\n");
        for (auto& kv : per_cu)
            fprintf(f, "%u,%u,%u,%llu\n", std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first),
                    (unsigned long long)kv.second);
        fclose(f);
        return GCL_OK;
    }

    // Per size class: bound, graphs, workers, last run's kernel time (events
    // on the class's stream) and the class's totals from the result records.
    gcl_status class_stats(gcl_class_stats* out, int32_t cap, int32_t* count) {
        GCL_HIP(hipSetDevice(0));
        std::vector<DeviceGraphResult> res((size_t)std::max(G, 1));
        if (G > 0)
            GCL_HIP(hipMemcpy(res.data(), this->out.p, sizeof(DeviceGraphResult) * (size_t)G,
                              hipMemcpyDeviceToHost));
        const int32_t k = (int32_t)classes.size();
        if (count) *count = k;
        for (int32_t i = 0; i < std::min(k, cap); ++i) {
            const Cls& c = classes[(size_t)i];
            gcl_class_stats& o = out[i];
            memset(&o, 0, sizeof(o));
            o.max_n = (int32_t)c.P.maxn;
            o.graphs = c.count;
            o.workers = (int32_t)c.grid;
            float t = -1.f;
            if (timed_run) GCL_HIP(hipEventElapsedTime(&t, c.t0, c.t1));
            o.kernel_ms = t;
            for (int32_t g : c.graphs) {
                o.nodes += res[(size_t)g].nodes;
                o.splitter_steps += res[(size_t)g].splitter_steps;
                o.csr_bytes += res[(size_t)g].csr_bytes;
                o.alg_bytes += res[(size_t)g].csr_bytes + res[(size_t)g].part_bytes;
            }
        }
        return GCL_OK;
    }

    gcl_status automorphisms(int32_t g, int32_t* dst, int64_t cap_rows, int64_t* rows) {
        if (g < 0 || g >= G) return fail(GCL_E_INVALID, "graph %d out of range", g);
        GCL_HIP(hipSetDevice(0));
        DeviceGraphResult r;
        GCL_HIP(hipMemcpy(&r, (DeviceGraphResult*)out.p + g, sizeof(r), hipMemcpyDeviceToHost));
        for (auto& c : classes) {
            for (size_t item = 0; item < c.graphs.size(); ++item) {
                if (c.graphs[item] != g) continue;
                if (c.aut_stride == 0) return fail(GCL_E_INVALID, "batch was not created with keep_auts");
                const int64_t n = vtx_h[(size_t)g + 1] - vtx_h[(size_t)g];
                const int64_t stored = std::min<int64_t>((int64_t)r.automorphisms, (int64_t)ring);
                const int64_t k = std::min<int64_t>(stored, cap_rows);
                if (k > 0 && n > 0) {
                    std::vector<ix_t> buf((size_t)(k * n));
                    const uint8_t* base = (const uint8_t*)c.autpool.p + item * c.aut_stride + aut_fix_bytes(ring);
                    GCL_HIP(hipMemcpy(buf.data(), base, sizeof(ix_t) * buf.size(), hipMemcpyDeviceToHost));
                    for (size_t i = 0; i < buf.size(); ++i) dst[i] = (int32_t)buf[i];
                }
                if (rows) *rows = stored;
                return GCL_OK;
            }
        }
        return fail(GCL_E_INTERNAL, "graph %d is in no class", g);
    }

    gcl_status results(int32_t* out_label, DeviceGraphResult* res_out, gcl_batch_stats* st) {
        GCL_HIP(hipSetDevice(0));
        std::vector<DeviceGraphResult> res((size_t)std::max(G, 1));
        if (G > 0)
            GCL_HIP(hipMemcpy(res.data(), out.p, sizeof(DeviceGraphResult) * (size_t)G, hipMemcpyDeviceToHost));
        if (out_label && total_n > 0)
            GCL_HIP(hipMemcpy(out_label, label.p, sizeof(int32_t) * (size_t)total_n, hipMemcpyDeviceToHost));
        int rate_khz = 0;
        GCL_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0));
        const double ns_per_tick = rate_khz > 0 ? 1e6 / (double)rate_khz : 10.0;
        for (auto& r : res) r.device_ns = (uint64_t)((double)r.device_ns * ns_per_tick);
        uint64_t nodes = 0, auts = 0, steps = 0, bytes = 0, alg = 0;
        gcl_status worst = GCL_OK;
        for (int32_t g = 0; g < G; ++g) {
            nodes += res[(size_t)g].nodes;
            auts += res[(size_t)g].automorphisms;
            steps += res[(size_t)g].splitter_steps;
            bytes += res[(size_t)g].csr_bytes;
            alg += res[(size_t)g].csr_bytes + res[(size_t)g].part_bytes;
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
            st->alg_bytes = alg;
            st->workers = workers;
        }
        return worst;
    }
};

// ----------------------------------------------------------------- serving
// Persistent serving kernels: one LDS kernel for every graph up to
// min(max_n, 1024) vertices and, when max_n is larger, a second one for the
// rest.  A persistent kernel holds its hardware queue for its whole life and
// kernels that share a queue run one after the other, so the server keeps to
// two (HIP gives a process 4 queues by default, one of them taken by the
// null stream, where the runtime's copy and fill kernels run).  For the same
// reason every buffer is allocated and initialised before the first launch:
// host-coherent queue and request slots (written by the host, read by the
// kernels with system-scope loads; results written back the same way), and
// the device arrays the claiming worker copies a request into.
template <int MAXN>
hipError_t launch_serve_lds(const KParams* P, int grid, hipStream_t s) {
    hipLaunchKernelGGL(serve_kernel_lds<MAXN>, dim3(grid), dim3(kWave), 0, s, P);
    return hipGetLastError();
}

hipError_t launch_serve(int cls, const KParams* P, int grid, hipStream_t s) {
    switch (cls) {
        case 0: return launch_serve_lds<64>(P, grid, s);
        case 1: return launch_serve_lds<128>(P, grid, s);
        case 2: return launch_serve_lds<256>(P, grid, s);
        case 3: return launch_serve_lds<512>(P, grid, s);
        case 4: return launch_serve_lds<1024>(P, grid, s);
        case 5: return launch_serve_lds<2048>(P, grid, s);
        case 6: return launch_serve_lds<4096>(P, grid, s);
        case 7: return launch_serve_lds<8192>(P, grid, s);
        default:
            hipLaunchKernelGGL(serve_kernel_global, dim3(grid), dim3(kWave), 0, s, P);
            return hipGetLastError();
    }
}

const void* serve_kernel(int cls) {
    switch (cls) {
        case 0: return (const void*)serve_kernel_lds<64>;
        case 1: return (const void*)serve_kernel_lds<128>;
        case 2: return (const void*)serve_kernel_lds<256>;
        case 3: return (const void*)serve_kernel_lds<512>;
        case 4: return (const void*)serve_kernel_lds<1024>;
        case 5: return (const void*)serve_kernel_lds<2048>;
        case 6: return (const void*)serve_kernel_lds<4096>;
        case 7: return (const void*)serve_kernel_lds<8192>;
        default: return (const void*)serve_kernel_global;
    }
}

struct HostBuf {
    void* p = nullptr;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    HostBuf(HostBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
};

struct ServeEngine {
    struct Cls {
        int cls = 0;
        uint32_t maxn = 0, slots = 0, ring = 0;
        uint64_t max_nnz = 0;
        int64_t grid = 0;
        HostBuf q, ringb, hn, hrow, hcol, hcolor, hinit, hcolored, cq, out, label;
        DevBuf vtx, row, col, color, init, scratch, spec, board, ctr;
        hipStream_t stream = nullptr;
        uint32_t tail = 0;     // host copy of the published count
        uint32_t cq_head = 0;  // completions consumed
    };
    std::vector<Cls> classes;
    DevBuf params;  // the serving kernels' launch parameters (KParams)
    bool running = false;

    ~ServeEngine() {
        stop();
        (void)wait();
        for (auto& c : classes)
            if (c.stream) (void)hipStreamDestroy(c.stream);
    }

    template <class T>
    static gcl_status hmalloc(HostBuf& b, size_t count, T** ptr) {
        GCL_HIP(hipHostMalloc(&b.p, sizeof(T) * std::max<size_t>(count, 1), hipHostMallocCoherent | hipHostMallocMapped));
        memset(b.p, 0, sizeof(T) * std::max<size_t>(count, 1));
        *ptr = (T*)b.p;
        return GCL_OK;
    }

    gcl_status create(const ServeConfig& cfg) {
        if (cfg.max_n <= 0 || cfg.max_n > kMaxVertices || cfg.max_graph_nnz < 0 || cfg.max_aut == 0)
            return fail(GCL_E_INVALID, "invalid server configuration");
        if (device_count() <= 0) return fail(GCL_E_NODEVICE, "no HIP device visible");
        GCL_HIP(hipSetDevice(0));
        hipDeviceProp_t prop;
        GCL_HIP(hipGetDeviceProperties(&prop, 0));
        const uint32_t ring_slots = std::max(1u, std::min(cfg.max_aut, cfg.ring_cap));
        // serving kernels: the size class of min(max_n, 1024), and for larger
        // graphs the class of max_n (2048, 4096, 8192 or the global one)
        auto class_for = [](uint32_t n) {
            for (int k = 0; k < kClasses; ++k)
                if (n <= kClassN[k]) return k;
            return (int)kClasses;
        };
        std::vector<int> ks = {class_for(std::min<uint32_t>((uint32_t)cfg.max_n, 1024u))};
        if (cfg.max_n > 1024) ks.push_back(class_for((uint32_t)cfg.max_n));
        const double big = std::min(0.8, std::max(0.05, cfg.large_share));
        const double share[2] = {ks.size() == 1 ? 0.9 : 0.9 - big, big};
        const uint32_t smin = spec_min_n(cfg.spec_min_n);
        std::vector<KParams> params;
        for (size_t ki = 0; ki < ks.size(); ++ki) {
            const int k = ks[ki];
            Cls C;
            C.cls = k;
            C.maxn = k < kClasses ? kClassN[k] : (uint32_t)cfg.max_n;
            // the first kernel's slots take up to 24 edges per vertex (twice the
            // densest graphs of the benchmark mix); the second one max_graph_nnz
            C.max_nnz = ki == 0 ? std::min<uint64_t>((uint64_t)std::max<int64_t>(cfg.max_graph_nnz, 1), 24ull * C.maxn)
                                : (uint64_t)std::max<int64_t>(cfg.max_graph_nnz, 1);
            const uint64_t arena = std::min<uint64_t>((uint64_t)C.maxn * (C.maxn + 1) / 2, 64ull * C.maxn) + 64;
            const uint32_t levels = std::min<uint32_t>(C.maxn + 2, 4096u);
            const Layout L = make_layout(C.maxn, C.max_nnz, ring_slots, arena, levels, k == kClasses);
            int per_cu = 0;
            GCL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_kernel(k), kWave, 0));
            if (per_cu <= 0) per_cu = 1;
            int64_t grid = (int64_t)((double)per_cu * prop.multiProcessorCount * share[ki]);
            const uint64_t scratch_budget = ki == 0 ? cfg.scratch_bytes_per_class : cfg.scratch_bytes_large;
            grid = std::max<int64_t>(1, std::min<int64_t>(grid, (int64_t)(scratch_budget / L.total)));
            C.grid = grid;
            // request slots: twice the workers in the first class; a large
            // graph keeps many workers busy (its root children), so the large
            // class holds a few dozen requests at a time
            C.slots = cfg.slots > 0 ? (uint32_t)cfg.slots
                      : ki == 0     ? (uint32_t)std::min<int64_t>(4096, std::max<int64_t>(64, 2 * grid))
                                    : (uint32_t)std::min<int64_t>(64, std::max<int64_t>(8, grid));
            C.ring = 1;
            while (C.ring < C.slots) C.ring <<= 1;
            const size_t S = C.slots, n1 = (size_t)C.maxn + 1;
            uint32_t* q;
            uint32_t* ringp;
            int32_t *hn, *hrow, *hcol, *hcolor, *hinit, *hcolored, *label;
            unsigned long long* cq;
            DeviceGraphResult* out;
            gcl_status st;
            if ((st = hmalloc(C.q, 64, &q)) || (st = hmalloc(C.ringb, C.ring, &ringp)) || (st = hmalloc(C.hn, S, &hn)) ||
                (st = hmalloc(C.hrow, S * n1, &hrow)) || (st = hmalloc(C.hcol, S * C.max_nnz, &hcol)) ||
                (st = hmalloc(C.hcolor, S * C.maxn, &hcolor)) || (st = hmalloc(C.hinit, S * C.maxn, &hinit)) ||
                (st = hmalloc(C.hcolored, S, &hcolored)) || (st = hmalloc(C.cq, C.ring, &cq)) ||
                (st = hmalloc(C.out, 2 * S, &out)) || (st = hmalloc(C.label, S * n1, &label)))
                return st;
            GCL_HIP(hipMalloc(&C.vtx.p, sizeof(int64_t) * 2 * S));
            GCL_HIP(hipMalloc(&C.row.p, sizeof(int64_t) * S * n1));
            GCL_HIP(hipMalloc(&C.col.p, sizeof(int32_t) * S * C.max_nnz));
            GCL_HIP(hipMalloc(&C.color.p, sizeof(int32_t) * S * C.maxn));
            GCL_HIP(hipMalloc(&C.init.p, sizeof(int32_t) * S * C.maxn));
            GCL_HIP(hipMalloc(&C.scratch.p, (size_t)(L.total * (uint64_t)grid)));
            // work head, pending counter (never 0 while serving), the host
            // queue's mirror (tail, stop), the next host-poll tick and the
            // completion count
            unsigned int ctr_h[8] = {0u, 0x40000000u, 0u, 0u, 0u, 0u, 0u, 0u};
            GCL_HIP(hipMalloc(&C.ctr.p, sizeof(ctr_h)));
            GCL_HIP(hipMemcpy(C.ctr.p, ctr_h, sizeof(ctr_h), hipMemcpyHostToDevice));
            KParams P;
            memset(&P, 0, sizeof(P));
            P.vtx_off = (const int64_t*)C.vtx.p;
            P.row_ptr = (const int64_t*)C.row.p;
            P.col = (const int32_t*)C.col.p;
            P.color = (const int32_t*)C.color.p;
            P.init_order = (const int32_t*)C.init.p;
            void* dout = nullptr;
            void* dlabel = nullptr;
            GCL_HIP(hipHostGetDevicePointer(&dout, C.out.p, 0));
            GCL_HIP(hipHostGetDevicePointer(&dlabel, C.label.p, 0));
            P.out = (DeviceGraphResult*)dout;
            P.out_label = (int32_t*)dlabel;
            P.next_item = (unsigned int*)C.ctr.p;
            P.pending = (unsigned int*)C.ctr.p + 1;
            P.scratch = (uint8_t*)C.scratch.p;
            P.scratch_stride = L.total;
            P.maxn = C.maxn;
            P.max_nnz = C.max_nnz;
            P.ring = ring_slots;
            P.max_aut = cfg.max_aut;
            P.arena_cap = arena;
            P.levels = levels;
            P.step_budget = step_budget();
            // a serving worker runs until the host stops the queue (or its
            // lifetime passes); each request's search has its own limit
            const TimeLimits lim = time_limits(0.0);
            P.time_limit = cfg.lifetime_ticks;
            P.graph_ticks_base = lim.base;
            P.graph_ticks_per_n2 = lim.per_n2;
            if (smin > 0 && C.maxn >= smin) {
                // one record per worker in the first class; the large class
                // keeps fewer, larger ones, so that each holds every root
                // child of the graphs that publish
                const uint32_t records =
                    (uint32_t)std::min<int64_t>(grid, ki == 0 ? (int64_t)kMaxRecords : (int64_t)C.slots);
                const uint64_t budget = ki == 0 ? cfg.spec_bytes_per_class : cfg.spec_bytes_large;
                const uint32_t spec_slots = C.maxn;
                const uint64_t stride = align256(std::min<uint64_t>(
                    spec_record_bytes(C.maxn, spec_slots, spec_slots),
                    std::max<uint64_t>(spec_record_bytes(C.maxn, spec_slots, 64), budget / records)));
                GCL_HIP(hipMalloc(&C.spec.p, (size_t)(stride * records)));
                GCL_HIP(hipMemset(C.spec.p, 0, (size_t)(stride * records)));
                const uint32_t nwr = (records + 63) / 64;
                std::vector<unsigned long long> masks(2 * (size_t)nwr, 0ull);
                for (uint32_t r = 0; r < records; ++r) masks[r / 64] |= 1ull << (r % 64);
                GCL_HIP(hipMalloc(&C.board.p, sizeof(unsigned long long) * masks.size()));
                GCL_HIP(hipMemcpy(C.board.p, masks.data(), sizeof(unsigned long long) * masks.size(),
                                  hipMemcpyHostToDevice));
                P.spec = (uint8_t*)C.spec.p;
                P.spec_stride = stride;
                P.spec_slots = spec_slots;
                P.spec_records = records;
                P.spec_min_n = smin;
                P.slot_free = (unsigned long long*)C.board.p;
                P.board_mask = (unsigned long long*)C.board.p + nwr;
            }
            void* dp = nullptr;
            ServeParams& V = P.sv;
            GCL_HIP(hipHostGetDevicePointer(&dp, C.q.p, 0));
            V.tail = (const uint32_t*)dp;
            V.stop = (const uint32_t*)dp + 16;
            GCL_HIP(hipHostGetDevicePointer(&dp, C.ringb.p, 0));
            V.ring = (const uint32_t*)dp;
            V.ring_mask = C.ring - 1;
            V.slots = C.slots;
            GCL_HIP(hipHostGetDevicePointer(&dp, C.hn.p, 0));
            V.h_n = (const int32_t*)dp;
            GCL_HIP(hipHostGetDevicePointer(&dp, C.hrow.p, 0));
            V.h_row = (const int32_t*)dp;
            GCL_HIP(hipHostGetDevicePointer(&dp, C.hcol.p, 0));
            V.h_col = (const int32_t*)dp;
            GCL_HIP(hipHostGetDevicePointer(&dp, C.hcolor.p, 0));
            V.h_color = (const int32_t*)dp;
            GCL_HIP(hipHostGetDevicePointer(&dp, C.hinit.p, 0));
            V.h_init = (const int32_t*)dp;
            GCL_HIP(hipHostGetDevicePointer(&dp, C.hcolored.p, 0));
            V.h_colored = (const int32_t*)dp;
            GCL_HIP(hipHostGetDevicePointer(&dp, C.cq.p, 0));
            V.cq = (unsigned long long*)dp;
            V.cq_mask = C.ring - 1;  // C.ring >= slots: a slot has one completion pending at most
            V.cq_count = (unsigned int*)C.ctr.p + 6;
            V.tail_seen = (uint32_t*)C.ctr.p + 2;
            V.stop_seen = (uint32_t*)C.ctr.p + 3;
            V.poll_at = (unsigned long long*)((uint32_t*)C.ctr.p + 4);
            // one host read per 20 us at most
            V.poll_ticks = (uint32_t)std::max<uint64_t>(1, device_wall_ticks_per_ms() / 50);
            GCL_HIP(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
            params.push_back(P);
            classes.push_back(std::move(C));
        }
        GCL_HIP(hipMalloc(&this->params.p, sizeof(KParams) * params.size()));
        GCL_HIP(hipMemcpy(this->params.p, params.data(), sizeof(KParams) * params.size(), hipMemcpyHostToDevice));
        // everything is in place before the first persistent launch
        GCL_HIP(hipDeviceSynchronize());
        for (size_t ki = 0; ki < classes.size(); ++ki) {
            const KParams* dP = (const KParams*)this->params.p + ki;
            hipError_t e = launch_serve(classes[ki].cls, dP, (int)classes[ki].grid, classes[ki].stream);
            if (e != hipSuccess) return fail(GCL_E_DEVICE, "serving kernel launch failed: %s", hipGetErrorString(e));
            running = true;
        }
        return GCL_OK;
    }

    void stop() {
        for (auto& c : classes)
            if (c.q.p) __atomic_store_n((uint32_t*)c.q.p + 16, 1u, __ATOMIC_RELEASE);
    }

    gcl_status wait() {
        gcl_status r = GCL_OK;
        for (auto& c : classes)
            if (c.stream && hipStreamSynchronize(c.stream) != hipSuccess && r == GCL_OK)
                r = fail(GCL_E_DEVICE, "serving kernel failed");
        running = false;
        return r;
    }

    // GCL_OK while every serving kernel runs; GCL_E_TIMEOUT once one has
    // returned (its lifetime passed: it takes no more requests), GCL_E_DEVICE
    // when one failed.  Only meaningful before stop().
    gcl_status check() {
        for (auto& c : classes) {
            const hipError_t e = hipStreamQuery(c.stream);
            if (e == hipErrorNotReady) continue;
            if (e == hipSuccess) return fail(GCL_E_TIMEOUT, "serving kernel stopped: the server's lifetime passed");
            return fail(GCL_E_DEVICE, "serving kernel failed: %s", hipGetErrorString(e));
        }
        return GCL_OK;
    }
};

gcl_status serve_create(const ServeConfig& c, ServeEngine** out) {
    ServeEngine* s = new (std::nothrow) ServeEngine();
    if (!s) return fail(GCL_E_NOMEM, "out of host memory");
    gcl_status st = s->create(c);
    if (st != GCL_OK) {
        delete s;
        return st;
    }
    *out = s;
    return GCL_OK;
}
void serve_destroy(ServeEngine* s) { delete s; }
int serve_classes(const ServeEngine* s) { return (int)s->classes.size(); }
ServeClassInfo serve_class_info(const ServeEngine* s, int k) {
    const auto& c = s->classes[(size_t)k];
    ServeClassInfo i;
    i.maxn = c.maxn;
    i.max_nnz = c.max_nnz;
    i.slots = c.slots;
    i.workers = c.grid;
    return i;
}
int serve_class_of(const ServeEngine* s, int64_t n, int64_t nnz) {
    for (size_t k = 0; k < s->classes.size(); ++k)
        if ((uint64_t)n <= s->classes[k].maxn && (uint64_t)nnz <= s->classes[k].max_nnz) return (int)k;
    return -1;
}
ServeSlotView serve_slot(ServeEngine* s, int k, uint32_t slot) {
    auto& c = s->classes[(size_t)k];
    const size_t n1 = (size_t)c.maxn + 1;
    ServeSlotView v;
    v.n = (int32_t*)c.hn.p + slot;
    v.row = (int32_t*)c.hrow.p + slot * n1;
    v.col = (int32_t*)c.hcol.p + slot * c.max_nnz;
    v.color = (int32_t*)c.hcolor.p + slot * (size_t)c.maxn;
    v.init = (int32_t*)c.hinit.p + slot * (size_t)c.maxn;
    v.colored = (int32_t*)c.hcolored.p + slot;
    v.label = (const int32_t*)c.label.p + slot * n1;
    v.out = (const DeviceGraphResult*)c.out.p + 2 * (size_t)slot;
    return v;
}
int serve_completions(ServeEngine* s, int k, uint32_t* out, int cap) {
    auto& c = s->classes[(size_t)k];
    const unsigned long long* cq = (const unsigned long long*)c.cq.p;
    int got = 0;
    while (got < cap) {
        const unsigned long long e = __atomic_load_n(&cq[c.cq_head & (c.ring - 1)], __ATOMIC_ACQUIRE);
        if ((uint32_t)(e >> 32) != c.cq_head + 1u) break;  // not written yet
        out[got++] = (uint32_t)e;
        ++c.cq_head;
    }
    return got;
}
void serve_publish(ServeEngine* s, int k, uint32_t slot) {
    auto& c = s->classes[(size_t)k];
    uint32_t* ring = (uint32_t*)c.ringb.p;
    __atomic_store_n(&ring[c.tail & (c.ring - 1)], slot, __ATOMIC_RELAXED);
    ++c.tail;
    // the slot's data and ring entry before the count that makes them visible
    __atomic_store_n((uint32_t*)c.q.p, c.tail, __ATOMIC_RELEASE);
}
void serve_stop(ServeEngine* s) { s->stop(); }
gcl_status serve_wait(ServeEngine* s) { return s->wait(); }
gcl_status serve_check(ServeEngine* s) { return s->check(); }

// ----------------------------------------------------- allocation-free path
// One graph searched with no device allocation: every buffer the search
// needs beyond the caller's host input and output (launch parameters, work
// counters, the speculation board, the staged CSR, labels and result record,
// the owner's search scratch and one speculation record) is carved out of the
// file-scope array `scratch`.  Each device that runs the path holds its own
// copy of the array for the life of the process; its address is taken with
// hipGetSymbolAddress on the first call on that device, and only that address
// is used from the host.  Calls in one process are serialized (a host mutex):
// each call overwrites the array, and the kernels of two overlapping calls on
// one device would race on it.
namespace {
constexpr uint32_t kNoAllocMaxN = 1024;
constexpr uint64_t kNoAllocMaxNnz = 32768;
constexpr uint64_t kNoAllocBytes = 8ull << 20;
constexpr uint32_t kNoAllocRing = 64;  // automorphism slots (the batch ring)
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

std::mutex g_noalloc_mu;
uint8_t* g_noalloc_ptr[64] = {};  // per device: the address of `scratch`
struct NoAllocDev {
    hipDeviceProp_t prop;
    int per_cu[kClasses];
};
NoAllocDev g_noalloc_dev[64];

// Offsets inside the array for a graph of size class maxn.
struct NoAllocLayout {
    uint64_t params, ctr, board, vtx, row, col, color, init, label, out, items, spec, work, lists, total;
    uint64_t spec_stride;
    Layout L;
};
NoAllocLayout noalloc_layout(uint32_t maxn, uint64_t max_nnz) {
    NoAllocLayout A;
    uint64_t o = 0;
    A.params = o; o = align256(o + sizeof(KParams));
    A.ctr = o; o = align256(o + 64 * sizeof(unsigned int));
    A.board = o; o = align256(o + 2 * sizeof(unsigned long long));
    A.vtx = o; o = align256(o + 2 * sizeof(int64_t));
    A.row = o; o = align256(o + ((uint64_t)maxn + 1) * sizeof(int64_t));
    A.col = o; o = align256(o + std::max<uint64_t>(max_nnz, 1) * sizeof(int32_t));
    A.color = o; o = align256(o + (uint64_t)maxn * sizeof(int32_t));
    A.init = o; o = align256(o + (uint64_t)maxn * sizeof(int32_t));
    A.label = o; o = align256(o + (uint64_t)maxn * sizeof(int32_t));
    A.out = o; o = align256(o + sizeof(DeviceGraphResult));
    A.items = o; o = align256(o + sizeof(int32_t));
    A.spec_stride = spec_record_bytes(maxn, maxn, maxn);  // every root child
    A.spec = o; o = align256(o + A.spec_stride);
    const uint64_t arena = std::min<uint64_t>((uint64_t)maxn * (maxn + 1) / 2, 64ull * maxn) + 64;
    A.L = make_layout(maxn, max_nnz, kNoAllocRing, arena, maxn + 2, false);
    A.work = o; o = align256(o + A.L.total);
    // every worker shares the one search slice (only the owner searches),
    // but helpers refine root children with touched lists of their own
    A.lists = o; if (maxn >= kListsGlobalN) o += lists_bytes(maxn) * ((uint64_t)kNoAllocMaxN + 1);
    A.total = o;
    return A;
}
}  // namespace

gcl_status noalloc_canon(const gcl_csr& g, uint32_t max_aut, int32_t* out_label, DeviceGraphResult* res) {
    const int32_t n = g.n;
    if (n > (int32_t)kNoAllocMaxN || g.nnz > (int64_t)kNoAllocMaxNnz)
        return fail(GCL_E_CAPACITY, "graph (n=%d, nnz=%lld) exceeds the allocation-free path's limit (n <= %u, nnz <= %llu)",
                    n, (long long)g.nnz, kNoAllocMaxN, (unsigned long long)kNoAllocMaxNnz);
    if (device_count() <= 0) return fail(GCL_E_NODEVICE, "no HIP device visible");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    GCL_HIP(hipSetDevice(0));
    int dev = 0;
    GCL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(GCL_E_DEVICE, "device %d out of range", dev);
    if (!g_noalloc_ptr[dev]) {
        void* ptr = nullptr;
        GCL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        NoAllocDev& d = g_noalloc_dev[dev];
        GCL_HIP(hipGetDeviceProperties(&d.prop, dev));
        for (int c = 0; c < kClasses; ++c) {
            d.per_cu[c] = 0;
            GCL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&d.per_cu[c], class_kernel(c), kWave, 0));
            if (d.per_cu[c] <= 0) d.per_cu[c] = 1;
        }
        g_noalloc_ptr[dev] = (uint8_t*)ptr;
    }
    uint8_t* base = g_noalloc_ptr[dev];
    const NoAllocDev& d = g_noalloc_dev[dev];
    int cls = 0;
    while (cls < kClasses - 1 && (uint32_t)n > kClassN[cls]) ++cls;
    const uint32_t maxn = kClassN[cls];
    const NoAllocLayout A = noalloc_layout(maxn, kNoAllocMaxNnz);
    if (A.total > kNoAllocBytes) return fail(GCL_E_INTERNAL, "allocation-free layout exceeds its array");
    // host-side staging of the graph (the device copies go into the array)
    std::vector<int64_t> row((size_t)n + 1);
    for (int32_t i = 0; i <= n; ++i) row[(size_t)i] = g.row_ptr[i];
    std::vector<int32_t> init;
    if (g.color) {
        init.resize((size_t)n);
        std::iota(init.begin(), init.end(), 0);
        std::stable_sort(init.begin(), init.end(), [&](int32_t x, int32_t y) { return g.color[x] < g.color[y]; });
    }
    const int64_t vtx[2] = {0, n};
    const int32_t item0 = 0;
    const uint32_t smin = spec_min_n();
    const bool spec = smin > 0 && (uint32_t)n >= smin;
    // counters: work head 0; pending (graphs that may publish) 1 or 0
    unsigned int ctr[64] = {};
    ctr[1] = spec ? 1u : 0u;
    const unsigned long long board[2] = {1ull, 0ull};  // record 0 free, none published
    const int64_t resident = (int64_t)d.per_cu[cls] * d.prop.multiProcessorCount;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(spec ? 1 + (int64_t)n : 1, resident));
    const TimeLimits lim = time_limits((double)n * (double)n);
    KParams P;
    memset(&P, 0, sizeof(P));
    P.num_items = 1;
    P.items = (const int32_t*)(base + A.items);
    P.vtx_off = (const int64_t*)(base + A.vtx);
    P.row_ptr = (const int64_t*)(base + A.row);
    P.col = (const int32_t*)(base + A.col);
    P.init_order = g.color ? (const int32_t*)(base + A.init) : nullptr;
    P.color = g.color ? (const int32_t*)(base + A.color) : nullptr;
    P.out_label = (int32_t*)(base + A.label);
    P.out = (DeviceGraphResult*)(base + A.out);
    P.next_item = (unsigned int*)(base + A.ctr);
    P.pending = (unsigned int*)(base + A.ctr) + 1;
    // the launch holds one graph: only the worker that takes it searches and
    // uses search scratch; the others only refine root children in LDS and
    // write the speculation record, so every worker is given the same slice
    P.scratch = base + A.work;
    P.scratch_stride = 0;
    if (maxn >= kListsGlobalN) {
        P.lists = base + A.lists;
        P.lists_stride = lists_bytes(maxn);
    }
    P.maxn = maxn;
    P.max_nnz = kNoAllocMaxNnz;
    P.ring = std::max(1u, std::min(max_aut, kNoAllocRing));
    P.max_aut = max_aut;
    P.arena_cap = std::min<uint64_t>((uint64_t)maxn * (maxn + 1) / 2, 64ull * maxn) + 64;
    P.levels = maxn + 2;
    P.step_budget = step_budget();
    P.time_limit = lim.launch;
    P.graph_ticks_base = lim.base;
    P.graph_ticks_per_n2 = lim.per_n2;
    if (spec) {
        P.spec = base + A.spec;
        P.spec_stride = A.spec_stride;
        P.spec_slots = maxn;
        P.spec_records = 1;
        P.spec_min_n = smin;
        P.slot_free = (unsigned long long*)(base + A.board);
        P.board_mask = (unsigned long long*)(base + A.board) + 1;
    }
    hipStream_t st = 0;
    GCL_HIP(hipMemcpyAsync(base + A.params, &P, sizeof(P), hipMemcpyHostToDevice, st));
    GCL_HIP(hipMemcpyAsync(base + A.ctr, ctr, sizeof(ctr), hipMemcpyHostToDevice, st));
    GCL_HIP(hipMemcpyAsync(base + A.board, board, sizeof(board), hipMemcpyHostToDevice, st));
    GCL_HIP(hipMemcpyAsync(base + A.vtx, vtx, sizeof(vtx), hipMemcpyHostToDevice, st));
    GCL_HIP(hipMemcpyAsync(base + A.items, &item0, sizeof(item0), hipMemcpyHostToDevice, st));
    GCL_HIP(hipMemcpyAsync(base + A.row, row.data(), sizeof(int64_t) * row.size(), hipMemcpyHostToDevice, st));
    if (g.nnz > 0)
        GCL_HIP(hipMemcpyAsync(base + A.col, g.col, sizeof(int32_t) * (size_t)g.nnz, hipMemcpyHostToDevice, st));
    if (g.color) {
        GCL_HIP(hipMemcpyAsync(base + A.color, g.color, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
        GCL_HIP(hipMemcpyAsync(base + A.init, init.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
    }
    hipError_t e = launch_class(cls, false, (const KParams*)(base + A.params), (int)grid, st);
    if (e != hipSuccess) return fail(GCL_E_DEVICE, "kernel launch failed: %s", hipGetErrorString(e));
    GCL_HIP(hipMemcpyAsync(out_label, base + A.label, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, st));
    DeviceGraphResult r;
    GCL_HIP(hipMemcpyAsync(&r, base + A.out, sizeof(r), hipMemcpyDeviceToHost, st));
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(GCL_E_DEVICE, "search kernel failed: %s", hipGetErrorString(e));
    r.device_ns = (uint64_t)((double)r.device_ns * 1e6 / (double)device_wall_ticks_per_ms());
    if (res) *res = r;
    if (r.status != 0) return fail((gcl_status)r.status, "device search failed (%s)", gcl_status_string((gcl_status)r.status));
    return GCL_OK;
}

BatchTuning default_batch_tuning() {
    BatchTuning t;
    t.spec_min_n = spec_min_n();
    t.min_class_n = kClassN[0];
    return t;
}

gcl_status device_batch_create(const gcl_batch& b, uint32_t max_aut, uint32_t ring_cap, DeviceBatch** out,
                               const BatchTuning* tuning) {
    const BatchTuning t = tuning ? *tuning : default_batch_tuning();
    bool class_ok = false;
    for (int c = 0; c < kClasses; ++c) class_ok |= t.min_class_n == kClassN[c];
    if (!class_ok) return fail(GCL_E_INVALID, "min_class_n %u is not a size class bound", t.min_class_n);
    DeviceBatch* db = new (std::nothrow) DeviceBatch();
    if (!db) return fail(GCL_E_NOMEM, "out of host memory");
    gcl_status s = db->create(b, max_aut, ring_cap, t);
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

gcl_status device_batch_class_stats(DeviceBatch* db, gcl_class_stats* out, int32_t cap, int32_t* count) {
    return db->class_stats(out, cap, count);
}

void device_batch_destroy(DeviceBatch* db) { delete db; }

gcl_status device_batch_automorphisms(DeviceBatch* db, int32_t g, int32_t* out, int64_t cap_rows, int64_t* rows) {
    return db->automorphisms(g, out, cap_rows, rows);
}

int32_t device_batch_size(const DeviceBatch* db) { return db->G; }

gcl_status batch_autotune(const gcl_batch& b, uint32_t max_aut, uint32_t ring_cap, int32_t sample, BatchTuning* best,
                          double* best_ms, int32_t* tried) {
    if (sample < 1) return fail(GCL_E_INVALID, "sample must be >= 1");
    // evenly strided sample: same mix of sizes and kinds as the batch
    const int32_t G = b.num_graphs;
    const int32_t stride = std::max<int32_t>(1, (G + sample - 1) / sample);
    std::vector<int64_t> vtx{0}, row{0};
    std::vector<int32_t> col, color;
    for (int32_t g = 0; g < G; g += stride) {
        const int64_t v0 = b.vtx_off[g], v1 = b.vtx_off[g + 1];
        const int64_t base = b.row_ptr[v0];
        for (int64_t v = v0; v < v1; ++v) {
            row.push_back(row.back() + (b.row_ptr[v + 1] - b.row_ptr[v]));
            if (b.color) color.push_back(b.color[v]);
        }
        col.insert(col.end(), b.col + base, b.col + b.row_ptr[v1]);
        vtx.push_back(vtx.back() + (v1 - v0));
    }
    gcl_batch s;
    s.num_graphs = (int32_t)vtx.size() - 1;
    s.vtx_off = vtx.data();
    s.row_ptr = row.data();
    s.col = col.empty() ? nullptr : col.data();
    s.color = b.color ? color.data() : nullptr;
    // candidates: speculation threshold x smallest size class
    const uint32_t spec_c[] = {0u, 64u, 128u, 256u};
    const uint32_t cls_c[] = {kClassN[0], kClassN[1], kClassN[2]};
    double best_t = -1.0;
    int32_t n_tried = 0;
    for (uint32_t sc : spec_c)
        for (uint32_t cc : cls_c) {
            BatchTuning t;
            t.spec_min_n = sc;
            t.min_class_n = cc;
            DeviceBatch* db = nullptr;
            gcl_status st = device_batch_create(s, max_aut, ring_cap, &db, &t);
            if (st != GCL_OK) return st;
            float ms = 0.f;
            double tb = 1e300;
            for (int r = 0; r < 3 && st == GCL_OK; ++r) {
                st = db->run(false, &ms);
                if (r > 0) tb = std::min(tb, (double)ms);
            }
            delete db;
            if (st != GCL_OK) return st;
            ++n_tried;
            if (best_t < 0 || tb < best_t) {
                best_t = tb;
                *best = t;
            }
        }
    if (best_ms) *best_ms = best_t;
    if (tried) *tried = n_tried;
    return GCL_OK;
}

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

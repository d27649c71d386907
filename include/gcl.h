/*
 * gcl.h -- C ABI of libgcl, MI355X-native canonical labeling.
 *
 * Drop-in for the canonical-labeling path of rana-dbouk/gpu-canonical-labeling.
 * Every entry point takes plain pointers and sizes, returns a gcl_status and
 * never exits or throws across the ABI.  The message of the last failure on
 * the calling thread is available from gcl_last_error().
 *
 * Graph convention (CSRGraph, src/CSRGraph.h:6-39 of the reference): an
 * undirected simple graph on vertices 0..n-1 stored as a symmetric CSR,
 * row_ptr[n+1] and col[row_ptr[n]], each row sorted ascending, no self loops,
 * no repeated neighbours; optional vertex colours (the DIMACS "n" lines).
 */
#ifndef GCL_H
#define GCL_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCL_VERSION_MAJOR 0
#define GCL_VERSION_MINOR 1

typedef enum gcl_status {
    GCL_OK = 0,
    GCL_E_INVALID = -1,   /* malformed graph or argument */
    GCL_E_NOMEM = -2,     /* host or device allocation failed */
    GCL_E_DEVICE = -3,    /* HIP runtime error */
    GCL_E_NODEVICE = -4,  /* no usable GPU */
    GCL_E_IO = -5,        /* file could not be read or written */
    GCL_E_CAPACITY = -6,  /* input beyond a documented size limit */
    GCL_E_INTERNAL = -7,  /* consistency check failed inside the library */
    GCL_E_BUSY = -8,      /* admission refused (server full) */
    GCL_E_TIMEOUT = -9    /* a search or a device wait passed its time limit */
} gcl_status;

typedef enum gcl_engine {
    GCL_ENGINE_GPU = 0,   /* --parallel: HIP search on device 0 */
    GCL_ENGINE_CPU = 1    /* --sequential: host search, same traversal */
} gcl_engine;

typedef struct gcl_csr {
    int32_t n;
    int64_t nnz;             /* == row_ptr[n] == 2 * edges */
    const int32_t* row_ptr;  /* n + 1 */
    const int32_t* col;      /* nnz */
    const int32_t* color;    /* n, or NULL for a single colour */
} gcl_csr;

typedef struct gcl_stats {
    uint64_t nodes;          /* search-tree nodes (root + one per individualize+refine) */
    uint64_t automorphisms;  /* automorphisms discovered */
    uint64_t refinements;    /* refinement calls */
    uint64_t splitter_steps; /* splitter iterations over all refinements */
    uint64_t csr_bytes;      /* algorithmic CSR bytes read by refinement */
    double kernel_ms;        /* device time of the search launch(es) */
    double total_ms;         /* wall time of the call */
    uint64_t alg_bytes;      /* all algorithmic bytes of refinement: csr_bytes
                                plus 8 per non-singleton cell hashed under a
                                non-singleton splitter, 8 per vertex of a split
                                cell and 8n per refinement call */
} gcl_stats;

/* ---------------------------------------------------------------- errors */
const char* gcl_last_error(void);
const char* gcl_status_string(gcl_status s);

/* ------------------------------------------------------ graph validation */
/* Checks the CSR convention above (sorted rows, symmetry, no loops, no
 * repeated neighbours).  Returns GCL_OK or GCL_E_INVALID with a message. */
gcl_status gcl_csr_validate(const gcl_csr* g);

/* ------------------------------------------- reference-shaped entry points */
/*
 * Replaces Nauty_Parallel (reference src/nauty_parallel.h:9-13,
 * nauty_parallel.cu:640-1031): canonical labeling on the GPU.  out_canon
 * receives the n*n row-major 0/1 canonical adjacency matrix (the
 * std::vector<unsigned char> the reference returns); node_count, if not NULL,
 * receives the search-tree node count.  max_aut must be >= 1.
 */
gcl_status gcl_nauty_parallel(const gcl_csr* g, int32_t max_aut,
                              uint8_t* out_canon, uint32_t* node_count);

/*
 * Side files, the counterpart of Nauty_Parallel's output_prefix (reference
 * src/nauty_parallel.h:9-13) and of the per-block CSVs it writes
 * (nauty_parallel.cu:882-981).  With a prefix set, every GPU search launch
 * started afterwards (gcl_nauty_parallel, gcl_canon_label, the batch entry
 * points) writes, after each run:
 *   <prefix>_counters.csv      one row per worker of each size-class kernel:
 *                              XCC/SE/CU, graphs, nodes, splitter steps, root
 *                              children refined for others, and the share of
 *                              its time spent searching, waiting for root
 *                              children, helping others and idle
 *   <prefix>_nodes_per_cu.csv  search-tree nodes per CU
 * NULL or "" (the default) writes none.  The setting is process-wide; calls
 * that run at the same time with one prefix overwrite each other's files.
 */
gcl_status gcl_set_side_files(const char* prefix);

/*
 * Replaces Nauty_Sequential + best_leaf.hash_of_perm_graph (reference
 * src/sequential.h:14,30; read at main.cpp:133-142): the same search on the
 * host.  Unlike the reference it does not free the caller's graph.
 */
gcl_status gcl_nauty_sequential(const gcl_csr* g, uint32_t max_aut,
                                uint8_t* out_canon, uint32_t* node_count);

/* ---------------------------------------------------- general entry points */
/*
 * Canonical labeling with either engine.
 *   out_label  n ints or NULL: out_label[i] = input vertex placed at canonical
 *              position i (the best leaf's element order)
 *   out_keys   nnz uint32 or NULL: sorted keys i*n+j of the canonical graph's
 *              directed edges (the sparse form of the canonical adjacency)
 *   out_dense  n*n bytes or NULL: canonical adjacency matrix
 *   stats      or NULL
 */
gcl_status gcl_canon_label(const gcl_csr* g, uint32_t max_aut, gcl_engine engine,
                           int32_t* out_label, uint32_t* out_keys,
                           uint8_t* out_dense, gcl_stats* stats);

/*
 * gcl_canon_label on the GPU with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every device buffer the search needs (launch parameters, counters, the
 * staged CSR, labels, result record, search scratch, one speculation
 * record and the helpers' touched lists) lives in a static device array of
 * 8 MiB that each device running
 * this call holds until the process ends; the call makes no hipMalloc,
 * hipFree or other allocating device call.  Limits: n <= 1024 and
 * nnz <= 32768 (GCL_E_CAPACITY beyond); automorphism ring of
 * min(max_aut, 64) slots, as the batch entry points (labels are those of
 * gcl_canon_label, node counts may be lower).  Calls in one process are
 * serialized; the array keeps the last call's contents.
 *   out_label  n ints (required), out_keys nnz uint32 or NULL, stats or NULL
 */
gcl_status gcl_canon_label_noalloc(const gcl_csr* g, uint32_t max_aut, int32_t* out_label,
                                   uint32_t* out_keys, gcl_stats* stats);

/* The automorphisms the GPU search stored for one graph, for callers that
 * check or reuse them.  No counterpart in the reference: its store is
 * internal (sequential.cpp:515,552; nauty_parallel.cu).  Labels as
 * gcl_canon_label (out_label may be NULL); out_auts receives up to cap_rows
 * rows of g->n ints, row r holding the image a(v) of every vertex v; *rows
 * (optional) the number stored, min(automorphisms found, max_aut) (beyond
 * max_aut the ring keeps the latest).  Each is a leaf-to-leaf map whose
 * certificates compared equal, hence an automorphism of g preserving
 * colours. */
gcl_status gcl_canon_automorphisms(const gcl_csr* g, uint32_t max_aut, int32_t* out_label, int32_t* out_auts,
                                   int64_t cap_rows, int64_t* rows, gcl_stats* stats);

/*
 * Isomorphism verdict of the two-graph CLI mode (main.cpp:173-198):
 * *verdict = 1 when the canonical forms agree, 0 otherwise.  Graphs with
 * different n or edge count are reported non-isomorphic (the reference
 * compares with graph 1's n only).
 */
gcl_status gcl_is_isomorphic(const gcl_csr* a, const gcl_csr* b, uint32_t max_aut,
                             gcl_engine engine, int* verdict);

/* ------------------------------------------------------------- batching */
/*
 * A batch of graphs in one flat CSR:
 *   vtx_off[num_graphs+1]   graph g owns vertices vtx_off[g] .. vtx_off[g+1]-1
 *   row_ptr[vtx_off[G]+1]   global offsets into col, row_ptr[0] == 0 (row of graph g's local
 *                           vertex i is row_ptr[vtx_off[g]+i] .. +1)
 *   col                     LOCAL vertex ids (0 .. n_g-1)
 *   color                   vtx_off[G] ints or NULL
 */
typedef struct gcl_batch {
    int32_t num_graphs;
    const int64_t* vtx_off;
    const int64_t* row_ptr;
    const int32_t* col;
    const int32_t* color;
} gcl_batch;

typedef struct gcl_batch_stats {
    uint64_t nodes;
    uint64_t automorphisms;
    uint64_t splitter_steps;
    uint64_t csr_bytes;
    double kernel_ms;
    double total_ms;
    int32_t workers;         /* concurrent search workers of the launch */
    int32_t pad;
    uint64_t alg_bytes;      /* as in gcl_stats */
} gcl_batch_stats;

/*
 * Canonical labeling of every graph of a batch in one persistent GPU launch
 * (or, with GCL_ENGINE_CPU, a host loop).  out_label has vtx_off[G] entries
 * (per graph as in gcl_canon_label); out_nodes (G entries) is optional.
 * Host pointers; the call copies to the device, runs, and copies back.
 */
gcl_status gcl_batch_canon(const gcl_batch* b, uint32_t max_aut, gcl_engine engine,
                           int32_t* out_label, uint64_t* out_nodes,
                           gcl_batch_stats* stats);

/* Plain baseline: the same graphs, one graph per launch (gcl_canon_label on
 * each), for the batched-vs-plain comparison. */
gcl_status gcl_batch_canon_plain(const gcl_batch* b, uint32_t max_aut,
                                 int32_t* out_label, uint64_t* out_nodes,
                                 gcl_batch_stats* stats);

/*
 * Batch resident on the device: the CSR arrays, per-worker scratch and
 * outputs stay in HBM between runs, so a run measures the search alone.
 *   create   validates, uploads and sizes the launch (max_aut as above)
 *   run      one persistent launch per size class, concurrent on their own
 *            streams; use_graph != 0 captures the run as a hipGraph on first
 *            use and replays it afterwards.  stats gets kernel_ms and workers
 *   results  copies labels (vtx_off[G] ints) and per-graph node counts back
 *            and fills the totals of stats; either pointer may be NULL
 * Not thread-safe per handle.
 */
typedef struct gcl_device_batch gcl_device_batch;
gcl_status gcl_device_batch_create(const gcl_batch* b, uint32_t max_aut, gcl_device_batch** out);
gcl_status gcl_device_batch_run(gcl_device_batch* db, int32_t use_graph, gcl_batch_stats* stats);
gcl_status gcl_device_batch_results(gcl_device_batch* db, int32_t* out_label, uint64_t* out_nodes,
                                    gcl_batch_stats* stats);
/* Per-graph records of the last run (G entries each; any pointer may be
 * NULL): search-tree nodes, automorphisms, splitter steps and the device wall
 * time of the graph's search in nanoseconds. */
gcl_status gcl_device_batch_graph_stats(gcl_device_batch* db, uint64_t* nodes, uint64_t* automorphisms,
                                        uint64_t* splitter_steps, uint64_t* device_ns);
/* Per size class of the batch (one persistent kernel each), for the last
 * run: the class bound on n, graphs, workers, the kernel's time from HIP
 * events recorded on the class's own stream (-1 after a hipGraph replay,
 * which carries no events), and the class's search nodes,
 * splitter steps and algorithmic CSR bytes.  Fills min(cap, classes)
 * entries; *count (optional) receives the number of classes. */
typedef struct gcl_class_stats {
    int32_t max_n;
    int32_t graphs;
    int32_t workers;
    int32_t pad;
    double kernel_ms;
    uint64_t nodes;
    uint64_t splitter_steps;
    uint64_t csr_bytes;
    uint64_t alg_bytes;      /* as in gcl_stats */
} gcl_class_stats;
gcl_status gcl_device_batch_class_stats(gcl_device_batch* db, gcl_class_stats* out, int32_t cap,
                                        int32_t* count);
void gcl_device_batch_destroy(gcl_device_batch* db);

/* Launch tuning of a device batch (no counterpart in the reference, whose one
 * kernel shape is fixed at nauty_parallel.cu:640-760; this is the project's
 * own autotuner over the batched launch):
 *   spec_min_n   graphs with at least this many vertices publish their root
 *                children for idle workers (0: no speculation); default 128,
 *                or GCL_SPEC_MIN_N
 *   min_class_n  smallest size class used (64, 128 or 256 ...): smaller
 *                graphs run in it, so there are fewer, larger class kernels;
 *                default 64
 * gcl_device_batch_create_tuned is gcl_device_batch_create with a tuning
 * (NULL: the defaults).  gcl_batch_autotune times every candidate (spec_min_n
 * in {0, 64, 128, 256} x min_class_n in {64, 128, 256}) on an evenly strided
 * sample of at most `sample` graphs of b: one untimed run, then the best of
 * two.  It returns the fastest in *best, its kernel time in *best_ms and the
 * number of candidates in *tried (both optional).  Results do not depend on
 * the tuning: labels and node counts are the same under every candidate. */
typedef struct gcl_batch_tuning {
    int32_t spec_min_n;
    int32_t min_class_n;
} gcl_batch_tuning;
gcl_status gcl_batch_tuning_default(gcl_batch_tuning* t);
gcl_status gcl_device_batch_create_tuned(const gcl_batch* b, uint32_t max_aut, const gcl_batch_tuning* t,
                                         gcl_device_batch** out);
gcl_status gcl_batch_autotune(const gcl_batch* b, uint32_t max_aut, int32_t sample, gcl_batch_tuning* best,
                              double* best_ms, int32_t* tried);

/* Canonical keys of one graph from a labeling: out_keys[nnz] = sorted
 * i*n+j over directed edges of the relabeled graph. */
gcl_status gcl_canonical_keys(const gcl_csr* g, const int32_t* label, uint32_t* out_keys);

/* --------------------------------------------------------------- DIMACS */
/* Owned graph returned by the reader (CSRGraph::fromDIMACS,
 * reference src/CSRGraph.cpp:40-119).  Free with gcl_graph_free. */
typedef struct gcl_graph {
    gcl_csr csr;
    int32_t* row_ptr_buf;
    int32_t* col_buf;
    int32_t* color_buf;
} gcl_graph;

gcl_status gcl_read_dimacs(const char* path, gcl_graph** out);
gcl_status gcl_write_dimacs(const char* path, const gcl_csr* g);
void gcl_graph_free(gcl_graph* g);

/*
 * Canonical adjacency list text of main.cpp:33-70 (print_canonical_adjacency_list):
 * "%3d : " then " j" per neighbour, wrapped to "\n    " at 80 columns, ";\n"
 * per row.  Writes at most cap bytes (NUL terminated when room) and returns
 * the full length, so a first call with cap 0 sizes the buffer.
 */
int64_t gcl_format_adjacency_keys(const uint32_t* keys, int64_t nnz, int32_t n,
                                  char* buf, int64_t cap);

/* ------------------------------------------------------------ generators */
/* Seeded synthetic graphs; each returns an owned gcl_graph. */
gcl_status gcl_gen_erdos_renyi(int32_t n, double p, uint64_t seed, gcl_graph** out);
gcl_status gcl_gen_erdos_renyi_mean_degree(int32_t n, double mean_degree, uint64_t seed,
                                           gcl_graph** out);
gcl_status gcl_gen_random_regular(int32_t n, int32_t d, uint64_t seed, gcl_graph** out);
gcl_status gcl_gen_torus(int32_t a, int32_t b, gcl_graph** out);  /* C_a x C_b */
gcl_status gcl_gen_cfi(int32_t base_n, int32_t twisted, uint64_t seed, gcl_graph** out);
gcl_status gcl_gen_permute(const gcl_csr* g, uint64_t seed, gcl_graph** out);
gcl_status gcl_gen_edge_swap(const gcl_csr* g, uint64_t seed, gcl_graph** out);

/* Owned flat batch returned by the workload generator. */
typedef struct gcl_owned_batch {
    gcl_batch batch;
    int64_t* vtx_buf;
    int64_t* row_buf;
    int32_t* col_buf;
    int32_t* kind;        /* per graph: 0 G(n,p), 1 random regular, 2 C_a x C_b */
} gcl_owned_batch;

/*
 * The batched benchmark mix: `count` graphs; n log-uniform in [n_min, n_max];
 * 50% G(n,p) with mean degree uniform in [3,12], 30% random d-regular with d
 * in {3,4,5} (n raised by one when n*d is odd), 20% C_a x C_b with a, b
 * uniform in [4,32]; every graph under a seeded random relabeling.  Graph i
 * depends only on (seed, i); `threads` <= 0 uses the hardware concurrency.
 */
gcl_status gcl_gen_mix_batch(int32_t count, uint64_t seed, int32_t n_min, int32_t n_max,
                             int32_t threads, gcl_owned_batch** out);
void gcl_owned_batch_free(gcl_owned_batch* b);

/* ---------------------------------------------------------------- server */
/*
 * Request server: canonicalize and isomorphism requests submitted as they
 * arrive, run with continuous batching on device 0.  At create, one
 * persistent search kernel per size class is started, each on a fixed share
 * of the chip, and every buffer is allocated once: per class, request slots
 * in host-coherent memory (graph in, labels and result out), the device
 * arrays a worker copies a claimed request into, worker scratch and
 * speculation records.  A scheduler thread writes each request's graphs into
 * free slots of their size classes and publishes them to the class queues,
 * in arrival order; idle workers claim them while the others keep searching,
 * and a request completes as soon as its own graphs are done.  Admission is
 * by a footprint estimated from n and m: a graph larger than every class
 * (max_n, max_graph_nnz) is refused with GCL_E_CAPACITY, a request that would
 * take the queued footprint past queue_bytes with GCL_E_BUSY (retry later);
 * requests wait in the host queue while their class has no free slot.  No
 * counterpart in the reference (its only entry point is the per-graph
 * Nauty_Parallel call, src/nauty_parallel.h:9-13).
 */
typedef struct gcl_server_config {
    int32_t max_n;             /* largest graph accepted; default 4096 */
    int64_t max_graph_nnz;     /* directed edges of the largest graph; default 2^20 */
    uint32_t max_aut;          /* automorphism ring per search (>= 1); default 10000 */
    int32_t slots;             /* request slots per size class; 0 (default): twice the
                                  workers of the first class, up to 64 in the large class */
    uint64_t queue_bytes;      /* footprint budget of queued + running requests; default 2^30 */
    int32_t lifetime_ms;       /* serving kernels stop after this long; 0 (default): no limit.
                                  Each request's search has its own limit either way
                                  (60 s + n^2 / 10^4 ms, or GCL_TIME_LIMIT_MS) */
    int32_t large_share_pct;   /* with max_n > 1024: percent of the chip's worker slots for
                                  the class above 1024 vertices (5..80); the first class
                                  gets 90 minus this.  Default 40 */
    uint64_t max_search_cost;  /* admission cap on a graph's estimated search cost in
                                  splitter steps: n^2 for a graph whose vertices all have
                                  one degree and one colour (refinement cannot split its
                                  root; without automorphisms every one of its n root
                                  children then refines to about n steps), n otherwise.
                                  A graph above it is refused with GCL_E_CAPACITY.
                                  0: no cap.  Default 1.5e8 (regular graphs up to
                                  n = 12,247) */
} gcl_server_config;

/* Fills the defaults listed above. */
void gcl_server_config_default(gcl_server_config* cfg);

typedef struct gcl_server gcl_server;

typedef struct gcl_server_result {
    uint64_t ticket;
    int32_t status;            /* gcl_status of the request */
    int32_t verdict;           /* pair: 1 isomorphic, 0 not; canonicalize: -1 */
    int32_t n;                 /* vertices of the first graph */
    int32_t pad;
    double latency_ms;         /* submit (or scheduled arrival) to completion */
    double queue_ms;           /* submit to publication on the device queue */
    uint64_t nodes;            /* search-tree nodes (both graphs of a pair) */
    double device_ms;          /* search time on the device (both graphs of a pair) */
} gcl_server_result;

typedef struct gcl_server_stats {
    uint64_t submitted, completed, refused_busy;
    uint64_t launches;         /* serving kernels (one per size class) */
    uint64_t graphs_launched;  /* graphs published to the device queues */
    uint64_t pending, in_flight, queued_bytes;
    double sched_publish_ms;   /* scheduler thread: writing requests into slots and publishing them */
    double sched_collect_ms;   /* scheduler thread: reading results, pair verdicts */
    uint64_t refused_capacity; /* requests refused with GCL_E_CAPACITY (a graph larger than
                                  every class, or above max_search_cost) */
} gcl_server_stats;

gcl_status gcl_server_create(const gcl_server_config* cfg, gcl_server** out);

/*
 * Queues a request and returns its ticket.  b == NULL: canonicalize a;
 * otherwise the isomorphism test of a and b.  The graphs are copied.
 * out_label (n_a ints, or NULL) receives a's canonical labeling and must stay
 * valid until the request's result has been polled.
 */
gcl_status gcl_server_submit(gcl_server* s, const gcl_csr* a, const gcl_csr* b,
                             int32_t* out_label, uint64_t* ticket);

/* Up to cap completed results, waiting up to timeout_ms (0: no wait) for
 * the first.  Returns the number written, or a negative gcl_status. */
int32_t gcl_server_poll(gcl_server* s, gcl_server_result* out, int32_t cap, int32_t timeout_ms);

gcl_status gcl_server_stats_get(gcl_server* s, gcl_server_stats* st);

/*
 * Load replay: request i is graph i of `graphs`, paired with graph
 * pair_with[i] when pair_with[i] >= 0 (NULL: all canonicalize), submitted at
 * arrival_s[i] seconds after the call starts (non-decreasing); a request
 * refused as busy is retried and its latency counts from the scheduled
 * arrival; a request refused with GCL_E_CAPACITY gets that status as its
 * result (latency 0).  Blocks until every request has completed; out[i] gets
 * the result of request i.
 */
gcl_status gcl_server_replay(gcl_server* s, const gcl_batch* graphs, const int32_t* pair_with,
                             const double* arrival_s, int32_t count, gcl_server_result* out);

/* Completes the queued and running requests, then frees everything. */
void gcl_server_destroy(gcl_server* s);

/* ----------------------------------------------------------- device info */
/* Number of visible GPUs (0 when none); never initialises a context when
 * there is no driver. */
int32_t gcl_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* GCL_H */

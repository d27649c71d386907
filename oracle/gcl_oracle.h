/*
 * gcl_oracle.h -- CPU oracle for canonical labeling (TEST INFRASTRUCTURE).
 *
 * This is the checker, not the product.  Only tests/, the smoke() entry and
 * bench.py's cpu_baseline leg may load it.  It restates the reference's
 * sequential individualization-refinement search (src/sequential.cpp,
 * src/partition_refinement.cpp, src/perm_group.cpp, src/CSRGraph.cpp of
 * rana-dbouk/gpu-canonical-labeling) with the same data structures and the
 * same asymptotic cost, so that it is traversal-exact: same canonical
 * adjacency bytes, same node count and same automorphism count.
 */
#ifndef GCL_ORACLE_H
#define GCL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_stats {
    uint64_t nodes;           /* search-tree nodes (sequential.cpp:243,371,631) */
    uint64_t automorphisms;   /* num_aut at the end (sequential.cpp:83) */
    uint64_t refinements;     /* refinement calls, root included */
    uint64_t splitter_steps;  /* splitter iterations over all refinements */
    uint64_t stale_classify;  /* classify_node calls that reused an old code */
    uint64_t alg_bytes;       /* algorithmic bytes of all refinement calls:
                                 12 per splitter vertex, 4 per neighbour id,
                                 8 per non-singleton cell hashed under a
                                 non-singleton splitter, 8 per vertex of a
                                 split cell, 8n per call */
} oracle_stats;

/*
 * Canonical labeling of an undirected simple graph in CSR form.
 *   n, row_ptr[n+1], col[row_ptr[n]]  -- symmetric adjacency, rows sorted
 *   color                              -- n ints or NULL (all 0)
 *   max_aut                            -- automorphism ring size (>= 1)
 *   out_dense                          -- n*n bytes or NULL: canonical adjacency
 *   out_label                          -- n ints or NULL: best leaf element order
 *   out_auts / aut_cap                 -- optional: first min(aut_cap, stored)
 *                                         automorphism rows (n ints each), ring order
 * Returns 0 on success, -1 on bad arguments / allocation failure.
 */
int oracle_canon(int32_t n, const int32_t* row_ptr, const int32_t* col,
                 const int32_t* color, uint32_t max_aut,
                 uint8_t* out_dense, int32_t* out_label,
                 int32_t* out_auts, uint32_t aut_cap,
                 oracle_stats* stats);

/*
 * One refinement call on a caller-held partition, for per-call checks.
 *   elem[n]                         -- vertex at each position (in/out)
 *   cell_first/len/inlev, *ncells   -- cells in position order (in/out, cap n+1)
 *   level                           -- partition level (in/out, incremented)
 *   act_first / nact                -- initial active cells by first position;
 *                                      nact < 0 means "all cells" (root call)
 * Returns the 15-bit refinement code.
 */
int oracle_refine(int32_t n, const int32_t* row_ptr, const int32_t* col,
                  uint32_t* elem, uint32_t* cell_first, uint32_t* cell_len,
                  uint32_t* cell_inlev, uint32_t* ncells, uint32_t* level,
                  const uint32_t* act_first, int32_t nact);

/* Text writer of the canonical adjacency list (main.cpp:33-70). Returns bytes
 * written (excluding NUL) or the size needed when buf is too small. */
int64_t oracle_format_adjacency(const uint8_t* dense, int32_t n,
                                char* buf, int64_t cap);

/* Upper-triangle comparison (main.cpp:73-90): 1, -1 or 0. */
int oracle_compare_upper(const uint8_t* a, const uint8_t* b, int32_t n);

#ifdef __cplusplus
}
#endif

#endif

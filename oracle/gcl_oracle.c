/*
 * gcl_oracle.c -- CPU oracle for canonical labeling.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
 * entry and bench.py's cpu_baseline leg load it; no product code links it.
 *
 * It is a plain-C restatement of the sequential search of the reference
 * (rana-dbouk/gpu-canonical-labeling), keeping the reference's data layout
 * (cells as {first,length,in_level} in position order, an (n+1) x n table of
 * target cells, n x n leaf matrices, an O(n^2 deg) automorphism test) so that
 * its node/automorphism counts and its cost match the reference binary.
 * Each function names the reference lines it follows.
 *
 * Pinning: the canonical adjacency bytes and node / automorphism counts are
 * checked against golden fixtures written by the reference's own sequential
 * binary (tests/golden/), and against invariants (relabeling invariance,
 * brute-force maximum leaf key on tiny graphs, automorphism validity).
 */
#include "gcl_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* nauty hash macros (partition_refinement.h:9-19) */
#define O_MASH(l, i) ((((l) ^ 065435ULL) + (uint64_t)(i)) & 077777ULL)
#define O_CLEANUP(l) ((int)((l) % 077777ULL))
#define O_SENTINEL 077777

typedef struct {
    uint32_t first, length, in_level;
} ocell;

typedef struct {
    uint32_t n;
    const uint32_t* rp;
    const uint32_t* col;
} ograph;

typedef struct {
    uint32_t* elem;     /* element_vec */
    ocell* lcs;         /* cells in position order, capacity n+1 */
    uint32_t lcs_size;
    uint32_t level;
    uint32_t* seq;      /* current_vertex_sequence */
    uint32_t seq_size;
} opart;

typedef struct {
    unsigned char* active; /* n+1 */
    uint32_t* cnt;         /* all_degrees, n */
    uint32_t* work;        /* work_cell, n */
    uint32_t* bucket;      /* n+1 */
    uint64_t steps;
    /* algorithmic bytes of refinement: 12 per splitter vertex, 4 per
     * neighbour id read, 8 per non-singleton cell hashed under a
     * non-singleton splitter, 8 per vertex of a cell that splits, 8n per
     * call (one load and one store of the element vector) */
    uint64_t bytes;
} oscratch;

/* cell_index_by_first (partition_refinement.cpp:12-25) */
static int o_cell_index(const opart* P, uint32_t first)
{
    int lo = 0, hi = (int)P->lcs_size - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        uint32_t f = P->lcs[mid].first;
        if (first < f) hi = mid - 1;
        else if (first > f) lo = mid + 1;
        else return mid;
    }
    return -1;
}

/* is_discrete (partition_refinement.cpp:70-78) */
static int o_discrete(const opart* P)
{
    for (uint32_t i = 0; i < P->lcs_size; ++i)
        if (P->lcs[i].length != 1) return 0;
    return 1;
}

/* initializePartition_s (partition_refinement.cpp:28-68): stable order by
 * (colour, id), one cell per colour run, in_level 0. */
static void o_init_partition(opart* P, uint32_t n, const int32_t* color)
{
    P->lcs_size = 0;
    P->level = 0;
    P->seq_size = 0;
    for (uint32_t v = 0; v < n; ++v) P->elem[v] = v;
    if (color) {
        /* insertion-free counting: stable sort by colour using a merge sort */
        uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
        for (uint32_t width = 1; width < n; width *= 2) {
            for (uint32_t lo = 0; lo < n; lo += 2 * width) {
                uint32_t mid = lo + width < n ? lo + width : n;
                uint32_t hi = lo + 2 * width < n ? lo + 2 * width : n;
                uint32_t a = lo, b = mid, k = lo;
                while (a < mid && b < hi) {
                    if (color[P->elem[b]] < color[P->elem[a]]) tmp[k++] = P->elem[b++];
                    else tmp[k++] = P->elem[a++];
                }
                while (a < mid) tmp[k++] = P->elem[a++];
                while (b < hi) tmp[k++] = P->elem[b++];
            }
            memcpy(P->elem, tmp, sizeof(uint32_t) * n);
        }
        free(tmp);
    }
    uint32_t start = 0;
    while (start < n) {
        int c = color ? color[P->elem[start]] : 0;
        uint32_t end = start + 1;
        while (end < n && (color ? color[P->elem[end]] : 0) == c) ++end;
        P->lcs[P->lcs_size].first = start;
        P->lcs[P->lcs_size].length = end - start;
        P->lcs[P->lcs_size].in_level = 0;
        P->lcs_size++;
        start = end;
    }
}

/* Partition_s_refinement (partition_refinement.cpp:150-485). */
static int o_refine(opart* P, const ograph* g, const uint32_t* act_first,
                    int nact, oscratch* s)
{
    const uint32_t n = g->n;
    unsigned char* active = s->active;
    uint32_t* cnt = s->cnt;
    uint32_t* work = s->work;
    uint32_t* bucket = s->bucket;
    uint32_t active_count = 0;

    memset(active, 0, n + 1);
    memset(cnt, 0, sizeof(uint32_t) * n);
    memset(bucket, 0, sizeof(uint32_t) * (n + 1));
    s->bytes += 8ull * n;

#define ADD_ACTIVE(f)                                  \
    do {                                               \
        uint32_t f_ = (f);                             \
        if (f_ < n && active[f_] == 0) {               \
            active[f_] = 1;                            \
            ++active_count;                            \
        }                                              \
    } while (0)
#define REMOVE_ACTIVE(f)                               \
    do {                                               \
        uint32_t f_ = (f);                             \
        if (f_ < n && active[f_] != 0) {               \
            active[f_] = 0;                            \
            --active_count;                            \
        }                                              \
    } while (0)

    if (nact < 0) {
        for (uint32_t i = 0; i < P->lcs_size; ++i) ADD_ACTIVE(P->lcs[i].first);
    } else {
        for (int i = 0; i < nact; ++i) ADD_ACTIVE(act_first[i]);
    }

    uint64_t longcode = (uint64_t)P->lcs_size;
    int hint = 0;

    while (!o_discrete(P) && active_count > 0) {
        /* splitter: hint, else first active after hint, else wrap (197-227) */
        int split1_int = -1;
        if (hint >= 0 && (uint32_t)hint < n && active[hint] != 0) {
            split1_int = hint;
        } else {
            uint32_t start_pos = (hint >= 0) ? (uint32_t)hint + 1 : 0;
            for (uint32_t p = start_pos; p < n; ++p)
                if (active[p]) { split1_int = (int)p; break; }
            if (split1_int < 0)
                for (uint32_t p = 0; p < n; ++p)
                    if (active[p]) { split1_int = (int)p; break; }
        }
        uint32_t split1 = (uint32_t)split1_int;
        REMOVE_ACTIVE(split1);
        s->steps++;

        int sidx = o_cell_index(P, split1);
        ocell W = P->lcs[sidx];
        const uint32_t* wverts = &P->elem[W.first];
        uint32_t split2 = split1 + W.length - 1;
        longcode = O_MASH(longcode, (uint64_t)(split1 + split2));
        if (W.length > 1) longcode = O_MASH(longcode, (uint64_t)W.length);

        /* neighbour counts of every vertex into W (244-251) */
        s->bytes += 12ull * W.length;
        for (uint32_t i = 0; i < W.length; ++i) {
            uint32_t w = wverts[i];
            s->bytes += 4ull * (g->rp[w + 1] - g->rp[w]);
            for (uint32_t e = g->rp[w]; e < g->rp[w + 1]; ++e) ++cnt[g->col[e]];
        }

        /* scan every cell in position order (253-468) */
        for (int i = 0; i < (int)P->lcs_size; ++i) {
            int ci = i;
            ocell C = P->lcs[ci];
            if (C.length == 1) continue;
            if (W.length > 1) s->bytes += 8;  /* the cell header read for the hash */
            uint32_t* cv = &P->elem[C.first];
            uint32_t L = C.length;
            uint32_t common = cnt[cv[0]];
            uint32_t bmin = common, bmax = common;
            int is_common = 1;
            for (uint32_t j = 1; j < L; ++j) {
                uint32_t d = cnt[cv[j]];
                if (d != common) is_common = 0;
                if (d < bmin) bmin = d;
                if (d > bmax) bmax = d;
            }
            if (is_common) {
                if (W.length > 1)
                    longcode = O_MASH(longcode, (uint64_t)(common + C.first));
                continue;
            }
            s->bytes += 8ull * L;  /* the split cell's ids, read and written */
            for (uint32_t d = bmin; d <= bmax; ++d) bucket[d] = 0;
            for (uint32_t j = 0; j < L; ++j) ++bucket[cnt[cv[j]]];

            uint32_t nnew = 0;
            if (W.length == 1) {
                if (bmin <= 1 && bmax >= 1 && bucket[1] > 0) ++nnew;
                if (bmin <= 0 && bmax >= 0 && bucket[0] > 0) ++nnew;
            } else {
                for (uint32_t d = bmin; d <= bmax; ++d)
                    if (bucket[d] > 0) ++nnew;
            }
            /* make room for the fragments (316-320) */
            if (nnew > 1)
                for (int j = (int)P->lcs_size - 1; j > ci; --j)
                    P->lcs[j + (int)nnew - 1] = P->lcs[j];

            /* fragment descriptors + write offsets (326-373) */
            uint32_t fidx = 0, first = C.first, max_size = 0, max_index = 0;
#define ADD_FRAGMENT(dv)                                                     \
            do {                                                             \
                uint32_t dv_ = (dv);                                         \
                uint32_t size_ = bucket[dv_];                                \
                if (size_ != 0) {                                            \
                    ocell* nc = &P->lcs[ci + (int)fidx];                     \
                    nc->first = first;                                       \
                    nc->length = size_;                                      \
                    nc->in_level = P->level + 1;                             \
                    if (W.length > 1)                                        \
                        longcode = O_MASH(longcode, (uint64_t)(dv_ + nc->first)); \
                    if (size_ > max_size) { max_size = size_; max_index = fidx; } \
                    bucket[dv_] = first - C.first;                           \
                    first += size_;                                          \
                    ++fidx;                                                  \
                }                                                            \
            } while (0)
            if (W.length == 1) {
                if (bmin <= 1 && bmax >= 1) ADD_FRAGMENT(1);
                if (bmin <= 0 && bmax >= 0) ADD_FRAGMENT(0);
            } else {
                for (uint32_t d = bmin; d <= bmax; ++d) ADD_FRAGMENT(d);
            }
#undef ADD_FRAGMENT
            P->lcs_size += nnew - 1;

            /* stable regroup by count (381-391) */
            for (uint32_t j = 0; j < L; ++j) {
                uint32_t v = cv[j];
                work[bucket[cnt[v]]++] = v;
            }
            for (uint32_t j = 0; j < L; ++j) P->elem[C.first + j] = work[j];

            int old_active = active[C.first] != 0;

            /* singleton-splitter boundary hash (399-409) */
            if (W.length == 1) {
                uint32_t nb = P->lcs[ci].length;
                uint32_t nn = (nnew > 1) ? P->lcs[ci + 1].length : 0;
                if (nb > 0 && nn > 0)
                    longcode = O_MASH(longcode, (uint64_t)(C.first + nb - 1));
            }

            /* active set (414-460) */
            if (W.length == 1) {
                uint32_t left = P->lcs[ci].length;
                uint32_t right = (nnew > 1) ? P->lcs[ci + 1].length : 0;
                if (old_active || left >= right) {
                    if (nnew > 1) {
                        uint32_t fs = P->lcs[ci + 1].first;
                        ADD_ACTIVE(fs);
                        if (P->lcs[ci + 1].length == 1) hint = (int)fs;
                    }
                } else {
                    uint32_t fs = P->lcs[ci].first;
                    ADD_ACTIVE(fs);
                    if (P->lcs[ci].length == 1) hint = (int)fs;
                }
            } else {
                for (uint32_t f = 0; f < nnew; ++f) {
                    ocell* fr = &P->lcs[ci + (int)f];
                    if (fr->first != C.first) {
                        ADD_ACTIVE(fr->first);
                        if (fr->length == 1) hint = (int)fr->first;
                    }
                }
                if (!old_active) {
                    ADD_ACTIVE(C.first);
                    REMOVE_ACTIVE(P->lcs[ci + (int)max_index].first);
                }
            }

            /* last fragment keeps the old in_level (463-465) */
            if (C.in_level != P->level + 1 && nnew > 0)
                P->lcs[ci + (int)nnew - 1].in_level = C.in_level;

            i = ci + (int)nnew - 1;
        }
        memset(cnt, 0, sizeof(uint32_t) * n);
    }
#undef ADD_ACTIVE
#undef REMOVE_ACTIVE

    longcode = O_MASH(longcode, (uint64_t)P->lcs_size);
    int code = O_CLEANUP(longcode);
    P->level++;
    return code;
}

/* split_by_and_refine (partition_refinement.cpp:486-539) */
static void o_split_and_refine(opart* P, const ograph* g, uint32_t vertex,
                               int* code, int tc_first, oscratch* s)
{
    int ci = o_cell_index(P, (uint32_t)tc_first);
    ocell cell = P->lcs[ci];
    if (cell.length == 1) return;
    uint32_t start = cell.first, end = cell.first + cell.length;
    uint32_t vi = start;
    while (vi < end && P->elem[vi] != vertex) vi++;
    if (vi > start) {
        for (uint32_t i = vi; i > start; --i) P->elem[i] = P->elem[i - 1];
        P->elem[start] = vertex;
    }
    ocell rest;
    rest.first = cell.first + 1;
    rest.length = cell.length - 1;
    rest.in_level = cell.in_level;
    for (int i = (int)P->lcs_size - 1; i > ci; --i) P->lcs[i + 1] = P->lcs[i];
    P->lcs[ci + 1] = rest;
    P->lcs_size++;
    cell.length = 1;
    cell.in_level = P->level + 1;
    P->lcs[ci] = cell;
    uint32_t a = cell.first;
    *code = o_refine(P, g, &a, 1, s);
}

static int o_cmp_u32(const void* a, const void* b)
{
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}

/* reconstruct_at_level (partition_refinement.cpp:541-604) */
static void o_reconstruct(opart* P, uint32_t rl, uint32_t* eqlev_first,
                          int* eqlev_canon, int* comp_canon)
{
    if (rl < 1) return;
    for (uint32_t i = 0; i < P->lcs_size;) {
        if (P->lcs[i].in_level <= rl) { ++i; continue; }
        uint32_t first = P->lcs[i].first;
        uint32_t j = i + 1;
        uint32_t len = P->lcs[i].length;
        while (j < P->lcs_size && P->lcs[j].in_level > rl) {
            len += P->lcs[j].length;
            ++j;
        }
        if (j < P->lcs_size) len += P->lcs[j].length;
        qsort(P->elem + first, len, sizeof(uint32_t), o_cmp_u32);
        P->lcs[i].first = first;
        P->lcs[i].length = len;
        P->lcs[i].in_level = P->lcs[j].in_level;
        uint32_t drop = j - i;
        if (drop > 0) {
            for (uint32_t k = j + 1; k < P->lcs_size; ++k) P->lcs[k - drop] = P->lcs[k];
            P->lcs_size -= drop;
        }
        ++i;
    }
    if (rl < *eqlev_first) *eqlev_first = rl;
    if ((int)rl <= *eqlev_canon) {
        *eqlev_canon = (int)rl;
        *comp_canon = 0;
    }
    P->level = rl;
}

/* TargetCell (structs.h:10-17) */
typedef struct {
    int32_t* elems;
    int32_t first;
    uint32_t length;
    uint32_t counter;
    int32_t last_num_aut;
} otc;

/* Leaf_s (structs.h:34-61) */
typedef struct {
    uint32_t* seq;
    uint32_t seq_size;
    unsigned char* hash; /* n*n */
    int discovered;
    int32_t* label;
} oleaf;

typedef struct {
    ograph g;
    uint32_t n, max_aut;
    opart P;
    oscratch s;
    otc* tcal;          /* target_cell_at_level, n+1 entries */
    int tc_level_reached;
    int* firstcode;     /* n+2 */
    int* canon_code;    /* n+2 */
    uint32_t eqlevfirst;
    int eqlevcanon, comp_canon, canon_level;
    int code;
    uint32_t* auts;     /* max_aut * n ring */
    int num_aut;
    oleaf first_leaf, best_leaf;
    uint64_t nodes, refinements, stale;
    /* scratch */
    uint32_t* perm;
    unsigned char* tmp_hash;
    uint32_t* dsu;
    int* subgroup;
} osearch;

/* target_cell_selector with hint -1 (partition_refinement.cpp:607-655) */
static int o_target_cell(osearch* S)
{
    opart* P = &S->P;
    int tc = -1;
    for (uint32_t i = 0; i < P->lcs_size; ++i)
        if (P->lcs[i].length > 1) { tc = (int)i; break; }
    if (tc >= 0) {
        otc* out = &S->tcal[S->tc_level_reached];
        ocell c = P->lcs[tc];
        for (uint32_t v = 0; v < S->n; ++v) out->elems[v] = -1;
        for (uint32_t k = 0; k < c.length; ++k) {
            int32_t vtx = (int32_t)P->elem[c.first + k];
            out->elems[vtx] = vtx;
        }
        out->length = c.length;
        out->counter = 0;
        out->last_num_aut = 0;
        out->first = (int32_t)c.first;
        S->tc_level_reached++;
    }
    return tc;
}

/* discrete_partition_to_perm (perm_group.cpp:118-127) */
static void o_leaf_perm(const osearch* S, uint32_t* perm)
{
    for (uint32_t i = 0; i < S->n; ++i) perm[S->P.elem[i]] = i;
}

/* buildPermutedFlatAdjacencyMatrixArray (CSRGraph.cpp:332-351) */
static void o_perm_matrix(const osearch* S, const uint32_t* perm, unsigned char* m)
{
    uint32_t n = S->n;
    memset(m, 0, (size_t)n * n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t pi = perm[i];
        for (uint32_t k = S->g.rp[i]; k < S->g.rp[i + 1]; ++k) {
            uint32_t pj = perm[S->g.col[k]];
            m[(size_t)pi * n + pj] = 1;
            m[(size_t)pj * n + pi] = 1;
        }
    }
}

/* compare_adj_upper (sequential.cpp:114-124) */
static int o_compare_upper(const unsigned char* a, const unsigned char* b, uint32_t n)
{
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = i + 1; j < n; ++j) {
            unsigned char x = a[(size_t)i * n + j] != 0, y = b[(size_t)i * n + j] != 0;
            if (x != y) return x ? 1 : -1;
        }
    return 0;
}

/* isAutomorphism (sequential.cpp:127-167), O(n^2 deg) as in the reference */
static int o_is_automorphism(const osearch* S, const uint32_t* perm)
{
    uint32_t n = S->n;
    const uint32_t* rp = S->g.rp;
    const uint32_t* col = S->g.col;
    for (uint32_t u = 0; u < n; ++u) {
        uint32_t pu = perm[u];
        if (rp[u + 1] - rp[u] != rp[pu + 1] - rp[pu]) return 0;
    }
    for (uint32_t u = 0; u < n; ++u) {
        for (uint32_t v = 0; v < n; ++v) {
            int a = 0, b = 0;
            for (uint32_t j = rp[u]; j < rp[u + 1]; ++j)
                if (col[j] == v) { a = 1; break; }
            uint32_t pu = perm[u], pv = perm[v];
            for (uint32_t j = rp[pu]; j < rp[pu + 1]; ++j)
                if (col[j] == pv) { b = 1; break; }
            if (a != b) return 0;
        }
    }
    return 1;
}

/* get_gca_level (sequential.cpp:183-213) */
static int o_gca_level(const uint32_t* a, uint32_t na, const uint32_t* b, uint32_t nb)
{
    int m = (int)(na < nb ? na : nb);
    for (int k = 0; k < m; ++k)
        if (a[k] != b[k]) return k + 1;
    return -1;
}

/* subgroup_fixing_sequence + compute_orbits_fast (perm_group.cpp:5-116) */
static void o_compute_orbits(osearch* S, otc* tc)
{
    uint32_t n = S->n;
    int num_aut = S->num_aut < (int)S->max_aut ? S->num_aut : (int)S->max_aut;
    int start = tc->last_num_aut == -1 ? 0 : tc->last_num_aut;
    int nsub = 0;
    for (int i = start; i < num_aut; ++i) {
        const uint32_t* perm = &S->auts[(size_t)i * n];
        int fixes = 1;
        for (uint32_t j = 0; j < S->P.seq_size; ++j)
            if (perm[S->P.seq[j]] != S->P.seq[j]) { fixes = 0; break; }
        if (fixes) S->subgroup[nsub++] = i;
    }
    if (nsub == 0) return;
    uint32_t* parent = S->dsu;
    for (uint32_t v = 0; v < n; ++v)
        parent[v] = tc->elems[v] != -1 ? (uint32_t)tc->elems[v] : v;
    for (int s = 0; s < nsub; ++s) {
        const uint32_t* perm = &S->auts[(size_t)S->subgroup[s] * n];
        for (uint32_t v = 0; v < n; ++v) {
            uint32_t img = perm[v];
            if (v == img) continue;
            uint32_t a = v, b = img;
            while (parent[a] != a) a = parent[a];
            while (parent[b] != b) b = parent[b];
            if (a == b) continue;
            if (a < b) parent[b] = a;
            else parent[a] = b;
        }
    }
    for (uint32_t v = 0; v < n; ++v)
        if (tc->elems[v] != -1) {
            uint32_t x = v;
            while (parent[x] != x) x = parent[x];
            tc->elems[v] = (int32_t)x;
        }
}

/* backtrack_to (sequential.cpp:650-670) */
static void o_backtrack(osearch* S, uint32_t level)
{
    if (level == 0) {
        S->P.level = 0;
        return;
    }
    o_reconstruct(&S->P, level, &S->eqlevfirst, &S->eqlevcanon, &S->comp_canon);
    S->P.seq_size = level - 1;
    S->tc_level_reached = (int)level;
    S->P.level = level;
}

/* first_terminal (sequential.cpp:259-303) */
static void o_first_terminal(osearch* S)
{
    uint32_t n = S->n;
    opart* P = &S->P;
    S->eqlevfirst = P->level;
    S->firstcode[P->level] = O_SENTINEL;
    o_leaf_perm(S, S->perm);
    o_perm_matrix(S, S->perm, S->first_leaf.hash);
    memcpy(S->best_leaf.hash, S->first_leaf.hash, (size_t)n * n);
    S->first_leaf.seq_size = P->seq_size;
    memcpy(S->first_leaf.seq, P->seq, sizeof(uint32_t) * P->seq_size);
    S->best_leaf.seq_size = P->seq_size;
    memcpy(S->best_leaf.seq, P->seq, sizeof(uint32_t) * P->seq_size);
    for (uint32_t j = 0; j < n; ++j)
        S->first_leaf.label[j] = S->best_leaf.label[j] = (int32_t)P->elem[j];
    S->first_leaf.discovered = 1;
    S->canon_level = S->eqlevcanon = (int)P->level;
    S->comp_canon = 0;
    for (uint32_t i = 0; i < P->level; ++i) S->canon_code[i] = S->firstcode[i];
    S->canon_code[P->level] = O_SENTINEL;
}

/* first_path_node (sequential.cpp:216-257) */
static void o_first_path_node(osearch* S)
{
    opart* P = &S->P;
    if (o_discrete(P)) {
        o_first_terminal(S);
        o_backtrack(S, P->level - 1);
        return;
    }
    o_target_cell(S);
    otc* tc = &S->tcal[S->tc_level_reached - 1];
    while (tc->counter < S->n && tc->elems[tc->counter] != (int32_t)tc->counter) ++tc->counter;
    uint32_t vertex = tc->counter++;
    S->nodes++;
    P->seq[P->seq_size++] = vertex;
    o_split_and_refine(P, &S->g, vertex, &S->code, tc->first, &S->s);
    S->firstcode[P->level - 1] = S->code;
    S->refinements++;
}

static void o_store_aut(osearch* S, const int32_t* ref_label)
{
    uint32_t n = S->n;
    uint32_t idx = (uint32_t)S->num_aut % S->max_aut;
    uint32_t* dst = &S->auts[(size_t)idx * n];
    for (uint32_t j = 0; j < n; ++j) dst[ref_label[j]] = S->P.elem[j];
    S->num_aut++;
}

/* classify_node (sequential.cpp:396-624) */
static void o_classify(osearch* S)
{
    opart* P = &S->P;
    uint32_t n = S->n;
    uint32_t L = P->level;
    if (S->eqlevfirst == L - 1 && S->code == S->firstcode[L - 1]) S->eqlevfirst = L;
    if (S->eqlevcanon == (int)L - 1) {
        if (S->code < S->canon_code[L - 1]) S->comp_canon = -1;
        else if (S->code > S->canon_code[L - 1]) S->comp_canon = 1;
        else {
            S->comp_canon = 0;
            S->eqlevcanon = (int)L;
        }
    }
    if (S->comp_canon > 0) S->canon_code[L - 1] = S->code;

    int decision = 0;
    if (S->eqlevfirst != L && S->comp_canon < 0) {
        decision = 4;
    } else if (o_discrete(P)) {
        if (S->eqlevfirst == L) {
            for (uint32_t j = 0; j < n; ++j) S->perm[S->first_leaf.label[j]] = P->elem[j];
            if (o_is_automorphism(S, S->perm)) decision = 1;
        }
        if (decision == 0) {
            if (S->comp_canon == 0) {
                if ((int)L < S->canon_level) {
                    S->comp_canon = 1;
                } else {
                    o_leaf_perm(S, S->perm);
                    o_perm_matrix(S, S->perm, S->tmp_hash);
                    S->comp_canon = o_compare_upper(S->tmp_hash, S->best_leaf.hash, n);
                }
            }
            if (S->comp_canon == 0) decision = 2;
            else if (S->comp_canon > 0) decision = 3;
            else decision = 4;
        }
    }

    switch (decision) {
    case 0:
        break;
    case 1: {
        o_store_aut(S, S->first_leaf.label);
        int gl = o_gca_level(S->first_leaf.seq, S->first_leaf.seq_size, P->seq, P->seq_size);
        o_backtrack(S, gl >= 0 ? (uint32_t)gl : L - 1);
        break;
    }
    case 2: {
        o_store_aut(S, S->best_leaf.label);
        int gl = o_gca_level(S->best_leaf.seq, S->best_leaf.seq_size, P->seq, P->seq_size);
        o_backtrack(S, gl >= 0 ? (uint32_t)gl : L - 1);
        break;
    }
    case 3: {
        S->canon_level = S->eqlevcanon = (int)L;
        S->comp_canon = 0;
        S->canon_code[L] = O_SENTINEL;
        for (uint32_t j = 0; j < n; ++j) S->best_leaf.label[j] = (int32_t)P->elem[j];
        o_leaf_perm(S, S->perm);
        o_perm_matrix(S, S->perm, S->best_leaf.hash);
        memcpy(S->best_leaf.seq, P->seq, sizeof(uint32_t) * P->seq_size);
        S->best_leaf.seq_size = P->seq_size;
        o_backtrack(S, L - 1);
        break;
    }
    case 4:
        o_backtrack(S, L - 1);
        break;
    }
}

/* other_path_node (sequential.cpp:304-392) */
static void o_other_path_node(osearch* S)
{
    opart* P = &S->P;
    if (!o_discrete(P) && (S->eqlevfirst == P->level || S->comp_canon >= 0)) {
        if (S->tc_level_reached < (int)P->level) o_target_cell(S);
        otc* tc = &S->tcal[P->level - 1];
        if (tc->counter >= S->n) {
            o_backtrack(S, P->level - 1);
            return;
        }
        if (S->num_aut > 0 && S->num_aut > tc->last_num_aut) {
            o_compute_orbits(S, tc);
            tc->last_num_aut = S->num_aut;
        }
        while (tc->counter < S->n && tc->elems[tc->counter] != (int32_t)tc->counter) ++tc->counter;
        if (tc->counter >= S->n) {
            o_backtrack(S, P->level - 1);
            return;
        }
        uint32_t vertex = tc->counter++;
        S->nodes++;
        P->seq[P->seq_size++] = vertex;
        o_split_and_refine(P, &S->g, vertex, &S->code, tc->first, &S->s);
        S->refinements++;
    } else {
        S->stale++;
    }
    o_classify(S);
}

static void* o_alloc(size_t bytes)
{
    return malloc(bytes ? bytes : 1);
}

int oracle_canon(int32_t n_, const int32_t* row_ptr, const int32_t* col,
                 const int32_t* color, uint32_t max_aut,
                 uint8_t* out_dense, int32_t* out_label,
                 int32_t* out_auts, uint32_t aut_cap, oracle_stats* stats)
{
    if (n_ < 0 || max_aut == 0 || !row_ptr || (!col && n_ > 0 && row_ptr[n_] > 0)) return -1;
    uint32_t n = (uint32_t)n_;
    osearch S;
    memset(&S, 0, sizeof(S));
    S.n = n;
    S.max_aut = max_aut;
    S.g.n = n;
    S.g.rp = (const uint32_t*)row_ptr;
    S.g.col = (const uint32_t*)col;

    S.P.elem = (uint32_t*)o_alloc(sizeof(uint32_t) * n);
    S.P.lcs = (ocell*)o_alloc(sizeof(ocell) * (n + 1));
    S.P.seq = (uint32_t*)o_alloc(sizeof(uint32_t) * n);
    S.s.active = (unsigned char*)o_alloc(n + 1);
    S.s.cnt = (uint32_t*)o_alloc(sizeof(uint32_t) * n);
    S.s.work = (uint32_t*)o_alloc(sizeof(uint32_t) * n);
    S.s.bucket = (uint32_t*)o_alloc(sizeof(uint32_t) * (n + 1));
    S.tcal = (otc*)o_alloc(sizeof(otc) * (n + 1));
    int32_t* tc_pool = (int32_t*)o_alloc(sizeof(int32_t) * (size_t)(n + 1) * n);
    for (uint32_t i = 0; i <= n; ++i) {
        S.tcal[i].elems = tc_pool + (size_t)i * n;
        S.tcal[i].length = 0;
        S.tcal[i].counter = 0;
        S.tcal[i].last_num_aut = 0;
    }
    S.firstcode = (int*)o_alloc(sizeof(int) * (n + 2));
    S.canon_code = (int*)o_alloc(sizeof(int) * (n + 2));
    S.auts = (uint32_t*)o_alloc(sizeof(uint32_t) * (size_t)max_aut * n);
    S.first_leaf.seq = (uint32_t*)o_alloc(sizeof(uint32_t) * n);
    S.best_leaf.seq = (uint32_t*)o_alloc(sizeof(uint32_t) * n);
    S.first_leaf.hash = (unsigned char*)calloc((size_t)n * n + 1, 1);
    S.best_leaf.hash = (unsigned char*)calloc((size_t)n * n + 1, 1);
    S.first_leaf.label = (int32_t*)calloc(n + 1, sizeof(int32_t));
    S.best_leaf.label = (int32_t*)calloc(n + 1, sizeof(int32_t));
    S.perm = (uint32_t*)o_alloc(sizeof(uint32_t) * n);
    S.tmp_hash = (unsigned char*)o_alloc((size_t)n * n);
    S.dsu = (uint32_t*)o_alloc(sizeof(uint32_t) * n);
    S.subgroup = (int*)o_alloc(sizeof(int) * max_aut);

    int rc = 0;
    if (!S.P.elem || !S.P.lcs || !S.P.seq || !S.s.active || !S.s.cnt || !S.s.work ||
        !S.s.bucket || !S.tcal || !tc_pool || !S.firstcode || !S.canon_code || !S.auts ||
        !S.first_leaf.seq || !S.best_leaf.seq || !S.first_leaf.hash || !S.best_leaf.hash ||
        !S.first_leaf.label || !S.best_leaf.label || !S.perm || !S.tmp_hash || !S.dsu ||
        !S.subgroup) {
        rc = -1;
        goto done;
    }
    for (uint32_t i = 0; i < n + 2; ++i) S.firstcode[i] = S.canon_code[i] = 0;

    o_init_partition(&S.P, n, color);

    /* search_tree_traversal (sequential.cpp:626-648) */
    S.code = o_refine(&S.P, &S.g, NULL, -1, &S.s);
    S.firstcode[S.P.level - 1] = S.code;
    S.nodes++;
    S.refinements++;
    while (S.P.level >= 1) {
        if (!S.first_leaf.discovered) o_first_path_node(&S);
        else o_other_path_node(&S);
    }

    if (out_dense) memcpy(out_dense, S.best_leaf.hash, (size_t)n * n);
    if (out_label) memcpy(out_label, S.best_leaf.label, sizeof(int32_t) * n);
    if (out_auts && aut_cap) {
        uint32_t stored = (uint32_t)S.num_aut < max_aut ? (uint32_t)S.num_aut : max_aut;
        if (stored > aut_cap) stored = aut_cap;
        memcpy(out_auts, S.auts, sizeof(uint32_t) * (size_t)stored * n);
    }
    if (stats) {
        stats->nodes = S.nodes;
        stats->automorphisms = (uint64_t)S.num_aut;
        stats->refinements = S.refinements;
        stats->splitter_steps = S.s.steps;
        stats->stale_classify = S.stale;
        stats->alg_bytes = S.s.bytes;
    }

done:
    free(S.P.elem); free(S.P.lcs); free(S.P.seq);
    free(S.s.active); free(S.s.cnt); free(S.s.work); free(S.s.bucket);
    free(S.tcal); free(tc_pool); free(S.firstcode); free(S.canon_code); free(S.auts);
    free(S.first_leaf.seq); free(S.best_leaf.seq); free(S.first_leaf.hash);
    free(S.best_leaf.hash); free(S.first_leaf.label); free(S.best_leaf.label);
    free(S.perm); free(S.tmp_hash); free(S.dsu); free(S.subgroup);
    return rc;
}

int oracle_refine(int32_t n_, const int32_t* row_ptr, const int32_t* col,
                  uint32_t* elem, uint32_t* cell_first, uint32_t* cell_len,
                  uint32_t* cell_inlev, uint32_t* ncells, uint32_t* level,
                  const uint32_t* act_first, int32_t nact)
{
    uint32_t n = (uint32_t)n_;
    ograph g = {n, (const uint32_t*)row_ptr, (const uint32_t*)col};
    opart P;
    oscratch s;
    memset(&s, 0, sizeof(s));
    P.elem = elem;
    P.lcs = (ocell*)o_alloc(sizeof(ocell) * (n + 1));
    P.seq = NULL;
    P.seq_size = 0;
    P.lcs_size = *ncells;
    P.level = *level;
    for (uint32_t i = 0; i < P.lcs_size; ++i) {
        P.lcs[i].first = cell_first[i];
        P.lcs[i].length = cell_len[i];
        P.lcs[i].in_level = cell_inlev[i];
    }
    s.active = (unsigned char*)o_alloc(n + 1);
    s.cnt = (uint32_t*)o_alloc(sizeof(uint32_t) * n);
    s.work = (uint32_t*)o_alloc(sizeof(uint32_t) * n);
    s.bucket = (uint32_t*)o_alloc(sizeof(uint32_t) * (n + 1));
    int code = o_refine(&P, &g, act_first, nact, &s);
    for (uint32_t i = 0; i < P.lcs_size; ++i) {
        cell_first[i] = P.lcs[i].first;
        cell_len[i] = P.lcs[i].length;
        cell_inlev[i] = P.lcs[i].in_level;
    }
    *ncells = P.lcs_size;
    *level = P.level;
    free(P.lcs); free(s.active); free(s.cnt); free(s.work); free(s.bucket);
    return code;
}

/* print_canonical_adjacency_list (main.cpp:33-70) */
int64_t oracle_format_adjacency(const uint8_t* dense, int32_t n_, char* buf, int64_t cap)
{
    int64_t len = 0;
    char tmp[64];
#define EMIT(str, k)                                            \
    do {                                                        \
        int64_t k_ = (k);                                       \
        if (buf && len + k_ < cap) memcpy(buf + len, (str), k_); \
        len += k_;                                              \
    } while (0)
    uint32_t n = (uint32_t)n_;
    for (uint32_t i = 0; i < n; ++i) {
        int w = snprintf(tmp, sizeof(tmp), "%3u : ", i);
        EMIT(tmp, w);
        int width = w;
        int firstnb = 1;
        for (uint32_t j = 0; j < n; ++j) {
            if (!dense[(size_t)i * n + j]) continue;
            int t = snprintf(tmp, sizeof(tmp), " %u", j);
            if (!firstnb && width + t + 1 > 80) {
                EMIT("\n    ", 5);
                width = 4;
            }
            EMIT(tmp, t);
            width += t;
            firstnb = 0;
        }
        EMIT(";\n", 2);
    }
#undef EMIT
    if (buf && len < cap) buf[len] = 0;
    return len;
}

/* compare_canonical_adjacency_upper (main.cpp:73-90) */
int oracle_compare_upper(const uint8_t* a, const uint8_t* b, int32_t n)
{
    return o_compare_upper(a, b, (uint32_t)n);
}

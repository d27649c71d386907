// gcl_internal.h -- shared host-side helpers of libgcl (not part of the ABI).
#pragma once

#include "gcl.h"

#include <cstdint>
#include <string>
#include <vector>

namespace gcl {

// Error text of the last failing call on this thread.
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
gcl_status fail(gcl_status s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Owned CSR used inside the library.
struct Graph {
    int32_t n = 0;
    std::vector<int32_t> row_ptr;  // n + 1
    std::vector<int32_t> col;      // row_ptr[n]
    std::vector<int32_t> color;    // empty or n
    gcl_csr view() const {
        gcl_csr c;
        c.n = n;
        c.nnz = row_ptr.empty() ? 0 : row_ptr[n];
        c.row_ptr = row_ptr.data();
        c.col = col.data();
        c.color = color.empty() ? nullptr : color.data();
        return c;
    }
};

// Builds a CSR from undirected pairs (u != v, no repeats); rows sorted.
Graph graph_from_pairs(int32_t n, const std::vector<std::pair<int32_t, int32_t>>& pairs);
gcl_status graph_to_owned(Graph&& g, gcl_graph** out);

// Deterministic 64-bit generator (splitmix64 seeding a xoshiro256**).
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed);
    uint64_t next();
    double uniform();                 // [0, 1)
    uint64_t below(uint64_t bound);   // [0, bound)
};

// Sorted canonical keys i*n+j of the graph relabeled by `label`
// (label[i] = vertex at canonical position i).
void canonical_keys(const gcl_csr& g, const int32_t* label, uint32_t* keys);

// Dense n*n matrix from sorted keys.
void keys_to_dense(const uint32_t* keys, int64_t nnz, int32_t n, uint8_t* dense);

// Edges of a uniform random simple d-regular graph (configuration model).
gcl_status random_regular_pairs(int32_t n, int32_t d, uint64_t seed,
                                std::vector<std::pair<int32_t, int32_t>>& pairs);

// Largest n the library accepts (keys i*n+j must fit in 32 bits).
constexpr int32_t kMaxVertices = 65535;

}  // namespace gcl

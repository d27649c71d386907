// gcl_search.h -- constants and helpers shared by the host and device search.
//
// The search is individualization-refinement with the exact traversal of the
// reference's sequential engine (reference src/sequential.cpp:216-670): the
// same refinement codes, target cells, branch order, pruning and leaf order,
// hence the same canonical form and the same node / automorphism counts.
//
// Partition representation (both engines):
//   elem[p]    vertex at position p
//   pos[v]     position of vertex v
//   cend[p]    last position of the cell holding position p; a cell is named
//              by its last position ("cell id"), so when a cell splits the
//              last fragment keeps the name and only the other fragments'
//              positions are relabeled
//   cfirst[c]  first position of cell c
//   nsmask     bit p set <=> p is the first position of a non-singleton cell
// Backtracking replays an undo log of created cell ids in reverse (a merge
// per entry), which restores every cell as a set of positions; the order of
// vertices inside a cell is never observed by codes, target cells or leaves.
#pragma once

#include <cstdint>

namespace gcl {

// nauty's hash macros (reference src/partition_refinement.h:9-19)
#define GCL_MASH(l, i) ((((l) ^ 065435u) + (uint32_t)(i)) & 077777u)
#define GCL_CLEANUP(l) ((int)((l) % 077777u))
constexpr int kSentinel = 077777;  // trail terminator, larger than any code

// Leaf-decision codes of classify_node (reference src/sequential.cpp:437-489).
enum Decision : int {
    kNormal = 0,       // interior node, continue
    kAutFirst = 1,     // leaf equivalent to the first leaf
    kAutBest = 2,      // leaf equivalent to the best leaf
    kNewBest = 3,      // leaf better than the best leaf
    kBad = 4           // prune
};

}  // namespace gcl

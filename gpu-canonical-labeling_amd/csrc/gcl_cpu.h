// gcl_cpu.h -- host engine entry (internal).
#pragma once

#include "gcl_internal.h"
#include "gcl_search.h"

namespace gcl {

// Exact-traversal canonical labeling on the host.  out_label: n ints,
// out_keys: nnz sorted canonical keys; either may be NULL.
gcl_status cpu_canon_label(const gcl_csr& g, uint32_t max_aut, int32_t* out_label, uint32_t* out_keys,
                           gcl_stats* st);

}  // namespace gcl

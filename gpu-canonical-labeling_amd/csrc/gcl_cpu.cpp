// gcl_cpu.cpp -- host engine of libgcl (the --sequential path).
//
// Same traversal as the reference's sequential engine (Nauty_Sequential,
// reference src/sequential.cpp:13-670) on the data structures described in
// gcl_search.h: O(deg) singleton splitter steps instead of an O(#cells) scan,
// sorted edge keys instead of n x n leaf matrices, an undo log instead of
// in_level reconstruction.  It is the blueprint of the device engine in
// gcl_device.hip and is checked against the oracle on every test graph.
#include "gcl_cpu.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <numeric>

namespace gcl {

namespace {

struct Split {
    uint32_t f, k, L;
};

struct Frag {
    uint32_t val, first, size;
};

struct LevelRec {
    uint32_t tc_first = 0, tc_len = 0, cursor = 0, arena_off = 0;
    int64_t last_num_aut = 0;
};

class CpuSearch {
   public:
    CpuSearch(const gcl_csr& g, uint32_t max_aut) : g_(g), n_((uint32_t)g.n), max_aut_(max_aut) {}
    gcl_status run(int32_t* out_label, uint32_t* out_keys, gcl_stats* st);

   private:
    void init_partition();
    int refine(int32_t init_active, uint32_t tag);
    void singleton_step(uint32_t w, uint32_t tag, uint32_t& longcode, int& hint);
    void multi_step(uint32_t p, uint32_t e, uint32_t tag, uint32_t& longcode, int& hint);
    void individualize_and_refine(uint32_t v, uint32_t tc_first, uint32_t tag);
    void undo_to(uint32_t level);
    bool test_bit(const std::vector<uint64_t>& m, uint32_t p) const { return (m[p >> 6] >> (p & 63)) & 1; }
    void add_active(uint32_t p) {
        if (p < n_ && !test_bit(active_, p)) {
            active_[p >> 6] |= 1ull << (p & 63);
            ++nactive_;
        }
    }
    void remove_active(uint32_t p) {
        if (p < n_ && test_bit(active_, p)) {
            active_[p >> 6] &= ~(1ull << (p & 63));
            --nactive_;
        }
    }
    void set_ns(uint32_t p, bool on) {
        if (on) nsmask_[p >> 6] |= 1ull << (p & 63);
        else nsmask_[p >> 6] &= ~(1ull << (p & 63));
    }
    int32_t first_set_from(const std::vector<uint64_t>& m, uint32_t start) const;

    void select_target();
    void compute_orbits(LevelRec& r);
    void leaf_keys(uint32_t* keys) const;
    static int compare_keys(const uint32_t* a, const uint32_t* b, uint64_t nnz);
    void store_aut(const int32_t* ref_label);
    static int gca_level(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b, uint32_t bsize);
    void backtrack(uint32_t level);
    void classify();
    void first_path_node();
    void other_path_node();
    void first_terminal();

    const gcl_csr& g_;
    const uint32_t n_, max_aut_;
    uint64_t nnz_ = 0;

    // partition
    std::vector<uint32_t> elem_, pos_, cend_, cfirst_;
    std::vector<uint64_t> nsmask_, active_;
    uint32_t ncells_ = 0, nactive_ = 0;
    // refinement scratch
    std::vector<uint32_t> cnt_, hits_, cmin_, cmax_, hcur_, tv_, tcells_, cval_, frag_off_, frag_n_;
    std::vector<uint8_t> mark_, cstate_;
    std::vector<Split> splits_;
    std::vector<Frag> frags_;
    std::vector<uint32_t> split_firsts_;
    std::vector<std::pair<uint32_t, uint32_t>> sortbuf_;
    std::vector<uint32_t> undo_id_, undo_tag_;

    // search state (reference src/sequential.h:24-58)
    uint32_t level_ = 0;
    std::vector<uint32_t> seq_;
    uint32_t seq_size_ = 0;
    std::vector<LevelRec> lv_;  // lv_[L-1] = target cell of the level-L node
    std::vector<uint32_t> arena_mem_, arena_rep_;
    uint32_t tc_level_reached_ = 0;
    std::vector<int> firstcode_, canon_code_;
    uint32_t eqlevfirst_ = 0;
    int eqlevcanon_ = 0, comp_canon_ = 0, canon_level_ = 0, code_ = 0;
    bool first_discovered_ = false;
    std::vector<int32_t> first_label_, best_label_;
    std::vector<uint32_t> first_seq_, best_seq_;
    uint32_t first_seq_size_ = 0, best_seq_size_ = 0;
    std::vector<uint32_t> first_keys_, best_keys_, cur_keys_;
    std::vector<uint32_t> auts_;
    int64_t num_aut_ = 0;
    std::vector<uint32_t> dsu_;
    std::vector<int64_t> subgroup_;
    uint64_t nodes_ = 0, refinements_ = 0, steps_ = 0, csr_bytes_ = 0;
    // algorithmic bytes beyond the CSR reads (as the device counts them)
    uint64_t part_bytes_ = 0;
};

void CpuSearch::init_partition() {
    // initializePartition_s (reference src/partition_refinement.cpp:28-68):
    // stable order by (colour, id), one cell per colour run.
    const uint32_t n = n_;
    elem_.resize(n);
    std::iota(elem_.begin(), elem_.end(), 0u);
    if (g_.color)
        std::stable_sort(elem_.begin(), elem_.end(),
                         [&](uint32_t a, uint32_t b) { return g_.color[a] < g_.color[b]; });
    pos_.assign(n, 0);
    cend_.assign(n, 0);
    cfirst_.assign(n, 0);
    nsmask_.assign((n + 63) / 64 + 1, 0);
    ncells_ = 0;
    for (uint32_t p = 0; p < n; ++p) pos_[elem_[p]] = p;
    uint32_t start = 0;
    while (start < n) {
        int32_t c = g_.color ? g_.color[elem_[start]] : 0;
        uint32_t end = start + 1;
        while (end < n && (g_.color ? g_.color[elem_[end]] : 0) == c) ++end;
        for (uint32_t p = start; p < end; ++p) cend_[p] = end - 1;
        cfirst_[end - 1] = start;
        set_ns(start, end - start > 1);
        ++ncells_;
        start = end;
    }
}

int32_t CpuSearch::first_set_from(const std::vector<uint64_t>& m, uint32_t start) const {
    const uint32_t nw = (n_ + 63) / 64;
    for (uint32_t w = start >> 6; w < nw; ++w) {
        uint64_t bits = m[w];
        if (w == (start >> 6)) bits &= ~0ull << (start & 63);
        if (bits) return (int32_t)(w * 64 + (uint32_t)__builtin_ctzll(bits));
    }
    return -1;
}

void CpuSearch::singleton_step(uint32_t w, uint32_t tag, uint32_t& longcode, int& hint) {
    // Refinement step for a splitter of size 1 (reference
    // src/partition_refinement.cpp:253-468, splitter.length == 1): only cells
    // holding a neighbour of w can split, each into (neighbours, others), and
    // uniform cells add nothing to the code.
    const int32_t* rp = g_.row_ptr;
    const int32_t* col = g_.col;
    tcells_.clear();
    csr_bytes_ += 12 + 4ull * (uint64_t)(rp[w + 1] - rp[w]);
    for (int32_t e = rp[w]; e < rp[w + 1]; ++e) {
        uint32_t u = (uint32_t)col[e];
        uint32_t c = cend_[pos_[u]];
        if (cfirst_[c] == c) continue;
        mark_[u] = 1;
        if (hits_[c]++ == 0) {
            tcells_.push_back(c);
            hcur_[c] = cfirst_[c];
        }
    }
    for (int32_t e = rp[w]; e < rp[w + 1]; ++e) {
        uint32_t u = (uint32_t)col[e];
        if (!mark_[u]) continue;
        uint32_t q = pos_[u];
        uint32_t c = cend_[q];
        uint32_t f = cfirst_[c], k = hits_[c], L = c - f + 1;
        if (k == L || q < f + k) continue;
        // move u into the front block [f, f+k) over the next non-neighbour
        uint32_t h = hcur_[c];
        while (mark_[elem_[h]]) ++h;
        uint32_t x = elem_[h];
        elem_[h] = u;
        pos_[u] = h;
        elem_[q] = x;
        pos_[x] = q;
        hcur_[c] = h + 1;
    }
    splits_.clear();
    for (uint32_t c : tcells_) {
        uint32_t f = cfirst_[c], k = hits_[c], L = c - f + 1;
        hits_[c] = 0;
        if (k == L) continue;
        uint32_t e0 = f + k - 1;
        for (uint32_t q = f; q <= e0; ++q) cend_[q] = e0;
        cfirst_[e0] = f;
        cfirst_[c] = f + k;
        undo_id_.push_back(e0);
        undo_tag_.push_back(tag);
        ++ncells_;
        splits_.push_back({f, k, L});
        part_bytes_ += 8ull * L;
    }
    for (int32_t e = rp[w]; e < rp[w + 1]; ++e) mark_[(uint32_t)col[e]] = 0;
    std::sort(splits_.begin(), splits_.end(), [](const Split& a, const Split& b) { return a.f < b.f; });
    for (const Split& s : splits_) {
        const uint32_t f = s.f, k = s.k, L = s.L;
        longcode = GCL_MASH(longcode, f + k - 1);  // boundary hash (399-409)
        if (test_bit(active_, f) || k >= L - k) {  // active-set rule (414-435)
            add_active(f + k);
            if (L - k == 1) hint = (int)(f + k);
        } else {
            add_active(f);
            if (k == 1) hint = (int)f;
        }
        set_ns(f, k > 1);
        set_ns(f + k, L - k > 1);
    }
}

void CpuSearch::multi_step(uint32_t p, uint32_t e, uint32_t tag, uint32_t& longcode, int& hint) {
    // Refinement step for a splitter of size > 1 (reference
    // src/partition_refinement.cpp:253-468): every non-singleton cell hashes
    // in position order; touched cells split into fragments of increasing
    // neighbour count.
    const int32_t* rp = g_.row_ptr;
    const int32_t* col = g_.col;
    tv_.clear();
    tcells_.clear();
    frags_.clear();
    split_firsts_.clear();
    for (uint32_t q = p; q <= e; ++q) {
        uint32_t w = elem_[q];
        csr_bytes_ += 12 + 4ull * (uint64_t)(rp[w + 1] - rp[w]);
        for (int32_t x = rp[w]; x < rp[w + 1]; ++x) {
            uint32_t u = (uint32_t)col[x];
            if (cnt_[u]++ == 0) tv_.push_back(u);
        }
    }
    for (uint32_t u : tv_) {
        uint32_t c = cend_[pos_[u]];
        if (cfirst_[c] == c) continue;
        uint32_t v = cnt_[u];
        if (hits_[c]++ == 0) {
            tcells_.push_back(c);
            cmin_[c] = cmax_[c] = v;
        } else {
            cmin_[c] = std::min(cmin_[c], v);
            cmax_[c] = std::max(cmax_[c], v);
        }
    }
    for (uint32_t c : tcells_) {
        const uint32_t f = cfirst_[c], L = c - f + 1, h = hits_[c];
        hits_[c] = 0;
        if (h == L && cmin_[c] == cmax_[c]) {
            cstate_[f] = 1;
            cval_[f] = cmin_[c];
            continue;
        }
        cstate_[f] = 2;
        split_firsts_.push_back(f);
        part_bytes_ += 8ull * L;
        sortbuf_.clear();
        for (uint32_t q = f; q <= c; ++q) sortbuf_.push_back({cnt_[elem_[q]], elem_[q]});
        std::sort(sortbuf_.begin(), sortbuf_.end());
        frag_off_[f] = (uint32_t)frags_.size();
        uint32_t q = f, nf = 0;
        size_t i = 0;
        while (i < sortbuf_.size()) {
            size_t j = i;
            while (j < sortbuf_.size() && sortbuf_[j].first == sortbuf_[i].first) ++j;
            const uint32_t first = q, last = q + (uint32_t)(j - i) - 1;
            for (size_t t = i; t < j; ++t, ++q) {
                elem_[q] = sortbuf_[t].second;
                pos_[sortbuf_[t].second] = q;
                cend_[q] = last;
            }
            cfirst_[last] = first;
            frags_.push_back({sortbuf_[i].first, first, last - first + 1});
            ++nf;
            if (last != c) {
                undo_id_.push_back(last);
                undo_tag_.push_back(tag);
                ++ncells_;
            }
            i = j;
        }
        frag_n_[f] = nf;
    }
    // serial pass over the pre-step non-singleton cells in position order
    const uint32_t nw = (n_ + 63) / 64;
    for (uint32_t wi = 0; wi < nw; ++wi) {
        uint64_t bits = nsmask_[wi];
        while (bits) {
            const uint32_t f = wi * 64 + (uint32_t)__builtin_ctzll(bits);
            bits &= bits - 1;
            part_bytes_ += 8;  // the cell header hashed
            const uint8_t st = cstate_[f];
            cstate_[f] = 0;
            if (st == 0) {
                longcode = GCL_MASH(longcode, f);  // untouched: common count 0
            } else if (st == 1) {
                longcode = GCL_MASH(longcode, cval_[f] + f);
            } else {
                const Frag* fr = &frags_[frag_off_[f]];
                const uint32_t nf = frag_n_[f];
                uint32_t max_size = 0, max_idx = 0;
                for (uint32_t r = 0; r < nf; ++r) {
                    longcode = GCL_MASH(longcode, fr[r].val + fr[r].first);
                    if (fr[r].size > max_size) {
                        max_size = fr[r].size;
                        max_idx = r;
                    }
                }
                const bool was_active = test_bit(active_, f);
                for (uint32_t r = 1; r < nf; ++r) {
                    add_active(fr[r].first);
                    if (fr[r].size == 1) hint = (int)fr[r].first;
                }
                if (!was_active) {
                    add_active(f);
                    remove_active(fr[max_idx].first);
                }
            }
        }
    }
    for (uint32_t f : split_firsts_) {
        const Frag* fr = &frags_[frag_off_[f]];
        for (uint32_t r = 0; r < frag_n_[f]; ++r) set_ns(fr[r].first, fr[r].size > 1);
    }
    for (uint32_t u : tv_) cnt_[u] = 0;
}

int CpuSearch::refine(int32_t init_active, uint32_t tag) {
    // Partition_s_refinement (reference src/partition_refinement.cpp:150-485).
    // init_active < 0: every cell starts active (root call).
    if (init_active < 0) {
        for (uint32_t p = 0; p < n_; ++p)
            if (cfirst_[cend_[p]] == p) add_active(p);
    } else {
        add_active((uint32_t)init_active);
    }
    uint32_t longcode = ncells_;
    int hint = 0;
    while (ncells_ < n_ && nactive_ > 0) {
        int32_t sp = first_set_from(active_, (uint32_t)hint);
        if (sp < 0) sp = first_set_from(active_, 0);
        const uint32_t p = (uint32_t)sp;
        remove_active(p);
        ++steps_;
        const uint32_t e = cend_[p];
        longcode = GCL_MASH(longcode, p + e);
        if (e > p) longcode = GCL_MASH(longcode, e - p + 1);
        if (e == p) singleton_step(elem_[p], tag, longcode, hint);
        else multi_step(p, e, tag, longcode, hint);
    }
    if (nactive_) {
        std::fill(active_.begin(), active_.end(), 0ull);
        nactive_ = 0;
    }
    longcode = GCL_MASH(longcode, ncells_);
    ++refinements_;
    part_bytes_ += 8ull * n_;
    return GCL_CLEANUP(longcode);
}

void CpuSearch::individualize_and_refine(uint32_t v, uint32_t tc_first, uint32_t tag) {
    // split_by_and_refine (reference src/partition_refinement.cpp:486-539):
    // {v} becomes a singleton at the target cell's first position, the rest
    // follows, refinement starts from {v}.
    const uint32_t f = tc_first, c = cend_[f];
    const uint32_t q = pos_[v];
    const uint32_t x = elem_[f];
    elem_[f] = v;
    pos_[v] = f;
    elem_[q] = x;
    pos_[x] = q;
    cend_[f] = f;
    cfirst_[f] = f;
    cfirst_[c] = f + 1;
    set_ns(f, false);
    set_ns(f + 1, c > f + 1);
    ++ncells_;
    undo_id_.push_back(f);
    undo_tag_.push_back(tag);
    code_ = refine((int32_t)f, tag);
}

void CpuSearch::undo_to(uint32_t level) {
    // Merges every cell created above `level`, newest first: the cell
    // [cfirst[x], x] rejoins the cell that starts at x+1.  Equivalent to
    // reconstruct_at_level (reference src/partition_refinement.cpp:541-604)
    // up to the unobservable order inside cells.
    while (!undo_id_.empty() && undo_tag_.back() > level) {
        const uint32_t x = undo_id_.back();
        undo_id_.pop_back();
        undo_tag_.pop_back();
        const uint32_t f = cfirst_[x];
        const uint32_t y = cend_[x + 1];
        for (uint32_t q = f; q <= x; ++q) cend_[q] = y;
        cfirst_[y] = f;
        set_ns(f, true);
        set_ns(x + 1, false);
        --ncells_;
    }
}

void CpuSearch::select_target() {
    // target_cell_selector with hint -1 (reference
    // src/partition_refinement.cpp:607-655): first non-singleton cell; the
    // branch order is increasing vertex id.
    const int32_t f = first_set_from(nsmask_, 0);
    if (f < 0) return;
    LevelRec& r = lv_[tc_level_reached_];
    const uint32_t c = cend_[(uint32_t)f];
    r.tc_first = (uint32_t)f;
    r.tc_len = c - (uint32_t)f + 1;
    r.cursor = 0;
    r.last_num_aut = 0;
    r.arena_off = tc_level_reached_ == 0 ? 0 : lv_[tc_level_reached_ - 1].arena_off + lv_[tc_level_reached_ - 1].tc_len;
    if (r.arena_off + r.tc_len > arena_mem_.size()) {
        arena_mem_.resize((size_t)(r.arena_off + r.tc_len) * 2);
        arena_rep_.resize(arena_mem_.size());
    }
    uint32_t* mem = &arena_mem_[r.arena_off];
    for (uint32_t i = 0; i < r.tc_len; ++i) mem[i] = elem_[(uint32_t)f + i];
    std::sort(mem, mem + r.tc_len);
    uint32_t* rep = &arena_rep_[r.arena_off];
    for (uint32_t i = 0; i < r.tc_len; ++i) rep[i] = mem[i];
    ++tc_level_reached_;
}

void CpuSearch::compute_orbits(LevelRec& r) {
    // subgroup_fixing_sequence + compute_orbits_fast (reference
    // src/perm_group.cpp:5-116): ring slots [last, min(num_aut, max_aut))
    // whose automorphism fixes the current sequence pointwise are merged into
    // the orbit partition of the target cell; representatives are minima.
    const int64_t upto = std::min<int64_t>(num_aut_, (int64_t)max_aut_);
    const int64_t from = r.last_num_aut < 0 ? 0 : r.last_num_aut;
    subgroup_.clear();
    for (int64_t i = from; i < upto; ++i) {
        const uint32_t* a = &auts_[(size_t)i * n_];
        bool fixes = true;
        for (uint32_t j = 0; j < seq_size_ && fixes; ++j)
            if (a[seq_[j]] != seq_[j]) fixes = false;
        if (fixes) subgroup_.push_back(i);
    }
    if (subgroup_.empty()) return;
    uint32_t* mem = &arena_mem_[r.arena_off];
    uint32_t* rep = &arena_rep_[r.arena_off];
    auto find = [&](uint32_t x) {
        while (dsu_[x] != x) x = dsu_[x];
        return x;
    };
    auto unite = [&](uint32_t v, uint32_t w) {
        const uint32_t x = find(v), y = find(w);
        if (x < y) dsu_[y] = x;
        else if (y < x) dsu_[x] = y;
    };
    // An automorphism that fixes the sequence maps this node's partition to
    // itself (refinement is label-invariant), so it permutes the target
    // cell: unions over the cell's members give the members' orbits, whose
    // minima are the representatives.  An image outside the cell would mean
    // otherwise; then every vertex takes part, as in the reference.
    bool inside = true;
    for (int64_t s : subgroup_) {
        const uint32_t* a = &auts_[(size_t)s * n_];
        for (uint32_t i = 0; i < r.tc_len && inside; ++i)
            inside = pos_[a[mem[i]]] - r.tc_first < r.tc_len;
    }
    if (inside) {
        for (uint32_t i = 0; i < r.tc_len; ++i) dsu_[mem[i]] = rep[i];
        for (int64_t s : subgroup_) {
            const uint32_t* a = &auts_[(size_t)s * n_];
            for (uint32_t i = 0; i < r.tc_len; ++i)
                if (a[mem[i]] != mem[i]) unite(mem[i], a[mem[i]]);
        }
    } else {
        for (uint32_t v = 0; v < n_; ++v) dsu_[v] = v;
        for (uint32_t i = 0; i < r.tc_len; ++i) dsu_[mem[i]] = rep[i];
        for (int64_t s : subgroup_) {
            const uint32_t* a = &auts_[(size_t)s * n_];
            for (uint32_t v = 0; v < n_; ++v)
                if (a[v] != v) unite(v, a[v]);
        }
    }
    for (uint32_t i = 0; i < r.tc_len; ++i) rep[i] = find(mem[i]);
}

void CpuSearch::leaf_keys(uint32_t* keys) const {
    // Sorted keys i*n+j of the leaf's relabeled graph: the sparse form of
    // buildPermutedFlatAdjacencyMatrixArray (reference src/CSRGraph.cpp:332-351).
    uint64_t k = 0;
    for (uint32_t i = 0; i < n_; ++i) {
        const uint32_t v = elem_[i];
        const uint64_t b = k;
        for (int32_t e = g_.row_ptr[v]; e < g_.row_ptr[v + 1]; ++e)
            keys[k++] = i * n_ + pos_[(uint32_t)g_.col[e]];
        std::sort(keys + b, keys + k);
    }
}

int CpuSearch::compare_keys(const uint32_t* a, const uint32_t* b, uint64_t nnz) {
    // compare_adj_upper (reference src/sequential.cpp:114-124) on sorted keys:
    // at the first difference the smaller key marks the matrix with the
    // earlier 1, which is the greater matrix.
    for (uint64_t k = 0; k < nnz; ++k)
        if (a[k] != b[k]) return a[k] < b[k] ? 1 : -1;
    return 0;
}

void CpuSearch::store_aut(const int32_t* ref_label) {
    const size_t slot = (size_t)(num_aut_ % (int64_t)max_aut_);
    uint32_t* dst = &auts_[slot * n_];
    for (uint32_t j = 0; j < n_; ++j) dst[(uint32_t)ref_label[j]] = elem_[j];
    ++num_aut_;
}

int CpuSearch::gca_level(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b, uint32_t m) {
    // get_gca_level (reference src/sequential.cpp:183-213); m = min length
    for (uint32_t k = 0; k < m; ++k)
        if (a[k] != b[k]) return (int)k + 1;
    return -1;
}

void CpuSearch::backtrack(uint32_t level) {
    // backtrack_to (reference src/sequential.cpp:650-670)
    if (level == 0) {
        level_ = 0;
        return;
    }
    undo_to(level);
    if (level < eqlevfirst_) eqlevfirst_ = level;
    if ((int)level <= eqlevcanon_) {
        eqlevcanon_ = (int)level;
        comp_canon_ = 0;
    }
    seq_size_ = level - 1;
    tc_level_reached_ = level;
    level_ = level;
}

void CpuSearch::first_terminal() {
    // first_terminal (reference src/sequential.cpp:259-303)
    eqlevfirst_ = level_;
    firstcode_[level_] = kSentinel;
    leaf_keys(first_keys_.data());
    best_keys_ = first_keys_;
    first_seq_size_ = best_seq_size_ = seq_size_;
    std::copy(seq_.begin(), seq_.begin() + seq_size_, first_seq_.begin());
    std::copy(seq_.begin(), seq_.begin() + seq_size_, best_seq_.begin());
    for (uint32_t j = 0; j < n_; ++j) first_label_[j] = best_label_[j] = (int32_t)elem_[j];
    first_discovered_ = true;
    canon_level_ = eqlevcanon_ = (int)level_;
    comp_canon_ = 0;
    for (uint32_t i = 0; i < level_; ++i) canon_code_[i] = firstcode_[i];
    canon_code_[level_] = kSentinel;
}

void CpuSearch::first_path_node() {
    // first_path_node (reference src/sequential.cpp:216-257)
    if (ncells_ == n_) {
        first_terminal();
        backtrack(level_ - 1);
        return;
    }
    select_target();
    LevelRec& r = lv_[tc_level_reached_ - 1];
    const uint32_t v = arena_mem_[r.arena_off + r.cursor++];
    ++nodes_;
    seq_[seq_size_++] = v;
    individualize_and_refine(v, r.tc_first, level_ + 1);
    ++level_;
    firstcode_[level_ - 1] = code_;
}

void CpuSearch::classify() {
    // classify_node (reference src/sequential.cpp:396-624)
    const uint32_t L = level_;
    if (eqlevfirst_ == L - 1 && code_ == firstcode_[L - 1]) eqlevfirst_ = L;
    if (eqlevcanon_ == (int)L - 1) {
        if (code_ < canon_code_[L - 1]) comp_canon_ = -1;
        else if (code_ > canon_code_[L - 1]) comp_canon_ = 1;
        else {
            comp_canon_ = 0;
            eqlevcanon_ = (int)L;
        }
    }
    if (comp_canon_ > 0) canon_code_[L - 1] = code_;

    int decision = kNormal;
    bool keys_ready = false;
    if (eqlevfirst_ != L && comp_canon_ < 0) {
        decision = kBad;
    } else if (ncells_ == n_) {
        if (eqlevfirst_ == L) {
            // isAutomorphism(first -> current) <=> equal relabeled graphs
            leaf_keys(cur_keys_.data());
            keys_ready = true;
            if (compare_keys(cur_keys_.data(), first_keys_.data(), nnz_) == 0) decision = kAutFirst;
        }
        if (decision == kNormal) {
            if (comp_canon_ == 0) {
                if ((int)L < canon_level_) {
                    comp_canon_ = 1;
                } else {
                    if (!keys_ready) leaf_keys(cur_keys_.data());
                    keys_ready = true;
                    comp_canon_ = compare_keys(cur_keys_.data(), best_keys_.data(), nnz_);
                }
            }
            if (comp_canon_ == 0) decision = kAutBest;
            else if (comp_canon_ > 0) decision = kNewBest;
            else decision = kBad;
        }
    }
    switch (decision) {
        case kNormal:
            break;
        case kAutFirst: {
            store_aut(first_label_.data());
            int gl = gca_level(first_seq_, seq_, std::min(first_seq_size_, seq_size_));
            backtrack(gl >= 0 ? (uint32_t)gl : L - 1);
            break;
        }
        case kAutBest: {
            store_aut(best_label_.data());
            int gl = gca_level(best_seq_, seq_, std::min(best_seq_size_, seq_size_));
            backtrack(gl >= 0 ? (uint32_t)gl : L - 1);
            break;
        }
        case kNewBest: {
            canon_level_ = eqlevcanon_ = (int)L;
            comp_canon_ = 0;
            canon_code_[L] = kSentinel;
            for (uint32_t j = 0; j < n_; ++j) best_label_[j] = (int32_t)elem_[j];
            if (!keys_ready) leaf_keys(cur_keys_.data());
            best_keys_.swap(cur_keys_);
            std::copy(seq_.begin(), seq_.begin() + seq_size_, best_seq_.begin());
            best_seq_size_ = seq_size_;
            backtrack(L - 1);
            break;
        }
        case kBad:
            backtrack(L - 1);
            break;
    }
}

void CpuSearch::other_path_node() {
    // other_path_node (reference src/sequential.cpp:304-392)
    if (ncells_ < n_ && (eqlevfirst_ == level_ || comp_canon_ >= 0)) {
        if (tc_level_reached_ < level_) select_target();
        LevelRec& r = lv_[level_ - 1];
        if (r.cursor >= r.tc_len) {
            backtrack(level_ - 1);
            return;
        }
        if (num_aut_ > 0 && num_aut_ > r.last_num_aut) {
            compute_orbits(r);
            r.last_num_aut = num_aut_;
        }
        const uint32_t* mem = &arena_mem_[r.arena_off];
        const uint32_t* rep = &arena_rep_[r.arena_off];
        while (r.cursor < r.tc_len && rep[r.cursor] != mem[r.cursor]) ++r.cursor;
        if (r.cursor >= r.tc_len) {
            backtrack(level_ - 1);
            return;
        }
        const uint32_t v = mem[r.cursor++];
        ++nodes_;
        seq_[seq_size_++] = v;
        individualize_and_refine(v, r.tc_first, level_ + 1);
        ++level_;
    }
    classify();
}

gcl_status CpuSearch::run(int32_t* out_label, uint32_t* out_keys, gcl_stats* st) {
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = n_;
    nnz_ = (uint64_t)g_.nnz;
    try {
        cnt_.assign(n, 0);
        hits_.assign(n, 0);
        cmin_.assign(n, 0);
        cmax_.assign(n, 0);
        hcur_.assign(n, 0);
        cval_.assign(n, 0);
        frag_off_.assign(n, 0);
        frag_n_.assign(n, 0);
        mark_.assign(n, 0);
        cstate_.assign(n, 0);
        active_.assign((n + 63) / 64 + 1, 0);
        seq_.assign(n + 1, 0);
        lv_.assign(n + 1, LevelRec());
        arena_mem_.assign((size_t)4 * n + 64, 0);
        arena_rep_.assign(arena_mem_.size(), 0);
        firstcode_.assign(n + 2, 0);
        canon_code_.assign(n + 2, 0);
        first_label_.assign(n, 0);
        best_label_.assign(n, 0);
        first_seq_.assign(n + 1, 0);
        best_seq_.assign(n + 1, 0);
        first_keys_.assign(nnz_, 0);
        best_keys_.assign(nnz_, 0);
        cur_keys_.assign(nnz_, 0);
        auts_.assign((size_t)max_aut_ * n, 0);
        dsu_.assign(n, 0);
    } catch (const std::bad_alloc&) {
        return fail(GCL_E_NOMEM, "host allocation failed for n=%u max_aut=%u", n, max_aut_);
    }
    init_partition();
    // search_tree_traversal (reference src/sequential.cpp:626-648)
    code_ = refine(-1, 1);
    level_ = 1;
    firstcode_[0] = code_;
    ++nodes_;
    while (level_ >= 1) {
        if (!first_discovered_) first_path_node();
        else other_path_node();
    }
    if (out_label)
        for (uint32_t i = 0; i < n; ++i) out_label[i] = best_label_[i];
    if (out_keys && nnz_) std::copy(best_keys_.begin(), best_keys_.end(), out_keys);
    if (st) {
        st->nodes = nodes_;
        st->automorphisms = (uint64_t)num_aut_;
        st->refinements = refinements_;
        st->splitter_steps = steps_;
        st->csr_bytes = csr_bytes_;
        st->alg_bytes = csr_bytes_ + part_bytes_;
        st->kernel_ms = 0.0;
        st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return GCL_OK;
}

}  // namespace

gcl_status cpu_canon_label(const gcl_csr& g, uint32_t max_aut, int32_t* out_label, uint32_t* out_keys,
                           gcl_stats* st) {
    CpuSearch s(g, max_aut);
    return s.run(out_label, out_keys, st);
}

}  // namespace gcl

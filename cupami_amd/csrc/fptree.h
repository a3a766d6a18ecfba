// FP-tree internals shared by the tree build and projection (fptree.hip)
// and the FP-growth search (fpgrowth.hip).
#pragma once

#include "sort.h"

#include <algorithm>
#include <vector>

// A tree in the layout of _fptree.build_tree (_fptree.py:108-177): node 0 is
// the root (parent -1, item -1, count 0), items are the caller's ids.
struct cpm_fptree {
    cpm_ctx *ctx = nullptr;
    uint64_t n_nodes = 0;
    uint32_t depth = 0;
    bool single_path = true;
    uint32_t item_count = 0;
    cpm::DBuf parents, items, counts;  // int64, int64, uint64
    std::vector<uint32_t> header_item;  // ascending (support, then descending id)
    std::vector<uint64_t> header_sup;
    bool links = false;
    cpm::DBuf link_nodes;               // node ids grouped by item, ascending within an item
    std::vector<uint32_t> link_off;     // item_count + 1
};

namespace cpm {

// Working buffers of a build / projection / search.
struct FpWs {
    SortWs sort;
    DBuf keys, vals, flag, pos, rows, cols, rows2, cols2, weights, rank, lengths, off, act, act2,
        pfr, sup, freq, total, nodes, scalar;
    HBuf host;
};

inline dim3 grid_of(uint64_t n) {
    return dim3((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 16384)));
}

// One u64 from the device (waits for the stream).
int read_u64(cpm_ctx *ctx, FpWs &ws, const void *d, uint64_t *out);
// Two device regions to host memory with one wait (through ws.host).
int read_two(cpm_ctx *ctx, FpWs &ws, const void *d0, size_t b0, void *h0, const void *d1, size_t b1, void *h1);

// Build a tree from weighted baskets given as (row, col) pairs in device
// memory; d_supports (item_count u64, or null) gives the item supports.  *out
// is null when no frequent pair remains.
int fptree_build(cpm_ctx *ctx, FpWs &ws, const uint64_t *d_rows, const uint32_t *d_cols, uint64_t nnz,
                 const uint64_t *d_weights, uint64_t n_rows, uint32_t item_count, uint64_t minsup,
                 const unsigned long long *d_supports, cpm_fptree **out);
// Node lists per item (t->link_nodes, t->link_off), made once per tree.
int build_links(cpm_ctx *ctx, FpWs &ws, cpm_fptree *t);
// Conditional tree of `item` (null when no ancestor item is frequent).
int fptree_conditional(cpm_ctx *ctx, FpWs &ws, cpm_fptree *t, uint32_t item, uint64_t minsup, cpm_fptree **out);
// Deduplicate (row, col) pairs, sorted by (row, col).
int unique_pairs(cpm_ctx *ctx, FpWs &ws, const uint64_t *d_rows, const uint32_t *d_cols, uint64_t n,
                 uint64_t *d_out_rows, uint32_t *d_out_cols, uint64_t *n_out, uint64_t max_row);

}  // namespace cpm

// FP-tree build, conditional projection and FP-growth on the device.
//
// Tree layout and numbering follow _fptree.build_tree (_fptree.py:108-177):
// node 0 is the root (parent -1, item -1, count 0); the nodes of depth L come
// next, ordered by (parent node, item id).  Each basket's frequent items are
// ordered by (weighted support desc, item id asc) and a basket of length m
// walks levels 0..m-1.  One level is built by
//
//   keys     : parent node << item_bits | item at that level, per active row
//   sort     : stable radix sort of the keys (sort.hip)
//   groups   : first-of-run flags + scan -> node id of every row
//   nodes    : parent/item of each new node, counts += row weights (u64 atomics)
//
// so the merging of equal prefixes is a sort and a segmented group-by rather
// than pointer insertion.  Conditional projection follows
// _fptree.conditional_tree (_fptree.py:180-231): ancestor walks add each node's
// count into its ancestors' items, frequent ancestors are emitted as a
// weighted basket per node, and the basket set is built into a tree again.
// The FP-growth search over these trees is fpgrowth.hip.
#include "fptree.h"

#include <string.h>

namespace cpm {

namespace {

__global__ void k_item_supports(const uint64_t *__restrict__ rows, const uint32_t *__restrict__ cols,
                                uint64_t n, const uint64_t *__restrict__ w,
                                unsigned long long *__restrict__ sup) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(sup + cols[i], (unsigned long long)w[rows[i]]);
}

__global__ void k_flag_frequent(const uint32_t *__restrict__ cols, uint64_t n,
                                const unsigned long long *__restrict__ sup, uint64_t minsup,
                                uint32_t *__restrict__ flag) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        flag[i] = sup[cols[i]] >= minsup ? 1u : 0u;
}

__global__ void k_compact_pairs(const uint64_t *__restrict__ rows, const uint32_t *__restrict__ cols,
                                const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
                                uint64_t n, uint64_t *__restrict__ orows, uint32_t *__restrict__ ocols) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (flag[i]) {
            orows[pos[i]] = rows[i];
            ocols[pos[i]] = cols[i];
        }
}

// item order: by support descending, ties by id (the sort is stable)
__global__ void k_rank_keys(const unsigned long long *__restrict__ sup, uint32_t n,
                            uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        keys[i] = ~(uint64_t)sup[i];
        vals[i] = i;
    }
}

__global__ void k_rank_scatter(const uint32_t *__restrict__ order, uint32_t n, uint32_t *__restrict__ rank) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        rank[order[i]] = i;
}

__global__ void k_pair_keys(const uint64_t *__restrict__ rows, const uint32_t *__restrict__ cols,
                            uint64_t n, const uint32_t *__restrict__ rank, uint64_t *__restrict__ keys,
                            uint32_t *__restrict__ vals, uint32_t *__restrict__ lengths) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = rows[i];
        keys[i] = (r << 32) | rank[cols[i]];
        vals[i] = cols[i];
        atomicAdd(lengths + r, 1u);
    }
}

// flag[i] = lengths[row(i)] > level; rows are ids[i] (or i when ids is null)
__global__ void k_flag_len(const uint32_t *__restrict__ ids, uint64_t n,
                           const uint32_t *__restrict__ lengths, uint32_t level,
                           uint32_t *__restrict__ flag) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = ids ? ids[i] : (uint32_t)i;
        flag[i] = lengths[r] > level ? 1u : 0u;
    }
}

__global__ void k_compact_ids(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ flag,
                              const uint32_t *__restrict__ pos, uint64_t n, uint32_t *__restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (flag[i]) out[pos[i]] = ids ? ids[i] : (uint32_t)i;
}

__global__ void k_level_keys(const uint32_t *__restrict__ act, uint64_t n,
                             const uint32_t *__restrict__ pfr, const uint32_t *__restrict__ cols,
                             const uint32_t *__restrict__ off, uint32_t level, int ib,
                             uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = act[i];
        keys[i] = ((uint64_t)pfr[r] << ib) | cols[off[r] + level];
        vals[i] = r;
    }
}

__global__ void k_group_flags(const uint64_t *__restrict__ keys, uint64_t n, uint32_t *__restrict__ flag) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_level_nodes(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                              const uint32_t *__restrict__ flag, const uint32_t *__restrict__ excl,
                              uint64_t n, int ib, uint64_t base, const uint64_t *__restrict__ w,
                              int64_t *__restrict__ parents, int64_t *__restrict__ items,
                              unsigned long long *__restrict__ counts, uint32_t *__restrict__ pfr) {
    const uint64_t mask = (1ull << ib) - 1;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t node = base + excl[i] + flag[i] - 1;
        const uint32_t r = vals[i];
        if (flag[i]) {
            parents[node] = (int64_t)(keys[i] >> ib);
            items[node] = (int64_t)(keys[i] & mask);
        }
        atomicAdd(counts + node, (unsigned long long)w[r]);
        pfr[r] = (uint32_t)node;
    }
}

// One level of the tree build for n <= kSmallLevel active rows, in one
// workgroup: the steps of k_level_keys, the stable sort by (parent node,
// item), k_group_flags and its scan, k_level_nodes, k_flag_len and its scan,
// and k_compact_ids, in the same order, so the tree is the same node for node.
// totals[0] = nodes added, totals[1] = rows that continue (now in act).
constexpr uint32_t kSmallLevel = 2048;
__global__ __launch_bounds__(1024) void k_level_small(uint32_t *__restrict__ act, uint32_t n,
                                                     uint32_t *__restrict__ pfr, const uint32_t *__restrict__ cols,
                                                     const uint32_t *__restrict__ off,
                                                     const uint32_t *__restrict__ lengths, uint32_t level, int ib,
                                                     uint64_t base, const uint64_t *__restrict__ w,
                                                     int64_t *__restrict__ parents, int64_t *__restrict__ items,
                                                     unsigned long long *__restrict__ counts,
                                                     uint64_t *__restrict__ totals) {
    __shared__ uint64_t sk[kSmallLevel];
    __shared__ uint32_t si[kSmallLevel];
    __shared__ uint32_t rows[kSmallLevel];
    __shared__ uint32_t lds[16];
    const uint32_t tid = threadIdx.x;
    uint32_t P = 1;
    while (P < n) P <<= 1;
    for (uint32_t i = tid; i < P; i += blockDim.x) {
        if (i < n) {
            const uint32_t r = act[i];
            rows[i] = r;
            sk[i] = ((uint64_t)pfr[r] << ib) | cols[off[r] + level];
        } else {
            sk[i] = ~0ull;
        }
        si[i] = i;
    }
    __syncthreads();
    bitonic_sort_lds(sk, si, P);
    // lane tid owns sorted positions 2 tid, 2 tid + 1
    const uint64_t mask = (1ull << ib) - 1;
    uint32_t f[2], r[2];
    uint64_t key[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const uint32_t i = 2 * tid + q;
        f[q] = 0;
        if (i < n) {
            key[q] = sk[i];
            r[q] = rows[si[i]];
            f[q] = (i == 0 || sk[i] != sk[i - 1]) ? 1u : 0u;
        }
    }
    uint32_t groups;
    uint32_t excl = block_excl_scan_u32(f[0] + f[1], lds, &groups);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const uint32_t i = 2 * tid + q;
        if (i < n) {
            const uint64_t node = base + excl + f[q] - 1;
            if (f[q]) {
                parents[node] = (int64_t)(key[q] >> ib);
                items[node] = (int64_t)(key[q] & mask);
            }
            atomicAdd(counts + node, (unsigned long long)w[r[q]]);
            pfr[r[q]] = (uint32_t)node;
        }
        excl += f[q];
    }
    // rows continuing to the next level, compacted in sorted order
    uint32_t g[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const uint32_t i = 2 * tid + q;
        g[q] = i < n && lengths[r[q]] > level + 1 ? 1u : 0u;
    }
    uint32_t next;
    uint32_t pos = block_excl_scan_u32(g[0] + g[1], lds, &next);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        if (g[q]) act[pos] = r[q];
        pos += g[q];
    }
    if (tid == 0) {
        totals[0] = groups;
        totals[1] = next;
    }
}

__global__ void k_root(int64_t *parents, int64_t *items, unsigned long long *counts) {
    parents[0] = -1;
    items[0] = -1;
    counts[0] = 0;
}

__global__ void k_max_u32(const uint32_t *__restrict__ v, uint64_t n, unsigned int *__restrict__ out) {
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        m = max(m, v[i]);
    atomicMax(out, m);
}

// ---- conditional projection (_fptree.py:21-71) ----------------------------------

__global__ void k_link_keys(const int64_t *__restrict__ items, uint64_t n_nodes,
                            uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
                            uint32_t *__restrict__ hist) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i + 1 < n_nodes;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t node = i + 1;
        keys[i] = (uint64_t)items[node];
        vals[i] = (uint32_t)node;
        atomicAdd(hist + items[node], 1u);
    }
}

__global__ void k_ancestor_supports(const uint32_t *__restrict__ nodes, uint64_t n,
                                    const int64_t *__restrict__ parents,
                                    const int64_t *__restrict__ items,
                                    const unsigned long long *__restrict__ counts,
                                    unsigned long long *__restrict__ sup) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t node = nodes[i];
        const unsigned long long w = counts[node];
        for (int64_t p = parents[node]; p > 0; p = parents[p]) atomicAdd(sup + items[p], w);
    }
}

__global__ void k_proj_lengths(const uint32_t *__restrict__ nodes, uint64_t n,
                               const int64_t *__restrict__ parents, const int64_t *__restrict__ items,
                               const uint32_t *__restrict__ frequent, uint32_t *__restrict__ len) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t c = 0;
        for (int64_t p = parents[nodes[i]]; p > 0; p = parents[p]) c += frequent[items[p]];
        len[i] = c;
    }
}

__global__ void k_proj_emit(const uint32_t *__restrict__ nodes, uint64_t n,
                            const int64_t *__restrict__ parents, const int64_t *__restrict__ items,
                            const uint32_t *__restrict__ frequent, const uint32_t *__restrict__ off,
                            const unsigned long long *__restrict__ counts,
                            uint64_t *__restrict__ rows, uint32_t *__restrict__ cols,
                            uint64_t *__restrict__ weights) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t o = off[i];
        for (int64_t p = parents[nodes[i]]; p > 0; p = parents[p]) {
            const int64_t it = items[p];
            if (frequent[it]) {
                rows[o] = i;
                cols[o] = (uint32_t)it;
                ++o;
            }
        }
        weights[i] = counts[nodes[i]];
    }
}

// ---- unique pairs (_fptree.py:97-105) ---------------------------------------------

__global__ void k_pair_rowcol_keys(const uint64_t *__restrict__ rows, const uint32_t *__restrict__ cols,
                                   uint64_t n, uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        keys[i] = (rows[i] << 32) | cols[i];
        vals[i] = 0;
    }
}

__global__ void k_unique_emit(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ flag,
                              const uint32_t *__restrict__ pos, uint64_t n,
                              uint64_t *__restrict__ rows, uint32_t *__restrict__ cols) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (flag[i]) {
            rows[pos[i]] = keys[i] >> 32;
            cols[pos[i]] = (uint32_t)keys[i];
        }
}

}  // namespace

int read_u64(cpm_ctx *ctx, FpWs &ws, const void *d, uint64_t *out) {
    CPM_CHECK(ws.host.ensure(64));
    CPM_HIP(hipMemcpyAsync(ws.host.p, d, 8, hipMemcpyDeviceToHost, ctx->stream));
    CPM_HIP(hipStreamSynchronize(ctx->stream));
    *out = *ws.host.as<uint64_t>();
    return CPM_OK;
}

// Two device regions to host memory with one wait (through the pinned buffer).
int read_two(cpm_ctx *ctx, FpWs &ws, const void *d0, size_t b0, void *h0, const void *d1, size_t b1, void *h1) {
    const size_t o1 = (b0 + 63) / 64 * 64;
    CPM_CHECK(ws.host.ensure(o1 + b1 + 64));
    unsigned char *h = ws.host.as<unsigned char>();
    if (b0) CPM_HIP(hipMemcpyAsync(h, d0, b0, hipMemcpyDeviceToHost, ctx->stream));
    if (b1) CPM_HIP(hipMemcpyAsync(h + o1, d1, b1, hipMemcpyDeviceToHost, ctx->stream));
    CPM_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(h0, h, b0);
    memcpy(h1, h + o1, b1);
    return CPM_OK;
}

// Build a tree from weighted baskets given as (row, col) pairs in device
// memory.  d_supports (item_count u64) skips the support count and the
// infrequent-pair filter (conditional bases hold frequent items only).
// *out is null when no frequent pair remains.
int fptree_build(cpm_ctx *ctx, FpWs &ws, const uint64_t *d_rows, const uint32_t *d_cols,
                 uint64_t nnz, const uint64_t *d_weights, uint64_t n_rows, uint32_t item_count,
                 uint64_t minsup, const unsigned long long *d_supports, cpm_fptree **out) {
    *out = nullptr;
    if (nnz == 0 || item_count == 0) return CPM_OK;
    if (nnz >= 0xffffffffull || n_rows >= 0xffffffffull)
        return fail(CPM_ERR_INVALID, "FP-tree input too large (2^32 pairs or rows)");
    hipStream_t st = ctx->stream;
    // supports and frequent pairs
    CPM_CHECK(ws.sup.ensure((size_t)item_count * 8));
    unsigned long long *sup = ws.sup.as<unsigned long long>();
    const uint64_t *rows = d_rows;
    const uint32_t *cols = d_cols;
    uint64_t m = nnz;
    if (!d_supports) {
        CPM_HIP(hipMemsetAsync(sup, 0, (size_t)item_count * 8, st));
        hipLaunchKernelGGL(k_item_supports, grid_of(nnz), dim3(256), 0, st, d_rows, d_cols, nnz,
                           d_weights, sup);
        CPM_CHECK(ws.flag.ensure((nnz + 1) * 4));
        CPM_CHECK(ws.pos.ensure((nnz + 1) * 4));
        CPM_CHECK(ws.total.ensure(16));
        hipLaunchKernelGGL(k_flag_frequent, grid_of(nnz), dim3(256), 0, st, d_cols, nnz, sup, minsup,
                           ws.flag.as<uint32_t>());
        CPM_CHECK(exclusive_scan_u32(ctx, ws.sort, ws.flag.as<uint32_t>(), ws.pos.as<uint32_t>(), nnz,
                                     ws.total.as<uint64_t>()));
        CPM_CHECK(read_u64(ctx, ws, ws.total.p, &m));
        if (m == 0) return CPM_OK;
        CPM_CHECK(ws.rows2.ensure(m * 8));
        CPM_CHECK(ws.cols2.ensure(m * 4));
        hipLaunchKernelGGL(k_compact_pairs, grid_of(nnz), dim3(256), 0, st, d_rows, d_cols,
                           ws.flag.as<uint32_t>(), ws.pos.as<uint32_t>(), nnz,
                           ws.rows2.as<uint64_t>(), ws.cols2.as<uint32_t>());
        CPM_HIP(hipGetLastError());
        rows = ws.rows2.as<uint64_t>();
        cols = ws.cols2.as<uint32_t>();
    } else {
        CPM_HIP(hipMemcpyAsync(sup, d_supports, (size_t)item_count * 8, hipMemcpyDeviceToDevice, st));
    }

    // item ranks: support descending, id ascending
    CPM_CHECK(ws.keys.ensure(std::max<uint64_t>(m, item_count) * 8));
    CPM_CHECK(ws.vals.ensure(std::max<uint64_t>(m, item_count) * 4));
    CPM_CHECK(ws.rank.ensure((size_t)item_count * 4));
    uint64_t *keys = ws.keys.as<uint64_t>();
    uint32_t *vals = ws.vals.as<uint32_t>();
    hipLaunchKernelGGL(k_rank_keys, grid_of(item_count), dim3(256), 0, st, sup, item_count, keys, vals);
    CPM_CHECK(radix_sort_pairs(ctx, ws.sort, keys, vals, item_count, 64));
    hipLaunchKernelGGL(k_rank_scatter, grid_of(item_count), dim3(256), 0, st, vals, item_count,
                       ws.rank.as<uint32_t>());
    // header: frequent items in ascending (support, then descending id)
    std::vector<uint32_t> order(item_count);
    std::vector<uint64_t> hsup(item_count);
    CPM_CHECK(read_two(ctx, ws, vals, (size_t)item_count * 4, order.data(), sup, (size_t)item_count * 8,
                       hsup.data()));

    // baskets sorted by (row, rank) with their lengths
    CPM_CHECK(ws.lengths.ensure((n_rows + 1) * 4));
    CPM_CHECK(ws.off.ensure((n_rows + 1) * 4));
    CPM_HIP(hipMemsetAsync(ws.lengths.p, 0, (n_rows + 1) * 4, st));
    hipLaunchKernelGGL(k_pair_keys, grid_of(m), dim3(256), 0, st, rows, cols, m,
                       ws.rank.as<uint32_t>(), keys, vals, ws.lengths.as<uint32_t>());
    CPM_HIP(hipGetLastError());
    CPM_CHECK(radix_sort_pairs(ctx, ws.sort, keys, vals, m, 32 + bits_for(n_rows - 1)));
    CPM_CHECK(exclusive_scan_u32(ctx, ws.sort, ws.lengths.as<uint32_t>(), ws.off.as<uint32_t>(), n_rows,
                                 nullptr));
    // sorted columns live in vals; keep them
    CPM_CHECK(ws.cols.ensure(m * 4));
    CPM_HIP(hipMemcpyAsync(ws.cols.p, vals, m * 4, hipMemcpyDeviceToDevice, st));
    const uint32_t *scols = ws.cols.as<uint32_t>();
    // depth
    CPM_CHECK(ws.scalar.ensure(16));
    CPM_HIP(hipMemsetAsync(ws.scalar.p, 0, 16, st));
    hipLaunchKernelGGL(k_max_u32, grid_of(n_rows), dim3(256), 0, st, ws.lengths.as<uint32_t>(), n_rows,
                       ws.scalar.as<unsigned int>());

    cpm_fptree *t = new cpm_fptree();
    t->ctx = ctx;
    t->item_count = item_count;
    int s = CPM_OK;
    auto done = [&](int status) {
        if (status != CPM_OK) delete t;
        return status;
    };
    const uint64_t cap = m + 1;
    if ((s = t->parents.ensure(cap * 8)) != CPM_OK) return done(s);
    if ((s = t->items.ensure(cap * 8)) != CPM_OK) return done(s);
    if ((s = t->counts.ensure(cap * 8)) != CPM_OK) return done(s);
    CPM_HIP(hipMemsetAsync(t->counts.p, 0, cap * 8, st));
    hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, st, t->parents.as<int64_t>(), t->items.as<int64_t>(),
                       t->counts.as<unsigned long long>());

    // rows active at level 0
    CPM_CHECK(ws.pfr.ensure((n_rows + 1) * 4));
    CPM_HIP(hipMemsetAsync(ws.pfr.p, 0, (n_rows + 1) * 4, st));
    CPM_CHECK(ws.flag.ensure((std::max(n_rows, m) + 1) * 4));
    CPM_CHECK(ws.pos.ensure((std::max(n_rows, m) + 1) * 4));
    CPM_CHECK(ws.act.ensure((n_rows + 1) * 4));
    CPM_CHECK(ws.total.ensure(16));
    hipLaunchKernelGGL(k_flag_len, grid_of(n_rows), dim3(256), 0, st, nullptr, n_rows,
                       ws.lengths.as<uint32_t>(), 0u, ws.flag.as<uint32_t>());
    CPM_CHECK(exclusive_scan_u32(ctx, ws.sort, ws.flag.as<uint32_t>(), ws.pos.as<uint32_t>(), n_rows,
                                 ws.total.as<uint64_t>()));
    hipLaunchKernelGGL(k_compact_ids, grid_of(n_rows), dim3(256), 0, st, nullptr, ws.flag.as<uint32_t>(),
                       ws.pos.as<uint32_t>(), n_rows, ws.act.as<uint32_t>());
    // the depth (largest basket) and the rows active at level 0, read together
    uint64_t n_act = 0, depth64 = 0;
    CPM_CHECK(read_two(ctx, ws, ws.total.p, 8, &n_act, ws.scalar.p, 8, &depth64));
    const uint32_t depth = (uint32_t)(depth64 & 0xffffffffu);
    t->depth = depth;

    const int ib = std::max(1, bits_for(item_count - 1));
    uint64_t node_count = 1;
    bool single = true;
    for (uint32_t level = 0; level < depth && n_act; ++level) {
        if (n_act <= kSmallLevel) {
            hipLaunchKernelGGL(k_level_small, dim3(1), dim3(1024), 0, st, ws.act.as<uint32_t>(), (uint32_t)n_act,
                               ws.pfr.as<uint32_t>(), scols, ws.off.as<uint32_t>(), ws.lengths.as<uint32_t>(), level,
                               ib, node_count, d_weights, t->parents.as<int64_t>(), t->items.as<int64_t>(),
                               t->counts.as<unsigned long long>(), ws.total.as<uint64_t>());
            CPM_HIP(hipGetLastError());
            uint64_t totals[2] = {0, 0};
            CPM_CHECK(read_two(ctx, ws, ws.total.p, 8, &totals[0], ws.total.as<uint64_t>() + 1, 8, &totals[1]));
            single = single && totals[0] == 1;
            node_count += totals[0];
            n_act = totals[1];
            continue;
        }
        hipLaunchKernelGGL(k_level_keys, grid_of(n_act), dim3(256), 0, st, ws.act.as<uint32_t>(), n_act,
                           ws.pfr.as<uint32_t>(), scols, ws.off.as<uint32_t>(), level, ib, keys, vals);
        CPM_HIP(hipGetLastError());
        CPM_CHECK(radix_sort_pairs(ctx, ws.sort, keys, vals, n_act, ib + bits_for(node_count - 1)));
        hipLaunchKernelGGL(k_group_flags, grid_of(n_act), dim3(256), 0, st, keys, n_act,
                           ws.flag.as<uint32_t>());
        CPM_CHECK(exclusive_scan_u32(ctx, ws.sort, ws.flag.as<uint32_t>(), ws.pos.as<uint32_t>(), n_act,
                                     ws.total.as<uint64_t>()));
        hipLaunchKernelGGL(k_level_nodes, grid_of(n_act), dim3(256), 0, st, keys, vals,
                           ws.flag.as<uint32_t>(), ws.pos.as<uint32_t>(), n_act, ib, node_count,
                           d_weights, t->parents.as<int64_t>(), t->items.as<int64_t>(),
                           t->counts.as<unsigned long long>(), ws.pfr.as<uint32_t>());
        CPM_HIP(hipGetLastError());
        // rows continuing to the next level (their count in the second total)
        hipLaunchKernelGGL(k_flag_len, grid_of(n_act), dim3(256), 0, st, vals, n_act,
                           ws.lengths.as<uint32_t>(), level + 1, ws.flag.as<uint32_t>());
        CPM_CHECK(exclusive_scan_u32(ctx, ws.sort, ws.flag.as<uint32_t>(), ws.pos.as<uint32_t>(), n_act,
                                     ws.total.as<uint64_t>() + 1));
        hipLaunchKernelGGL(k_compact_ids, grid_of(n_act), dim3(256), 0, st, vals, ws.flag.as<uint32_t>(),
                           ws.pos.as<uint32_t>(), n_act, ws.act.as<uint32_t>());
        CPM_HIP(hipGetLastError());
        // the level's new nodes and the next level's rows, read together
        uint64_t totals[2] = {0, 0};
        CPM_CHECK(read_two(ctx, ws, ws.total.p, 8, &totals[0], ws.total.as<uint64_t>() + 1, 8, &totals[1]));
        const uint64_t groups = totals[0];
        single = single && groups == 1;
        node_count += groups;
        n_act = totals[1];
    }
    t->n_nodes = node_count;
    t->single_path = single;
    CPM_HIP(hipStreamSynchronize(st));
    for (uint32_t i = item_count; i-- > 0;) {
        const uint32_t it = order[i];
        if (hsup[it] >= minsup) {
            t->header_item.push_back(it);
            t->header_sup.push_back(hsup[it]);
        }
    }
    *out = t;
    return CPM_OK;
}

int build_links(cpm_ctx *ctx, FpWs &ws, cpm_fptree *t) {
    if (t->links) return CPM_OK;
    hipStream_t st = ctx->stream;
    const uint64_t n = t->n_nodes > 0 ? t->n_nodes - 1 : 0;
    CPM_CHECK(ws.keys.ensure(std::max<uint64_t>(n, 1) * 8));
    CPM_CHECK(ws.vals.ensure(std::max<uint64_t>(n, 1) * 4));
    CPM_CHECK(ws.lengths.ensure(((uint64_t)t->item_count + 1) * 4));
    CPM_CHECK(ws.off.ensure(((uint64_t)t->item_count + 1) * 4));
    CPM_HIP(hipMemsetAsync(ws.lengths.p, 0, ((uint64_t)t->item_count + 1) * 4, st));
    hipLaunchKernelGGL(k_link_keys, grid_of(n), dim3(256), 0, st, t->items.as<int64_t>(), t->n_nodes,
                       ws.keys.as<uint64_t>(), ws.vals.as<uint32_t>(), ws.lengths.as<uint32_t>());
    CPM_HIP(hipGetLastError());
    CPM_CHECK(radix_sort_pairs(ctx, ws.sort, ws.keys.as<uint64_t>(), ws.vals.as<uint32_t>(), n,
                               std::max(1, bits_for(t->item_count - 1))));
    CPM_CHECK(exclusive_scan_u32(ctx, ws.sort, ws.lengths.as<uint32_t>(), ws.off.as<uint32_t>(),
                                 t->item_count, nullptr));
    CPM_CHECK(t->link_nodes.ensure(std::max<uint64_t>(n, 1) * 4));
    if (n)
        CPM_HIP(hipMemcpyAsync(t->link_nodes.p, ws.vals.p, n * 4, hipMemcpyDeviceToDevice, st));
    t->link_off.resize((size_t)t->item_count + 1);
    CPM_CHECK(copy_to_host(ctx, t->link_off.data(), ws.off.p, ((size_t)t->item_count + 1) * 4));
    t->links = true;
    return CPM_OK;
}

static int conditional_from(cpm_ctx *ctx, FpWs &ws, cpm_fptree *t, uint32_t item, uint64_t minsup,
                            const unsigned long long *csup, const uint64_t *hs, cpm_fptree **out);

// Conditional tree of `item` (null when no ancestor item is frequent).
int fptree_conditional(cpm_ctx *ctx, FpWs &ws, cpm_fptree *t, uint32_t item, uint64_t minsup,
                       cpm_fptree **out) {
    *out = nullptr;
    if (item >= t->item_count) return fail(CPM_ERR_INVALID, "item %u >= %u", item, t->item_count);
    CPM_CHECK(build_links(ctx, ws, t));
    hipStream_t st = ctx->stream;
    const uint32_t a = t->link_off[item], b = t->link_off[item + 1];
    const uint64_t nn = b - a;
    if (nn == 0) return CPM_OK;
    const uint32_t *nodes = t->link_nodes.as<uint32_t>() + a;
    const uint32_t I = t->item_count;
    CPM_CHECK(ws.freq.ensure((size_t)I * 8));
    unsigned long long *csup = ws.freq.as<unsigned long long>();  // conditional supports
    CPM_HIP(hipMemsetAsync(csup, 0, (size_t)I * 8, st));
    hipLaunchKernelGGL(k_ancestor_supports, grid_of(nn), dim3(256), 0, st, nodes, nn,
                       t->parents.as<int64_t>(), t->items.as<int64_t>(),
                       t->counts.as<unsigned long long>(), csup);
    CPM_HIP(hipGetLastError());
    std::vector<uint64_t> hs(I);
    CPM_CHECK(copy_to_host(ctx, hs.data(), csup, (size_t)I * 8));
    return conditional_from(ctx, ws, t, item, minsup, csup, hs.data(), out);
}

// The conditional tree of `item` from its conditional supports, on the device
// (csup) and on the host (hs), I values each.
static int conditional_from(cpm_ctx *ctx, FpWs &ws, cpm_fptree *t, uint32_t item, uint64_t minsup,
                            const unsigned long long *csup, const uint64_t *hs, cpm_fptree **out) {
    *out = nullptr;
    CPM_CHECK(build_links(ctx, ws, t));
    hipStream_t st = ctx->stream;
    const uint32_t a = t->link_off[item], b = t->link_off[item + 1];
    const uint64_t nn = b - a;
    if (nn == 0) return CPM_OK;
    const uint32_t *nodes = t->link_nodes.as<uint32_t>() + a;
    const uint32_t I = t->item_count;
    std::vector<uint32_t> header;
    for (uint32_t i = 0; i < I; ++i)
        if (hs[i] >= minsup) header.push_back(i);
    if (header.empty()) return CPM_OK;
    if (header.size() == 1) {
        // one frequent ancestor item: a single-node tree (_fptree.py:210-215)
        cpm_fptree *c = new cpm_fptree();
        c->ctx = ctx;
        c->item_count = I;
        c->depth = 1;
        c->single_path = true;
        c->n_nodes = 2;
        int s = c->parents.ensure(16);
        if (s == CPM_OK) s = c->items.ensure(16);
        if (s == CPM_OK) s = c->counts.ensure(16);
        if (s != CPM_OK) { delete c; return s; }
        int64_t hp[2] = {-1, 0}, hi[2] = {-1, (int64_t)header[0]};
        uint64_t hc[2] = {0, hs[header[0]]};
        CPM_HIP(hipMemcpyAsync(c->parents.p, hp, 16, hipMemcpyHostToDevice, st));
        CPM_HIP(hipMemcpyAsync(c->items.p, hi, 16, hipMemcpyHostToDevice, st));
        CPM_HIP(hipMemcpyAsync(c->counts.p, hc, 16, hipMemcpyHostToDevice, st));
        CPM_HIP(hipStreamSynchronize(st));
        c->header_item.push_back(header[0]);
        c->header_sup.push_back(hs[header[0]]);
        *out = c;
        return CPM_OK;
    }
    // frequent-ancestor flags, projected basket lengths, offsets
    std::vector<uint32_t> fr(I, 0);
    for (uint32_t i : header) fr[i] = 1;
    CPM_CHECK(ws.nodes.ensure((size_t)I * 4));  // frequent flags
    CPM_HIP(hipMemcpyAsync(ws.nodes.p, fr.data(), (size_t)I * 4, hipMemcpyHostToDevice, st));
    CPM_CHECK(ws.flag.ensure((nn + 1) * 4));
    CPM_CHECK(ws.pos.ensure((nn + 1) * 4));
    CPM_CHECK(ws.total.ensure(16));
    hipLaunchKernelGGL(k_proj_lengths, grid_of(nn), dim3(256), 0, st, nodes, nn, t->parents.as<int64_t>(),
                       t->items.as<int64_t>(), ws.nodes.as<uint32_t>(), ws.flag.as<uint32_t>());
    CPM_CHECK(exclusive_scan_u32(ctx, ws.sort, ws.flag.as<uint32_t>(), ws.pos.as<uint32_t>(), nn,
                                 ws.total.as<uint64_t>()));
    uint64_t total = 0;
    CPM_CHECK(read_u64(ctx, ws, ws.total.p, &total));
    DBuf prow, pcol, pw;
    CPM_CHECK(prow.ensure(std::max<uint64_t>(total, 1) * 8));
    CPM_CHECK(pcol.ensure(std::max<uint64_t>(total, 1) * 4));
    CPM_CHECK(pw.ensure(nn * 8));
    hipLaunchKernelGGL(k_proj_emit, grid_of(nn), dim3(256), 0, st, nodes, nn, t->parents.as<int64_t>(),
                       t->items.as<int64_t>(), ws.nodes.as<uint32_t>(), ws.pos.as<uint32_t>(),
                       t->counts.as<unsigned long long>(), prow.as<uint64_t>(), pcol.as<uint32_t>(),
                       pw.as<uint64_t>());
    CPM_HIP(hipGetLastError());
    // the conditional supports become the new tree's supports
    DBuf given;
    CPM_CHECK(given.ensure((size_t)I * 8));
    CPM_HIP(hipMemcpyAsync(given.p, csup, (size_t)I * 8, hipMemcpyDeviceToDevice, st));
    return fptree_build(ctx, ws, prow.as<uint64_t>(), pcol.as<uint32_t>(), total, pw.as<uint64_t>(), nn,
                        I, minsup, given.as<unsigned long long>(), out);
}

// Deduplicate (row, col) pairs, sorted by (row, col) (_fptree.py:97-105).
int unique_pairs(cpm_ctx *ctx, FpWs &ws, const uint64_t *d_rows, const uint32_t *d_cols, uint64_t n,
                 uint64_t *d_out_rows, uint32_t *d_out_cols, uint64_t *n_out, uint64_t max_row) {
    *n_out = 0;
    if (n == 0) return CPM_OK;
    if (max_row >= 0xffffffffull) return fail(CPM_ERR_INVALID, "rows must be < 2^32");
    hipStream_t st = ctx->stream;
    CPM_CHECK(ws.keys.ensure(n * 8));
    CPM_CHECK(ws.vals.ensure(n * 4));
    CPM_CHECK(ws.flag.ensure((n + 1) * 4));
    CPM_CHECK(ws.pos.ensure((n + 1) * 4));
    CPM_CHECK(ws.total.ensure(16));
    hipLaunchKernelGGL(k_pair_rowcol_keys, grid_of(n), dim3(256), 0, st, d_rows, d_cols, n,
                       ws.keys.as<uint64_t>(), ws.vals.as<uint32_t>());
    CPM_CHECK(radix_sort_pairs(ctx, ws.sort, ws.keys.as<uint64_t>(), ws.vals.as<uint32_t>(), n,
                               32 + bits_for(max_row)));
    hipLaunchKernelGGL(k_group_flags, grid_of(n), dim3(256), 0, st, ws.keys.as<uint64_t>(), n,
                       ws.flag.as<uint32_t>());
    CPM_CHECK(exclusive_scan_u32(ctx, ws.sort, ws.flag.as<uint32_t>(), ws.pos.as<uint32_t>(), n,
                                 ws.total.as<uint64_t>()));
    hipLaunchKernelGGL(k_unique_emit, grid_of(n), dim3(256), 0, st, ws.keys.as<uint64_t>(),
                       ws.flag.as<uint32_t>(), ws.pos.as<uint32_t>(), n, d_out_rows, d_out_cols);
    CPM_HIP(hipGetLastError());
    return read_u64(ctx, ws, ws.total.p, n_out);
}

}  // namespace cpm

using namespace cpm;

extern "C" {

int cpm_unique_pairs(cpm_ctx *ctx, const uint64_t *rows, const uint32_t *cols, uint64_t nnz,
                     uint64_t *out_rows, uint32_t *out_cols, uint64_t *out_n) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out_n) return fail(CPM_ERR_INVALID, "null out_n");
    *out_n = 0;
    if (nnz == 0) return CPM_OK;
    if (!rows || !cols || !out_rows || !out_cols) return fail(CPM_ERR_INVALID, "null argument");
    uint64_t max_row = 0;
    for (uint64_t i = 0; i < nnz; ++i) max_row = std::max(max_row, rows[i]);
    FpWs ws;
    DBuf dr, dc, orow, ocol;
    CPM_CHECK(dr.ensure(nnz * 8));
    CPM_CHECK(dc.ensure(nnz * 4));
    CPM_CHECK(orow.ensure(nnz * 8));
    CPM_CHECK(ocol.ensure(nnz * 4));
    CPM_HIP(hipMemcpyAsync(dr.p, rows, nnz * 8, hipMemcpyHostToDevice, ctx->stream));
    CPM_HIP(hipMemcpyAsync(dc.p, cols, nnz * 4, hipMemcpyHostToDevice, ctx->stream));
    uint64_t n = 0;
    CPM_CHECK(unique_pairs(ctx, ws, dr.as<uint64_t>(), dc.as<uint32_t>(), nnz, orow.as<uint64_t>(),
                           ocol.as<uint32_t>(), &n, max_row));
    CPM_CHECK(copy_to_host(ctx, out_rows, orow.p, n * 8));
    CPM_CHECK(copy_to_host(ctx, out_cols, ocol.p, n * 4));
    *out_n = n;
    return CPM_OK;
}

int cpm_fptree_build(cpm_ctx *ctx, const uint64_t *rows, const uint32_t *cols, uint64_t nnz,
                     const uint64_t *weights, uint64_t n_rows, uint32_t item_count, uint64_t minsup,
                     const uint64_t *supports, cpm_fptree **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "null out");
    *out = nullptr;
    if (nnz == 0) return CPM_OK;
    if (!rows || !cols || !weights) return fail(CPM_ERR_INVALID, "null argument");
    for (uint64_t i = 0; i < nnz; ++i)
        if (rows[i] >= n_rows || cols[i] >= item_count)
            return fail(CPM_ERR_INVALID, "pair %llu outside %llu rows x %u items",
                        (unsigned long long)i, (unsigned long long)n_rows, item_count);
    FpWs ws;
    DBuf dr, dc, dw, ds;
    CPM_CHECK(dr.ensure(nnz * 8));
    CPM_CHECK(dc.ensure(nnz * 4));
    CPM_CHECK(dw.ensure(std::max<uint64_t>(n_rows, 1) * 8));
    CPM_HIP(hipMemcpyAsync(dr.p, rows, nnz * 8, hipMemcpyHostToDevice, ctx->stream));
    CPM_HIP(hipMemcpyAsync(dc.p, cols, nnz * 4, hipMemcpyHostToDevice, ctx->stream));
    if (n_rows)
        CPM_HIP(hipMemcpyAsync(dw.p, weights, n_rows * 8, hipMemcpyHostToDevice, ctx->stream));
    if (supports) {
        CPM_CHECK(ds.ensure((size_t)item_count * 8));
        CPM_HIP(hipMemcpyAsync(ds.p, supports, (size_t)item_count * 8, hipMemcpyHostToDevice,
                               ctx->stream));
    }
    return fptree_build(ctx, ws, dr.as<uint64_t>(), dc.as<uint32_t>(), nnz, dw.as<uint64_t>(), n_rows,
                        item_count, minsup, supports ? ds.as<unsigned long long>() : nullptr, out);
}

int cpm_fptree_conditional(cpm_ctx *ctx, cpm_fptree *tree, uint32_t item, uint64_t minsup,
                           cpm_fptree **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!tree || !out) return fail(CPM_ERR_INVALID, "null argument");
    FpWs ws;
    return fptree_conditional(ctx, ws, tree, item, minsup, out);
}

int cpm_fptree_info(const cpm_fptree *t, uint64_t *n_nodes, uint32_t *depth, int *single_path,
                    uint32_t *header_len) {
    if (!t) return fail(CPM_ERR_INVALID, "null tree");
    if (n_nodes) *n_nodes = t->n_nodes;
    if (depth) *depth = t->depth;
    if (single_path) *single_path = t->single_path ? 1 : 0;
    if (header_len) *header_len = (uint32_t)t->header_item.size();
    return CPM_OK;
}

int cpm_fptree_download(const cpm_fptree *t, int64_t *parents, int64_t *items, uint64_t *counts,
                        uint32_t *header_items, uint64_t *header_supports) {
    if (!t) return fail(CPM_ERR_INVALID, "null tree");
    CPM_CHECK(bind(t->ctx));
    hipStream_t st = t->ctx->stream;
    if (parents) CPM_CHECK(copy_to_host(t->ctx, parents, t->parents.p, t->n_nodes * 8));
    if (items) CPM_CHECK(copy_to_host(t->ctx, items, t->items.p, t->n_nodes * 8));
    if (counts) CPM_CHECK(copy_to_host(t->ctx, counts, t->counts.p, t->n_nodes * 8));
    CPM_HIP(hipStreamSynchronize(st));
    if (header_items && !t->header_item.empty())
        memcpy(header_items, t->header_item.data(), t->header_item.size() * 4);
    if (header_supports && !t->header_sup.empty())
        memcpy(header_supports, t->header_sup.data(), t->header_sup.size() * 8);
    return CPM_OK;
}

int cpm_fptree_destroy(cpm_fptree *t) {
    if (!t) return CPM_OK;
    if (t->ctx) {
        (void)hipSetDevice(t->ctx->device);
        (void)hipStreamSynchronize(t->ctx->stream);
    }
    delete t;
    return CPM_OK;
}

}  // extern "C"

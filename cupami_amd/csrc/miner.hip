// Device-resident frequent-itemset search.
//
// The search is level-synchronous over prefix tasks.  A task is a frequent
// prefix P (|P| = d) with a list of extension items E; every entry e of E is
// the candidate P + {e}.  One level runs:
//
//   support  : support (and periodicity) of every entry        (support.hip)
//   plan/1   : per task, compact the kept entries K; count the patterns,
//              the child tasks (K_j, j < |K|-1) and the child entries
//              {K_l : l > j, pair (K_j, K_l) frequent}
//   scan     : offsets of patterns / child tasks / child entries
//   plan/2   : write the patterns (parent pattern id, item, support) and the
//              next level's tasks
//
// so candidate generation, pruning and counting never leave the device; the
// host reads three totals per level to size the next level's buffers.
//
// The children of a task are exactly cuEclat's child prefix classes
// (cuEclat.py:211-245): the entries of child j are the later kept siblings.
// They are pruned with the level-2 pair table (a subset {K_j, K_l} must be
// frequent), which removes candidates only when they cannot be frequent, so
// the itemsets found are the reference's.  Level 1 is one task with the empty
// prefix and every item; level 2 is built over the frequent items with pair
// supports from the INT8 Gram (gram.hip) or from the support kernel.
#include "support_kernels.h"
#include "sort.h"

#include <string.h>
#include <algorithm>

namespace cpm {

namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kGroup = 32;  // extension entries per support work item

struct Keep {
    const uint32_t *sup;
    const int64_t *per;
    uint64_t minsup;
    int64_t maxper;
    int periodic;
    __device__ __forceinline__ bool operator()(uint32_t e) const {
        const uint64_t s = sup[e];
        if (s < minsup) return false;
        if (periodic) return s > 0 && per[e] <= maxper;
        return true;
    }
};

struct PairTable {
    const uint32_t *bits;  // nf x words32, bit (a, b) set when pair a<b is kept
    uint32_t words32;
    __device__ __forceinline__ bool ok(uint32_t a, uint32_t b) const {
        if (!bits) return true;
        return (bits[(uint64_t)a * words32 + (b >> 5)] >> (b & 31)) & 1u;
    }
};

struct Level {
    const uint32_t *prefix;
    uint32_t d;
    const uint32_t *pid;
    const uint32_t *ext_off;
    const uint32_t *ext;
    uint32_t n_tasks;
};

// exclusive scan of one value per lane over a 256-lane block; returns the
// offset, *total receives the block sum.  Uses lds[0..4].
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *lds, uint32_t *total) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // inclusive wave scan
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += lds[w];
    *total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return base + x - v;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *lds) {
    uint32_t t;
    (void)block_scan(v, lds, &t);
    return t;
}

// The kept set of a task as a bit mask over the frequent items, in LDS: a
// child task K_j takes the kept items g > K_j whose pair (K_j, g) is kept,
// which is the AND of this mask with row K_j of the pair table.
constexpr uint32_t kMaskWords = 2048;  // up to 65536 frequent items

__device__ __forceinline__ uint32_t above(uint32_t f) { return ~((2u << (f & 31)) - 1u); }

__device__ __forceinline__ void mask_set(uint32_t *lmask, const uint32_t *__restrict__ kept,
                                         const uint32_t *__restrict__ ext, uint32_t a, uint32_t nk) {
    for (uint32_t j = threadIdx.x; j < nk; j += blockDim.x) {
        const uint32_t f = ext[kept[a + j]];
        atomicOr(&lmask[f >> 5], 1u << (f & 31));
    }
}

__device__ __forceinline__ void mask_clear(uint32_t *lmask, const uint32_t *__restrict__ kept,
                                           const uint32_t *__restrict__ ext, uint32_t a, uint32_t nk) {
    for (uint32_t j = threadIdx.x; j < nk; j += blockDim.x) lmask[ext[kept[a + j]] >> 5] = 0;
}

// number of child entries of kept item fj (mask path)
__device__ __forceinline__ uint32_t child_count(const uint32_t *lmask, const uint32_t *__restrict__ prow,
                                                uint32_t fj, uint32_t wlast) {
    uint32_t w = fj >> 5;
    uint32_t c = __popc(lmask[w] & prow[w] & above(fj));
    for (++w; w <= wlast; ++w) c += __popc(lmask[w] & prow[w]);
    return c;
}

// plan/1 -------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_plan_count(Level L, Keep keep, PairTable pairs,
                                                   int emit_children, uint32_t *__restrict__ kept,
                                                   uint32_t *__restrict__ cnt_pat,
                                                   uint32_t *__restrict__ cnt_child,
                                                   uint32_t *__restrict__ cnt_cext) {
    __shared__ uint32_t lds[8];
    __shared__ uint32_t lmask[kMaskWords];
    const uint32_t tid = threadIdx.x;
    for (uint32_t w = tid; w < kMaskWords; w += blockDim.x) lmask[w] = 0;
    __syncthreads();
    for (uint32_t t = blockIdx.x; t < L.n_tasks; t += gridDim.x) {
        const uint32_t a = L.ext_off[t], b = L.ext_off[t + 1];
        uint32_t nk = 0;
        for (uint32_t i0 = a; i0 < b; i0 += 256) {
            const uint32_t e = i0 + tid;
            const uint32_t f = (e < b && keep(e)) ? 1u : 0u;
            uint32_t tot;
            const uint32_t pos = block_scan(f, lds, &tot);
            if (f) kept[a + nk + pos] = e;
            nk += tot;
        }
        __syncthreads();
        uint32_t cext = 0;
        if (emit_children && nk >= 2) {
            uint32_t c = 0;
            if (pairs.bits) {
                mask_set(lmask, kept, L.ext, a, nk);
                __syncthreads();
                const uint32_t wlast = L.ext[kept[a + nk - 1]] >> 5;
                for (uint32_t j = tid; j + 1 < nk; j += 256) {
                    const uint32_t fj = L.ext[kept[a + j]];
                    c += child_count(lmask, pairs.bits + (uint64_t)fj * pairs.words32, fj, wlast);
                }
            } else {
                for (uint32_t j = tid; j + 1 < nk; j += 256) c += nk - 1 - j;
            }
            cext = block_sum(c, lds);
            if (pairs.bits) mask_clear(lmask, kept, L.ext, a, nk);
        }
        if (tid == 0) {
            cnt_pat[t] = nk;
            cnt_child[t] = (emit_children && nk >= 2) ? nk - 1 : 0;
            cnt_cext[t] = cext;
        }
        __syncthreads();
    }
}

// plan/2 -------------------------------------------------------------------------
struct PlanOut {
    // patterns
    uint32_t *pat_parent, *pat_item, *pat_sup;
    int64_t *pat_per;
    uint32_t pat_base;  // patterns before this level
    const uint32_t *fcol;
    // next level
    uint32_t *prefix, *pid, *ext_off, *ext, *sup, *ngrp;
    // this level's frequent itemsets, (d+1) items each, in pattern order
    // (lexicographic), for the subset test of the next level; may be null
    uint32_t *lv;
};

__global__ __launch_bounds__(256) void k_plan_write(Level L, Keep keep, PairTable pairs,
                                                   int emit_children,
                                                   const uint32_t *__restrict__ kept,
                                                   const uint32_t *__restrict__ cnt_pat,
                                                   const uint32_t *__restrict__ off_pat,
                                                   const uint32_t *__restrict__ off_child,
                                                   const uint32_t *__restrict__ off_cext,
                                                   PlanOut o) {
    __shared__ uint32_t lds[8];
    __shared__ uint32_t lmask[kMaskWords];
    const uint32_t tid = threadIdx.x;
    const uint32_t d1 = L.d + 1;
    for (uint32_t w = tid; w < kMaskWords; w += blockDim.x) lmask[w] = 0;
    __syncthreads();
    for (uint32_t t = blockIdx.x; t < L.n_tasks; t += gridDim.x) {
        const uint32_t a = L.ext_off[t];
        const uint32_t nk = cnt_pat[t];
        const uint32_t pb = o.pat_base + off_pat[t];
        const uint32_t parent = L.pid[t];
        for (uint32_t j = tid; j < nk; j += 256) {
            const uint32_t e = kept[a + j];
            const uint32_t f = L.ext[e];
            o.pat_parent[pb + j] = parent;
            o.pat_item[pb + j] = o.fcol ? o.fcol[f] : f;
            o.pat_sup[pb + j] = keep.sup[e];
            if (o.pat_per) o.pat_per[pb + j] = keep.per[e];
            if (o.lv) {
                uint32_t *dst = o.lv + (uint64_t)(off_pat[t] + j) * d1;
                const uint32_t *src = L.prefix + (uint64_t)t * L.d;
                for (uint32_t i = 0; i < L.d; ++i) dst[i] = src[i];
                dst[L.d] = f;
            }
        }
        if (!(emit_children && nk >= 2)) continue;
        if (pairs.bits) {
            mask_set(lmask, kept, L.ext, a, nk);
            __syncthreads();
        }
        const uint32_t wlast = L.ext[kept[a + nk - 1]] >> 5;
        const uint32_t cb = off_child[t];
        const uint32_t xb = off_cext[t];
        uint32_t carry = 0;
        for (uint32_t j0 = 0; j0 + 1 < nk; j0 += 256) {
            const uint32_t j = j0 + tid;
            const bool live = j + 1 < nk;
            uint32_t cj = 0, fj = 0;
            const uint32_t *prow = nullptr;
            if (live) {
                fj = L.ext[kept[a + j]];
                if (pairs.bits) {
                    prow = pairs.bits + (uint64_t)fj * pairs.words32;
                    cj = child_count(lmask, prow, fj, wlast);
                } else {
                    cj = nk - 1 - j;
                }
            }
            uint32_t tot;
            const uint32_t off = block_scan(cj, lds, &tot);
            if (live) {
                const uint32_t ct = cb + j;
                const uint32_t *src = L.prefix + (uint64_t)t * L.d;
                uint32_t *dst = o.prefix + (uint64_t)ct * d1;
                for (uint32_t i = 0; i < L.d; ++i) dst[i] = src[i];
                dst[L.d] = fj;
                o.pid[ct] = pb + j;
                uint32_t w = xb + carry + off;
                o.ext_off[ct] = w;
                o.ngrp[ct] = (cj + kGroup - 1) / kGroup;
                if (pairs.bits) {
                    for (uint32_t q = fj >> 5; q <= wlast; ++q) {
                        uint32_t m = lmask[q] & prow[q];
                        if (q == (fj >> 5)) m &= above(fj);
                        while (m) {
                            o.ext[w] = q * 32 + (uint32_t)__builtin_ctz(m);
                            o.sup[w] = 0;
                            ++w;
                            m &= m - 1;
                        }
                    }
                } else {
                    for (uint32_t l = j + 1; l < nk; ++l) {
                        o.ext[w] = L.ext[kept[a + l]];
                        o.sup[w] = 0;
                        ++w;
                    }
                }
            }
            carry += tot;
        }
        if (pairs.bits) {
            __syncthreads();
            mask_clear(lmask, kept, L.ext, a, nk);
            __syncthreads();
        }
    }
}

// Subset test (cuApriori.py:487-490): candidate prefix(ct) + {x} of the next
// level survives only if every subset that drops one item of the parent
// prefix P is a frequent itemset of this level (the two subsets that drop the
// last prefix item or x are frequent by construction).  lv holds this
// level's itemsets in lexicographic order, d1 items each.
// Is (prefix without item `drop`) + {x} -- d1 items, ascending -- in lv?
__device__ __forceinline__ bool lv_contains(const uint32_t *__restrict__ lv, uint32_t n_lv, uint32_t d1,
                                            const uint32_t *__restrict__ pre, uint32_t drop,
                                            uint32_t x) {
    uint32_t lo = 0, hi = n_lv;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t *m = lv + (uint64_t)mid * d1;
        int c = 0;
        for (uint32_t i = 0; i < d1 && c == 0; ++i) {
            const uint32_t k = i < drop ? pre[i] : (i + 1 < d1 ? pre[i + 1] : x);
            c = m[i] < k ? -1 : (m[i] > k ? 1 : 0);
        }
        if (c == 0) return true;
        if (c < 0) lo = mid + 1; else hi = mid;
    }
    return false;
}

constexpr uint32_t kMaxPrefix = 64;

__global__ __launch_bounds__(256) void k_prune_flags(const uint32_t *__restrict__ prefix, uint32_t d1,
                                                    const uint32_t *__restrict__ ext_off,
                                                    uint32_t n_child, uint64_t n_ext,
                                                    const uint32_t *__restrict__ ext,
                                                    const uint32_t *__restrict__ lv, uint32_t n_lv,
                                                    uint32_t *__restrict__ live) {
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n_ext;
         w += (uint64_t)gridDim.x * blockDim.x) {
        // task of entry w: largest ct with ext_off[ct] <= w
        uint32_t lo = 0, hi = n_child;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ext_off[mid] <= w) lo = mid; else hi = mid;
        }
        const uint32_t *pre = prefix + (uint64_t)lo * d1;
        const uint32_t x = ext[w];
        bool ok = true;
        for (uint32_t drop = 0; ok && drop + 1 < d1; ++drop) ok = lv_contains(lv, n_lv, d1, pre, drop, x);
        live[w] = ok ? 1u : 0u;
    }
}

// new entry offsets and group counts after the prune
__global__ void k_prune_offsets(const uint32_t *__restrict__ old_off, uint32_t n_child,
                                const uint32_t *__restrict__ pos, uint32_t *__restrict__ new_off,
                                uint32_t *__restrict__ ngrp) {
    for (uint32_t ct = blockIdx.x * blockDim.x + threadIdx.x; ct <= n_child;
         ct += gridDim.x * blockDim.x) {
        const uint32_t a = pos[old_off[ct]];
        new_off[ct] = a;
        if (ct < n_child) ngrp[ct] = (pos[old_off[ct + 1]] - a + kGroup - 1) / kGroup;
    }
}

__global__ void k_prune_compact(const uint32_t *__restrict__ ext, const uint32_t *__restrict__ live,
                                const uint32_t *__restrict__ pos, uint64_t n,
                                uint32_t *__restrict__ out, uint32_t *__restrict__ sup) {
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n;
         w += (uint64_t)gridDim.x * blockDim.x)
        if (live[w]) {
            out[pos[w]] = ext[w];
            sup[pos[w]] = 0;
        }
}

// Exclusive scans of up to three uint32 arrays of length n (one array per
// block); out[i] for i <= n, out[n] = total; totals[blockIdx] = 64-bit total.
struct ScanArgs {
    const uint32_t *in[3];
    uint32_t *out[3];
};

__global__ __launch_bounds__(1024) void k_scan(ScanArgs args, uint32_t n,
                                              unsigned long long *__restrict__ totals) {
    __shared__ uint32_t wsum[16];
    __shared__ unsigned long long carry_s;
    const uint32_t *in = args.in[blockIdx.x];
    uint32_t *out = args.out[blockIdx.x];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) carry_s = 0;
    __syncthreads();
    for (uint32_t i0 = 0; i0 < n; i0 += 1024) {
        const uint32_t i = i0 + tid;
        const uint32_t v = i < n ? in[i] : 0;
        uint32_t x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(x, off, 64);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        uint32_t base = 0, tot = 0;
        for (uint32_t w = 0; w < 16; ++w) {
            if (w < wave) base += wsum[w];
            tot += wsum[w];
        }
        const unsigned long long carry = carry_s;
        if (i < n) out[i] = (uint32_t)(carry + base + x - v);
        __syncthreads();
        if (tid == 0) carry_s = carry + tot;
        __syncthreads();
    }
    if (tid == 0) {
        out[n] = (uint32_t)carry_s;
        totals[blockIdx.x] = carry_s;
    }
}

// Level-1 task: the empty prefix with every item column.
__global__ void k_iota(uint32_t *p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = i;
}

// Level-2 tasks over nf frequent items: task a = prefix {a}, entries a+1..nf-1.
__global__ void k_level2(uint32_t nf, uint32_t *prefix, uint32_t *pid, uint32_t *ext_off,
                         uint32_t *ext, uint32_t *sup, uint32_t *ngrp) {
    for (uint32_t a = blockIdx.x; a + 1 < nf; a += gridDim.x) {
        const uint32_t base = a * (nf - 1) - (uint32_t)(((uint64_t)a * (a - 1)) / 2);
        const uint32_t k = nf - 1 - a;
        if (threadIdx.x == 0) {
            prefix[a] = a;
            pid[a] = a;
            ext_off[a] = base;
            ngrp[a] = (k + kGroup - 1) / kGroup;
            if (a + 2 == nf) ext_off[a + 1] = base + k;
        }
        for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
            ext[base + j] = a + 1 + j;
            sup[base + j] = 0;
        }
    }
}

// Level-2 supports from the Gram matrix (upper triangle, leading dim ld).
__global__ void k_gram_to_level2(const int32_t *gram, uint32_t ld, uint32_t nf, uint32_t *sup) {
    for (uint32_t a = blockIdx.x; a + 1 < nf; a += gridDim.x) {
        const uint32_t base = a * (nf - 1) - (uint32_t)(((uint64_t)a * (a - 1)) / 2);
        for (uint32_t b = a + 1 + threadIdx.x; b < nf; b += blockDim.x)
            sup[base + (b - a - 1)] = (uint32_t)gram[(uint64_t)a * ld + b];
    }
}

// Pair table from the level-2 entries: bit (a, b) = entry (a, b) kept.
__global__ void k_pair_table(uint32_t nf, const uint32_t *ext_off, const uint32_t *ext, Keep keep,
                             uint32_t *bits, uint32_t words32) {
    for (uint32_t a = blockIdx.x; a + 1 < nf; a += gridDim.x) {
        for (uint32_t e = ext_off[a] + threadIdx.x; e < ext_off[a + 1]; e += blockDim.x)
            if (keep(e)) {
                const uint32_t b = ext[e];
                atomicOr(bits + (uint64_t)a * words32 + (b >> 5), 1u << (b & 31));
            }
    }
}

__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }

// ---- host engine -----------------------------------------------------------------

struct LevelBufs {
    DBuf prefix, pid, ext_off, ext, grp_off, ngrp, sup, per;
    uint32_t d = 0, n_tasks = 0;
    uint64_t n_ext = 0;
};

struct Arena {
    DBuf parent, item, sup, per;
    uint64_t cap = 0, n = 0;
    bool periodic = false;
};

int grow_copy(cpm_ctx *ctx, DBuf &b, size_t keep_bytes, size_t want) {
    if (want <= b.cap && b.p) return CPM_OK;
    DBuf nb;
    CPM_CHECK(nb.ensure(want));
    if (keep_bytes)
        CPM_HIP(hipMemcpyAsync(nb.p, b.p, keep_bytes, hipMemcpyDeviceToDevice, ctx->stream));
    CPM_HIP(hipStreamSynchronize(ctx->stream));
    std::swap(b.p, nb.p);
    std::swap(b.cap, nb.cap);
    return CPM_OK;
}

int arena_reserve(cpm_ctx *ctx, Arena &A, uint64_t need) {
    uint64_t cap = A.cap;
    if (need > cap) cap = std::max<uint64_t>(need, std::max<uint64_t>(cap * 2, 4096));
    CPM_CHECK(grow_copy(ctx, A.parent, A.n * 4, cap * 4));
    CPM_CHECK(grow_copy(ctx, A.item, A.n * 4, cap * 4));
    CPM_CHECK(grow_copy(ctx, A.sup, A.n * 4, cap * 4));
    if (A.periodic) CPM_CHECK(grow_copy(ctx, A.per, A.n * 8, cap * 8));
    A.cap = cap;
    return CPM_OK;
}

// Search buffers, kept by the context between searches so that repeated
// searches allocate nothing once the buffers have reached their sizes.
struct SearchWs {
    LevelBufs lv[2];
    Arena arena;
    DBuf kept, cnt[3], off[3], totals, fcol, pairbits, gram, parts, ts;
    DBuf lvitems, live, pos, ext_tmp, off_tmp, ptotal;
    SortWs sort;
    HBuf host;
};

void free_search_ws(void *p) { delete static_cast<SearchWs *>(p); }

SearchWs &search_ws(cpm_ctx *ctx) {
    if (!ctx->search_ws) {
        ctx->search_ws = new SearchWs();
        ctx->search_ws_free = free_search_ws;
    }
    return *static_cast<SearchWs *>(ctx->search_ws);
}

struct Miner {
    cpm_ctx *ctx;
    const cpm_bits *bits;
    uint64_t minsup;
    int64_t maxper;
    int64_t end;
    bool periodic;
    bool use_gram;
    const int64_t *d_ts = nullptr;

    LevelBufs (&lv)[2];
    Arena &arena;
    DBuf &kept;
    DBuf (&cnt)[3];
    DBuf (&off)[3];
    DBuf &totals, &fcol, &pairbits, &gram, &parts;
    HBuf &host;
    SearchWs &ws;
    uint32_t nf = 0;
    bool subset_prune = true;

    Miner(cpm_ctx *c, SearchWs &w)
        : ctx(c), lv(w.lv), arena(w.arena), kept(w.kept), cnt(w.cnt), off(w.off),
          totals(w.totals), fcol(w.fcol), pairbits(w.pairbits), gram(w.gram), parts(w.parts),
          host(w.host), ws(w) {}

    uint32_t grid_tasks(uint32_t n) const {
        return std::max<uint32_t>(1, std::min<uint32_t>(n, (uint32_t)ctx->cu_count * 8));
    }

    Keep keep_of(const LevelBufs &L) const {
        Keep k;
        k.sup = L.sup.as<uint32_t>();
        k.per = L.per.as<int64_t>();
        k.minsup = minsup;
        k.maxper = maxper;
        k.periodic = periodic ? 1 : 0;
        return k;
    }

    Level view(const LevelBufs &L) const {
        Level v;
        v.prefix = L.prefix.as<uint32_t>();
        v.d = L.d;
        v.pid = L.pid.as<uint32_t>();
        v.ext_off = L.ext_off.as<uint32_t>();
        v.ext = L.ext.as<uint32_t>();
        v.n_tasks = L.n_tasks;
        return v;
    }

    // support (and periodicity) of every entry of level L
    int count(LevelBufs &L, const uint32_t *fc) {
        TaskView tv;
        tv.prefix = L.prefix.as<uint32_t>();
        tv.d = L.d;
        tv.ext_off = L.ext_off.as<uint32_t>();
        tv.ext = L.ext.as<uint32_t>();
        tv.grp_off = L.grp_off.as<uint32_t>();
        tv.n_tasks = L.n_tasks;
        tv.fcol = fc;
        tv.G = kGroup;
        // algorithmic bytes: every task reads its d prefix rows once per
        // word, every entry its extension row
        ctx->stats[2] += 8.0 * (double)bits->words * ((double)L.n_tasks * L.d + (double)L.n_ext);
        if (!periodic) return launch_support_count(ctx, bits, tv, L.n_ext, L.sup.as<uint32_t>());
        const uint32_t tiles = tiles_per_row(bits->stride);
        CPM_CHECK(parts.ensure((size_t)L.n_ext * tiles * sizeof(PeriodPart)));
        return launch_support_periodic(ctx, bits, tv, L.n_ext, d_ts, end, parts.as<PeriodPart>(),
                                       L.sup.as<uint32_t>(), L.per.as<int64_t>(), nullptr, 0);
    }

    int scan(const uint32_t *in0, uint32_t *out0, const uint32_t *in1, uint32_t *out1,
             const uint32_t *in2, uint32_t *out2, int arrays, uint32_t n) {
        ScanArgs a;
        a.in[0] = in0; a.out[0] = out0;
        a.in[1] = in1; a.out[1] = out1;
        a.in[2] = in2; a.out[2] = out2;
        hipLaunchKernelGGL(k_scan, dim3(arrays), dim3(1024), 0, ctx->stream, a, n,
                           totals.as<unsigned long long>());
        CPM_HIP(hipGetLastError());
        return CPM_OK;
    }

    // plan the children of level L into level N; returns the number of child
    // tasks through *n_child.
    int plan(LevelBufs &L, LevelBufs &N, bool children, const uint32_t *fc, PairTable pt,
             uint32_t *n_child) {
        *n_child = 0;
        const uint32_t nt = L.n_tasks;
        CPM_CHECK(kept.ensure(std::max<uint64_t>(L.n_ext, 1) * 4));
        for (int i = 0; i < 3; ++i) {
            CPM_CHECK(cnt[i].ensure((size_t)(nt + 1) * 4));
            CPM_CHECK(off[i].ensure((size_t)(nt + 1) * 4));
        }
        Keep kp = keep_of(L);
        Level v = view(L);
        hipLaunchKernelGGL(k_plan_count, dim3(grid_tasks(nt)), dim3(256), 0, ctx->stream, v, kp, pt,
                           children ? 1 : 0, kept.as<uint32_t>(), cnt[0].as<uint32_t>(),
                           cnt[1].as<uint32_t>(), cnt[2].as<uint32_t>());
        CPM_HIP(hipGetLastError());
        CPM_CHECK(scan(cnt[0].as<uint32_t>(), off[0].as<uint32_t>(), cnt[1].as<uint32_t>(),
                       off[1].as<uint32_t>(), cnt[2].as<uint32_t>(), off[2].as<uint32_t>(), 3, nt));
        unsigned long long *ht = host.as<unsigned long long>();
        CPM_HIP(hipMemcpyAsync(ht, totals.p, 3 * 8, hipMemcpyDeviceToHost, ctx->stream));
        CPM_HIP(hipStreamSynchronize(ctx->stream));
        const uint64_t npat = ht[0], nchild = ht[1];
        uint64_t ncext = ht[2];
        if (arena.n + npat >= 0xffffffffull) return fail(CPM_ERR_NOMEM, "more than 2^32 patterns");
        if (ncext >= 0xffffffffull) return fail(CPM_ERR_NOMEM, "more than 2^32 candidates in a level");
        CPM_CHECK(arena_reserve(ctx, arena, arena.n + npat));
        PlanOut o;
        o.pat_parent = arena.parent.as<uint32_t>() ;
        o.pat_item = arena.item.as<uint32_t>();
        o.pat_sup = arena.sup.as<uint32_t>();
        o.pat_per = periodic ? arena.per.as<int64_t>() : nullptr;
        o.pat_base = (uint32_t)arena.n;
        o.fcol = fc;
        const uint32_t d1 = L.d + 1;
        if (children && nchild) {
            CPM_CHECK(N.prefix.ensure(nchild * d1 * 4));
            CPM_CHECK(N.pid.ensure(nchild * 4));
            CPM_CHECK(N.ext_off.ensure((nchild + 1) * 4));
            CPM_CHECK(N.grp_off.ensure((nchild + 1) * 4));
            CPM_CHECK(N.ngrp.ensure((nchild + 1) * 4));
            CPM_CHECK(N.ext.ensure(std::max<uint64_t>(ncext, 1) * 4));
            CPM_CHECK(N.sup.ensure(std::max<uint64_t>(ncext, 1) * 4));
            if (periodic) CPM_CHECK(N.per.ensure(std::max<uint64_t>(ncext, 1) * 8));
        }
        // subset test for candidates of size >= 4 (size 3 is the pair table)
        const bool prune = subset_prune && children && nchild && L.d >= 2 && d1 <= kMaxPrefix;
        o.lv = nullptr;
        if (prune) {
            CPM_CHECK(ws.lvitems.ensure(std::max<uint64_t>(npat, 1) * d1 * 4));
            o.lv = ws.lvitems.as<uint32_t>();
        }
        o.prefix = N.prefix.as<uint32_t>();
        o.pid = N.pid.as<uint32_t>();
        o.ext_off = N.ext_off.as<uint32_t>();
        o.ext = N.ext.as<uint32_t>();
        o.sup = N.sup.as<uint32_t>();
        o.ngrp = N.ngrp.as<uint32_t>();
        hipLaunchKernelGGL(k_plan_write, dim3(grid_tasks(nt)), dim3(256), 0, ctx->stream, v, kp, pt,
                           (children && nchild) ? 1 : 0, kept.as<uint32_t>(), cnt[0].as<uint32_t>(),
                           off[0].as<uint32_t>(), off[1].as<uint32_t>(), off[2].as<uint32_t>(), o);
        CPM_HIP(hipGetLastError());
        arena.n += npat;
        if (children && nchild) {
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, ctx->stream,
                               N.ext_off.as<uint32_t>() + nchild, (uint32_t)ncext);
            CPM_HIP(hipGetLastError());
            uint64_t live_ext = ncext;
            if (prune && ncext) {
                CPM_CHECK(ws.live.ensure((ncext + 1) * 4));
                CPM_CHECK(ws.pos.ensure((ncext + 1) * 4));
                CPM_CHECK(ws.ptotal.ensure(16));
                CPM_CHECK(ws.off_tmp.ensure((nchild + 1) * 4));
                CPM_CHECK(ws.ext_tmp.ensure(ncext * 4));
                hipLaunchKernelGGL(k_prune_flags,
                                   dim3((unsigned)std::min<uint64_t>((ncext + 255) / 256, 16384)),
                                   dim3(256), 0, ctx->stream, N.prefix.as<uint32_t>(), d1,
                                   N.ext_off.as<uint32_t>(), (uint32_t)nchild, ncext,
                                   N.ext.as<uint32_t>(), ws.lvitems.as<uint32_t>(), (uint32_t)npat,
                                   ws.live.as<uint32_t>());
                CPM_HIP(hipGetLastError());
                CPM_CHECK(exclusive_scan_u32(ctx, ws.sort, ws.live.as<uint32_t>(), ws.pos.as<uint32_t>(),
                                             ncext, ws.ptotal.as<uint64_t>()));
                hipLaunchKernelGGL(k_prune_offsets, dim3((unsigned)((nchild + 256) / 256)), dim3(256), 0,
                                   ctx->stream, N.ext_off.as<uint32_t>(), (uint32_t)nchild,
                                   ws.pos.as<uint32_t>(), ws.off_tmp.as<uint32_t>(),
                                   N.ngrp.as<uint32_t>());
                hipLaunchKernelGGL(k_prune_compact, dim3((unsigned)std::min<uint64_t>((ncext + 255) / 256, 16384)),
                                   dim3(256), 0, ctx->stream, N.ext.as<uint32_t>(), ws.live.as<uint32_t>(),
                                   ws.pos.as<uint32_t>(), ncext, ws.ext_tmp.as<uint32_t>(),
                                   N.sup.as<uint32_t>());
                CPM_HIP(hipGetLastError());
                unsigned long long *ht = host.as<unsigned long long>();
                CPM_HIP(hipMemcpyAsync(ht, ws.ptotal.p, 8, hipMemcpyDeviceToHost, ctx->stream));
                CPM_HIP(hipStreamSynchronize(ctx->stream));
                live_ext = ht[0];
                std::swap(N.ext.p, ws.ext_tmp.p);
                std::swap(N.ext.cap, ws.ext_tmp.cap);
                std::swap(N.ext_off.p, ws.off_tmp.p);
                std::swap(N.ext_off.cap, ws.off_tmp.cap);
            }
            ncext = live_ext;
            CPM_CHECK(scan(N.ngrp.as<uint32_t>(), N.grp_off.as<uint32_t>(), nullptr, nullptr,
                           nullptr, nullptr, 1, (uint32_t)nchild));
            N.d = d1;
            N.n_tasks = (uint32_t)nchild;
            N.n_ext = ncext;
            *n_child = (uint32_t)nchild;
        }
        return CPM_OK;
    }

    int run() {
        const uint32_t n_items = bits->n_items;
        arena.periodic = periodic;
        arena.n = 0;
        CPM_CHECK(host.ensure(64));
        CPM_CHECK(totals.ensure(64));
        if (n_items == 0) return CPM_OK;
        ctx->stats[7] = 0;

        // level 1: one task, empty prefix, every column
        LevelBufs &L1 = lv[0];
        L1.d = 0;
        L1.n_tasks = 1;
        L1.n_ext = n_items;
        CPM_CHECK(L1.prefix.ensure(4));
        CPM_CHECK(L1.pid.ensure(4));
        CPM_CHECK(L1.ext_off.ensure(8));
        CPM_CHECK(L1.grp_off.ensure(8));
        CPM_CHECK(L1.ext.ensure((size_t)n_items * 4));
        CPM_CHECK(L1.sup.ensure((size_t)n_items * 4));
        if (periodic) CPM_CHECK(L1.per.ensure((size_t)n_items * 8));
        {
            uint32_t *h = host.as<uint32_t>();
            h[0] = kNone; h[1] = 0; h[2] = n_items; h[3] = 0; h[4] = (n_items + kGroup - 1) / kGroup;
            CPM_HIP(hipMemcpyAsync(L1.pid.p, h, 4, hipMemcpyHostToDevice, ctx->stream));
            CPM_HIP(hipMemcpyAsync(L1.ext_off.p, h + 1, 8, hipMemcpyHostToDevice, ctx->stream));
            CPM_HIP(hipMemcpyAsync(L1.grp_off.p, h + 3, 8, hipMemcpyHostToDevice, ctx->stream));
            CPM_HIP(hipStreamSynchronize(ctx->stream));
        }
        hipLaunchKernelGGL(k_iota, dim3((n_items + 255) / 256), dim3(256), 0, ctx->stream,
                           L1.ext.as<uint32_t>(), n_items);
        CPM_HIP(hipGetLastError());
        CPM_HIP(hipMemsetAsync(L1.sup.p, 0, (size_t)n_items * 4, ctx->stream));
        CPM_CHECK(count(L1, nullptr));
        ctx->stats[7] += 1;
        uint32_t unused = 0;
        PairTable none{nullptr, 0};
        CPM_CHECK(plan(L1, lv[1], false, nullptr, none, &unused));
        nf = (uint32_t)arena.n;
        if (nf < 2) return CPM_OK;
        // frequent item index -> column: the kept entries of the single task
        CPM_CHECK(fcol.ensure((size_t)nf * 4));
        CPM_HIP(hipMemcpyAsync(fcol.p, kept.p, (size_t)nf * 4, hipMemcpyDeviceToDevice,
                               ctx->stream));

        // level 2 over the frequent items
        LevelBufs &L2 = lv[1];
        const uint64_t n2 = (uint64_t)nf * (nf - 1) / 2;
        if (n2 >= 0xffffffffull) return fail(CPM_ERR_NOMEM, "too many frequent items");
        L2.d = 1;
        L2.n_tasks = nf - 1;
        L2.n_ext = n2;
        CPM_CHECK(L2.prefix.ensure((size_t)nf * 4));
        CPM_CHECK(L2.pid.ensure((size_t)nf * 4));
        CPM_CHECK(L2.ext_off.ensure((size_t)(nf + 1) * 4));
        CPM_CHECK(L2.grp_off.ensure((size_t)(nf + 1) * 4));
        CPM_CHECK(L2.ngrp.ensure((size_t)(nf + 1) * 4));
        CPM_CHECK(L2.ext.ensure(n2 * 4));
        CPM_CHECK(L2.sup.ensure(n2 * 4));
        if (periodic) CPM_CHECK(L2.per.ensure(n2 * 8));
        hipLaunchKernelGGL(k_level2, dim3(std::min<uint32_t>(nf, 4096)), dim3(256), 0, ctx->stream,
                           nf, L2.prefix.as<uint32_t>(), L2.pid.as<uint32_t>(),
                           L2.ext_off.as<uint32_t>(), L2.ext.as<uint32_t>(), L2.sup.as<uint32_t>(),
                           L2.ngrp.as<uint32_t>());
        CPM_HIP(hipGetLastError());
        CPM_CHECK(scan(L2.ngrp.as<uint32_t>(), L2.grp_off.as<uint32_t>(), nullptr, nullptr,
                       nullptr, nullptr, 1, nf - 1));
        if (use_gram && !periodic && nf <= 16384 && bits->n_transactions < (1ull << 31)) {
            CPM_CHECK(gram.ensure((size_t)nf * nf * 4));
            CPM_CHECK(launch_gram_i8(ctx, bits, fcol.as<uint32_t>(), nf, gram.as<int32_t>(), nf));
            hipLaunchKernelGGL(k_gram_to_level2, dim3(std::min<uint32_t>(nf, 4096)), dim3(256), 0,
                               ctx->stream, gram.as<int32_t>(), nf, nf, L2.sup.as<uint32_t>());
            CPM_HIP(hipGetLastError());
        } else {
            CPM_CHECK(count(L2, fcol.as<uint32_t>()));
        }
        ctx->stats[7] += 1;
        // pair table from the level-2 outcome (the plan kernels hold a task's
        // kept set in an LDS mask of 65536 bits)
        PairTable pt{nullptr, 0};
        const uint32_t w32 = (nf + 31) / 32;
        if (nf <= 65536) {
        CPM_CHECK(pairbits.ensure((size_t)nf * w32 * 4));
        CPM_HIP(hipMemsetAsync(pairbits.p, 0, (size_t)nf * w32 * 4, ctx->stream));
        hipLaunchKernelGGL(k_pair_table, dim3(std::min<uint32_t>(nf, 4096)), dim3(256), 0,
                           ctx->stream, nf, L2.ext_off.as<uint32_t>(), L2.ext.as<uint32_t>(),
                           keep_of(L2), pairbits.as<uint32_t>(), w32);
        CPM_HIP(hipGetLastError());
        pt = PairTable{pairbits.as<uint32_t>(), w32};
        }

        int cur = 1;
        for (;;) {
            uint32_t nchild = 0;
            CPM_CHECK(plan(lv[cur], lv[cur ^ 1], true, fcol.as<uint32_t>(), pt, &nchild));
            if (nchild == 0) break;
            cur ^= 1;
            CPM_CHECK(count(lv[cur], fcol.as<uint32_t>()));
            ctx->stats[7] += 1;
        }
        return CPM_OK;
    }

    int download(cpm_patterns **out) {
        cpm_patterns *p = new cpm_patterns();
        p->n = arena.n;
        p->periodic = periodic;
        p->parent.resize(arena.n);
        p->item.resize(arena.n);
        p->support.resize(arena.n);
        std::vector<uint32_t> s32(arena.n);
        if (periodic) p->period.resize(arena.n);
        if (arena.n) {
            hipError_t e = hipMemcpyAsync(p->parent.data(), arena.parent.p, arena.n * 4,
                                          hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(p->item.data(), arena.item.p, arena.n * 4,
                                   hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(s32.data(), arena.sup.p, arena.n * 4, hipMemcpyDeviceToHost,
                                   ctx->stream);
            if (e == hipSuccess && periodic)
                e = hipMemcpyAsync(p->period.data(), arena.per.p, arena.n * 8,
                                   hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) {
                delete p;
                return fail(CPM_ERR_RUNTIME, "pattern download failed: %s", hipGetErrorString(e));
            }
            for (uint64_t i = 0; i < arena.n; ++i) p->support[i] = s32[i];
        }
        *out = p;
        return CPM_OK;
    }
};

}  // namespace

}  // namespace cpm

using namespace cpm;

extern "C" int cpm_eclat_mine(cpm_ctx *ctx, const cpm_bits *bits, uint64_t minsup, int use_gram,
                              cpm_patterns **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!bits || !out) return fail(CPM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (bits->n_transactions >= 0xffffffffull)
        return fail(CPM_ERR_INVALID, "supports are counted in 32 bits: too many transactions");
    Miner m(ctx, search_ws(ctx));
    m.bits = bits;
    m.minsup = minsup;
    m.maxper = 0;
    m.end = 0;
    m.periodic = false;
    m.use_gram = use_gram != 0;
    CPM_CHECK(m.run());
    return m.download(out);
}

extern "C" int cpm_pfeclat_mine(cpm_ctx *ctx, const cpm_bits *bits, const int64_t *timestamps,
                                int64_t end, uint64_t minsup, int64_t maxper_floor,
                                cpm_patterns **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!bits || !out) return fail(CPM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (bits->n_transactions >= 0xffffffffull)
        return fail(CPM_ERR_INVALID, "supports are counted in 32 bits: too many transactions");
    if (bits->n_transactions && !timestamps) return fail(CPM_ERR_INVALID, "null timestamps");
    Miner m(ctx, search_ws(ctx));
    m.bits = bits;
    m.minsup = minsup;
    m.maxper = maxper_floor;
    m.end = end;
    m.periodic = true;
    m.use_gram = false;
    DBuf &ts = search_ws(ctx).ts;
    const uint64_t n = bits->n_transactions;
    CPM_CHECK(ts.ensure((n ? n : 1) * 8));
    if (n) CPM_HIP(hipMemcpyAsync(ts.p, timestamps, n * 8, hipMemcpyHostToDevice, ctx->stream));
    m.d_ts = ts.as<int64_t>();
    CPM_CHECK(m.run());
    return m.download(out);
}

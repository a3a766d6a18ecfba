// Device-resident frequent-itemset search over one or many databases.
//
// The search is level-synchronous over prefix tasks.  A task is a frequent
// prefix P (|P| = d) of one database with a list of extension items E; entry
// e of E is the candidate P + {e}.  After levels 1 and 2 are set up, one
// level step is a fixed sequence of launches whose sizes all live in device
// memory:
//
//   plan/1   per task: compact the kept entries K; count patterns, child
//            tasks (K_j, j < |K|-1) and child entries {K_l : l > j, pair
//            (K_j, K_l) kept}
//   scan     offsets of patterns / child tasks / child entries
//   alloc    one lane: check the capacities, size the next level
//   plan/2   write the patterns (parent id, item, support) and next level
//   prune    drop child candidates with an infrequent subset (binary search
//            in this level's sorted itemsets), compact the survivors
//   scan     support work offsets
//   support  count the next level (support.hip)
//
// so the step can be captured once as a hipGraph and replayed level after
// level with no host round trip.  The children of a task are cuEclat's child
// prefix classes (cuEclat.py:211-245); the join/prune is cuApriori's
// (cuApriori.py:473-492).  Pruning removes only candidates that cannot be
// frequent, so the itemsets are the reference's.  When a level outgrows a
// capacity, the step raises an overflow flag, later steps do nothing, and the
// host reruns the search with larger buffers.
#include "support_kernels.h"
#include "sort.h"

#include <string.h>
#include <algorithm>

namespace cpm {

namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kMaskWords = 2048;   // LDS kept mask: up to 65536 frequent items per database
constexpr uint32_t kScanBlocks = 256;

enum Overflow : uint32_t {
    kOvTasks = 1, kOvEntries = 2, kOvPatterns = 4, kOvDepth = 8, kOvLevelSets = 16, kOvParts = 32
};

// Control block in device memory.
struct Ctl {
    LevelSizes sz[2];
    unsigned long long n_patterns;   // patterns written so far
    unsigned long long pat_base;     // first pattern id of the level being planned
    unsigned int overflow;           // Overflow bits
    unsigned int pad;
    unsigned long long live_total;   // survivors of the subset test
    unsigned long long clock[2];     // support launch span (earliest start, latest end)
    unsigned long long clock_ticks;  // accumulated spans
    unsigned long long clock_launches;
    unsigned long long alg_bytes;    // algorithmic bitset bytes of the support launches
    unsigned long long candidates;
    unsigned long long levels;
};

struct Caps {
    uint64_t tasks, ext, patterns, lv, parts;
};

struct PlanArgs {
    LevelPtrs L, N;
    const DbDev *dbs;
    const uint32_t *fcol;       // null at level 1 (entries are columns)
    const uint32_t *pairbits;   // pair tables, null for none
    uint32_t *kept, *cnt_pat, *cnt_child, *cnt_cext, *off_pat, *off_child, *off_cext;
    uint32_t *lvitems;          // this level's itemsets for the subset test
    uint32_t *off2, *ext2, *live, *pos;  // next level's prune buffers
    uint32_t *pat_parent, *pat_item, *pat_sup;
    int64_t *pat_per;
    Ctl *ctl;
    Caps caps;                  // capacities of the next level set
    uint32_t G;
    int emit_children;
    int periodic;
    int prune;                  // run the subset test
    uint32_t parts_tiles;
};

__device__ __forceinline__ bool keep_entry(const PlanArgs &p, const DbDev &D, uint32_t e) {
    const uint64_t s = p.L.sup[e];
    if (s < D.minsup) return false;
    if (p.periodic) return s > 0 && p.L.per[e] <= D.maxper;
    return true;
}

// exclusive scan of one value per lane over a 256-lane block; *total gets
// the block sum.  Uses lds[0..4].
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *lds, uint32_t *total) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
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

__device__ __forceinline__ uint32_t above(uint32_t f) { return ~((2u << (f & 31)) - 1u); }

// A task's kept set as a bit mask over its database's frequent items (local
// index f - fbase), in LDS; a child K_j takes the kept items above it whose
// pair with K_j is kept: the AND of the mask and K_j's pair-table row.
__device__ __forceinline__ void mask_set(uint32_t *lmask, const uint32_t *kept, const uint32_t *ext,
                                         uint32_t a, uint32_t nk, uint32_t fbase) {
    for (uint32_t j = threadIdx.x; j < nk; j += blockDim.x) {
        const uint32_t f = ext[kept[a + j]] - fbase;
        atomicOr(&lmask[f >> 5], 1u << (f & 31));
    }
}

__device__ __forceinline__ void mask_clear(uint32_t *lmask, const uint32_t *kept, const uint32_t *ext,
                                           uint32_t a, uint32_t nk, uint32_t fbase) {
    for (uint32_t j = threadIdx.x; j < nk; j += blockDim.x) lmask[(ext[kept[a + j]] - fbase) >> 5] = 0;
}

__device__ __forceinline__ uint32_t child_count(const uint32_t *lmask, const uint32_t *prow, uint32_t f,
                                                uint32_t wlast) {
    uint32_t w = f >> 5;
    uint32_t c = __popc(lmask[w] & prow[w] & above(f));
    for (++w; w <= wlast; ++w) c += __popc(lmask[w] & prow[w]);
    return c;
}

// plan/1 -------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_plan_count(PlanArgs p) {
    __shared__ uint32_t lds[8];
    __shared__ uint32_t lmask[kMaskWords];
    if (p.ctl->overflow) return;
    const uint32_t tid = threadIdx.x;
    const uint32_t n_tasks = (uint32_t)p.L.sz->n_tasks;
    for (uint32_t w = tid; w < kMaskWords; w += blockDim.x) lmask[w] = 0;
    __syncthreads();
    for (uint32_t t = blockIdx.x; t < n_tasks; t += gridDim.x) {
        const DbDev &D = p.dbs[p.L.db[t]];
        const uint32_t a = p.L.ext_off[t], b = p.L.ext_off[t + 1];
        uint32_t nk = 0;
        for (uint32_t i0 = a; i0 < b; i0 += 256) {
            const uint32_t e = i0 + tid;
            const uint32_t f = (e < b && keep_entry(p, D, e)) ? 1u : 0u;
            uint32_t tot;
            const uint32_t pos = block_scan(f, lds, &tot);
            if (f) p.kept[a + nk + pos] = e;
            nk += tot;
        }
        __syncthreads();
        uint32_t cext = 0;
        if (p.emit_children && nk >= 2) {
            uint32_t c = 0;
            if (p.pairbits) {
                mask_set(lmask, p.kept, p.L.ext, a, nk, D.fbase);
                __syncthreads();
                const uint32_t wlast = (p.L.ext[p.kept[a + nk - 1]] - D.fbase) >> 5;
                const uint32_t *table = p.pairbits + D.pair_off;
                for (uint32_t j = tid; j + 1 < nk; j += 256) {
                    const uint32_t f = p.L.ext[p.kept[a + j]] - D.fbase;
                    c += child_count(lmask, table + (uint64_t)f * D.pair_w32, f, wlast);
                }
            } else {
                for (uint32_t j = tid; j + 1 < nk; j += 256) c += nk - 1 - j;
            }
            (void)block_scan(c, lds, &cext);
            if (p.pairbits) mask_clear(lmask, p.kept, p.L.ext, a, nk, D.fbase);
        }
        if (tid == 0) {
            p.cnt_pat[t] = nk;
            p.cnt_child[t] = (p.emit_children && nk >= 2) ? nk - 1 : 0;
            p.cnt_cext[t] = cext;
        }
        __syncthreads();
    }
}

// alloc (one lane) ------------------------------------------------------------------
__global__ void k_alloc(PlanArgs p) {
    Ctl *c = p.ctl;
    LevelSizes *L = p.L.sz, *N = p.N.sz;
    if (c->overflow) {
        N->n_tasks = N->n_ext = N->n_work = 0;
        return;
    }
    const uint64_t n_pat = L->n_pat;
    const uint64_t n_child = p.emit_children ? L->n_child : 0;
    const uint64_t n_cext = p.emit_children ? L->n_cext : 0;
    const uint64_t d1 = L->d + 1;
    uint32_t ov = 0;
    if (n_child > p.caps.tasks) ov |= kOvTasks;
    if (n_cext > p.caps.ext) ov |= kOvEntries;
    if (c->n_patterns + n_pat > p.caps.patterns) ov |= kOvPatterns;
    if (n_child && d1 > kMaxDepth) ov |= kOvDepth;
    if (p.prune && n_child && L->d >= 2 && n_pat * d1 > p.caps.lv) ov |= kOvLevelSets;
    if (p.periodic && n_cext * p.parts_tiles > p.caps.parts) ov |= kOvParts;
    if (ov) {
        c->overflow |= ov;
        N->n_tasks = N->n_ext = N->n_work = 0;
        return;
    }
    c->pat_base = c->n_patterns;
    c->n_patterns += n_pat;
    N->n_tasks = n_child;
    N->d = d1;
    N->n_ext = n_cext;
    N->n_work = 0;
    if (n_child) {
        p.N.ext_off[n_child] = (uint32_t)n_cext;
        c->levels += 1;
    }
}

// plan/2 -------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_plan_write(PlanArgs p) {
    __shared__ uint32_t lds[8];
    __shared__ uint32_t lmask[kMaskWords];
    if (p.ctl->overflow) return;
    const uint32_t tid = threadIdx.x;
    const uint32_t n_tasks = (uint32_t)p.L.sz->n_tasks;
    const uint32_t d = (uint32_t)p.L.sz->d;
    const uint32_t d1 = d + 1;
    const uint64_t pat0 = p.ctl->pat_base;
    const bool children = p.emit_children && p.N.sz->n_tasks > 0;
    // this level's itemsets feed the subset test of its children (the
    // allocator sized lvitems for exactly this case)
    const bool write_lv = p.prune && p.lvitems && d >= 2 && children;
    for (uint32_t w = tid; w < kMaskWords; w += blockDim.x) lmask[w] = 0;
    __syncthreads();
    for (uint32_t t = blockIdx.x; t < n_tasks; t += gridDim.x) {
        const uint32_t dbi = p.L.db[t];
        const DbDev &D = p.dbs[dbi];
        const uint32_t a = p.L.ext_off[t];
        const uint32_t nk = p.cnt_pat[t];
        const uint64_t pb = pat0 + p.off_pat[t];
        const uint32_t parent = p.L.pid[t];
        const uint32_t *src = p.L.prefix + (uint64_t)t * kMaxDepth;
        for (uint32_t j = tid; j < nk; j += 256) {
            const uint32_t e = p.kept[a + j];
            const uint32_t f = p.L.ext[e];
            p.pat_parent[pb + j] = parent;
            p.pat_item[pb + j] = p.fcol ? p.fcol[f] : f;
            p.pat_sup[pb + j] = p.L.sup[e];
            if (p.pat_per) p.pat_per[pb + j] = p.L.per[e];
            if (write_lv) {
                uint32_t *dst = p.lvitems + (uint64_t)(p.off_pat[t] + j) * d1;
                for (uint32_t i = 0; i < d; ++i) dst[i] = src[i];
                dst[d] = f;
            }
        }
        if (!(children && nk >= 2)) continue;
        if (p.pairbits) {
            mask_set(lmask, p.kept, p.L.ext, a, nk, D.fbase);
            __syncthreads();
        }
        const uint32_t *table = p.pairbits ? p.pairbits + D.pair_off : nullptr;
        const uint32_t wlast = (p.L.ext[p.kept[a + nk - 1]] - D.fbase) >> 5;
        const uint32_t cb = p.off_child[t];
        const uint32_t xb = p.off_cext[t];
        uint32_t carry = 0;
        for (uint32_t j0 = 0; j0 + 1 < nk; j0 += 256) {
            const uint32_t j = j0 + tid;
            const bool live = j + 1 < nk;
            uint32_t cj = 0, fj = 0;
            const uint32_t *prow = nullptr;
            if (live) {
                fj = p.L.ext[p.kept[a + j]];
                if (table) {
                    prow = table + (uint64_t)(fj - D.fbase) * D.pair_w32;
                    cj = child_count(lmask, prow, fj - D.fbase, wlast);
                } else {
                    cj = nk - 1 - j;
                }
            }
            uint32_t tot;
            const uint32_t off = block_scan(cj, lds, &tot);
            if (live) {
                const uint32_t ct = cb + j;
                uint32_t *dst = p.N.prefix + (uint64_t)ct * kMaxDepth;
                for (uint32_t i = 0; i < d; ++i) dst[i] = src[i];
                dst[d] = fj;
                p.N.pid[ct] = (uint32_t)(pb + j);
                p.N.db[ct] = dbi;
                uint32_t w = xb + carry + off;
                p.N.ext_off[ct] = w;
                p.N.nwork[ct] = task_work(cj, p.G, D.tiles_eff);
                if (table) {
                    const uint32_t fl = fj - D.fbase;
                    for (uint32_t q = fl >> 5; q <= wlast; ++q) {
                        uint32_t m = lmask[q] & prow[q];
                        if (q == (fl >> 5)) m &= above(fl);
                        while (m) {
                            p.N.ext[w] = D.fbase + q * 32 + (uint32_t)__builtin_ctz(m);
                            p.N.sup[w] = 0;
                            ++w;
                            m &= m - 1;
                        }
                    }
                } else {
                    for (uint32_t l = j + 1; l < nk; ++l) {
                        p.N.ext[w] = p.L.ext[p.kept[a + l]];
                        p.N.sup[w] = 0;
                        ++w;
                    }
                }
            }
            carry += tot;
        }
        if (p.pairbits) {
            __syncthreads();
            mask_clear(lmask, p.kept, p.L.ext, a, nk, D.fbase);
            __syncthreads();
        }
    }
}

// subset test (cuApriori.py:487-490) ----------------------------------------------
// Is (prefix without item `drop`) + {x} -- d1 items, ascending -- in lv?
__device__ __forceinline__ bool lv_contains(const uint32_t *__restrict__ lv, uint64_t n_lv, uint32_t d1,
                                            const uint32_t *__restrict__ pre, uint32_t drop,
                                            uint32_t x) {
    uint64_t lo = 0, hi = n_lv;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        const uint32_t *m = lv + mid * d1;
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

__device__ __forceinline__ bool prune_active(const PlanArgs &p) {
    return p.prune && !p.ctl->overflow && p.emit_children && p.L.sz->d >= 2 && p.N.sz->n_tasks > 0;
}

// largest t < n with off[t] <= i
__device__ __forceinline__ uint32_t upper_task(const uint32_t *__restrict__ off, uint32_t n, uint64_t i) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_prune_flags(PlanArgs p) {
    if (!prune_active(p)) return;
    const uint32_t n_child = (uint32_t)p.N.sz->n_tasks;
    const uint64_t n_ext = p.N.sz->n_ext;
    const uint32_t d1 = (uint32_t)p.N.sz->d;  // prefix length of the children
    const uint64_t n_lv = p.L.sz->n_pat;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n_ext;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t ct = upper_task(p.N.ext_off, n_child, w);
        const uint32_t *pre = p.N.prefix + (uint64_t)ct * kMaxDepth;
        const uint32_t x = p.N.ext[w];
        bool ok = true;
        // the subsets without the last prefix item or without x are kept by
        // construction; test the ones without each earlier prefix item
        for (uint32_t drop = 0; ok && drop + 1 < d1; ++drop) ok = lv_contains(p.lvitems, n_lv, d1, pre, drop, x);
        p.live[w] = ok ? 1u : 0u;
    }
}

__global__ __launch_bounds__(256) void k_prune_offsets(PlanArgs p) {
    if (!prune_active(p)) return;
    const uint32_t n_child = (uint32_t)p.N.sz->n_tasks;
    for (uint32_t ct = blockIdx.x * blockDim.x + threadIdx.x; ct <= n_child; ct += gridDim.x * blockDim.x) {
        const uint32_t a = p.pos[p.N.ext_off[ct]];
        p.off2[ct] = a;
        if (ct < n_child) {
            const DbDev &D = p.dbs[p.N.db[ct]];
            p.N.nwork[ct] = task_work(p.pos[p.N.ext_off[ct + 1]] - a, p.G, D.tiles_eff);
        }
    }
}

__global__ __launch_bounds__(256) void k_prune_compact(PlanArgs p) {
    if (!prune_active(p)) return;
    const uint64_t n_ext = p.N.sz->n_ext;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n_ext;
         w += (uint64_t)gridDim.x * blockDim.x)
        if (p.live[w]) p.ext2[p.pos[w]] = p.N.ext[w];
}

__global__ __launch_bounds__(256) void k_prune_copyback(PlanArgs p) {
    if (!prune_active(p)) return;
    const uint32_t n_child = (uint32_t)p.N.sz->n_tasks;
    const uint64_t live = p.ctl->live_total;
    // nothing in this launch reads the entry count
    if (blockIdx.x == 0 && threadIdx.x == 0) p.N.sz->n_ext = live;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < live;
         w += (uint64_t)gridDim.x * blockDim.x) {
        p.N.ext[w] = p.ext2[w];
        p.N.sup[w] = 0;
    }
    for (uint32_t ct = blockIdx.x * blockDim.x + threadIdx.x; ct <= n_child; ct += gridDim.x * blockDim.x)
        p.N.ext_off[ct] = p.off2[ct];
}

// Algorithmic bitset bytes of the next support launch: every task reads its
// d prefix rows and one row per entry, W words each.
__global__ __launch_bounds__(256) void k_level_stats(LevelPtrs N, const DbDev *dbs, Ctl *ctl) {
    __shared__ unsigned long long red[4];
    if (ctl->overflow) return;
    const uint32_t n = (uint32_t)N.sz->n_tasks;
    const uint64_t d = N.sz->d;
    unsigned long long s = 0, c = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint64_t k = N.ext_off[t + 1] - N.ext_off[t];
        s += 8ull * dbs[N.db[t]].words * (d + k);
        c += k;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long tot = red[0] + red[1] + red[2] + red[3];
        if (tot) atomicAdd(&ctl->alg_bytes, tot);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long tot = red[0] + red[1] + red[2] + red[3];
        if (tot) atomicAdd(&ctl->candidates, tot);
    }
}

// Fold the last support launch's per-block clock stamps (earliest start,
// latest end) into the totals.  One block; stamps of blocks that did not
// run this launch are older than its start and are skipped.
__global__ __launch_bounds__(256) void k_clock_acc(Ctl *ctl, unsigned long long *stamps, uint32_t blocks,
                                                  const LevelSizes *sz) {
    __shared__ unsigned long long lo[4], hi[4];
    unsigned long long a = ~0ull, b = 0;
    for (uint32_t i = threadIdx.x; i < blocks; i += blockDim.x) {
        const unsigned long long s = stamps[2 * i], e = stamps[2 * i + 1];
        if (e >= s && s > ctl->clock[0]) {
            a = s < a ? s : a;
            b = e > b ? e : b;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long x = __shfl_xor(a, off, 64), y = __shfl_xor(b, off, 64);
        a = x < a ? x : a;
        b = y > b ? y : b;
    }
    if ((threadIdx.x & 63) == 0) {
        lo[threadIdx.x >> 6] = a;
        hi[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            a = lo[w] < a ? lo[w] : a;
            b = hi[w] > b ? hi[w] : b;
        }
        a = lo[0] < a ? lo[0] : a;
        b = hi[0] > b ? hi[0] : b;
        if (sz->n_work && !ctl->overflow && b > a) {
            ctl->clock_ticks += b - a;
            ctl->clock_launches += 1;
            ctl->clock[0] = b;  // later launches start after this end
        }
    }
}

// ---- device-sized scans ------------------------------------------------------------
// Exclusive scan of up to three uint32 arrays whose length is *n (device);
// out[i] for i <= n, out[n] = total, *total = total.  blockIdx.y = array.
struct DScan {
    const uint32_t *in[3];
    uint32_t *out[3];
    unsigned long long *total[3];
    const unsigned long long *n;
    unsigned long long *part;  // 3 x (kScanBlocks + 1)
};

__device__ __forceinline__ void dscan_range(uint64_t n, uint32_t b, uint64_t *lo, uint64_t *hi) {
    const uint64_t chunk = (n + kScanBlocks - 1) / kScanBlocks;
    *lo = std::min<uint64_t>(n, (uint64_t)b * chunk);
    *hi = std::min<uint64_t>(n, *lo + chunk);
}

__device__ __forceinline__ unsigned long long block_sum64(unsigned long long v, unsigned long long *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const unsigned long long t = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(256) void k_dscan_part(DScan s) {
    __shared__ unsigned long long red[4];
    const uint32_t a = blockIdx.y;
    if (!s.in[a]) return;
    uint64_t lo, hi;
    dscan_range(*s.n, blockIdx.x, &lo, &hi);
    unsigned long long v = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) v += s.in[a][i];
    v = block_sum64(v, red);
    if (threadIdx.x == 0) s.part[a * (kScanBlocks + 1) + blockIdx.x] = v;
}

__global__ __launch_bounds__(256) void k_dscan_top(DScan s) {
    __shared__ unsigned long long lds[4];
    const uint32_t a = blockIdx.x;
    if (!s.in[a]) return;
    unsigned long long *part = s.part + a * (kScanBlocks + 1);
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long v = part[tid];
    unsigned long long x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    unsigned long long base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += lds[w];
    const unsigned long long total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    part[tid] = base + x - v;
    if (tid == 0) {
        part[kScanBlocks] = total;
        if (s.total[a]) *s.total[a] = total;
        s.out[a][*s.n] = (uint32_t)total;
    }
}

__global__ __launch_bounds__(256) void k_dscan_down(DScan s) {
    __shared__ uint32_t lds[4];
    const uint32_t a = blockIdx.y;
    if (!s.in[a]) return;
    uint64_t lo, hi;
    dscan_range(*s.n, blockIdx.x, &lo, &hi);
    unsigned long long carry = s.part[a * (kScanBlocks + 1) + blockIdx.x];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint64_t i0 = lo; i0 < hi; i0 += 256) {
        const uint64_t i = i0 + tid;
        const uint32_t v = i < hi ? s.in[a][i] : 0;
        uint32_t x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane == 63) lds[wave] = x;
        __syncthreads();
        uint32_t base = 0;
        for (uint32_t w = 0; w < wave; ++w) base += lds[w];
        const uint32_t tot = lds[0] + lds[1] + lds[2] + lds[3];
        __syncthreads();
        if (i < hi) s.out[a][i] = (uint32_t)(carry + base + x - v);
        carry += tot;
    }
}

// ---- level 1 and level 2 setup -----------------------------------------------------

// Level 1: one task per database, every column of it.
__global__ void k_level1(uint32_t n_db, const uint32_t *ext_off, uint32_t *ext, uint32_t *sup,
                         uint32_t *db, uint32_t *pid) {
    for (uint32_t t = blockIdx.x; t < n_db; t += gridDim.x) {
        if (threadIdx.x == 0) {
            db[t] = t;
            pid[t] = kNone;
        }
        for (uint32_t e = ext_off[t] + threadIdx.x; e < ext_off[t + 1]; e += blockDim.x) {
            ext[e] = e - ext_off[t];
            sup[e] = 0;
        }
    }
}

// fcol[fbase_i + j] = column of database i's j-th frequent item
__global__ void k_fcol(uint32_t n_db, const uint32_t *ext_off, const uint32_t *kept, const uint32_t *ext,
                       const uint32_t *off_pat, const uint32_t *cnt_pat, uint32_t *fcol) {
    for (uint32_t t = blockIdx.x; t < n_db; t += gridDim.x)
        for (uint32_t j = threadIdx.x; j < cnt_pat[t]; j += blockDim.x)
            fcol[off_pat[t] + j] = ext[kept[ext_off[t] + j]];
}

// Level 2 of every database: task = (database i, frequent item a < nf_i - 1),
// prefix {fbase + a}, entries fbase + a + 1 .. fbase + nf - 1.
struct L2Db {
    uint32_t task0, n_tasks;  // first task, nf - 1
    uint32_t ext0, pad;       // first entry
};

__global__ void k_level2(const L2Db *l2, uint32_t n_db, uint32_t n_tasks, LevelPtrs L, const DbDev *dbs,
                         uint32_t G) {
    for (uint32_t t = blockIdx.x; t < n_tasks; t += gridDim.x) {
        uint32_t lo = 0, hi = n_db;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (l2[mid].task0 <= t) lo = mid; else hi = mid;
        }
        // a database without level-2 tasks shares task0 with the next one, so
        // the largest database with task0 <= t always owns task t
        const uint32_t i = lo;
        const DbDev &D = dbs[i];
        const uint32_t a = t - l2[i].task0;
        const uint32_t nf = D.nf;
        const uint32_t base = l2[i].ext0 + a * (nf - 1) - (uint32_t)(((uint64_t)a * (a - 1)) / 2);
        const uint32_t k = nf - 1 - a;
        if (threadIdx.x == 0) {
            L.prefix[(uint64_t)t * kMaxDepth] = D.fbase + a;
            L.pid[t] = D.fbase + a;
            L.db[t] = i;
            L.ext_off[t] = base;
            L.nwork[t] = task_work(k, G, D.tiles_eff);
        }
        for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
            L.ext[base + j] = D.fbase + a + 1 + j;
            L.sup[base + j] = 0;
        }
    }
}

// level-2 supports from the per-database Gram matrices
__global__ void k_gram_to_level2(LevelPtrs L, const DbDev *dbs, const int32_t *gram, const uint64_t *goff) {
    const uint32_t n = (uint32_t)L.sz->n_tasks;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) {
        const uint32_t i = L.db[t];
        const DbDev &D = dbs[i];
        const uint32_t a = L.prefix[(uint64_t)t * kMaxDepth] - D.fbase;
        const int32_t *g = gram + goff[i] + (uint64_t)a * D.nf;
        for (uint32_t e = L.ext_off[t] + threadIdx.x; e < L.ext_off[t + 1]; e += blockDim.x)
            L.sup[e] = (uint32_t)g[L.ext[e] - D.fbase];
    }
}

// pair tables from the level-2 outcome: bit (a, b) = pair kept
__global__ void k_pair_table(PlanArgs p, uint32_t *bits) {
    const uint32_t n = (uint32_t)p.L.sz->n_tasks;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) {
        const DbDev &D = p.dbs[p.L.db[t]];
        if (D.pair_w32 == 0) continue;
        const uint32_t a = p.L.prefix[(uint64_t)t * kMaxDepth] - D.fbase;
        uint32_t *row = bits + D.pair_off + (uint64_t)a * D.pair_w32;
        for (uint32_t e = p.L.ext_off[t] + threadIdx.x; e < p.L.ext_off[t + 1]; e += blockDim.x)
            if (keep_entry(p, D, e)) {
                const uint32_t b = p.L.ext[e] - D.fbase;
                atomicOr(row + (b >> 5), 1u << (b & 31));
            }
    }
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- host side ---------------------------------------------------------------------

struct LevelSet {
    DBuf prefix, pid, db, ext_off, work_off, nwork, ext, sup, per;
    DBuf kept, cnt[3], off[3], off2, ext2, live, pos, lvitems, parts;
    LevelPtrs ptrs(LevelSizes *sz) const {
        LevelPtrs p;
        p.prefix = prefix.as<uint32_t>();
        p.pid = pid.as<uint32_t>();
        p.db = db.as<uint32_t>();
        p.ext_off = ext_off.as<uint32_t>();
        p.work_off = work_off.as<uint32_t>();
        p.nwork = nwork.as<uint32_t>();
        p.ext = ext.as<uint32_t>();
        p.sup = sup.as<uint32_t>();
        p.per = per.as<int64_t>();
        p.sz = sz;
        return p;
    }
    int reserve(const Caps &c, bool periodic, uint32_t parts_tiles) {
        const uint64_t t1 = c.tasks + 1, e1 = c.ext + 1;
        CPM_CHECK(prefix.ensure(t1 * kMaxDepth * 4));
        CPM_CHECK(pid.ensure(t1 * 4));
        CPM_CHECK(db.ensure(t1 * 4));
        CPM_CHECK(ext_off.ensure(t1 * 4));
        CPM_CHECK(work_off.ensure(t1 * 4));
        CPM_CHECK(nwork.ensure(t1 * 4));
        CPM_CHECK(off2.ensure(t1 * 4));
        for (int i = 0; i < 3; ++i) {
            CPM_CHECK(cnt[i].ensure(t1 * 4));
            CPM_CHECK(off[i].ensure(t1 * 4));
        }
        CPM_CHECK(ext.ensure(e1 * 4));
        CPM_CHECK(ext2.ensure(e1 * 4));
        CPM_CHECK(sup.ensure(e1 * 4));
        CPM_CHECK(kept.ensure(e1 * 4));
        CPM_CHECK(live.ensure(e1 * 4));
        CPM_CHECK(pos.ensure(e1 * 4));
        CPM_CHECK(lvitems.ensure((c.lv + 1) * 4));
        if (periodic) {
            CPM_CHECK(per.ensure(e1 * 8));
            CPM_CHECK(parts.ensure((c.parts + 1) * sizeof(PeriodPart)));
        }
        (void)parts_tiles;
        return CPM_OK;
    }
};

struct GraphKey {
    const void *ptrs[64];
    uint64_t caps[5];
    uint32_t G, tile_words, parts_tiles;
    int periodic, prune;
    bool operator==(const GraphKey &o) const { return memcmp(this, &o, sizeof(GraphKey)) == 0; }
};

struct SearchWs {
    LevelSet set[2];
    DBuf ctl, dbs, fcol, pairbits, gram, goff, gjobs, gdbs, l2, scanpart, pat_parent, pat_item, pat_sup,
        pat_per, ts, stamps;
    HBuf host;
    Caps caps{};
    bool graphs_ok[2] = {false, false};
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    GraphKey key[2];
    ~SearchWs() {
        for (auto &g : graph)
            if (g) (void)hipGraphExecDestroy(g);
    }
};

void free_search_ws(void *p) { delete static_cast<SearchWs *>(p); }

SearchWs &search_ws(cpm_ctx *ctx) {
    if (!ctx->search_ws) {
        ctx->search_ws = new SearchWs();
        ctx->search_ws_free = free_search_ws;
        memset(&static_cast<SearchWs *>(ctx->search_ws)->key, 0, sizeof(GraphKey) * 2);
    }
    return *static_cast<SearchWs *>(ctx->search_ws);
}

// One database handed to a search.
struct DbIn {
    const cpm_bits *bits;
    uint64_t minsup;
    int64_t maxper, end;
    const int64_t *d_ts;  // device timestamps (periodic)
};

struct Search {
    cpm_ctx *ctx;
    SearchWs &w;
    std::vector<DbIn> in;
    bool periodic = false, use_gram = true, prune = true, use_graph = true;
    uint32_t G = 32, tile_words = kTileWords;
    std::vector<DbDev> dbs;          // host copy
    std::vector<uint32_t> nf;        // frequent items per database
    uint32_t parts_tiles = 0;

    Search(cpm_ctx *c, SearchWs &ws) : ctx(c), w(ws) {}

    Ctl *ctl() const { return w.ctl.as<Ctl>(); }
    LevelSizes *sz(int i) const { return &ctl()->sz[i]; }

    PlanArgs plan_args(int c, int emit_children) const {
        const LevelSet &L = w.set[c], &N = w.set[c ^ 1];
        PlanArgs p;
        memset(&p, 0, sizeof(p));
        p.L = L.ptrs(sz(c));
        p.N = N.ptrs(sz(c ^ 1));
        p.dbs = w.dbs.as<DbDev>();
        p.fcol = emit_children ? w.fcol.as<uint32_t>() : nullptr;
        p.pairbits = emit_children && w.pairbits.p ? w.pairbits.as<uint32_t>() : nullptr;
        p.kept = L.kept.as<uint32_t>();
        p.cnt_pat = L.cnt[0].as<uint32_t>();
        p.cnt_child = L.cnt[1].as<uint32_t>();
        p.cnt_cext = L.cnt[2].as<uint32_t>();
        p.off_pat = L.off[0].as<uint32_t>();
        p.off_child = L.off[1].as<uint32_t>();
        p.off_cext = L.off[2].as<uint32_t>();
        p.lvitems = L.lvitems.as<uint32_t>();
        p.off2 = N.off2.as<uint32_t>();
        p.ext2 = N.ext2.as<uint32_t>();
        p.live = N.live.as<uint32_t>();
        p.pos = N.pos.as<uint32_t>();
        p.pat_parent = w.pat_parent.as<uint32_t>();
        p.pat_item = w.pat_item.as<uint32_t>();
        p.pat_sup = w.pat_sup.as<uint32_t>();
        p.pat_per = periodic ? w.pat_per.as<int64_t>() : nullptr;
        p.ctl = ctl();
        p.caps = w.caps;
        p.G = G;
        p.emit_children = emit_children;
        p.periodic = periodic ? 1 : 0;
        p.prune = prune ? 1 : 0;
        p.parts_tiles = parts_tiles;
        return p;
    }

    SupportArgs support_args(int c, bool level1) const {
        const LevelSet &L = w.set[c];
        SupportArgs a;
        a.L = L.ptrs(sz(c));
        a.dbs = w.dbs.as<DbDev>();
        a.fcol = level1 ? nullptr : w.fcol.as<uint32_t>();
        a.G = G;
        a.stop = &ctl()->overflow;
        a.parts = L.parts.as<PeriodPart>();
        a.parts_tiles = parts_tiles;
        a.clock = w.stamps.as<unsigned long long>();
        return a;
    }

    uint32_t grid() const { return support_grid(ctx); }
    uint32_t grid_tasks() const { return (uint32_t)ctx->cu_count * 4; }

    int dscan(const uint32_t *in0, uint32_t *out0, unsigned long long *t0, const uint32_t *in1,
              uint32_t *out1, unsigned long long *t1, const uint32_t *in2, uint32_t *out2,
              unsigned long long *t2, const unsigned long long *n) {
        DScan s;
        s.in[0] = in0; s.out[0] = out0; s.total[0] = t0;
        s.in[1] = in1; s.out[1] = out1; s.total[1] = t1;
        s.in[2] = in2; s.out[2] = out2; s.total[2] = t2;
        s.n = n;
        s.part = w.scanpart.as<unsigned long long>();
        const uint32_t arrays = in2 ? 3 : (in1 ? 2 : 1);
        hipLaunchKernelGGL(k_dscan_part, dim3(kScanBlocks, arrays), dim3(256), 0, ctx->stream, s);
        hipLaunchKernelGGL(k_dscan_top, dim3(arrays), dim3(256), 0, ctx->stream, s);
        hipLaunchKernelGGL(k_dscan_down, dim3(kScanBlocks, arrays), dim3(256), 0, ctx->stream, s);
        CPM_HIP(hipGetLastError());
        return CPM_OK;
    }

    int count(int c, bool level1) {
        SupportArgs a = support_args(c, level1);
        if (periodic) CPM_CHECK(launch_support_periodic(ctx, a, grid()));
        else CPM_CHECK(launch_support_count(ctx, a, tile_words, grid()));
        hipLaunchKernelGGL(k_clock_acc, dim3(1), dim3(256), 0, ctx->stream, ctl(),
                           w.stamps.as<unsigned long long>(), grid(), sz(c));
        CPM_HIP(hipGetLastError());
        return CPM_OK;
    }

    // plan level c and produce + count level c^1 (one level step)
    int step(int c) {
        PlanArgs p = plan_args(c, 1);
        const LevelSizes *L = sz(c);
        LevelSizes *N = sz(c ^ 1);
        hipLaunchKernelGGL(k_plan_count, dim3(grid_tasks()), dim3(256), 0, ctx->stream, p);
        CPM_CHECK(dscan(p.cnt_pat, p.off_pat, &sz(c)->n_pat, p.cnt_child, p.off_child, &sz(c)->n_child,
                        p.cnt_cext, p.off_cext, &sz(c)->n_cext, &L->n_tasks));
        hipLaunchKernelGGL(k_alloc, dim3(1), dim3(1), 0, ctx->stream, p);
        hipLaunchKernelGGL(k_plan_write, dim3(grid_tasks()), dim3(256), 0, ctx->stream, p);
        if (prune) {
            hipLaunchKernelGGL(k_prune_flags, dim3(grid()), dim3(256), 0, ctx->stream, p);
            CPM_CHECK(dscan(p.live, p.pos, &ctl()->live_total, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr, &N->n_ext));
            hipLaunchKernelGGL(k_prune_offsets, dim3(grid()), dim3(256), 0, ctx->stream, p);
            hipLaunchKernelGGL(k_prune_compact, dim3(grid()), dim3(256), 0, ctx->stream, p);
            hipLaunchKernelGGL(k_prune_copyback, dim3(grid()), dim3(256), 0, ctx->stream, p);
        }
        CPM_CHECK(dscan(p.N.nwork, p.N.work_off, &N->n_work, nullptr, nullptr, nullptr, nullptr, nullptr,
                        nullptr, &N->n_tasks));
        hipLaunchKernelGGL(k_level_stats, dim3(grid_tasks()), dim3(256), 0, ctx->stream, p.N, p.dbs, ctl());
        CPM_HIP(hipGetLastError());
        return count(c ^ 1, false);
    }

    GraphKey graph_key(int c) const {
        GraphKey k;
        memset(&k, 0, sizeof(k));
        int i = 0;
        auto add = [&](const DBuf &b) { k.ptrs[i++] = b.p; };
        for (int j = 0; j < 2; ++j) {
            const LevelSet &S = w.set[(c + j) & 1];
            add(S.prefix); add(S.pid); add(S.db); add(S.ext_off); add(S.work_off); add(S.nwork);
            add(S.ext); add(S.sup); add(S.per); add(S.kept); add(S.off2); add(S.ext2); add(S.live);
            add(S.pos); add(S.lvitems); add(S.parts);
            for (int q = 0; q < 3; ++q) { add(S.cnt[q]); add(S.off[q]); }
        }
        add(w.ctl); add(w.dbs); add(w.fcol); add(w.pairbits); add(w.scanpart); add(w.stamps);
        add(w.pat_parent); add(w.pat_item); add(w.pat_sup); add(w.pat_per);
        k.caps[0] = w.caps.tasks;
        k.caps[1] = w.caps.ext;
        k.caps[2] = w.caps.patterns;
        k.caps[3] = w.caps.lv;
        k.caps[4] = w.caps.parts;
        k.G = G;
        k.tile_words = tile_words;
        k.parts_tiles = parts_tiles;
        k.periodic = periodic;
        k.prune = prune;
        return k;
    }

    int launch_step(int c) {
        if (!use_graph) return step(c);
        const GraphKey k = graph_key(c);
        if (!w.graphs_ok[c] || !(w.key[c] == k)) {
            if (w.graph[c]) {
                (void)hipGraphExecDestroy(w.graph[c]);
                w.graph[c] = nullptr;
            }
            hipGraph_t g = nullptr;
            CPM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int s = step(c);
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (s != CPM_OK) {
                if (g) (void)hipGraphDestroy(g);
                return s;
            }
            if (e != hipSuccess) return fail(CPM_ERR_RUNTIME, "graph capture failed: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&w.graph[c], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                w.graph[c] = nullptr;
                return fail(CPM_ERR_RUNTIME, "graph instantiation failed: %s", hipGetErrorString(e));
            }
            w.key[c] = k;
            w.graphs_ok[c] = true;
        }
        CPM_HIP(hipGraphLaunch(w.graph[c], ctx->stream));
        return CPM_OK;
    }

    int reserve() {
        for (int i = 0; i < 2; ++i) CPM_CHECK(w.set[i].reserve(w.caps, periodic, parts_tiles));
        CPM_CHECK(w.pat_parent.ensure((w.caps.patterns + 1) * 4));
        CPM_CHECK(w.pat_item.ensure((w.caps.patterns + 1) * 4));
        CPM_CHECK(w.pat_sup.ensure((w.caps.patterns + 1) * 4));
        if (periodic) CPM_CHECK(w.pat_per.ensure((w.caps.patterns + 1) * 8));
        CPM_CHECK(w.scanpart.ensure(3 * (kScanBlocks + 1) * 8));
        CPM_CHECK(w.ctl.ensure(sizeof(Ctl)));
        CPM_CHECK(w.stamps.ensure((size_t)grid() * 2 * 8));
        CPM_CHECK(w.host.ensure(std::max<size_t>(sizeof(Ctl), 4096)));
        return CPM_OK;
    }

    int read_ctl(Ctl *h) {
        CPM_HIP(hipMemcpyAsync(h, w.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream));
        CPM_HIP(hipStreamSynchronize(ctx->stream));
        return CPM_OK;
    }

    // One full search with the current capacities; *overflow receives the
    // capacity bits that were exceeded (0 when the search completed).
    int attempt(uint32_t *overflow) {
        *overflow = 0;
        const uint32_t n_db = (uint32_t)in.size();
        hipStream_t st = ctx->stream;
        CPM_CHECK(reserve());
        Ctl init;
        memset(&init, 0, sizeof(init));
        init.clock[0] = 0;
        Ctl *hc = w.host.as<Ctl>();
        *hc = init;
        CPM_HIP(hipMemcpyAsync(w.ctl.p, hc, sizeof(Ctl), hipMemcpyHostToDevice, st));
        CPM_HIP(hipMemsetAsync(w.stamps.p, 0, (size_t)grid() * 2 * 8, st));
        CPM_HIP(hipStreamSynchronize(st));

        // database table (frequent ranges filled after level 1)
        dbs.assign(n_db, DbDev{});
        uint64_t total_items = 0;
        for (uint32_t i = 0; i < n_db; ++i) {
            const cpm_bits *b = in[i].bits;
            DbDev &D = dbs[i];
            D.store = b->store.as<const uint64_t>();
            D.stride = b->stride;
            D.words = b->words;
            D.n = b->n_transactions;
            D.tiles = tiles_per_row(b->stride, periodic ? kTileWords : tile_words);
            D.tiles_eff = tiles_eff_of(D.tiles);
            D.minsup = in[i].minsup;
            D.maxper = in[i].maxper;
            D.end = in[i].end;
            D.ts = in[i].d_ts;
            total_items += b->n_items;
        }
        CPM_CHECK(w.dbs.ensure(std::max<uint32_t>(n_db, 1) * sizeof(DbDev)));
        CPM_HIP(hipMemcpyAsync(w.dbs.p, dbs.data(), n_db * sizeof(DbDev), hipMemcpyHostToDevice, st));

        // ---- level 1 (set 0) ----
        if (n_db > w.caps.tasks || total_items > w.caps.ext) {
            *overflow = (n_db > w.caps.tasks ? kOvTasks : 0) | (total_items > w.caps.ext ? kOvEntries : 0);
            return CPM_OK;
        }
        if (periodic && total_items * parts_tiles > w.caps.parts) {
            *overflow = kOvParts;
            return CPM_OK;
        }
        LevelSet &S0 = w.set[0];
        {
            std::vector<uint32_t> off(n_db + 1), work(n_db + 1);
            uint64_t e = 0, wk = 0;
            for (uint32_t i = 0; i < n_db; ++i) {
                off[i] = (uint32_t)e;
                work[i] = (uint32_t)wk;
                e += in[i].bits->n_items;
                wk += task_work(in[i].bits->n_items, G, dbs[i].tiles_eff);
            }
            off[n_db] = (uint32_t)e;
            work[n_db] = (uint32_t)wk;
            CPM_HIP(hipMemcpyAsync(S0.ext_off.p, off.data(), (n_db + 1) * 4, hipMemcpyHostToDevice, st));
            CPM_HIP(hipMemcpyAsync(S0.work_off.p, work.data(), (n_db + 1) * 4, hipMemcpyHostToDevice, st));
            LevelSizes l1;
            memset(&l1, 0, sizeof(l1));
            l1.n_tasks = n_db;
            l1.d = 0;
            l1.n_ext = e;
            l1.n_work = wk;
            CPM_HIP(hipMemcpyAsync(sz(0), &l1, sizeof(l1), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_level1, dim3(std::min<uint32_t>(n_db, 4096)), dim3(256), 0, st, n_db,
                               S0.ext_off.as<uint32_t>(), S0.ext.as<uint32_t>(), S0.sup.as<uint32_t>(),
                               S0.db.as<uint32_t>(), S0.pid.as<uint32_t>());
            CPM_HIP(hipGetLastError());
            // algorithmic bytes of level 1: every row once
            uint64_t bytes = 0;
            for (uint32_t i = 0; i < n_db; ++i) bytes += 8ull * dbs[i].words * in[i].bits->n_items;
            CPM_HIP(hipStreamSynchronize(st));
            ctx_bytes_level1 = bytes;
            CPM_CHECK(count(0, true));
        }
        // plan level 1: the frequent items (no children)
        {
            PlanArgs p = plan_args(0, 0);
            hipLaunchKernelGGL(k_plan_count, dim3(grid_tasks()), dim3(256), 0, st, p);
            CPM_CHECK(dscan(p.cnt_pat, p.off_pat, &sz(0)->n_pat, p.cnt_child, p.off_child, &sz(0)->n_child,
                            p.cnt_cext, p.off_cext, &sz(0)->n_cext, &sz(0)->n_tasks));
            hipLaunchKernelGGL(k_alloc, dim3(1), dim3(1), 0, st, p);
            hipLaunchKernelGGL(k_plan_write, dim3(grid_tasks()), dim3(256), 0, st, p);
            CPM_HIP(hipGetLastError());
        }
        nf.assign(n_db, 0);
        std::vector<uint32_t> fbase(n_db + 1);
        CPM_HIP(hipMemcpyAsync(nf.data(), S0.cnt[0].p, n_db * 4, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipMemcpyAsync(fbase.data(), S0.off[0].p, (n_db + 1) * 4, hipMemcpyDeviceToHost, st));
        CPM_CHECK(read_ctl(hc));
        if (hc->overflow) {
            *overflow = hc->overflow;
            return CPM_OK;
        }
        const uint64_t nf_total = fbase[n_db];
        if (nf_total == 0) return CPM_OK;
        CPM_CHECK(w.fcol.ensure(nf_total * 4));
        hipLaunchKernelGGL(k_fcol, dim3(std::min<uint32_t>(n_db, 4096)), dim3(256), 0, st, n_db,
                           S0.ext_off.as<uint32_t>(), S0.kept.as<uint32_t>(), S0.ext.as<uint32_t>(),
                           S0.off[0].as<uint32_t>(), S0.cnt[0].as<uint32_t>(), w.fcol.as<uint32_t>());
        CPM_HIP(hipGetLastError());

        // ---- level 2 (set 1) ----
        std::vector<L2Db> l2(n_db);
        uint64_t t2 = 0, e2 = 0, pair_words = 0;
        bool pairs = true;
        for (uint32_t i = 0; i < n_db; ++i) {
            DbDev &D = dbs[i];
            D.fbase = fbase[i];
            D.nf = nf[i];
            const uint32_t k = nf[i] >= 2 ? nf[i] - 1 : 0;
            l2[i].task0 = (uint32_t)t2;
            l2[i].n_tasks = k;
            l2[i].ext0 = (uint32_t)e2;
            t2 += k;
            e2 += (uint64_t)nf[i] * (nf[i] ? nf[i] - 1 : 0) / 2;
            if (nf[i] > 65536) pairs = false;
            D.pair_w32 = (nf[i] + 31) / 32;
            D.pair_off = pair_words;
            pair_words += (uint64_t)nf[i] * D.pair_w32;
        }
        if (!pairs) {
            for (auto &D : dbs) D.pair_w32 = 0;
            pair_words = 0;
        }
        CPM_HIP(hipMemcpyAsync(w.dbs.p, dbs.data(), n_db * sizeof(DbDev), hipMemcpyHostToDevice, st));
        if (t2 == 0) return CPM_OK;
        if (t2 > w.caps.tasks || e2 > w.caps.ext || e2 >= 0xffffffffull) {
            *overflow = (t2 > w.caps.tasks ? kOvTasks : 0) | (e2 > w.caps.ext ? kOvEntries : 0);
            if (e2 >= 0xffffffffull) return fail(CPM_ERR_NOMEM, "too many frequent item pairs");
            return CPM_OK;
        }
        if (periodic && e2 * parts_tiles > w.caps.parts) {
            *overflow = kOvParts;
            return CPM_OK;
        }
        LevelSet &S1 = w.set[1];
        CPM_CHECK(w.l2.ensure(n_db * sizeof(L2Db)));
        CPM_HIP(hipMemcpyAsync(w.l2.p, l2.data(), n_db * sizeof(L2Db), hipMemcpyHostToDevice, st));
        {
            LevelSizes ls;
            memset(&ls, 0, sizeof(ls));
            ls.n_tasks = t2;
            ls.d = 1;
            ls.n_ext = e2;
            CPM_HIP(hipMemcpyAsync(sz(1), &ls, sizeof(ls), hipMemcpyHostToDevice, st));
            const uint32_t e2u = (uint32_t)e2;
            CPM_HIP(hipMemcpyAsync(S1.ext_off.as<uint32_t>() + t2, &e2u, 4, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_level2, dim3((unsigned)std::min<uint64_t>(t2, 8192)), dim3(256), 0, st,
                               w.l2.as<L2Db>(), n_db, (uint32_t)t2, S1.ptrs(sz(1)), w.dbs.as<DbDev>(), G);
            CPM_HIP(hipGetLastError());
            CPM_CHECK(dscan(S1.nwork.as<uint32_t>(), S1.work_off.as<uint32_t>(), &sz(1)->n_work, nullptr,
                            nullptr, nullptr, nullptr, nullptr, nullptr, &sz(1)->n_tasks));
            CPM_HIP(hipStreamSynchronize(st));
        }
        bool gram_ok = use_gram && !periodic;
        for (uint32_t i = 0; i < n_db && gram_ok; ++i)
            if (nf[i] > 16384 || dbs[i].n >= (1ull << 31)) gram_ok = false;
        if (gram_ok) {
            std::vector<GramDb> gdb(n_db);
            std::vector<uint64_t> goff(n_db);
            uint64_t off = 0, ops = 0;
            for (uint32_t i = 0; i < n_db; ++i) {
                gdb[i].store = dbs[i].store;
                gdb[i].stride = dbs[i].stride;
                gdb[i].words = dbs[i].words;
                gdb[i].n_transactions = dbs[i].n;
                gdb[i].cols = w.fcol.as<uint32_t>() + fbase[i];
                gdb[i].n_items = nf[i] >= 2 ? nf[i] : 0;
                gdb[i].out_off = off;
                goff[i] = off;
                off += (uint64_t)gdb[i].n_items * gdb[i].n_items;
                ops += dbs[i].n * (uint64_t)gdb[i].n_items * (gdb[i].n_items + 1);
            }
            std::vector<GramJob> jobs;
            uint64_t elems = 0;
            CPM_CHECK(gram_plan(ctx, gdb, jobs, &elems));
            CPM_CHECK(w.gram.ensure(std::max<uint64_t>(elems, 1) * 4));
            CPM_CHECK(w.goff.ensure(n_db * 8));
            CPM_CHECK(w.gjobs.ensure(std::max<size_t>(jobs.size(), 1) * sizeof(GramJob)));
            CPM_CHECK(w.gdbs.ensure(n_db * sizeof(GramDb)));
            CPM_HIP(hipMemcpyAsync(w.goff.p, goff.data(), n_db * 8, hipMemcpyHostToDevice, st));
            if (!jobs.empty())
                CPM_HIP(hipMemcpyAsync(w.gjobs.p, jobs.data(), jobs.size() * sizeof(GramJob),
                                       hipMemcpyHostToDevice, st));
            CPM_HIP(hipMemcpyAsync(w.gdbs.p, gdb.data(), n_db * sizeof(GramDb), hipMemcpyHostToDevice, st));
            CPM_CHECK(launch_gram_i8(ctx, w.gjobs.as<GramJob>(), jobs.size(), w.gdbs.as<GramDb>(),
                                     w.gram.as<int32_t>(), elems, ops));
            hipLaunchKernelGGL(k_gram_to_level2, dim3((unsigned)std::min<uint64_t>(t2, 8192)), dim3(256), 0,
                               st, S1.ptrs(sz(1)), w.dbs.as<DbDev>(), w.gram.as<int32_t>(),
                               w.goff.as<uint64_t>());
            CPM_HIP(hipGetLastError());
        } else {
            hipLaunchKernelGGL(k_level_stats, dim3(grid_tasks()), dim3(256), 0, st, S1.ptrs(sz(1)),
                               w.dbs.as<DbDev>(), ctl());
            CPM_CHECK(count(1, false));
        }
        // pair tables
        if (pair_words) {
            CPM_CHECK(w.pairbits.ensure(pair_words * 4));
            hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<uint64_t>((pair_words + 255) / 256, 16384)),
                               dim3(256), 0, st, w.pairbits.as<uint32_t>(), pair_words, 0u);
            PlanArgs p = plan_args(1, 1);
            hipLaunchKernelGGL(k_pair_table, dim3((unsigned)std::min<uint64_t>(t2, 8192)), dim3(256), 0, st, p,
                               w.pairbits.as<uint32_t>());
            CPM_HIP(hipGetLastError());
        } else {
            w.pairbits.release();
        }

        // ---- levels >= 2: replayed steps ----
        int c = 1;
        for (;;) {
            for (int k = 0; k < 4; ++k) {
                CPM_CHECK(launch_step(c));
                c ^= 1;
            }
            CPM_CHECK(read_ctl(hc));
            if (hc->overflow) {
                *overflow = hc->overflow;
                return CPM_OK;
            }
            if (hc->sz[c].n_tasks == 0) break;
        }
        return CPM_OK;
    }

    uint64_t ctx_bytes_level1 = 0;

    int run(cpm_patterns **out) {
        // first capacities: generous for the first level set, grown on overflow
        uint64_t items = 0, e2 = 0;
        for (auto &d : in) {
            items += d.bits->n_items;
            e2 += (uint64_t)d.bits->n_items * (d.bits->n_items ? d.bits->n_items - 1 : 0) / 2;
        }
        parts_tiles = 0;
        if (periodic)
            for (auto &d : in) parts_tiles = std::max(parts_tiles, tiles_per_row(d.bits->stride));
        Caps want;
        want.tasks = std::max<uint64_t>(std::max<uint64_t>(items, in.size()), 1 << 16);
        want.ext = std::max<uint64_t>(std::max<uint64_t>(e2, items), 1 << 18);
        want.patterns = std::max<uint64_t>(items * 4, 1 << 18);
        want.lv = std::max<uint64_t>(want.ext * 2, 1 << 18);
        want.parts = periodic ? want.ext * std::max<uint32_t>(parts_tiles, 1) : 0;
        for (int i = 0; i < 5; ++i) {
            uint64_t *a = &w.caps.tasks + i;
            const uint64_t *b = &want.tasks + i;
            if (*a < *b) *a = *b;
        }
        for (int tries = 0;; ++tries) {
            uint32_t ov = 0;
            CPM_CHECK(attempt(&ov));
            if (!ov) break;
            if (tries > 24) return fail(CPM_ERR_NOMEM, "search buffers keep overflowing (%u)", ov);
            if (ov & kOvDepth) return fail(CPM_ERR_NOMEM, "an itemset longer than %u items", kMaxDepth + 1);
            if (ov & kOvTasks) w.caps.tasks *= 2;
            if (ov & kOvEntries) w.caps.ext *= 2;
            if (ov & kOvPatterns) w.caps.patterns *= 2;
            if (ov & kOvLevelSets) w.caps.lv *= 2;
            if (ov & kOvParts) w.caps.parts = std::max<uint64_t>(w.caps.parts * 2, 1 << 20);
        }
        return download(out);
    }

    int download(cpm_patterns **out) {
        Ctl *hc = w.host.as<Ctl>();
        CPM_CHECK(read_ctl(hc));
        const uint64_t n = hc->n_patterns;
        cpm_patterns *p = new cpm_patterns();
        p->n = n;
        p->periodic = periodic;
        p->parent.resize(n);
        p->item.resize(n);
        p->support.resize(n);
        p->db.resize(n);
        std::vector<uint32_t> s32(n);
        if (periodic) p->period.resize(n);
        hipStream_t st = ctx->stream;
        if (n) {
            hipError_t e = hipMemcpyAsync(p->parent.data(), w.pat_parent.p, n * 4, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipMemcpyAsync(p->item.data(), w.pat_item.p, n * 4, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipMemcpyAsync(s32.data(), w.pat_sup.p, n * 4, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess && periodic)
                e = hipMemcpyAsync(p->period.data(), w.pat_per.p, n * 8, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                delete p;
                return fail(CPM_ERR_RUNTIME, "pattern download failed: %s", hipGetErrorString(e));
            }
        }
        // database of each pattern: singletons by range, others by parent
        uint32_t di = 0;
        for (uint64_t i = 0; i < n; ++i) {
            p->support[i] = s32[i];
            if (p->parent[i] == kNone) {
                while (di + 1 < dbs.size() && i >= (uint64_t)dbs[di].fbase + dbs[di].nf) ++di;
                p->db[i] = di;
            } else {
                p->db[i] = p->db[p->parent[i]];
            }
        }
        // statistics
        int wall_khz = 0;
        (void)hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, ctx->device);
        const double tick_ms = wall_khz > 0 ? 1.0 / wall_khz : 0.0;
        ctx->stats[0] += (double)hc->clock_launches;
        ctx->stats[1] += (double)hc->clock_ticks * tick_ms;
        ctx->stats[2] += (double)hc->alg_bytes + (double)ctx_bytes_level1;
        ctx->stats[3] += (double)hc->candidates;
        ctx->stats[7] = (double)hc->levels + 2;
        *out = p;
        return CPM_OK;
    }
};

int check_bits(const cpm_bits *b, cpm_ctx *ctx) {
    if (!b) return fail(CPM_ERR_INVALID, "null store");
    if (b->ctx != ctx) return fail(CPM_ERR_INVALID, "store belongs to another context");
    if (b->n_transactions >= 0xffffffffull)
        return fail(CPM_ERR_INVALID, "supports are counted in 32 bits: too many transactions");
    return CPM_OK;
}

}  // namespace

}  // namespace cpm

using namespace cpm;

extern "C" {

int cpm_set_search_options(cpm_ctx *ctx, uint32_t group, uint32_t tile_words, int use_graph,
                           int subset_prune) {
    clear_error();
    if (!ctx) return fail(CPM_ERR_INVALID, "null context");
    if (group == 0 || group > 64) return fail(CPM_ERR_INVALID, "group must be 1..64");
    if (tile_words != 1024 && tile_words != 2048) return fail(CPM_ERR_INVALID, "tile_words must be 1024 or 2048");
    ctx->opt_group = group;
    ctx->opt_tile_words = tile_words;
    ctx->opt_graph = use_graph != 0;
    ctx->opt_prune = subset_prune != 0;
    return CPM_OK;
}

int cpm_eclat_mine_batch(cpm_ctx *ctx, const cpm_bits *const *bits, const uint64_t *minsup,
                         uint32_t n_db, int use_gram, cpm_patterns **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "null out");
    *out = nullptr;
    if (n_db && (!bits || !minsup)) return fail(CPM_ERR_INVALID, "null argument");
    for (uint32_t i = 0; i < n_db; ++i) CPM_CHECK(check_bits(bits[i], ctx));
    if (n_db == 0) {
        *out = new cpm_patterns();
        return CPM_OK;
    }
    Search s(ctx, search_ws(ctx));
    s.periodic = false;
    s.use_gram = use_gram != 0;
    s.G = ctx->opt_group;
    s.tile_words = ctx->opt_tile_words;
    s.use_graph = ctx->opt_graph;
    s.prune = ctx->opt_prune;
    for (uint32_t i = 0; i < n_db; ++i) s.in.push_back({bits[i], minsup[i], 0, 0, nullptr});
    return s.run(out);
}

int cpm_eclat_mine(cpm_ctx *ctx, const cpm_bits *bits, uint64_t minsup, int use_gram, cpm_patterns **out) {
    return cpm_eclat_mine_batch(ctx, &bits, &minsup, 1, use_gram, out);
}

int cpm_pfeclat_mine(cpm_ctx *ctx, const cpm_bits *bits, const int64_t *timestamps, int64_t end,
                     uint64_t minsup, int64_t maxper_floor, cpm_patterns **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "null out");
    *out = nullptr;
    CPM_CHECK(check_bits(bits, ctx));
    if (bits->n_transactions && !timestamps) return fail(CPM_ERR_INVALID, "null timestamps");
    SearchWs &w = search_ws(ctx);
    const uint64_t n = bits->n_transactions;
    CPM_CHECK(w.ts.ensure((n ? n : 1) * 8));
    if (n) CPM_HIP(hipMemcpyAsync(w.ts.p, timestamps, n * 8, hipMemcpyHostToDevice, ctx->stream));
    Search s(ctx, w);
    s.periodic = true;
    s.use_gram = false;
    s.G = ctx->opt_group;
    s.tile_words = kTileWords;
    s.use_graph = ctx->opt_graph;
    s.prune = ctx->opt_prune;
    s.in.push_back({bits, minsup, maxper_floor, end, w.ts.as<int64_t>()});
    return s.run(out);
}

// Complete-intersection supports (and periodicity) of explicit candidates:
// runs of candidates sharing their first width-1 columns become one task.
static int run_candidates(cpm_ctx *ctx, const cpm_bits *b, const uint32_t *cand, uint64_t n_cand,
                          uint32_t width, bool periodic, const int64_t *ts, int64_t end,
                          uint64_t *sup_out, int64_t *per_out) {
    if (width == 0) return fail(CPM_ERR_INVALID, "candidate width must be >= 1");
    if (width > kMaxDepth + 1) return fail(CPM_ERR_INVALID, "candidates wider than %u items", kMaxDepth + 1);
    if (n_cand >= 0xffffffffull) return fail(CPM_ERR_INVALID, "too many candidates in one call");
    for (uint64_t i = 0; i < n_cand * width; ++i)
        if (cand[i] >= b->n_items)
            return fail(CPM_ERR_INVALID, "candidate column %u >= %u items", cand[i], b->n_items);
    const uint32_t d = width - 1;
    const uint32_t G = 32;
    const uint32_t tw = kTileWords;
    DbDev D;
    memset(&D, 0, sizeof(D));
    D.store = b->store.as<const uint64_t>();
    D.stride = b->stride;
    D.words = b->words;
    D.n = b->n_transactions;
    D.tiles = tiles_per_row(b->stride, tw);
    D.tiles_eff = tiles_eff_of(D.tiles);
    D.end = end;
    std::vector<uint32_t> prefix, off, work, ext(n_cand);
    uint64_t wk = 0;
    for (uint64_t c = 0; c < n_cand; ++c) {
        const uint32_t *it = cand + c * width;
        const bool fresh = c == 0 || memcmp(it, cand + (c - 1) * width, (size_t)d * 4) != 0;
        if (fresh) {
            if (!off.empty()) wk += task_work((uint32_t)(c - off.back()), G, D.tiles_eff);
            off.push_back((uint32_t)c);
            work.push_back((uint32_t)wk);
            prefix.resize(prefix.size() + kMaxDepth, 0);
            memcpy(prefix.data() + prefix.size() - kMaxDepth, it, (size_t)d * 4);
        }
        ext[c] = it[d];
    }
    wk += task_work((uint32_t)(n_cand - off.back()), G, D.tiles_eff);
    const uint64_t nt = off.size();
    off.push_back((uint32_t)n_cand);
    work.push_back((uint32_t)wk);
    std::vector<uint32_t> dbz(nt, 0);
    DBuf dprefix, doff, dwork, dext, dsup, dper, ddb, dD, dsz, dparts, dts;
    hipStream_t st = ctx->stream;
    CPM_CHECK(dprefix.ensure(prefix.size() * 4));
    CPM_CHECK(doff.ensure(off.size() * 4));
    CPM_CHECK(dwork.ensure(work.size() * 4));
    CPM_CHECK(dext.ensure(n_cand * 4));
    CPM_CHECK(dsup.ensure(n_cand * 4));
    CPM_CHECK(ddb.ensure(nt * 4));
    CPM_CHECK(dD.ensure(sizeof(DbDev)));
    CPM_CHECK(dsz.ensure(sizeof(LevelSizes)));
    if (periodic) {
        CPM_CHECK(dper.ensure(n_cand * 8));
        CPM_CHECK(dparts.ensure(n_cand * D.tiles * sizeof(PeriodPart)));
        CPM_CHECK(dts.ensure(std::max<uint64_t>(D.n, 1) * 8));
        if (D.n) CPM_HIP(hipMemcpyAsync(dts.p, ts, D.n * 8, hipMemcpyHostToDevice, st));
        D.ts = dts.as<int64_t>();
    }
    LevelSizes ls;
    memset(&ls, 0, sizeof(ls));
    ls.n_tasks = nt;
    ls.d = d;
    ls.n_ext = n_cand;
    ls.n_work = wk;
    CPM_HIP(hipMemcpyAsync(dprefix.p, prefix.data(), prefix.size() * 4, hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(doff.p, off.data(), off.size() * 4, hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(dwork.p, work.data(), work.size() * 4, hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(dext.p, ext.data(), n_cand * 4, hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(ddb.p, dbz.data(), nt * 4, hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(dD.p, &D, sizeof(D), hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(dsz.p, &ls, sizeof(ls), hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemsetAsync(dsup.p, 0, n_cand * 4, st));
    SupportArgs a;
    a.L.prefix = dprefix.as<uint32_t>();
    a.L.db = ddb.as<uint32_t>();
    a.L.ext_off = doff.as<uint32_t>();
    a.L.work_off = dwork.as<uint32_t>();
    a.L.ext = dext.as<uint32_t>();
    a.L.sup = dsup.as<uint32_t>();
    a.L.per = dper.as<int64_t>();
    a.L.sz = dsz.as<LevelSizes>();
    a.dbs = dD.as<DbDev>();
    a.G = G;
    a.parts = dparts.as<PeriodPart>();
    a.parts_tiles = D.tiles;
    if (periodic) CPM_CHECK(launch_support_periodic(ctx, a, support_grid(ctx)));
    else CPM_CHECK(launch_support_count(ctx, a, tw, support_grid(ctx)));
    std::vector<uint32_t> s32(n_cand);
    CPM_HIP(hipMemcpyAsync(s32.data(), dsup.p, n_cand * 4, hipMemcpyDeviceToHost, st));
    if (periodic) CPM_HIP(hipMemcpyAsync(per_out, dper.p, n_cand * 8, hipMemcpyDeviceToHost, st));
    CPM_HIP(hipStreamSynchronize(st));
    for (uint64_t c = 0; c < n_cand; ++c) sup_out[c] = s32[c];
    return CPM_OK;
}

int cpm_count_supports(cpm_ctx *ctx, const cpm_bits *b, const uint32_t *cand, uint64_t n_cand,
                       uint32_t width, uint64_t *supports) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!b) return fail(CPM_ERR_INVALID, "null store");
    if (n_cand == 0) return CPM_OK;
    if (!cand || !supports) return fail(CPM_ERR_INVALID, "null candidates/supports");
    return run_candidates(ctx, b, cand, n_cand, width, false, nullptr, 0, supports, nullptr);
}

int cpm_periodic_eval(cpm_ctx *ctx, const cpm_bits *b, const int64_t *timestamps, int64_t end,
                      const uint32_t *cand, uint64_t n_cand, uint32_t width, uint64_t *supports,
                      int64_t *periods) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!b) return fail(CPM_ERR_INVALID, "null store");
    if (n_cand == 0) return CPM_OK;
    if (!cand || !supports || !periods) return fail(CPM_ERR_INVALID, "null argument");
    if (b->n_transactions && !timestamps) return fail(CPM_ERR_INVALID, "null timestamps");
    return run_candidates(ctx, b, cand, n_cand, width, true, timestamps, end, supports, periods);
}

int cpm_patterns_databases(const cpm_patterns *p, uint32_t *db) {
    if (!p) return fail(CPM_ERR_INVALID, "null patterns");
    if (p->n && db) {
        if (p->db.size() == p->n) memcpy(db, p->db.data(), p->n * 4);
        else memset(db, 0, p->n * 4);
    }
    return CPM_OK;
}

}  // extern "C"

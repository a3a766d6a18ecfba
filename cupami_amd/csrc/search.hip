// Device-resident frequent-itemset search over one or many databases.
//
// The search is level-synchronous over prefix tasks.  A task is a frequent
// prefix P (|P| = d) of one database with a list of extension items E; entry
// e of E is the candidate P + {e}.  After levels 1 and 2 are set up, one
// level step is a fixed sequence of six launches whose sizes all live in
// device memory:
//
//   plan/count    per task: compact the kept entries K; count patterns, child
//                 tasks (K_j, j < |K|-1) and child entries {K_l : l > j, pair
//                 (K_j, K_l) kept}; per-block sums over a contiguous chunk
//   plan/offsets  offsets of patterns / child tasks / child entries (a block
//                 adds the sums of the chunks before its own); the last block
//                 checks the capacities and sizes the next level
//   plan/write    write the patterns (parent id, item, support) and the next
//                 level (to staging when the subset test follows)
//   prune/flags   test each staged candidate's subsets (Bloom filter of this
//                 level's itemsets); per-block survivor counts
//   prune/compact move the survivors into place, child entry offsets
//   support       count the next level (support.hip)
//
// so a chunk of steps can be captured once as a hipGraph and replayed
// level after level with no host round trip, and no block ever waits for
// another.  The
// children of a task are cuEclat's child prefix classes (cuEclat.py:211-245);
// the join/prune is cuApriori's (cuApriori.py:473-492).  Pruning removes only
// candidates that cannot be frequent, so the itemsets are the reference's.
// When a level outgrows a capacity, the step raises an overflow flag, later
// steps do nothing, and the host reruns the search with larger buffers.
#include "support_kernels.h"
#include "sort.h"

#include <string.h>
#include <algorithm>
#include <mutex>

namespace cpm {

namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kMaskWords = 2048;   // LDS kept mask: up to 65536 frequent items per database
constexpr uint64_t kTaskMaskBudget = 256ull << 20;  // bytes of per-task kept masks
// Results whose columns are host blocks of at least this many bytes (pinned,
// recycled) are copied from the device straight into them; smaller ones go
// through a pinned stage, as plain host memory is slow to copy into.
constexpr uint64_t kDirectDownload = 1 << 20;

enum Overflow : uint32_t {
    kOvTasks = 1, kOvEntries = 2, kOvPatterns = 4, kOvDepth = 8, kOvParts = 32,
    kOvIndex = 64  // a level with 2^32 or more entries
};

// Control block in device memory.
struct Ctl {
    LevelSizes sz[2];
    SupportAcc acc;                  // support launch totals
    unsigned long long n_patterns;   // patterns written so far
    unsigned long long pat_base;     // first pattern id of the level being planned
    unsigned int overflow;           // Overflow bits
    unsigned int pad;
    unsigned long long levels;
};

struct Caps {
    uint64_t tasks, ext, patterns, parts;
};

struct PlanArgs {
    LevelPtrs L, N;
    const DbDev *dbs;
    const uint32_t *fcol;       // null at level 1 (entries are columns)
    const uint32_t *pairbits;   // pair tables, null for none
    uint32_t *kept, *cnt_pat, *cnt_child, *cnt_cext, *off_pat, *off_child, *off_cext;
    uint32_t *ccount;           // child entries of each kept entry (kept order; with a pair table)
    unsigned long long *filter; // Bloom filter of this level's itemsets for the subset test
    uint32_t *off2, *ext2, *live, *pos;  // next level's staging for the subset test
    unsigned long long *part;   // plan/count chunk sums: 3 x (plan grid)
    unsigned long long *partl;  // prune/flags chunk survivor counts: (prune grid)
    unsigned long long *stamps; // the last support launch's per-block stamps
    uint32_t stamp_blocks;
    uint32_t *tmask;            // kept-item mask of each task (tmask_w words apart), null:
    uint32_t tmask_w;           //   plan/write writes the child entries itself
    uint32_t *pat_parent, *pat_item, *pat_sup, *pat_db;
    int64_t *pat_per;
    Ctl *ctl;
    Caps caps;                  // capacities of the next level set
    uint32_t G, te;             // support work: largest entries per range, tiles per range
    uint64_t work_target;       // work items a short level is cut into
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
    const uint32_t x = dpp_wave_scan_incl(v);
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

// Pair-table rows are read 8 words per batch: the loads of a batch are in
// flight together (a task's largest children decide the plan launches' time,
// and their walks are latency-bound).
constexpr uint32_t kRowBatch = 8;

__device__ __forceinline__ uint32_t child_count(const uint32_t *lmask, const uint32_t *prow, uint32_t f,
                                                uint32_t wlast) {
    uint32_t w = f >> 5;
    uint32_t c = __popc(lmask[w] & prow[w] & above(f));
    for (++w; w + kRowBatch <= wlast + 1; w += kRowBatch) {
        uint32_t r[kRowBatch];
#pragma unroll
        for (uint32_t i = 0; i < kRowBatch; ++i) r[i] = prow[w + i];
#pragma unroll
        for (uint32_t i = 0; i < kRowBatch; ++i) c += __popc(lmask[w + i] & r[i]);
    }
    for (; w <= wlast; ++w) c += __popc(lmask[w] & prow[w]);
    return c;
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

// Contiguous chunk [lo, hi) of n items for block b of nb.
__device__ __forceinline__ void chunk_of(uint64_t n, uint32_t b, uint32_t nb, uint64_t *lo, uint64_t *hi) {
    const uint64_t c = (n + nb - 1) / nb;
    *lo = min(n, (uint64_t)b * c);
    *hi = min(n, *lo + c);
}

// Tasks with at most kSmallTask entries are planned by one wave each (no
// block barriers); larger ones by the whole block.
constexpr uint32_t kSmallTask = 64;

// orders a wave's LDS writes before its later LDS reads in other lanes
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// exclusive scan over the 64 lanes of a wave; *total gets the wave sum
__device__ __forceinline__ uint32_t wave_scan(uint32_t v, uint32_t *total) {
    const uint32_t x = dpp_wave_scan_incl(v);
    *total = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    return x - v;
}

// bit f of a pair-table row: the pair (row item, f) is frequent
__device__ __forceinline__ uint32_t pair_bit(const uint32_t *prow, uint32_t f) {
    return (prow[f >> 5] >> (f & 31)) & 1u;
}

// Child entry counts of a small task, one wave: wf[0 .. nk) are its kept
// items (local indexes, ascending); lane j < nk - 1 gets the number of kept
// items above wf[j] whose pair with wf[j] is kept.  The kept items are
// grouped into runs sharing a pair-table word (rw[r], bits rm[r]), so a lane
// reads one word of its pair row per run instead of one per item.  When trow
// is not null it receives the task's kept mask (pair_w32 words).  The whole
// wave calls it.
__device__ __forceinline__ uint32_t small_child_count(const uint32_t *wf, uint32_t *rw, uint32_t *rm, uint32_t nk,
                                                      const uint32_t *table, uint32_t pair_w32, uint32_t *trow) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t f = lane < nk ? wf[lane] : 0u;
    const uint32_t q = f >> 5;
    const bool head = lane < nk && (lane == 0 || (wf[lane - 1] >> 5) != q);
    const unsigned long long heads = __ballot(head);
    const uint32_t nrun = (uint32_t)__popcll(heads);
    const uint32_t run = (uint32_t)__popcll(heads & ((2ull << lane) - 1)) - 1u;  // lane < nk: its run
    if (head) {
        rw[run] = q;
        rm[run] = 0;
    }
    wave_sync();
    if (lane < nk) atomicOr(&rm[run], 1u << (f & 31));
    wave_sync();
    if (trow)
        for (uint32_t w = lane; w < pair_w32; w += 64) {
            uint32_t lo = 0, hi = nrun;  // the run of word w, if any
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (rw[mid] < w) lo = mid + 1; else hi = mid;
            }
            trow[w] = lo < nrun && rw[lo] == w ? rm[lo] : 0u;
        }
    uint32_t c = 0;
    if (lane + 1 < nk) {
        const uint32_t *prow = table + (uint64_t)f * pair_w32;
        for (uint32_t r = run; r < nrun; ++r) {
            uint32_t m = rm[r] & prow[rw[r]];
            if (r == run) m &= above(f);
            c += __popc(m);
        }
    }
    wave_sync();  // the runs are read before the wave's next task rewrites them
    return c;
}

// Tasks of at most kTinyTask entries are planned four to a wave: each row of
// 16 lanes (a DPP row) takes one task, so a wave walks a quarter of the
// dependent load chains it would walk one task at a time (the deep levels of
// a batch are hundreds of thousands of tasks of a few entries).
constexpr uint32_t kTinyTask = 16;

// small_child_count for the task of one row of 16 lanes: wf, rw, rm are the
// row's 16-entry slices, nk <= 16.  The row's lanes call it together (rows
// of a wave may take other branches).
__device__ __forceinline__ uint32_t tiny_child_count(const uint32_t *wf, uint32_t *rw, uint32_t *rm, uint32_t nk,
                                                     const uint32_t *table, uint32_t pair_w32, uint32_t *trow) {
    const uint32_t lane = threadIdx.x & 63, lg = lane & 15, sh = lane & 48;
    const uint32_t f = lg < nk ? wf[lg] : 0u;
    const uint32_t q = f >> 5;
    const bool head = lg < nk && (lg == 0 || (wf[lg - 1] >> 5) != q);
    const uint32_t heads = (uint32_t)(__ballot(head) >> sh) & 0xffffu;
    const uint32_t nrun = (uint32_t)__popc(heads);
    const uint32_t run = (uint32_t)__popc(heads & ((2u << lg) - 1)) - 1u;  // lg < nk: its run
    if (head) {
        rw[run] = q;
        rm[run] = 0;
    }
    wave_sync();
    if (lg < nk) atomicOr(&rm[run], 1u << (f & 31));
    wave_sync();
    if (trow)
        for (uint32_t w = lg; w < pair_w32; w += 16) {
            uint32_t lo = 0, hi = nrun;  // the run of word w, if any
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (rw[mid] < w) lo = mid + 1; else hi = mid;
            }
            trow[w] = lo < nrun && rw[lo] == w ? rm[lo] : 0u;
        }
    uint32_t c = 0;
    if (lg + 1 < nk) {
        const uint32_t *prow = table + (uint64_t)f * pair_w32;
        for (uint32_t r = run; r < nrun; ++r) {
            uint32_t m = rm[r] & prow[rw[r]];
            if (r == run) m &= above(f);
            c += __popc(m);
        }
    }
    wave_sync();  // the runs are read before the row's next task rewrites them
    return c;
}


// For two keys x: the first i <= n with off[i] >= x (off[0 .. n]
// non-decreasing; n + 1 when there is none), found by a whole 256-lane
// block in 256-ary rounds.  The answer lies in [lo, hi]; each lane tests one
// pivot lo + span t / 256 of [lo, hi) per key, and the pivots below the key
// are counted (ballots summed in LDS).  That leaves 1/256 of the range, and
// a range of at most 256 has every position tested: two rounds for 65,536
// children instead of 16 dependent loads of a binary search.  Every lane
// gets both answers.  cnt: 16 words of LDS.
__device__ void block_lower_bound2(const uint32_t *off, uint32_t n, uint64_t x0, uint64_t x1, uint32_t *cnt,
                                   uint32_t &r0, uint32_t &r1) {
    uint32_t lo[2] = {0, 0}, hi[2] = {n + 1, n + 1};
    const uint64_t key[2] = {x0, x1};
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t par = 0; hi[0] > lo[0] || hi[1] > lo[1]; par ^= 1) {  // uniform
        uint32_t *const c = cnt + par * 8;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint64_t span = hi[k] - lo[k];
            const bool below = span > 0 && off[lo[k] + (uint32_t)((span * threadIdx.x) >> 8)] < key[k];
            const unsigned long long b = __ballot(below);
            if (lane == 0) c[k * 4 + wave] = (uint32_t)__popcll(b);
        }
        __syncthreads();  // c is written again two rounds on, after the next barrier
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint64_t span = hi[k] - lo[k];
            if (span == 0) continue;
            // pivots t < m are below the key: the answer is past pivot m - 1
            // and at most pivot m
            const uint32_t m = c[k * 4] + c[k * 4 + 1] + c[k * 4 + 2] + c[k * 4 + 3];
            if (m == 0) {
                hi[k] = lo[k];
            } else {
                const uint32_t a = lo[k] + (uint32_t)((span * (m - 1)) >> 8) + 1;
                if (m < 256) hi[k] = lo[k] + (uint32_t)((span * m) >> 8);
                lo[k] = a;
            }
        }
    }
    r0 = lo[0];
    r1 = lo[1];
}

// Fold the per-block stamps a support launch left (start, end, bytes,
// candidates; start 0 = no launch since the last fold) into the totals and
// clear them.  One 256-lane block of a later launch runs it, so the support
// kernel itself needs no fence or shared counter.
__device__ void fold_stamps(unsigned long long *stamps, uint32_t blocks, SupportAcc *acc) {
    __shared__ unsigned long long red[4][4];
    const uint32_t tid = threadIdx.x;
    unsigned long long lo = ~0ull, hi = 0, by = 0, ca = 0;
    for (uint32_t b = tid; b < blocks; b += blockDim.x) {
        unsigned long long *s = stamps + 4ull * b;
        if (s[0]) {
            lo = min(lo, s[0]);
            hi = max(hi, s[1]);
            by += s[2];
            ca += s[3];
            s[0] = s[1] = s[2] = s[3] = 0;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (unsigned long long)__shfl_xor(lo, off, 64));
        hi = max(hi, (unsigned long long)__shfl_xor(hi, off, 64));
        by += __shfl_xor(by, off, 64);
        ca += __shfl_xor(ca, off, 64);
    }
    if ((tid & 63) == 0) {
        red[tid >> 6][0] = lo;
        red[tid >> 6][1] = hi;
        red[tid >> 6][2] = by;
        red[tid >> 6][3] = ca;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 0; w < 4; ++w) {
            lo = min(lo, red[w][0]);
            hi = max(hi, red[w][1]);
        }
        if (hi) {
            acc->ticks += hi > lo ? hi - lo : 0;
            acc->launches += 1;
            acc->bytes += red[0][2] + red[1][2] + red[2][2] + red[3][2];
            acc->candidates += red[0][3] + red[1][3] + red[2][3] + red[3][3];
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_fold_stamps(unsigned long long *stamps, uint32_t blocks, Ctl *ctl) {
    fold_stamps(stamps, blocks, &ctl->acc);
}

// plan/count -----------------------------------------------------------------------
// (its last block -- whose chunk is the shortest -- also folds the stamps of
// the support launch before it)
__global__ __launch_bounds__(256) void k_plan_count(PlanArgs p) {
    __shared__ uint32_t lds[8];
    __shared__ unsigned long long red[4];
    __shared__ uint32_t lmask[kMaskWords];
    __shared__ uint32_t wf[4][kSmallTask];  // per wave: the task's kept items (local indexes)
    __shared__ uint32_t rw[4][kSmallTask], rm[4][kSmallTask];  // per wave: their pair-table word runs
    if (p.stamps && blockIdx.x == gridDim.x - 1) fold_stamps(p.stamps, p.stamp_blocks, &p.ctl->acc);
    if (p.ctl->overflow) return;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t n_tasks = (uint32_t)p.L.sz->n_tasks;
    uint64_t lo, hi;
    chunk_of(n_tasks, blockIdx.x, gridDim.x, &lo, &hi);
    for (uint32_t w = tid; w < kMaskWords; w += blockDim.x) lmask[w] = 0;
    __syncthreads();
    unsigned long long s_pat = 0, s_child = 0, s_cext = 0;  // this lane's share of the block sums
    // tiny tasks, one row of 16 lanes each: the same steps as a small task
    // with the row's share of the ballot and a row scan
    {
        const uint32_t grp = lane >> 4, lg = lane & 15, sh = lane & 48;
        for (uint64_t tb = lo + 4ull * wave; tb < hi; tb += 16) {
            const uint64_t tt = tb + grp;
            const uint32_t t = (uint32_t)(tt < hi ? tt : lo);
            const uint32_t a = p.L.ext_off[t], b = p.L.ext_off[t + 1];
            const bool tiny = tt < hi && b - a <= kTinyTask;
            const DbDev &D = p.dbs[p.L.db[t]];
            const uint32_t e = a + lg;
            const bool k = tiny && e < b && keep_entry(p, D, e);
            const uint32_t bal = (uint32_t)(__ballot(k) >> sh) & 0xffffu;
            const uint32_t nk = (uint32_t)__popc(bal);
            uint32_t *const wfr = &wf[wave][sh], *const rwr = &rw[wave][sh], *const rmr = &rm[wave][sh];
            wave_sync();  // the previous tasks' reads of wf are done
            if (k) {
                const uint32_t pos = (uint32_t)__popc(bal & ((1u << lg) - 1));
                p.kept[a + pos] = e;
                wfr[pos] = p.L.ext[e] - D.fbase;
            }
            wave_sync();
            const uint32_t nc = (tiny && p.emit_children && nk >= 2) ? nk - 1 : 0;
            uint32_t c = 0;
            if (nc && p.pairbits) {
                uint32_t *trow = p.tmask ? p.tmask + (uint64_t)t * p.tmask_w : nullptr;
                c = tiny_child_count(wfr, rwr, rmr, nk, p.pairbits + D.pair_off, D.pair_w32, trow);
                if (lg < nc) p.ccount[a + lg] = c;
            } else if (lg < nc) {
                c = nk - 1 - lg;
            }
            const uint32_t cext = dpp_row_sum(c);
            if (tiny && lg == 0) {
                p.cnt_pat[t] = nk;
                p.cnt_child[t] = nc;
                p.cnt_cext[t] = cext;
                s_pat += nk;
                s_child += nc;
                s_cext += cext;
            }
        }
    }
    // small tasks, one wave each: keep flags by ballot, child counts from the
    // pair table bit by bit
    for (uint32_t t = (uint32_t)lo + wave; t < hi; t += 4) {
        const uint32_t a = p.L.ext_off[t], b = p.L.ext_off[t + 1];
        if (b - a > kSmallTask || b - a <= kTinyTask) continue;
        const DbDev &D = p.dbs[p.L.db[t]];
        const uint32_t e = a + lane;
        const bool k = e < b && keep_entry(p, D, e);
        const unsigned long long bal = __ballot(k);
        const uint32_t nk = (uint32_t)__popcll(bal);
        wave_sync();  // the previous task's reads of wf are done
        if (k) {
            const uint32_t pos = (uint32_t)__popcll(bal & ((1ull << lane) - 1));
            p.kept[a + pos] = e;
            wf[wave][pos] = p.L.ext[e] - D.fbase;
        }
        wave_sync();
        const uint32_t nc = (p.emit_children && nk >= 2) ? nk - 1 : 0;
        uint32_t c = 0;
        if (nc && p.pairbits) {
            uint32_t *trow = p.tmask ? p.tmask + (uint64_t)t * p.tmask_w : nullptr;
            c = small_child_count(wf[wave], rw[wave], rm[wave], nk, p.pairbits + D.pair_off, D.pair_w32, trow);
            if (lane < nc) p.ccount[a + lane] = c;
        } else if (lane < nc) {
            c = nk - 1 - lane;
        }
        uint32_t cext;
        (void)wave_scan(c, &cext);
        if (lane == 0) {
            p.cnt_pat[t] = nk;
            p.cnt_child[t] = nc;
            p.cnt_cext[t] = cext;
            s_pat += nk;
            s_child += nc;
            s_cext += cext;
        }
    }
    // large tasks: the whole block.  The chunk's large tasks are found 256 at
    // a time by all lanes (a scan by the block would wait on one load per task).
    __shared__ uint32_t big[256];
    __shared__ uint32_t n_big;
    for (uint64_t t0 = lo; t0 < hi; t0 += 256) {
    __syncthreads();  // every lane has read the previous n_big
    if (tid == 0) n_big = 0;
    __syncthreads();
    if (t0 + tid < hi && p.L.ext_off[t0 + tid + 1] - p.L.ext_off[t0 + tid] > kSmallTask)
        big[atomicAdd(&n_big, 1u)] = (uint32_t)(t0 + tid);
    __syncthreads();
    const uint32_t nb = n_big;
    // a large task's kept items (local indexes), when there are at most
    // kBigKept of them: the mask, child-count and clear passes then read them
    // here instead of through kept -> ext, two dependent loads each
    constexpr uint32_t kBigKept = 1024;
    __shared__ uint32_t kf[kBigKept];
    for (uint32_t bi = 0; bi < nb; ++bi) {
        const uint32_t t = big[bi];
        const uint32_t a = p.L.ext_off[t], b = p.L.ext_off[t + 1];
        const DbDev &D = p.dbs[p.L.db[t]];
        uint32_t nk = 0;
        for (uint32_t i0 = a; i0 < b; i0 += 256) {
            const uint32_t e = i0 + tid;
            const uint32_t x = e < b ? p.L.ext[e] : 0u;  // loaded beside the support
            const uint32_t f = (e < b && keep_entry(p, D, e)) ? 1u : 0u;
            uint32_t tot;
            const uint32_t pos = block_scan(f, lds, &tot);
            if (f) {
                p.kept[a + nk + pos] = e;
                if (nk + pos < kBigKept) kf[nk + pos] = x - D.fbase;
            }
            nk += tot;
        }
        __syncthreads();
        const bool cached = nk <= kBigKept;
        auto item = [&](uint32_t j) { return cached ? kf[j] : p.L.ext[p.kept[a + j]] - D.fbase; };
        uint32_t cext = 0;
        if (p.emit_children && nk >= 2) {
            uint32_t c = 0;
            if (p.pairbits) {
                for (uint32_t j = tid; j < nk; j += blockDim.x) {
                    const uint32_t f = item(j);
                    atomicOr(&lmask[f >> 5], 1u << (f & 31));
                }
                __syncthreads();
                if (p.tmask)
                    for (uint32_t q = tid; q < D.pair_w32; q += blockDim.x)
                        p.tmask[(uint64_t)t * p.tmask_w + q] = lmask[q];
                const uint32_t wlast = item(nk - 1) >> 5;
                const uint32_t *table = p.pairbits + D.pair_off;
                for (uint32_t j = tid; j + 1 < nk; j += 256) {
                    const uint32_t f = item(j);
                    const uint32_t cj = child_count(lmask, table + (uint64_t)f * D.pair_w32, f, wlast);
                    p.ccount[a + j] = cj;
                    c += cj;
                }
            } else {
                for (uint32_t j = tid; j + 1 < nk; j += 256) c += nk - 1 - j;
            }
            (void)block_scan(c, lds, &cext);
            if (p.pairbits)
                for (uint32_t j = tid; j < nk; j += blockDim.x) lmask[item(j) >> 5] = 0;
        }
        const uint32_t nc = (p.emit_children && nk >= 2) ? nk - 1 : 0;
        if (tid == 0) {
            p.cnt_pat[t] = nk;
            p.cnt_child[t] = nc;
            p.cnt_cext[t] = cext;
            s_pat += nk;
            s_child += nc;
            s_cext += cext;
        }
        __syncthreads();
    }
    }
    s_pat = block_sum64(s_pat, red);
    s_child = block_sum64(s_child, red);
    s_cext = block_sum64(s_cext, red);
    if (tid == 0) {
        p.part[blockIdx.x] = s_pat;
        p.part[gridDim.x + blockIdx.x] = s_child;
        p.part[2 * gridDim.x + blockIdx.x] = s_cext;
    }
}

// Size the next level from this level's totals (one lane).
__device__ void alloc_next(const PlanArgs &p) {
    Ctl *c = p.ctl;
    LevelSizes *L = p.L.sz, *N = p.N.sz;
    const uint64_t n_pat = L->n_pat;
    const uint64_t n_child = p.emit_children ? L->n_child : 0;
    const uint64_t n_cext = p.emit_children ? L->n_cext : 0;
    const uint64_t d1 = L->d + 1;
    uint32_t ov = 0;
    if (n_child > p.caps.tasks) ov |= kOvTasks;
    if (n_cext > p.caps.ext) ov |= kOvEntries;
    if (n_cext >= 0xffffffffull || n_child >= 0xffffffffull) ov |= kOvIndex;
    if (c->n_patterns + n_pat > p.caps.patterns) ov |= kOvPatterns;
    if (n_child && d1 > kMaxDepth) ov |= kOvDepth;
    if (p.periodic && n_cext * p.parts_tiles > p.caps.parts) ov |= kOvParts;
    if (ov) {
        c->overflow |= ov;
        N->n_tasks = N->n_ext = N->n_work = N->n_cand = 0;
        return;
    }
    c->pat_base = c->n_patterns;
    c->n_patterns += n_pat;
    N->n_tasks = n_child;
    N->d = d1;
    N->n_ext = n_cext;
    N->n_cand = n_cext;
    N->group = level_group(n_cext, p.G, p.te, p.work_target);
    N->n_work = n_child ? level_work(n_cext, (uint32_t)N->group, p.te) : 0;
    if (n_child) {
        p.N.ext_off[n_child] = (uint32_t)n_cext;
        p.off2[n_child] = (uint32_t)n_cext;
        c->levels += 1;
    }
}

// plan/offsets: same grid as plan/count -----------------------------------------------
// The three counts (patterns, children, child entries) are scanned together:
// their chunk-sum loads are in flight at once and their block scans share
// barriers, so the launch is two rounds of loads rather than six.
__global__ __launch_bounds__(256) void k_plan_offsets(PlanArgs p) {
    __shared__ uint32_t lds[3][4];
    __shared__ unsigned long long red[3][4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool last = blockIdx.x == gridDim.x - 1;
    if (p.ctl->overflow) {
        if (last && tid == 0) p.N.sz->n_tasks = p.N.sz->n_ext = p.N.sz->n_work = p.N.sz->n_cand = 0;
        return;
    }
    const uint32_t n = (uint32_t)p.L.sz->n_tasks;
    uint64_t lo, hi;
    chunk_of(n, blockIdx.x, gridDim.x, &lo, &hi);
    const uint32_t *cnt[3] = {p.cnt_pat, p.cnt_child, p.cnt_cext};
    uint32_t *off[3] = {p.off_pat, p.off_child, p.off_cext};
    // the sums of the chunks before this block's
    unsigned long long base[3] = {0, 0, 0};
    for (uint32_t j = tid; j < blockIdx.x; j += blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) base[a] += p.part[(uint64_t)a * gridDim.x + j];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        base[a] = dpp_wave_sum64(base[a]);
        if (lane == 0) red[a][wave] = base[a];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; ++a) base[a] = red[a][0] + red[a][1] + red[a][2] + red[a][3];
    for (uint64_t i0 = lo; i0 < hi; i0 += 256) {
        const uint64_t i = i0 + tid;
        uint32_t v[3], x[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = i < hi ? cnt[a][i] : 0u;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            x[a] = dpp_wave_scan_incl(v[a]);
            if (lane == 63) lds[a][wave] = x[a];
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t before = 0;
            for (uint32_t w = 0; w < wave; ++w) before += lds[a][w];
            if (i < hi) off[a][i] = (uint32_t)(base[a] + before + x[a] - v[a]);
            base[a] += lds[a][0] + lds[a][1] + lds[a][2] + lds[a][3];
        }
        __syncthreads();  // lds is rewritten by the next round
    }
    if (last && tid == 0) {  // base: the totals
        for (int a = 0; a < 3; ++a) off[a][n] = (uint32_t)base[a];
        p.L.sz->n_pat = base[0];
        p.L.sz->n_child = base[1];
        p.L.sz->n_cext = base[2];
        alloc_next(p);
    }
}

// subset test (cuApriori.py:487-490) ----------------------------------------------
//
// The itemsets of the level being planned go into a blocked Bloom filter:
// a 64-bit hash of the itemset picks one 64-bit word and four bits in it,
// set with one atomicOr (nothing waits for it).  The hash of a set is the
// sum of a mix of each item (then mixed once more), so a subset's hash is the
// whole set's minus one term: a candidate's d - 1 subset hashes cost O(d).  The filter has 32 bits per
// itemset (a power of two), which keeps false positives near 0.1%.  A child
// candidate survives when all its subsets test present.  A false positive can
// only keep a candidate that should have been pruned; its support, counted
// exactly, is then below minsup (it is at most the missing subset's), so
// results never change.

// finalizer of splitmix64 (Steele, Lea and Flood, 2014)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// one item's term of a set hash
__device__ __forceinline__ uint64_t item_term(uint32_t item) { return mix64(item + 0x9e3779b97f4a7c15ull); }
// filter hash of a set from the sum of its items' terms
__device__ __forceinline__ uint64_t set_hash(uint64_t sum) { return mix64(sum ^ 0x243f6a8885a308d3ull); }
// words of the filter for n_pat itemsets, less one: next_pow2(32 n_pat) bits
__device__ __forceinline__ uint64_t filter_mask(uint64_t n_pat) {
    const uint64_t bits = n_pat <= 2 ? 64 : 1ull << (64 - __clzll(32 * n_pat - 1));
    return bits / 64 - 1;
}
// word index from the low bits, four bit positions from the top 24
__device__ __forceinline__ unsigned long long filter_bits(uint64_t h) {
    return (1ull << ((h >> 40) & 63)) | (1ull << ((h >> 46) & 63)) | (1ull << ((h >> 52) & 63)) |
           (1ull << (h >> 58));
}
__device__ __forceinline__ void filter_insert(unsigned long long *f, uint64_t mask, uint64_t h) {
    atomicOr(f + (h & mask), filter_bits(h));
}
__device__ __forceinline__ bool prune_active(const PlanArgs &p) {
    return p.prune && !p.ctl->overflow && p.emit_children && p.L.sz->d >= 2 && p.N.sz->n_tasks > 0;
}

// plan/write -------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_plan_write(PlanArgs p) {
    __shared__ uint32_t lds[8];
    __shared__ uint32_t lmask[kMaskWords];
    __shared__ uint32_t wf[4][kSmallTask];  // per wave: the task's kept items
    __shared__ uint32_t spre[kMaxDepth];    // a large task's prefix
    if (p.ctl->overflow) return;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t n_tasks = (uint32_t)p.L.sz->n_tasks;
    const uint32_t d = (uint32_t)p.L.sz->d;
    const uint64_t pat0 = p.ctl->pat_base;
    const bool children = p.emit_children && p.N.sz->n_tasks > 0;
    // With the subset test, the children's entries go to staging and this
    // level's itemsets go into the filter; without it, straight into the
    // next level.
    const bool stage = prune_active(p);
    const uint64_t fmask = filter_mask(p.L.sz->n_pat);
    uint32_t *const xdst = stage ? p.ext2 : p.N.ext;
    uint32_t *const odst = stage ? p.off2 : p.N.ext_off;
    const bool fill = p.tmask && p.pairbits;  // k_child_fill writes the child entries
    for (uint32_t w = tid; w < kMaskWords; w += blockDim.x) lmask[w] = 0;
    __syncthreads();
    // tiny tasks, one row of 16 lanes each: lane j writes kept entry j and
    // its child; the prefix's d <= 64 items are held four per lane
    {
        const uint32_t grp = lane >> 4, lg = lane & 15, sh = lane & 48;
        for (uint64_t tb = 4ull * (blockIdx.x * 4 + wave); tb < n_tasks; tb += 16ull * gridDim.x) {
            const uint64_t tt = tb + grp;
            const uint32_t t = (uint32_t)(tt < n_tasks ? tt : tb);
            const uint32_t nk_t = p.cnt_pat[t];
            if (!(tt < n_tasks && nk_t <= kTinyTask)) continue;  // the row's lanes together
            const uint32_t nk = nk_t;
            const uint32_t dbi = p.L.db[t];
            const DbDev &D = p.dbs[dbi];
            const uint32_t a = p.L.ext_off[t];
            const uint64_t pb = pat0 + p.off_pat[t];
            uint32_t sv[4];  // prefix item lg + 16 m
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const uint32_t j = lg + 16 * m;
                sv[m] = j < d ? p.L.prefix[(uint64_t)t * kMaxDepth + j] : 0u;
            }
            unsigned long long psum = 0;  // sum of the prefix items' filter terms
            if (stage) {
                unsigned long long v = 0;
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    if (lg + 16 * m < d) v += item_term(sv[m]);
                psum = dpp_row_sum64(v);
            }
            uint32_t f = 0, e = 0;
            if (lg < nk) {
                e = p.kept[a + lg];
                f = p.L.ext[e];
                p.pat_parent[pb + lg] = p.L.pid[t];
                p.pat_item[pb + lg] = p.fcol ? p.fcol[f] : f;
                p.pat_sup[pb + lg] = p.L.sup[e];
                p.pat_db[pb + lg] = dbi;
                if (p.pat_per) p.pat_per[pb + lg] = p.L.per[e];
                if (stage) filter_insert(p.filter, fmask, set_hash(psum + item_term(f)));
            }
            if (!(children && nk >= 2)) continue;
            uint32_t *const wfr = &wf[wave][sh];
            if (!fill) {
                wave_sync();  // the previous task's reads of wf are done
                if (lg < nk) wfr[lg] = f - D.fbase;
                wave_sync();
            }
            const uint32_t *table = p.pairbits ? p.pairbits + D.pair_off : nullptr;
            const bool live = lg + 1 < nk;
            const uint32_t fl = f - D.fbase;
            const uint32_t *prow = table && live ? table + (uint64_t)fl * D.pair_w32 : nullptr;
            uint32_t cj = 0;
            if (live) cj = table ? p.ccount[a + lg] : nk - 1 - lg;
            const uint32_t off = dpp_row_scan_incl(cj) - cj;
            // child c's prefix row: this task's prefix and item f of the row's
            // lane c (d + 1 <= 64 when children are emitted)
            const uint32_t cb = p.off_child[t];
            for (uint32_t c = 0; c + 1 < nk; ++c) {
                const uint32_t fc = (uint32_t)__shfl((int)f, (int)(sh + c), 64);
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const uint32_t j = lg + 16 * m;
                    if (j <= d) p.N.prefix[(uint64_t)(cb + c) * kMaxDepth + j] = j < d ? sv[m] : fc;
                }
            }
            if (live) {
                const uint32_t ct = cb + lg;
                p.N.par[ct] = t;
                p.N.src[ct] = e;
                p.N.pid[ct] = (uint32_t)(pb + lg);
                p.N.db[ct] = dbi;
                uint32_t w = p.off_cext[t] + off;
                odst[ct] = w;
                if (!fill) {
                    for (uint32_t l = lg + 1; l < nk; ++l) {
                        const uint32_t g = wfr[l];
                        if (table && !pair_bit(prow, g)) continue;
                        xdst[w] = D.fbase + g;
                        if (stage) p.pos[w] = ct;
                        else p.N.sup[w] = 0;
                        ++w;
                    }
                }
            }
        }
    }
    // small tasks, one wave each: lane j writes kept entry j and its child
    for (uint32_t t = blockIdx.x * 4 + wave; t < n_tasks; t += gridDim.x * 4) {
        const uint32_t nk = p.cnt_pat[t];
        if (nk > kSmallTask || nk <= kTinyTask) continue;
        const uint32_t dbi = p.L.db[t];
        const DbDev &D = p.dbs[dbi];
        const uint32_t a = p.L.ext_off[t];
        const uint64_t pb = pat0 + p.off_pat[t];
        // the task's prefix, item i in lane i (d <= 64): one load for the
        // wave instead of a dependent load per item in every lane
        const uint32_t sv = lane < d ? p.L.prefix[(uint64_t)t * kMaxDepth + lane] : 0u;
        unsigned long long psum = 0;  // sum of the prefix items' filter terms
        if (stage) psum = dpp_wave_sum64(lane < d ? item_term(sv) : 0ull);
        uint32_t f = 0, e = 0;
        if (lane < nk) {
            e = p.kept[a + lane];
            f = p.L.ext[e];
            p.pat_parent[pb + lane] = p.L.pid[t];
            p.pat_item[pb + lane] = p.fcol ? p.fcol[f] : f;
            p.pat_sup[pb + lane] = p.L.sup[e];
            p.pat_db[pb + lane] = dbi;
            if (p.pat_per) p.pat_per[pb + lane] = p.L.per[e];
            if (stage) filter_insert(p.filter, fmask, set_hash(psum + item_term(f)));
        }
        if (!(children && nk >= 2)) continue;
        // plan/count left each kept entry's child count (and the kept mask
        // plan/fill reads); the entries are written here only without a mask
        if (!fill) {
            wave_sync();  // the previous task's reads of wf are done
            if (lane < nk) wf[wave][lane] = f - D.fbase;
            wave_sync();
        }
        const uint32_t *table = p.pairbits ? p.pairbits + D.pair_off : nullptr;
        const bool live = lane + 1 < nk;
        const uint32_t fl = f - D.fbase;
        const uint32_t *prow = table && live ? table + (uint64_t)fl * D.pair_w32 : nullptr;
        uint32_t cj = 0;
        if (live) cj = table ? p.ccount[a + lane] : nk - 1 - lane;
        uint32_t tot;
        const uint32_t off = wave_scan(cj, &tot);
        // child c's prefix row: this task's prefix and item f of lane c, one
        // coalesced store per child (d + 1 <= 64 when children are emitted)
        const uint32_t cb = p.off_child[t];
        for (uint32_t c = 0; c + 1 < nk; ++c) {
            const uint32_t fc = (uint32_t)__builtin_amdgcn_readlane((int)f, (int)c);
            if (lane <= d) p.N.prefix[(uint64_t)(cb + c) * kMaxDepth + lane] = lane < d ? sv : fc;
        }
        if (live) {
            const uint32_t ct = cb + lane;
            p.N.par[ct] = t;
            p.N.src[ct] = e;
            p.N.pid[ct] = (uint32_t)(pb + lane);
            p.N.db[ct] = dbi;
            uint32_t w = p.off_cext[t] + off;
            odst[ct] = w;
            if (!fill) {
                for (uint32_t l = lane + 1; l < nk; ++l) {
                    const uint32_t g = wf[wave][l];
                    if (table && !pair_bit(prow, g)) continue;
                    xdst[w] = D.fbase + g;
                    if (stage) p.pos[w] = ct;
                    else p.N.sup[w] = 0;
                    ++w;
                }
            }
        }
    }
    // large tasks: the whole block, found 256 at a time by all lanes
    __shared__ uint32_t big[256];
    __shared__ uint32_t n_big;
    for (uint64_t k0 = 0; (uint64_t)blockIdx.x + k0 * gridDim.x < n_tasks; k0 += 256) {
    __syncthreads();
    if (tid == 0) n_big = 0;
    __syncthreads();
    const uint64_t tt = (uint64_t)blockIdx.x + (k0 + tid) * gridDim.x;
    if (tt < n_tasks && p.cnt_pat[tt] > kSmallTask) big[atomicAdd(&n_big, 1u)] = (uint32_t)tt;
    __syncthreads();
    const uint32_t nb = n_big;
    for (uint32_t bi = 0; bi < nb; ++bi) {
        const uint32_t t = big[bi];
        const uint32_t nk = p.cnt_pat[t];
        const uint32_t dbi = p.L.db[t];
        const DbDev &D = p.dbs[dbi];
        const uint32_t a = p.L.ext_off[t];
        const uint64_t pb = pat0 + p.off_pat[t];
        const uint32_t parent = p.L.pid[t];
        __syncthreads();  // the previous task's reads of spre are done
        if (tid < d) spre[tid] = p.L.prefix[(uint64_t)t * kMaxDepth + tid];
        __syncthreads();
        unsigned long long psum = 0;  // sum of the prefix items' filter terms
        if (stage) psum = dpp_wave_sum64(lane < d ? item_term(spre[lane]) : 0ull);
        for (uint32_t j = tid; j < nk; j += 256) {
            const uint32_t e = p.kept[a + j];
            const uint32_t f = p.L.ext[e];
            p.pat_parent[pb + j] = parent;
            p.pat_item[pb + j] = p.fcol ? p.fcol[f] : f;
            p.pat_sup[pb + j] = p.L.sup[e];
            p.pat_db[pb + j] = dbi;
            if (p.pat_per) p.pat_per[pb + j] = p.L.per[e];
            if (stage) filter_insert(p.filter, fmask, set_hash(psum + item_term(f)));
        }
        if (!(children && nk >= 2)) continue;
        // the kept mask only for writing the entries here (plan/count wrote
        // the child counts and plan/fill's masks)
        const bool inline_mask = p.pairbits && !fill;
        if (inline_mask) {
            mask_set(lmask, p.kept, p.L.ext, a, nk, D.fbase);
            __syncthreads();
        }
        const uint32_t *table = p.pairbits ? p.pairbits + D.pair_off : nullptr;
        const uint32_t wlast = inline_mask ? (p.L.ext[p.kept[a + nk - 1]] - D.fbase) >> 5 : 0u;
        const uint32_t cb = p.off_child[t];
        const uint32_t xb = p.off_cext[t];
        uint32_t carry = 0;
        for (uint32_t j0 = 0; j0 + 1 < nk; j0 += 256) {
            const uint32_t j = j0 + tid;
            const bool live = j + 1 < nk;
            uint32_t cj = 0, fj = 0, ej = 0;
            const uint32_t *prow = nullptr;
            if (live) {
                ej = p.kept[a + j];
                fj = p.L.ext[ej];
                if (table) {
                    prow = table + (uint64_t)(fj - D.fbase) * D.pair_w32;
                    cj = p.ccount[a + j];
                } else {
                    cj = nk - 1 - j;
                }
            }
            uint32_t tot;
            const uint32_t off = block_scan(cj, lds, &tot);
            if (live) {
                const uint32_t ct = cb + j;
                uint32_t *dst = p.N.prefix + (uint64_t)ct * kMaxDepth;
                for (uint32_t i = 0; i < d; ++i) dst[i] = spre[i];
                dst[d] = fj;
                p.N.par[ct] = t;
                p.N.src[ct] = ej;
                p.N.pid[ct] = (uint32_t)(pb + j);
                p.N.db[ct] = dbi;
                uint32_t w = xb + carry + off;
                odst[ct] = w;
                if (fill) {
                    // entries: k_child_fill
                } else if (table) {
                    const uint32_t fl = fj - D.fbase;
                    for (uint32_t q0 = fl >> 5; q0 <= wlast; q0 += kRowBatch) {
                        uint32_t r[kRowBatch];
#pragma unroll
                        for (uint32_t i = 0; i < kRowBatch; ++i) r[i] = q0 + i <= wlast ? prow[q0 + i] : 0u;
#pragma unroll
                        for (uint32_t i = 0; i < kRowBatch; ++i) {
                            const uint32_t q = q0 + i;
                            uint32_t m = q <= wlast ? lmask[q] & r[i] : 0u;
                            if (q == (fl >> 5)) m &= above(fl);
                            while (m) {
                                xdst[w] = D.fbase + q * 32 + (uint32_t)__builtin_ctz(m);
                                if (stage) p.pos[w] = ct;
                                else p.N.sup[w] = 0;
                                ++w;
                                m &= m - 1;
                            }
                        }
                    }
                } else {
                    for (uint32_t l = j + 1; l < nk; ++l) {
                        xdst[w] = p.L.ext[p.kept[a + l]];
                        if (stage) p.pos[w] = ct;
                        else p.N.sup[w] = 0;
                        ++w;
                    }
                }
            }
            carry += tot;
        }
        if (inline_mask) {
            __syncthreads();
            mask_clear(lmask, p.kept, p.L.ext, a, nk, D.fbase);
            __syncthreads();
        }
    }
    }
}

// plan/fill: the child entries of plan/write's children, one wave per child
// across the whole grid (a task's children are not tied to one block: the
// largest tasks would otherwise decide the launch's time).  Child ct of task
// t with item f takes the kept items above f whose pair with f is kept: the
// set bits of tmask[t] AND f's pair-table row, in ascending order.
// Child entries: a group of S lanes per child (S = the widest kept-mask row,
// rounded up to 8, 16, 32 or 64 words), so that short rows do not leave most
// of a wave idle.  Lanes cover mask words; a scan within the group places the
// items.
template <uint32_t S>
__device__ __forceinline__ void child_fill_groups(const PlanArgs &p, uint32_t n_child, uint32_t n_tasks, uint32_t d,
                                                  bool stage) {
    uint32_t *const xdst = stage ? p.ext2 : p.N.ext;
    const uint32_t *const odst = stage ? p.off2 : p.N.ext_off;
    constexpr uint32_t G = 64 / S;  // children per wave
    const uint32_t lane = threadIdx.x & 63, sub = lane % S, g = lane / S;
    const uint32_t groups = gridDim.x * (blockDim.x >> 6) * G;
    for (uint32_t ct = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * G + g; ct < n_child; ct += groups) {
        const uint32_t t = min(p.N.par[ct], n_tasks - 1);  // written by plan/write
        const DbDev &D = p.dbs[p.L.db[t]];
        const uint32_t fl = p.N.prefix[(uint64_t)ct * kMaxDepth + d] - D.fbase;
        const uint32_t *mask = p.tmask + (uint64_t)t * p.tmask_w;
        const uint32_t *prow = p.pairbits + D.pair_off + (uint64_t)fl * D.pair_w32;
        uint32_t w = odst[ct];
        for (uint32_t q0 = fl >> 5; q0 < D.pair_w32; q0 += S) {
            const uint32_t q = q0 + sub;
            uint32_t m = 0;
            if (q < D.pair_w32) {
                m = mask[q] & prow[q];
                if (q == (fl >> 5)) m &= above(fl);
            }
            const uint32_t c = __popc(m);
            uint32_t x = c;
#pragma unroll
            for (uint32_t o = 1; o < S; o <<= 1) {
                const uint32_t y = __shfl_up(x, o, S);
                if (sub >= o) x += y;
            }
            uint32_t pos = w + x - c;
            while (m) {
                xdst[pos] = D.fbase + q * 32 + (uint32_t)__builtin_ctz(m);
                if (stage) p.pos[pos] = ct;  // the entry's child, for the subset test
                else p.N.sup[pos] = 0;
                ++pos;
                m &= m - 1;
            }
            w += __shfl(x, S - 1, S);
        }
    }
}

__global__ __launch_bounds__(256) void k_child_fill(PlanArgs p) {
    if (p.ctl->overflow || !p.tmask || !p.pairbits || !p.emit_children) return;
    const uint32_t n_child = (uint32_t)p.N.sz->n_tasks;
    if (n_child == 0) return;
    const uint32_t n_tasks = (uint32_t)p.L.sz->n_tasks;
    const uint32_t d = (uint32_t)p.L.sz->d;
    const bool stage = prune_active(p);
    if (p.tmask_w > 32) child_fill_groups<64>(p, n_child, n_tasks, d, stage);
    else if (p.tmask_w > 16) child_fill_groups<32>(p, n_child, n_tasks, d, stage);
    else if (p.tmask_w > 8) child_fill_groups<16>(p, n_child, n_tasks, d, stage);
    else child_fill_groups<8>(p, n_child, n_tasks, d, stage);
}

// Is (prefix without item `drop`) + {x} -- d1 items, ascending -- in lv?
__global__ __launch_bounds__(256) void k_prune_flags(PlanArgs p) {
    __shared__ unsigned long long red[4];
    if (!prune_active(p)) return;
    const uint32_t n_child = (uint32_t)p.N.sz->n_tasks;
    const uint64_t n_cand = p.N.sz->n_cand;
    const uint32_t d1 = (uint32_t)p.N.sz->d;  // prefix length of the children
    const uint64_t fmask = filter_mask(p.L.sz->n_pat);
    uint64_t lo, hi;
    chunk_of(n_cand, blockIdx.x, gridDim.x, &lo, &hi);
    unsigned long long c = 0;
    for (uint64_t w = lo + threadIdx.x; w < hi; w += blockDim.x) {
        // the entry's child, left in pos by whichever kernel wrote the entry
        // (clamped: an index out of range could only come from a missed write)
        const uint32_t ct = min(p.pos[w], n_child - 1);
        const uint32_t *pre = p.N.prefix + (uint64_t)ct * kMaxDepth;
        const uint32_t x = p.ext2[w];
        // the subsets without the last prefix item or without x are kept by
        // construction; test the ones without each earlier prefix item, four
        // probes in flight at a time
        uint64_t all = item_term(x);
        for (uint32_t i = 0; i < d1; ++i) all += item_term(pre[i]);
        bool ok = true;
        for (uint32_t d0 = 0; ok && d0 + 1 < d1; d0 += 4) {
            unsigned long long word[4], bits[4];
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                word[k] = bits[k] = 0;
                if (d0 + k + 1 < d1) {  // the same for every lane (one prefix length per level)
                    const uint64_t h = set_hash(all - item_term(pre[d0 + k]));
                    bits[k] = filter_bits(h);
                    word[k] = p.filter[h & fmask];
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) ok = ok && (word[k] & bits[k]) == bits[k];
        }
        p.live[w] = ok ? 1u : 0u;
        c += ok ? 1u : 0u;
    }
    c = block_sum64(c, red);
    if (threadIdx.x == 0) p.partl[blockIdx.x] = c;
}

// prune/compact: same grid as prune/flags ---------------------------------------------
__global__ __launch_bounds__(256) void k_prune_compact(PlanArgs p) {
    __shared__ uint32_t lds[8];
    __shared__ uint32_t cnt[16];
    __shared__ unsigned long long red[4];
    if (!prune_active(p)) return;
    const uint32_t n_child = (uint32_t)p.N.sz->n_tasks;
    const uint64_t n_cand = p.N.sz->n_cand;
    const bool last = blockIdx.x == gridDim.x - 1;
    uint64_t lo, hi;
    chunk_of(n_cand, blockIdx.x, gridDim.x, &lo, &hi);
    // children whose staged entries start in [lo, hi); the last block also
    // takes those starting at n_cand, and the end sentinel
    // (the loads of the survivor counts before this chunk are issued first,
    // so they are in flight during the search's rounds)
    unsigned long long before = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += blockDim.x) before += p.partl[j];
    uint32_t c_lo, c_hi;
    block_lower_bound2(p.off2, n_child, lo, hi, cnt, c_lo, c_hi);
    if (last) c_hi = n_child + 1;
    unsigned long long base = block_sum64(before, red);
    for (uint64_t i0 = lo; i0 < hi; i0 += 256) {
        const uint64_t w = i0 + threadIdx.x;
        const uint32_t f = w < hi ? p.live[w] : 0u;
        uint32_t tot;
        const uint32_t x = block_scan(f, lds, &tot);  // ends with a barrier
        if (w < hi) {
            const uint64_t q = base + x;
            p.pos[w] = (uint32_t)q;
            if (f) {
                p.N.ext[q] = p.ext2[w];
                p.N.sup[q] = 0;
            }
        }
        base += tot;
    }
    __syncthreads();  // pos of the chunk is visible to the whole block
    for (uint32_t ct = c_lo + threadIdx.x; ct < c_hi; ct += blockDim.x) {
        const uint32_t x = p.off2[ct];
        p.N.ext_off[ct] = x < hi ? p.pos[x] : (uint32_t)base;
    }
    if (last && threadIdx.x == 0) {
        p.N.sz->n_ext = base;
        p.N.sz->group = level_group(base, p.G, p.te, p.work_target);
        p.N.sz->n_work = level_work(base, (uint32_t)p.N.sz->group, p.te);
    }
    // k_prune_flags is done with the filter: clear the words this level used
    // for the next one
    const uint64_t fwords = filter_mask(p.L.sz->n_pat) + 1;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < fwords;
         i += (uint64_t)gridDim.x * blockDim.x)
        p.filter[i] = 0;
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
// Level-1 results for the host, placed right after the control block so
// that one read returns both: frequent items per database, then their
// exclusive offsets (n_db + 1 words).
constexpr size_t kCtlBytes = (sizeof(Ctl) + 255) / 256 * 256;

__global__ void k_level1_pack(uint32_t n_db, const uint32_t *cnt, const uint32_t *off, Ctl *ctl) {
    uint32_t *dst = reinterpret_cast<uint32_t *>(reinterpret_cast<unsigned char *>(ctl) + kCtlBytes);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n_db + 1; i += gridDim.x * blockDim.x)
        dst[i] = i < n_db ? cnt[i] : off[i - n_db];
}

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

__global__ void k_level2(const L2Db *l2, uint32_t n_db, uint32_t n_tasks, LevelPtrs L, const DbDev *dbs) {
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
        }
        for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
            L.ext[base + j] = D.fbase + a + 1 + j;
            L.sup[base + j] = 0;
        }
    }
}

// Periodic searches: level 2 holds only the pairs whose support (from the
// Gram) reaches minsup.  A pair below it cannot be periodic-frequent
// (keep_entry), so the periodic scan, which walks every occurrence of a pair,
// runs on the frequent pairs alone.  Task t (item a of its database) keeps
// the items b > a with G[a][b] >= minsup: counted, scanned into ext_off,
// then written.
__device__ __forceinline__ uint32_t l2_db_of(const L2Db *l2, uint32_t n_db, uint32_t t) {
    uint32_t lo = 0, hi = n_db;  // the largest database with task0 <= t owns t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (l2[mid].task0 <= t) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_level2_frequent_count(const L2Db *l2, uint32_t n_db, uint32_t n_tasks,
                                                              const DbDev *dbs, const int32_t *gram,
                                                              const uint64_t *goff, uint32_t *cnt) {
    __shared__ uint32_t lds[8];
    for (uint32_t t = blockIdx.x; t < n_tasks; t += gridDim.x) {
        const uint32_t i = l2_db_of(l2, n_db, t);
        const DbDev &D = dbs[i];
        const uint32_t a = t - l2[i].task0;
        const int32_t *g = gram + goff[i] + (uint64_t)a * D.nf;
        uint32_t c = 0;
        for (uint32_t b = a + 1 + threadIdx.x; b < D.nf; b += blockDim.x) c += (uint64_t)g[b] >= D.minsup;
        uint32_t tot;
        (void)block_scan(c, lds, &tot);
        if (threadIdx.x == 0) cnt[t] = tot;
    }
}

// One workgroup: ext_off = exclusive scan of cnt (ext_off[n_tasks] = total)
// and level 2's sizes from the total.
__global__ __launch_bounds__(1024) void k_level2_frequent_scan(const uint32_t *cnt, uint32_t n_tasks, uint32_t *ext_off,
                                                              LevelSizes *sz, uint32_t G, uint32_t te,
                                                              uint64_t work_target) {
    __shared__ uint32_t part[16];
    __shared__ uint32_t carry_s;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) carry_s = 0;
    __syncthreads();
    for (uint32_t i0 = 0; i0 < n_tasks; i0 += blockDim.x) {
        const uint32_t i = i0 + tid;
        const uint32_t v = i < n_tasks ? cnt[i] : 0;
        uint32_t x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (lane >= (uint32_t)d) x += y;
        }
        if (lane == 63) part[wave] = x;
        __syncthreads();
        uint32_t base = carry_s, all = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
            if (w < wave) base += part[w];
            all += part[w];
        }
        if (i < n_tasks) ext_off[i] = base + x - v;
        __syncthreads();
        if (tid == 0) carry_s += all;
        __syncthreads();
    }
    if (tid == 0) {
        const uint32_t total = carry_s;
        ext_off[n_tasks] = total;
        sz->n_ext = total;
        sz->n_cand = total;
        sz->group = level_group(total, G, te, work_target);
        sz->n_work = level_work(total, (uint32_t)sz->group, te);
    }
}

__global__ __launch_bounds__(256) void k_level2_frequent_write(const L2Db *l2, uint32_t n_db, uint32_t n_tasks,
                                                              LevelPtrs L, const DbDev *dbs, const int32_t *gram,
                                                              const uint64_t *goff) {
    __shared__ uint32_t lds[8];
    for (uint32_t t = blockIdx.x; t < n_tasks; t += gridDim.x) {
        const uint32_t i = l2_db_of(l2, n_db, t);
        const DbDev &D = dbs[i];
        const uint32_t a = t - l2[i].task0;
        const int32_t *g = gram + goff[i] + (uint64_t)a * D.nf;
        if (threadIdx.x == 0) {
            L.prefix[(uint64_t)t * kMaxDepth] = D.fbase + a;
            L.pid[t] = D.fbase + a;
            L.db[t] = i;
        }
        uint32_t base = L.ext_off[t];
        for (uint32_t b0 = a + 1; b0 < D.nf; b0 += blockDim.x) {
            const uint32_t b = b0 + threadIdx.x;
            const uint32_t f = b < D.nf && (uint64_t)g[b] >= D.minsup;
            uint32_t tot;
            const uint32_t x = block_scan(f, lds, &tot);
            if (f) {
                L.ext[base + x] = D.fbase + b;
                L.sup[base + x] = 0;
            }
            base += tot;
        }
    }
}

// level-2 supports from the per-database Gram matrices
// (with `pairbits`, also the pair tables of the frequent pairs: what
// k_pair_table would write from these supports, in the same pass)
__global__ void k_gram_to_level2(LevelPtrs L, const DbDev *dbs, const int32_t *gram, const uint64_t *goff,
                                 uint32_t *pairbits) {
    const uint32_t n = (uint32_t)L.sz->n_tasks;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) {
        const uint32_t i = L.db[t];
        const DbDev &D = dbs[i];
        const uint32_t a = L.prefix[(uint64_t)t * kMaxDepth] - D.fbase;
        const int32_t *g = gram + goff[i] + (uint64_t)a * D.nf;
        uint32_t *row = pairbits && D.pair_w32 ? pairbits + D.pair_off + (uint64_t)a * D.pair_w32 : nullptr;
        for (uint32_t e = L.ext_off[t] + threadIdx.x; e < L.ext_off[t + 1]; e += blockDim.x) {
            const uint32_t b = L.ext[e] - D.fbase;
            const uint32_t sup = (uint32_t)g[b];
            L.sup[e] = sup;
            if (row && sup >= D.minsup) atomicOr(row + (b >> 5), 1u << (b & 31));
        }
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

// Queued host -> device copies, applied by one launch: record r copies
// `bytes` (a multiple of 4) from the blob at `src` to `dst`, or zeroes them
// when src == kZeroFill.
constexpr uint64_t kZeroFill = ~0ull;
__global__ void k_widen_u32(const uint32_t *__restrict__ in, uint64_t n, unsigned long long *__restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

inline dim3 grid_of_n(uint64_t n) {
    return dim3((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 16384)));
}

struct PutRec {
    uint32_t *dst;
    uint64_t src;
    uint64_t bytes;
};

__global__ __launch_bounds__(256) void k_scatter(const PutRec *__restrict__ recs, uint32_t n,
                                                const uint32_t *__restrict__ blob) {
    for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {
        const PutRec rec = recs[r];
        const uint64_t words = rec.bytes / 4;
        if (rec.src == kZeroFill) {
            for (uint64_t i = threadIdx.x; i < words; i += blockDim.x) rec.dst[i] = 0;
        } else {
            const uint32_t *src = blob + rec.src / 4;
            for (uint64_t i = threadIdx.x; i < words; i += blockDim.x) rec.dst[i] = src[i];
        }
    }
}

// ---- host side ---------------------------------------------------------------------

struct LevelSet {
    DBuf prefix, pid, par, src, db, ext_off, ext, sup, per;
    DBuf kept, cnt[3], off[3], off2, ext2, live, pos, parts;
    DBuf rows;    // retained itemset rows of this set's entries
    DBuf ccount;  // child entries of each kept entry
    LevelPtrs ptrs(LevelSizes *sz) const {
        LevelPtrs p;
        p.prefix = prefix.as<uint32_t>();
        p.pid = pid.as<uint32_t>();
        p.par = par.as<uint32_t>();
        p.src = src.as<uint32_t>();
        p.db = db.as<uint32_t>();
        p.ext_off = ext_off.as<uint32_t>();
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
        CPM_CHECK(par.ensure(t1 * 4));
        CPM_CHECK(src.ensure(t1 * 4));
        CPM_CHECK(db.ensure(t1 * 4));
        CPM_CHECK(ext_off.ensure(t1 * 4));
        CPM_CHECK(off2.ensure(t1 * 4));
        for (int i = 0; i < 3; ++i) {
            CPM_CHECK(cnt[i].ensure(t1 * 4));
            CPM_CHECK(off[i].ensure(t1 * 4));
        }
        CPM_CHECK(ext.ensure(e1 * 4));
        CPM_CHECK(ext2.ensure(e1 * 4));
        CPM_CHECK(sup.ensure(e1 * 4));
        CPM_CHECK(kept.ensure(e1 * 4));
        CPM_CHECK(ccount.ensure(e1 * 4));
        CPM_CHECK(live.ensure(e1 * 4));
        CPM_CHECK(pos.ensure(e1 * 4));
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
    uint64_t caps[4], slots;
    uint32_t G, te, tmask_w, tile_words, parts_tiles;
    int periodic, prune, ranges;
    bool operator==(const GraphKey &o) const { return memcmp(this, &o, sizeof(GraphKey)) == 0; }
};

struct SearchWs {
    LevelSet set[2];
    DBuf ctl, dbs, fcol, pairbits, tmask, gram, goff, gjobs, gdbs, l2, part, partl, scatter, ts, stamps;
    DBuf gblocks, gparts;  // the pair matrix's sliced blocks and their partial slots
    DBuf pat;             // pattern forest: parent, item, support and database columns, pat_pitch() bytes apart
    DBuf filter;          // subset-test Bloom filter (filter_words() words)
    DBuf ranges;          // per work range of the level being counted: first task, most tiles
    bool filter_dirty = true;  // holds bits of an attempt that stopped early
    DBuf pat_per;
    HBuf host;
    HBuf stage;           // pinned staging of small host <-> device copies
    size_t stage_off = 0; // in use since the last stream sync
    DBuf sup64;           // supports widened to u64 for the download of large results
    HBuf down;            // pinned stage of the download of small results
    hipStream_t copy = nullptr;       // streams large results to the host (default priority)
    hipEvent_t copy_ev = nullptr;     // a chunk's patterns are final and widened
    Caps caps{};
    bool graphs_ok[2] = {false, false};
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    GraphKey key[2];
    ~SearchWs() {
        for (auto &g : graph)
            if (g) (void)hipGraphExecDestroy(g);
        if (copy) {
            (void)hipStreamSynchronize(copy);
            (void)hipStreamDestroy(copy);
        }
        if (copy_ev) (void)hipEventDestroy(copy_ev);
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
    uint32_t te = 1;                 // tiles per support work range
    uint32_t tmask_w = 0;            // words per task kept mask (0: no plan/fill)
    bool ctl_current = false;        // the host copy of Ctl is final (read after the last step)
    std::vector<DbDev> dbs;          // host copy
    std::vector<uint32_t> nf;        // frequent items per database
    uint32_t parts_tiles = 0;
    Caps need{};                     // level 2's exact sizes when they overflowed
    uint64_t slots = 0;              // retained rows per level set (0: none)
    // Work items on tiles some databases do not reach exist (a batch of
    // databases with different row lengths): count through a range table.
    bool use_ranges = false;
    // Large results are streamed: once a search has found enough patterns
    // for the direct download, each finished chunk's patterns are copied into
    // the result's pinned arrays (sized to the pattern capacity) on a second
    // stream while the next chunk runs; download() copies only the rest.
    cpm_patterns *res = nullptr;
    uint64_t copied = 0;             // patterns queued for copying into res

    Search(cpm_ctx *c, SearchWs &ws) : ctx(c), w(ws) {}
    Search(const Search &) = delete;
    Search &operator=(const Search &) = delete;
    ~Search() { drop_stream(); }

    // the streamed result of an attempt that will not be used (copies waited for)
    void drop_stream() {
        if (res) {
            (void)hipStreamSynchronize(w.copy);
            delete res;
            res = nullptr;
        }
        copied = 0;
    }

    // queue the copy of patterns [copied, n_now): the support column widened
    // on the search stream, then every column on the copy stream
    int queue_copy(uint64_t n_now) {
        if (n_now <= copied) return CPM_OK;
        if (!res) {
            if (n_now * 4 < kDirectDownload || n_now > w.caps.patterns) return CPM_OK;
            if (!w.copy) CPM_HIP(hipStreamCreateWithFlags(&w.copy, hipStreamNonBlocking));
            if (!w.copy_ev) CPM_HIP(hipEventCreateWithFlags(&w.copy_ev, hipEventDisableTiming));
            CPM_CHECK(w.sup64.ensure(w.caps.patterns * 8));
            const uint64_t cap = w.caps.patterns;
            res = new cpm_patterns();
            res->periodic = periodic;
            try {
                res->parent.resize(cap);
                res->item.resize(cap);
                res->support.resize(cap);
                res->db.resize(cap);
                if (periodic) res->period.resize(cap);
            } catch (const std::bad_alloc &) {
                delete res;
                res = nullptr;
                return fail(CPM_ERR_NOMEM, "host memory for %llu patterns", (unsigned long long)cap);
            }
        }
        const uint64_t a = copied, m = n_now - a;
        hipStream_t cs = w.copy;
        hipLaunchKernelGGL(k_widen_u32, grid_of_n(m), dim3(256), 0, ctx->stream, pat_col(2) + a, m,
                           w.sup64.as<unsigned long long>() + a);
        CPM_HIP(hipGetLastError());
        CPM_HIP(hipEventRecord(w.copy_ev, ctx->stream));
        CPM_HIP(hipStreamWaitEvent(cs, w.copy_ev, 0));
        CPM_HIP(hipMemcpyAsync(res->parent.data() + a, pat_col(0) + a, m * 4, hipMemcpyDeviceToHost, cs));
        CPM_HIP(hipMemcpyAsync(res->item.data() + a, pat_col(1) + a, m * 4, hipMemcpyDeviceToHost, cs));
        CPM_HIP(hipMemcpyAsync(res->db.data() + a, pat_col(3) + a, m * 4, hipMemcpyDeviceToHost, cs));
        CPM_HIP(hipMemcpyAsync(res->support.data() + a, w.sup64.as<unsigned long long>() + a, m * 8,
                               hipMemcpyDeviceToHost, cs));
        if (periodic)
            CPM_HIP(hipMemcpyAsync(res->period.data() + a, w.pat_per.as<int64_t>() + a, m * 8,
                                   hipMemcpyDeviceToHost, cs));
        copied = n_now;
        return CPM_OK;
    }

    Ctl *ctl() const { return w.ctl.as<Ctl>(); }
    // words of the subset-test filter: 32 bits per itemset of the pattern
    // capacity, rounded up to a power of two (a level holds at most that many)
    uint64_t filter_words() const {
        uint64_t bits = 64;
        while (bits < 32 * w.caps.patterns) bits *= 2;
        return bits / 64;
    }
    size_t pat_pitch() const { return ((w.caps.patterns + 1) * 4 + 255) / 256 * 256; }
    uint32_t *pat_col(int k) const {
        return reinterpret_cast<uint32_t *>(w.pat.as<unsigned char>() + (size_t)k * pat_pitch());
    }
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
        p.ccount = L.ccount.as<uint32_t>();
        p.cnt_pat = L.cnt[0].as<uint32_t>();
        p.cnt_child = L.cnt[1].as<uint32_t>();
        p.cnt_cext = L.cnt[2].as<uint32_t>();
        p.off_pat = L.off[0].as<uint32_t>();
        p.off_child = L.off[1].as<uint32_t>();
        p.off_cext = L.off[2].as<uint32_t>();
        p.filter = w.filter.as<unsigned long long>();
        p.off2 = N.off2.as<uint32_t>();
        p.ext2 = N.ext2.as<uint32_t>();
        p.live = N.live.as<uint32_t>();
        p.pos = N.pos.as<uint32_t>();
        p.part = w.part.as<unsigned long long>();
        p.partl = w.partl.as<unsigned long long>();
        p.stamps = w.stamps.as<unsigned long long>();
        p.stamp_blocks = grid();
        p.tmask = emit_children && tmask_w ? w.tmask.as<uint32_t>() : nullptr;
        p.tmask_w = tmask_w;
        p.pat_parent = pat_col(0);
        p.pat_item = pat_col(1);
        p.pat_sup = pat_col(2);
        p.pat_db = pat_col(3);
        p.pat_per = periodic ? w.pat_per.as<int64_t>() : nullptr;
        p.ctl = ctl();
        p.caps = w.caps;
        p.G = G;
        p.te = te;
        p.work_target = work_target();
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
        a.stop = &ctl()->overflow;
        a.parts = L.parts.as<PeriodPart>();
        a.parts_tiles = parts_tiles;
        a.te = te;
        a.stamps = w.stamps.as<unsigned long long>();
        a.siblings = 1;
        if (slots) {
            a.rin = w.set[c ^ 1].rows.as<uint64_t>();
            a.rout = L.rows.as<uint64_t>();
            a.slots = slots;
        }
        if (use_ranges) {
            a.ranges = w.ranges.as<uint32_t>();
            a.rcap = w.caps.ext + 1;   // a level has at most caps.ext entries, so as many ranges
        }
        return a;
    }

    uint32_t grid() const { return support_grid(ctx); }
    // work items a short level is cut into: three per block (four and six
    // gave the 1M x 1k search 2-3% more support time, two 1%)
    uint64_t work_target() const { return 3ull * grid(); }
    uint32_t grid_tasks() const { return (uint32_t)ctx->cu_count * 6; }

    int count(int c, bool level1) {
        SupportArgs a = support_args(c, level1);
        CPM_CHECK(launch_ranges(ctx, a, grid()));
        if (periodic) return launch_support_periodic(ctx, a, grid());
        return launch_support_count(ctx, a, tile_words, grid());
    }

    // plan: count, offsets (+ sizing of the next level), write
    void plan(const PlanArgs &p) {
        hipLaunchKernelGGL(k_plan_count, dim3(grid_tasks()), dim3(256), 0, ctx->stream, p);
        hipLaunchKernelGGL(k_plan_offsets, dim3(grid_tasks()), dim3(256), 0, ctx->stream, p);
        hipLaunchKernelGGL(k_plan_write, dim3(grid_tasks()), dim3(256), 0, ctx->stream, p);
        if (p.tmask) hipLaunchKernelGGL(k_child_fill, dim3(grid()), dim3(256), 0, ctx->stream, p);
    }

    // plan level c and produce + count level c^1 (one level step)
    int step(int c) {
        PlanArgs p = plan_args(c, 1);
        plan(p);
        if (prune) {
            hipLaunchKernelGGL(k_prune_flags, dim3(grid()), dim3(256), 0, ctx->stream, p);
            hipLaunchKernelGGL(k_prune_compact, dim3(grid()), dim3(256), 0, ctx->stream, p);
        }
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
            add(S.prefix); add(S.pid); add(S.par); add(S.src); add(S.rows); add(S.ccount); add(S.db); add(S.ext_off); add(S.ext); add(S.sup); add(S.per); add(S.kept); add(S.off2); add(S.ext2); add(S.live);
            add(S.pos); add(S.parts);
            for (int q = 0; q < 3; ++q) { add(S.cnt[q]); add(S.off[q]); }
        }
        add(w.ctl); add(w.dbs); add(w.fcol); add(w.pairbits); add(w.tmask); add(w.part); add(w.partl);
        add(w.stamps);
        add(w.pat); add(w.pat_per); add(w.filter); add(w.ranges);
        k.caps[0] = w.caps.tasks;
        k.caps[1] = w.caps.ext;
        k.caps[2] = w.caps.patterns;
        k.caps[3] = w.caps.parts;
        k.slots = slots;
        k.G = G;
        k.te = te;
        k.tmask_w = tmask_w;
        k.tile_words = tile_words;
        k.parts_tiles = parts_tiles;
        k.periodic = periodic;
        k.prune = prune;
        k.ranges = use_ranges;
        return k;
    }

    // kChunkSteps levels, starting with set c, as one graph launch; an even
    // count leaves the next level in set c again
    static constexpr int kChunkSteps = 4;
    static_assert(kChunkSteps % 2 == 0, "a chunk ends on the set it starts from");
    int chunk(int c) {
        for (int k = 0; k < kChunkSteps; ++k) CPM_CHECK(step(c ^ (k & 1)));
        return CPM_OK;
    }
    int launch_chunk(int c) {
        if (!use_graph) return chunk(c);
        const GraphKey k = graph_key(c);
        if (!w.graphs_ok[c] || !(w.key[c] == k)) {
            if (w.graph[c]) {
                (void)hipGraphExecDestroy(w.graph[c]);
                w.graph[c] = nullptr;
            }
            hipGraph_t g = nullptr;
            std::lock_guard<std::recursive_mutex> lock(alloc_capture_mutex());
            CPM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int s = chunk(c);
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
        // retained rows when a store has one-stride rows (counting only)
        bool one_stride = false;
        if (!periodic)
            for (auto &d : in) one_stride = one_stride || d.bits->stride == kRetainWords;
        slots = one_stride ? std::min<uint64_t>(w.caps.ext + 1, kRetainBytes / (kRetainWords * 8)) : 0;
        if (slots)
            for (int i = 0; i < 2; ++i) CPM_CHECK(w.set[i].rows.ensure(slots * kRetainWords * 8));
        CPM_CHECK(w.pat.ensure(4 * pat_pitch()));
        {
            void *before = w.filter.p;
            CPM_CHECK(w.filter.ensure(filter_words() * 8));
            if (w.filter.p != before) w.filter_dirty = true;
        }
        if (periodic) CPM_CHECK(w.pat_per.ensure((w.caps.patterns + 1) * 8));
        CPM_CHECK(w.part.ensure((size_t)3 * grid_tasks() * 8));
        CPM_CHECK(w.ranges.ensure((size_t)2 * (w.caps.ext + 1) * 4));
        CPM_CHECK(w.partl.ensure((size_t)grid() * 8));
        CPM_CHECK(w.ctl.ensure(kCtlBytes + (2 * in.size() + 1) * 4));
        CPM_CHECK(w.stamps.ensure((size_t)grid() * 4 * 8));
        CPM_CHECK(w.host.ensure(std::max<size_t>(kCtlBytes + (2 * in.size() + 1) * 4, 4096)));
        return CPM_OK;
    }

    // the control block (and `tail` bytes after it) into pinned host memory
    int read_ctl(Ctl *h, size_t tail = 0) {
        CPM_HIP(hipMemcpyAsync(h, w.ctl.p, tail ? kCtlBytes + tail : sizeof(Ctl), hipMemcpyDeviceToHost,
                               ctx->stream));
        CPM_HIP(hipStreamSynchronize(ctx->stream));
        w.stage_off = 0;
        return CPM_OK;
    }

    // n bytes of pinned staging, valid until the next stream sync.  (A copy
    // from or to pageable memory makes the runtime stage it synchronously.)
    int stage(size_t n, unsigned char **h) {
        const size_t end = (w.stage_off + n + 63) / 64 * 64;
        if (!w.stage.p || end > w.stage.cap) {
            CPM_HIP(hipStreamSynchronize(ctx->stream));
            w.stage_off = 0;
            CPM_CHECK(w.stage.ensure(std::max<size_t>(2 * w.stage.cap, std::max<size_t>(n + 64, 1 << 16))));
        }
        *h = w.stage.as<unsigned char>() + w.stage_off;
        w.stage_off = (w.stage_off + n + 63) / 64 * 64;
        return CPM_OK;
    }

    // Queued host -> device copies (qput / qzero), applied by flush() with
    // one copy of the packed records and data and one k_scatter launch
    // instead of a copy launch each.  Large records are cut into 64 KB
    // pieces so the launch spreads them over many blocks.
    std::vector<PutRec> recs;
    std::vector<unsigned char> blob;
    static constexpr uint64_t kPiece = 64 << 10;

    void qput(void *dst, const void *src, size_t n) {
        for (size_t off = 0; off < n; off += kPiece) {
            const size_t m = std::min<size_t>(kPiece, n - off);
            recs.push_back({reinterpret_cast<uint32_t *>(static_cast<unsigned char *>(dst) + off),
                            (uint64_t)blob.size(), (uint64_t)m});
            blob.insert(blob.end(), static_cast<const unsigned char *>(src) + off,
                        static_cast<const unsigned char *>(src) + off + m);
            blob.resize((blob.size() + 15) / 16 * 16, 0);
        }
    }

    void qzero(void *dst, size_t n) {
        for (size_t off = 0; off < n; off += kPiece)
            recs.push_back({reinterpret_cast<uint32_t *>(static_cast<unsigned char *>(dst) + off), kZeroFill,
                            (uint64_t)std::min<size_t>(kPiece, n - off)});
    }

    int flush() {
        if (recs.empty()) return CPM_OK;
        for (const PutRec &r : recs)
            if (r.bytes % 4 || reinterpret_cast<uintptr_t>(r.dst) % 4)
                return fail(CPM_ERR_INVALID, "queued copy of %llu bytes is not word aligned",
                            (unsigned long long)r.bytes);
        // the records are applied in no particular order: none may overlap
        std::vector<std::pair<uintptr_t, uintptr_t>> span(recs.size());
        for (size_t i = 0; i < recs.size(); ++i)
            span[i] = {reinterpret_cast<uintptr_t>(recs[i].dst), reinterpret_cast<uintptr_t>(recs[i].dst) + recs[i].bytes};
        std::sort(span.begin(), span.end());
        for (size_t i = 1; i < span.size(); ++i)
            if (span[i].first < span[i - 1].second)
                return fail(CPM_ERR_INVALID, "queued copies overlap");
        const size_t head = (recs.size() * sizeof(PutRec) + 255) / 256 * 256;
        const size_t total = head + blob.size();
        unsigned char *h = nullptr;
        CPM_CHECK(stage(total, &h));
        memcpy(h, recs.data(), recs.size() * sizeof(PutRec));
        if (!blob.empty()) memcpy(h + head, blob.data(), blob.size());
        CPM_CHECK(w.scatter.ensure(total));
        CPM_HIP(hipMemcpyAsync(w.scatter.p, h, total, hipMemcpyHostToDevice, ctx->stream));
        const uint32_t n = (uint32_t)recs.size();
        hipLaunchKernelGGL(k_scatter, dim3(std::min<uint32_t>(n, 4096)), dim3(256), 0, ctx->stream,
                           w.scatter.as<PutRec>(), n,
                           reinterpret_cast<const uint32_t *>(w.scatter.as<unsigned char>() + head));
        CPM_HIP(hipGetLastError());
        recs.clear();
        blob.clear();
        return CPM_OK;
    }

    // One full search with the current capacities; *overflow receives the
    // capacity bits that were exceeded (0 when the search completed).
    int attempt(uint32_t *overflow) {
        *overflow = 0;
        const uint32_t n_db = (uint32_t)in.size();
        hipStream_t st = ctx->stream;
        CPM_CHECK(reserve());
        Ctl *hc = w.host.as<Ctl>();
        recs.clear();
        blob.clear();
        ctl_current = false;
        // stamps left by an attempt that overflowed
        qzero(w.stamps.p, (size_t)grid() * 4 * 8);
        if (w.filter_dirty) {
            qzero(w.filter.p, filter_words() * 8);
            w.filter_dirty = false;
        }

        // database table (frequent ranges filled after level 1)
        dbs.assign(n_db, DbDev{});
        te = 1;
        uint64_t total_items = 0;
        for (uint32_t i = 0; i < n_db; ++i) {
            const cpm_bits *b = in[i].bits;
            DbDev &D = dbs[i];
            D.store = b->rows();
            D.stride = b->stride;
            D.words = b->words;
            D.n = b->n_transactions;
            D.tiles = tiles_per_row(b->stride, periodic ? kTileWords : tile_words);
            D.tiles_eff = tiles_eff_of(D.tiles);
            te = std::max(te, D.tiles_eff);
            D.minsup = in[i].minsup;
            D.maxper = in[i].maxper;
            D.end = in[i].end;
            D.ts = in[i].d_ts;
            total_items += b->n_items;
        }
        use_ranges = false;
        for (const DbDev &D : dbs) use_ranges = use_ranges || (n_db > 1 && D.tiles < te);
        CPM_CHECK(w.dbs.ensure(std::max<uint32_t>(n_db, 1) * sizeof(DbDev)));
        qput(w.dbs.p, dbs.data(), n_db * sizeof(DbDev));

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
            std::vector<uint32_t> off(n_db + 1);
            uint64_t e = 0;
            for (uint32_t i = 0; i < n_db; ++i) {
                off[i] = (uint32_t)e;
                e += in[i].bits->n_items;
            }
            off[n_db] = (uint32_t)e;
            qput(S0.ext_off.p, off.data(), (n_db + 1) * 4);
            LevelSizes l1;
            memset(&l1, 0, sizeof(l1));
            l1.n_tasks = n_db;
            l1.d = 0;
            l1.n_ext = e;
            l1.n_cand = e;
            l1.group = level_group(e, G, te, work_target());
            l1.n_work = level_work(e, (uint32_t)l1.group, te);
            // the whole control block, zeroed but for level 1's sizes
            Ctl c0;
            memset(&c0, 0, sizeof(c0));
            c0.sz[0] = l1;
            qput(w.ctl.p, &c0, sizeof(c0));
            CPM_CHECK(flush());
            hipLaunchKernelGGL(k_level1, dim3(std::min<uint32_t>(n_db, 4096)), dim3(256), 0, st, n_db,
                               S0.ext_off.as<uint32_t>(), S0.ext.as<uint32_t>(), S0.sup.as<uint32_t>(),
                               S0.db.as<uint32_t>(), S0.pid.as<uint32_t>());
            CPM_HIP(hipGetLastError());
            CPM_CHECK(count(0, true));
        }
        // plan level 1: the frequent items (no children)
        plan(plan_args(0, 0));
        CPM_HIP(hipGetLastError());
        nf.assign(n_db, 0);
        std::vector<uint32_t> fbase(n_db + 1);
        hipLaunchKernelGGL(k_level1_pack, dim3(std::min<uint32_t>((2 * n_db + 256) / 256, 64)), dim3(256), 0, st,
                           n_db, S0.cnt[0].as<uint32_t>(), S0.off[0].as<uint32_t>(), ctl());
        CPM_HIP(hipGetLastError());
        CPM_CHECK(read_ctl(hc, (size_t)(2 * n_db + 1) * 4));
        {
            const unsigned char *tail = reinterpret_cast<const unsigned char *>(hc) + kCtlBytes;
            memcpy(nf.data(), tail, n_db * 4);
            memcpy(fbase.data(), tail + (size_t)n_db * 4, (n_db + 1) * 4);
        }
        if (hc->overflow) {
            *overflow = hc->overflow;
            return CPM_OK;
        }
        // the support stamps are folded and nothing below changes the
        // control block unless a level-2 search follows
        ctl_current = true;
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
        qput(w.dbs.p, dbs.data(), n_db * sizeof(DbDev));
        if (t2 == 0) return flush();
        ctl_current = false;
        if (e2 >= 0xffffffffull) return fail(CPM_ERR_NOMEM, "too many frequent item pairs");
        if (t2 > w.caps.tasks || e2 > w.caps.ext || (periodic && e2 * parts_tiles > w.caps.parts)) {
            // level 2 does not fit: the retry starts with room for it
            *overflow = (t2 > w.caps.tasks ? kOvTasks : 0) | (e2 > w.caps.ext ? kOvEntries : 0) |
                        (periodic && e2 * parts_tiles > w.caps.parts ? kOvParts : 0);
            need.tasks = t2;
            need.ext = e2;
            need.parts = periodic ? e2 * parts_tiles : 0;
            return CPM_OK;
        }
        LevelSet &S1 = w.set[1];
        // Level 2's tables, the Gram's job tables and the cleared Gram output
        // and pair tables go to the device in one queued flush.
        CPM_CHECK(w.l2.ensure(n_db * sizeof(L2Db)));
        qput(w.l2.p, l2.data(), n_db * sizeof(L2Db));
        {
            LevelSizes ls;
            memset(&ls, 0, sizeof(ls));
            ls.n_tasks = t2;
            ls.d = 1;
            ls.n_ext = e2;
            ls.n_cand = e2;
            ls.group = level_group(e2, G, te, work_target());
            ls.n_work = level_work(e2, (uint32_t)ls.group, te);
            qput(sz(1), &ls, sizeof(ls));
            const uint32_t e2u = (uint32_t)e2;
            qput(S1.ext_off.as<uint32_t>() + t2, &e2u, 4);
        }
        bool gram_ok = use_gram;
        for (uint32_t i = 0; i < n_db && gram_ok; ++i)
            if (nf[i] > 16384 || dbs[i].n >= (1ull << 31)) gram_ok = false;
        GramPlan gplan;
        GramDev gdev;
        uint64_t gram_elems = 0, gram_ops = 0;
        if (gram_ok) {
            std::vector<GramDb> gdb(n_db);
            std::vector<uint64_t> goff(n_db);
            uint64_t off = 0;
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
                gram_ops += dbs[i].n * (uint64_t)gdb[i].n_items * (gdb[i].n_items + 1);
            }
            CPM_CHECK(gram_plan(ctx, gdb, gplan));
            gram_elems = gplan.out_elems;
            CPM_CHECK(w.gram.ensure(std::max<uint64_t>(gram_elems, 1) * 4));
            CPM_CHECK(w.goff.ensure(n_db * 8));
            CPM_CHECK(w.gjobs.ensure(std::max<size_t>(gplan.jobs.size(), 1) * sizeof(GramJob)));
            CPM_CHECK(w.gblocks.ensure(std::max<size_t>(gplan.blocks.size(), 1) * sizeof(GramBlock)));
            CPM_CHECK(w.gparts.ensure(std::max<uint64_t>(gplan.slots, 1) * 256 * 256 * 4));
            CPM_CHECK(w.gdbs.ensure(n_db * sizeof(GramDb)));
            qput(w.goff.p, goff.data(), n_db * 8);
            qput(w.gjobs.p, gplan.jobs.data(), gplan.jobs.size() * sizeof(GramJob));
            qput(w.gblocks.p, gplan.blocks.data(), gplan.blocks.size() * sizeof(GramBlock));
            qput(w.gdbs.p, gdb.data(), n_db * sizeof(GramDb));
            gdev.jobs = w.gjobs.as<GramJob>();
            gdev.n_jobs = gplan.jobs.size();
            gdev.blocks = w.gblocks.as<GramBlock>();
            gdev.n_blocks = gplan.blocks.size();
            gdev.partials = w.gparts.as<int32_t>();
            qzero(w.gram.p, gram_elems * 4);  // the Gram writes the upper triangles
        }
        tmask_w = 0;
        if (pair_words) {
            // per-task kept masks for plan/fill, when they fit a modest budget
            uint32_t wmax = 0;
            for (auto &D : dbs) wmax = std::max(wmax, D.pair_w32);
            if ((w.caps.tasks + 1) * wmax * 4 <= kTaskMaskBudget) {
                tmask_w = wmax;
                CPM_CHECK(w.tmask.ensure((w.caps.tasks + 1) * wmax * 4));
            }
            CPM_CHECK(w.pairbits.ensure(pair_words * 4));
            qzero(w.pairbits.p, pair_words * 4);
        } else {
            w.pairbits.release();
        }
        CPM_CHECK(flush());

        const unsigned g2 = (unsigned)std::min<uint64_t>(t2, 8192);
        bool pairs_done = false;  // the pair tables were written with the level-2 supports
        if (gram_ok && periodic) {
            // level 2 = the frequent pairs (Gram supports), then their periods
            CPM_CHECK(launch_gram(ctx, gdev, w.gdbs.as<GramDb>(), w.gram.as<int32_t>(), gram_elems, gram_ops, false));
            uint32_t *cnt = S1.cnt[0].as<uint32_t>();
            hipLaunchKernelGGL(k_level2_frequent_count, dim3(g2), dim3(256), 0, st, w.l2.as<L2Db>(), n_db,
                               (uint32_t)t2, w.dbs.as<DbDev>(), w.gram.as<int32_t>(), w.goff.as<uint64_t>(), cnt);
            hipLaunchKernelGGL(k_level2_frequent_scan, dim3(1), dim3(1024), 0, st, cnt, (uint32_t)t2,
                               S1.ext_off.as<uint32_t>(), sz(1), G, te, work_target());
            hipLaunchKernelGGL(k_level2_frequent_write, dim3(g2), dim3(256), 0, st, w.l2.as<L2Db>(), n_db,
                               (uint32_t)t2, S1.ptrs(sz(1)), w.dbs.as<DbDev>(), w.gram.as<int32_t>(),
                               w.goff.as<uint64_t>());
            CPM_HIP(hipGetLastError());
            CPM_CHECK(count(1, false));
        } else {
            hipLaunchKernelGGL(k_level2, dim3(g2), dim3(256), 0, st, w.l2.as<L2Db>(), n_db, (uint32_t)t2,
                               S1.ptrs(sz(1)), w.dbs.as<DbDev>());
            CPM_HIP(hipGetLastError());
            if (gram_ok) {
                CPM_CHECK(launch_gram(ctx, gdev, w.gdbs.as<GramDb>(), w.gram.as<int32_t>(), gram_elems, gram_ops,
                                      false));
                hipLaunchKernelGGL(k_gram_to_level2, dim3(g2), dim3(256), 0, st, S1.ptrs(sz(1)),
                                   w.dbs.as<DbDev>(), w.gram.as<int32_t>(), w.goff.as<uint64_t>(),
                                   pair_words ? w.pairbits.as<uint32_t>() : nullptr);
                CPM_HIP(hipGetLastError());
                pairs_done = true;
            } else {
                CPM_CHECK(count(1, false));
            }
        }
        if (pair_words && !pairs_done) {
            PlanArgs p = plan_args(1, 1);
            hipLaunchKernelGGL(k_pair_table, dim3((unsigned)std::min<uint64_t>(t2, 8192)), dim3(256), 0, st, p,
                               w.pairbits.as<uint32_t>());
            CPM_HIP(hipGetLastError());
        }

        // ---- levels >= 2: replayed steps ----
        const int c = 1;
        for (;;) {
            CPM_CHECK(launch_chunk(c));
            CPM_CHECK(read_ctl(hc));
            if (hc->overflow) {
                *overflow = hc->overflow;
                return CPM_OK;
            }
            CPM_CHECK(queue_copy(hc->n_patterns));
            if (hc->sz[c].n_tasks == 0) {
                // level c is empty: its support launch wrote no stamps, and
                // the plan that made it folded the previous launch's
                ctl_current = true;
                break;
            }
        }
        return CPM_OK;
    }

    int run(cpm_patterns **out) {
        // First capacities: level 1 (one entry per item of every store).
        // Level 2's sizes are known exactly once level 1 has found the
        // frequent items (attempt() records them in `need`), so an input
        // with many distinct but few frequent items never sizes anything by
        // the square of its item count.  Later levels grow by doubling.
        uint64_t items = 0;
        for (auto &d : in) items += d.bits->n_items;
        parts_tiles = 0;
        if (periodic)
            for (auto &d : in) parts_tiles = std::max(parts_tiles, tiles_per_row(d.bits->stride));
        Caps want;
        want.tasks = std::max<uint64_t>(std::max<uint64_t>(items, in.size()), 1 << 16);
        want.ext = std::max<uint64_t>(items, 1 << 18);
        want.patterns = std::max<uint64_t>(items * 4, 1 << 18);
        want.parts = periodic ? want.ext * std::max<uint32_t>(parts_tiles, 1) : 0;
        grow_caps(want);
        for (int tries = 0;; ++tries) {
            uint32_t ov = 0;
            need = Caps{};
            drop_stream();  // before buffers are resized
            const int rc = attempt(&ov);
            if (rc != CPM_OK || ov) w.filter_dirty = true;  // stopped between insert and clear
            CPM_CHECK(rc);
            if (!ov) break;
            if (tries > 24) return fail(CPM_ERR_NOMEM, "search buffers keep overflowing (%u)", ov);
            if (ov & kOvDepth) return fail(CPM_ERR_NOMEM, "an itemset longer than %u items", kMaxDepth + 1);
            if (ov & kOvIndex) return fail(CPM_ERR_NOMEM, "a level with 2^32 or more candidates");
            if (ov & kOvTasks) w.caps.tasks *= 2;
            if (ov & kOvEntries) w.caps.ext *= 2;
            if (ov & kOvPatterns) w.caps.patterns *= 2;
            if (ov & kOvParts) w.caps.parts = std::max<uint64_t>(w.caps.parts * 2, 1 << 20);
            grow_caps(need);
        }
        return download(out);
    }

    // capacities raised to at least `c` (field by field)
    void grow_caps(const Caps &c) {
        for (int i = 0; i < 4; ++i) {
            uint64_t *a = &w.caps.tasks + i;
            const uint64_t *b = &c.tasks + i;
            if (*a < *b) *a = *b;
        }
    }

    int download(cpm_patterns **out) {
        Ctl *hc = w.host.as<Ctl>();
        if (!ctl_current) {
            hipLaunchKernelGGL(k_fold_stamps, dim3(1), dim3(256), 0, ctx->stream,
                               w.stamps.as<unsigned long long>(), grid(), ctl());
            CPM_HIP(hipGetLastError());
            CPM_CHECK(read_ctl(hc));
        }
        const uint64_t n = hc->n_patterns;
        if (res) {
            // streamed: the last chunk's patterns, then one wait
            CPM_CHECK(queue_copy(n));
            const hipError_t e = hipStreamSynchronize(w.copy);
            if (e != hipSuccess) return fail(CPM_ERR_RUNTIME, "pattern download failed: %s", hipGetErrorString(e));
            res->n = n;
            res->parent.resize(n);
            res->item.resize(n);
            res->support.resize(n);
            res->db.resize(n);
            if (periodic) res->period.resize(n);
            record_stats(hc);
            *out = res;
            res = nullptr;
            copied = 0;
            return CPM_OK;
        }
        cpm_patterns *p = new cpm_patterns();
        p->n = n;
        p->periodic = periodic;
        hipStream_t st = ctx->stream;
        try {
            // the result's arrays come from recycled (pinned) host blocks and
            // are not cleared first: the copies below write every element
            p->parent.resize(n);
            p->item.resize(n);
            p->support.resize(n);
            p->db.resize(n);
            if (periodic) p->period.resize(n);
        } catch (const std::bad_alloc &) {
            delete p;
            return fail(CPM_ERR_NOMEM, "host memory for %llu patterns", (unsigned long long)n);
        }
        if (n && n * 4 >= kDirectDownload) {
            // large results: supports widened on the device, then every
            // column straight into the result's (pinned) arrays, one wait
            int sc = w.sup64.ensure(n * 8);
            if (sc != CPM_OK) {
                delete p;
                return sc;
            }
            hipLaunchKernelGGL(k_widen_u32, grid_of_n(n), dim3(256), 0, st, pat_col(2), n,
                               w.sup64.as<unsigned long long>());
            hipError_t e = hipGetLastError();
            if (e == hipSuccess) e = hipMemcpyAsync(p->parent.data(), pat_col(0), n * 4, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipMemcpyAsync(p->item.data(), pat_col(1), n * 4, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipMemcpyAsync(p->db.data(), pat_col(3), n * 4, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipMemcpyAsync(p->support.data(), w.sup64.p, n * 8, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess && periodic)
                e = hipMemcpyAsync(p->period.data(), w.pat_per.p, n * 8, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                delete p;
                return fail(CPM_ERR_RUNTIME, "pattern download failed: %s", hipGetErrorString(e));
            }
        } else if (n) {
            // small results (plain host memory): the four columns in one
            // strided copy to the pinned stage, then copied out
            const size_t row = periodic ? 24 : 16;
            int sc = w.down.ensure(n * row);
            if (sc != CPM_OK) {
                delete p;
                return sc;
            }
            unsigned char *hd = w.down.as<unsigned char>();
            hipError_t e = hipMemcpy2DAsync(hd, n * 4, w.pat.p, pat_pitch(), n * 4, 4, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess && periodic)
                e = hipMemcpyAsync(hd + n * 16, w.pat_per.p, n * 8, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                delete p;
                return fail(CPM_ERR_RUNTIME, "pattern download failed: %s", hipGetErrorString(e));
            }
            const uint32_t *hp = reinterpret_cast<const uint32_t *>(hd);
            memcpy(p->parent.data(), hp, n * 4);
            memcpy(p->item.data(), hp + n, n * 4);
            memcpy(p->db.data(), hp + 3 * n, n * 4);
            for (uint64_t i = 0; i < n; ++i) p->support[i] = hp[2 * n + i];
            if (periodic) memcpy(p->period.data(), hd + n * 16, n * 8);
        }
        record_stats(hc);
        *out = p;
        return CPM_OK;
    }

    void record_stats(const Ctl *hc) {
        int wall_khz = 0;
        (void)hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, ctx->device);
        const double tick_ms = wall_khz > 0 ? 1.0 / wall_khz : 0.0;
        ctx->stats[0] += (double)hc->acc.launches;
        ctx->stats[1] += (double)hc->acc.ticks * tick_ms;
        ctx->stats[2] += (double)hc->acc.bytes;
        ctx->stats[3] += (double)hc->acc.candidates;
        ctx->stats[7] = (double)hc->levels + 2;
    }
};

int check_bits(const cpm_bits *b, cpm_ctx *ctx) {
    if (!b) return fail(CPM_ERR_INVALID, "null store");
    if (!b->ctx || b->ctx->device != ctx->device)
        return fail(CPM_ERR_INVALID, "store belongs to another device");
    if (b->n_transactions >= 0xffffffffull)
        return fail(CPM_ERR_INVALID, "supports are counted in 32 bits: too many transactions");
    return CPM_OK;
}

}  // namespace

}  // namespace cpm

namespace cpm {

// ---- complete-intersection counting without device allocation ---------------
//
// cpm_count_supports_noalloc builds the support kernel's task tables from
// device-resident candidates into `scratch`, a file-scope device array of
// fixed size, so the call allocates nothing.  Layout (byte offsets):
//   [kNaDb]     DbDev of the store
//   [kNaSizes]  LevelSizes of the one level
//   [kNaFlag]   u32 error flag (a column >= the store's item count)
//   [kNaPrefix] u32 prefix items, kMaxDepth per task
//   [kNaOff]    u32 ext_off[tasks + 1]
//   [kNaExt]    u32 ext[candidates]
//   [kNaDbi]    u32 db[tasks] (all 0)
//   [kNaSup]    u32 sup[candidates] (cleared by the table kernel)
constexpr uint64_t kNoAllocCands = 65536;   // candidates per call
constexpr size_t kNaDb = 0, kNaSizes = 256, kNaFlag = 512, kNaPrefix = 768;
constexpr size_t kNaOff = kNaPrefix + kNoAllocCands * kMaxDepth * 4;
constexpr size_t kNaExt = kNaOff + (kNoAllocCands + 64) * 4;
constexpr size_t kNaDbi = kNaExt + kNoAllocCands * 4;
constexpr size_t kNaSup = kNaDbi + kNoAllocCands * 4;
constexpr size_t kNoAllocBytes = kNaSup + kNoAllocCands * 4;
static_assert(sizeof(DbDev) <= 256 && sizeof(LevelSizes) <= 256, "scratch header slots");
static_assert(kNaOff % 256 == 0 && kNaExt % 256 == 0 && kNaDbi % 256 == 0 && kNaSup % 256 == 0,
              "scratch sections stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on each device, from hipGetSymbolAddress at
// context creation.  Host code must use these addresses: the array's own
// name is a host-side stand-in, not device memory.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(CPM_ERR_INVALID, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return CPM_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(CPM_ERR_RUNTIME, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                    hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return CPM_OK;
}

namespace {

// One 1024-lane workgroup: runs of candidates with equal first width - 1
// columns become one task (as run_candidates groups them on the host); the
// tables, the level's sizes and the store's DbDev go to the scratch array,
// the support accumulators are cleared.  A column beyond the store raises
// the flag, and the support launch that follows then does nothing.
__global__ __launch_bounds__(1024) void k_na_tasks(const uint32_t *__restrict__ cand, uint32_t n, uint32_t width,
                                                    uint32_t n_items, DbDev D, uint32_t group,
                                                    unsigned long long n_work, unsigned char *ws) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t bad;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t d = width - 1;
    uint32_t *prefix = reinterpret_cast<uint32_t *>(ws + kNaPrefix);
    uint32_t *off = reinterpret_cast<uint32_t *>(ws + kNaOff);
    uint32_t *ext = reinterpret_cast<uint32_t *>(ws + kNaExt);
    uint32_t *dbi = reinterpret_cast<uint32_t *>(ws + kNaDbi);
    uint32_t *sup = reinterpret_cast<uint32_t *>(ws + kNaSup);
    if (tid == 0) bad = 0;
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t lo = min(n, tid * per), hi = min(n, lo + per);
    auto fresh = [&](uint32_t c) {
        if (c == 0) return true;
        const uint32_t *a = cand + (uint64_t)c * width, *b = a - width;
        for (uint32_t i = 0; i < d; ++i)
            if (a[i] != b[i]) return true;
        return false;
    };
    uint32_t mine = 0, badl = 0;
    for (uint32_t c = lo; c < hi; ++c) {
        mine += fresh(c) ? 1u : 0u;
        for (uint32_t i = 0; i < width; ++i) badl |= cand[(uint64_t)c * width + i] >= n_items ? 1u : 0u;
    }
    // exclusive scan of the per-lane run starts over the workgroup
    uint32_t x = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (badl) atomicOr(&bad, 1u);
    uint32_t base = 0, total = 0;
    for (uint32_t w = 0; w < 16; ++w) {
        if (w < wave) base += wsum[w];
        total += wsum[w];
    }
    base += x - mine;
    // task of candidate lo: the last run started at or before it
    uint32_t t = base - 1;
    for (uint32_t c = lo; c < hi; ++c) {
        const uint32_t *it = cand + (uint64_t)c * width;
        if (fresh(c)) {
            ++t;
            off[t] = c;
            dbi[t] = 0;
            for (uint32_t i = 0; i < d; ++i) prefix[(uint64_t)t * kMaxDepth + i] = it[i];
        }
        ext[c] = it[d];
        sup[c] = 0;
    }
    __syncthreads();
    if (tid == 0) {
        off[total] = n;
        *reinterpret_cast<DbDev *>(ws + kNaDb) = D;
        LevelSizes ls;
        memset(&ls, 0, sizeof(ls));
        ls.n_tasks = total;
        ls.d = d;
        ls.n_ext = n;
        ls.n_cand = n;
        ls.group = group;
        ls.n_work = n_work;
        *reinterpret_cast<LevelSizes *>(ws + kNaSizes) = ls;
        *reinterpret_cast<uint32_t *>(ws + kNaFlag) = bad;
    }
}

// u32 supports from the scratch accumulators into the caller's u64 output
__global__ void k_na_widen(const uint32_t *__restrict__ sup, uint32_t n, uint64_t *__restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = sup[i];
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

static int pfeclat_run(cpm_ctx *ctx, const cpm_bits *bits, const int64_t *d_ts, int64_t end, uint64_t minsup,
                       int64_t maxper_floor, cpm_patterns **out) {
    Search s(ctx, search_ws(ctx));
    s.periodic = true;
    s.use_gram = true;  // level 2: the Gram's supports select the pairs to scan
    s.G = ctx->opt_group;
    s.tile_words = kTileWords;
    s.use_graph = ctx->opt_graph;
    s.prune = ctx->opt_prune;
    s.in.push_back({bits, minsup, maxper_floor, end, d_ts});
    return s.run(out);
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
    return pfeclat_run(ctx, bits, w.ts.as<int64_t>(), end, minsup, maxper_floor, out);
}

int cpm_pfeclat_mine_device(cpm_ctx *ctx, const cpm_bits *bits, const int64_t *d_timestamps, int64_t end,
                            uint64_t minsup, int64_t maxper_floor, cpm_patterns **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "null out");
    *out = nullptr;
    CPM_CHECK(check_bits(bits, ctx));
    if (bits->n_transactions && !d_timestamps) return fail(CPM_ERR_INVALID, "null timestamps");
    return pfeclat_run(ctx, bits, d_timestamps, end, minsup, maxper_floor, out);
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
    D.store = b->rows();
    D.stride = b->stride;
    D.words = b->words;
    D.n = b->n_transactions;
    D.tiles = tiles_per_row(b->stride, tw);
    D.tiles_eff = tiles_eff_of(D.tiles);
    D.end = end;
    std::vector<uint32_t> prefix, off, ext(n_cand);
    for (uint64_t c = 0; c < n_cand; ++c) {
        const uint32_t *it = cand + c * width;
        const bool fresh = c == 0 || memcmp(it, cand + (c - 1) * width, (size_t)d * 4) != 0;
        if (fresh) {
            off.push_back((uint32_t)c);
            prefix.resize(prefix.size() + kMaxDepth, 0);
            memcpy(prefix.data() + prefix.size() - kMaxDepth, it, (size_t)d * 4);
        }
        ext[c] = it[d];
    }
    const uint64_t nt = off.size();
    off.push_back((uint32_t)n_cand);
    std::vector<uint32_t> dbz(nt, 0);
    DBuf dprefix, doff, dext, dsup, dper, ddb, dD, dsz, dparts, dts;
    hipStream_t st = ctx->stream;
    CPM_CHECK(dprefix.ensure(prefix.size() * 4));
    CPM_CHECK(doff.ensure(off.size() * 4));
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
    ls.n_cand = n_cand;
    ls.group = level_group(n_cand, G, D.tiles_eff, 4ull * support_grid(ctx));
    ls.n_work = level_work(n_cand, (uint32_t)ls.group, D.tiles_eff);
    CPM_HIP(hipMemcpyAsync(dprefix.p, prefix.data(), prefix.size() * 4, hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(doff.p, off.data(), off.size() * 4, hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(dext.p, ext.data(), n_cand * 4, hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(ddb.p, dbz.data(), nt * 4, hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(dD.p, &D, sizeof(D), hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(dsz.p, &ls, sizeof(ls), hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemsetAsync(dsup.p, 0, n_cand * 4, st));
    SupportArgs a;
    a.L.prefix = dprefix.as<uint32_t>();
    a.L.db = ddb.as<uint32_t>();
    a.L.ext_off = doff.as<uint32_t>();
    a.L.ext = dext.as<uint32_t>();
    a.L.sup = dsup.as<uint32_t>();
    a.L.per = dper.as<int64_t>();
    a.L.sz = dsz.as<LevelSizes>();
    a.dbs = dD.as<DbDev>();
    a.te = D.tiles_eff;
    a.parts = dparts.as<PeriodPart>();
    a.parts_tiles = D.tiles;
    if (periodic) CPM_CHECK(launch_support_periodic(ctx, a, support_grid(ctx)));
    else CPM_CHECK(launch_support_count(ctx, a, tw, support_grid(ctx)));
    std::vector<uint32_t> s32(n_cand);
    CPM_CHECK(copy_to_host(ctx, s32.data(), dsup.p, n_cand * 4));
    if (periodic) CPM_CHECK(copy_to_host(ctx, per_out, dper.p, n_cand * 8));
    for (uint64_t c = 0; c < n_cand; ++c) sup_out[c] = s32[c];
    return CPM_OK;
}

int cpm_count_supports_noalloc(cpm_ctx *ctx, const cpm_bits *b, const uint32_t *d_cand, uint64_t n_cand,
                               uint32_t width, uint64_t *d_supports) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!b) return fail(CPM_ERR_INVALID, "null store");
    if (!b->ctx || b->ctx->device != ctx->device)
        return fail(CPM_ERR_INVALID, "store belongs to another device");
    if (n_cand == 0) return CPM_OK;
    if (!d_cand || !d_supports) return fail(CPM_ERR_INVALID, "null candidates/supports");
    if (n_cand > kNoAllocCands)
        return fail(CPM_ERR_INVALID, "%llu candidates in one call: at most %llu without device allocation",
                    (unsigned long long)n_cand, (unsigned long long)kNoAllocCands);
    if (width == 0) return fail(CPM_ERR_INVALID, "candidate width must be >= 1");
    if (width > kMaxDepth + 1) return fail(CPM_ERR_INVALID, "candidates wider than %u items", kMaxDepth + 1);
    if (b->n_transactions >= 0xffffffffull)
        return fail(CPM_ERR_INVALID, "supports are counted in 32 bits: too many transactions");
    unsigned char *ws = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices)
            ws = static_cast<unsigned char *>(g_scratch_ptr[ctx->device]);
    }
    if (!ws) return fail(CPM_ERR_RUNTIME, "no scratch address for device %d", ctx->device);
    const uint32_t tw = kTileWords;
    DbDev D;
    memset(&D, 0, sizeof(D));
    D.store = b->rows();
    D.stride = b->stride;
    D.words = b->words;
    D.n = b->n_transactions;
    D.tiles = tiles_per_row(b->stride, tw);
    D.tiles_eff = tiles_eff_of(D.tiles);
    const uint32_t group = level_group(n_cand, 32, D.tiles_eff, 4ull * support_grid(ctx));
    const unsigned long long n_work = level_work(n_cand, group, D.tiles_eff);
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_na_tasks, dim3(1), dim3(1024), 0, st, d_cand, (uint32_t)n_cand, width, b->n_items, D,
                       group, n_work, ws);
    CPM_HIP(hipGetLastError());
    SupportArgs a;
    a.L.prefix = reinterpret_cast<uint32_t *>(ws + kNaPrefix);
    a.L.db = reinterpret_cast<uint32_t *>(ws + kNaDbi);
    a.L.ext_off = reinterpret_cast<uint32_t *>(ws + kNaOff);
    a.L.ext = reinterpret_cast<uint32_t *>(ws + kNaExt);
    a.L.sup = reinterpret_cast<uint32_t *>(ws + kNaSup);
    a.L.sz = reinterpret_cast<LevelSizes *>(ws + kNaSizes);
    a.dbs = reinterpret_cast<const DbDev *>(ws + kNaDb);
    a.te = D.tiles_eff;
    a.stop = reinterpret_cast<const uint32_t *>(ws + kNaFlag);
    CPM_CHECK(launch_support_count(ctx, a, tw, support_grid(ctx)));
    hipLaunchKernelGGL(k_na_widen, dim3((unsigned)std::min<uint64_t>((n_cand + 255) / 256, 256)), dim3(256), 0, st,
                       a.L.sup, (uint32_t)n_cand, d_supports);
    CPM_HIP(hipGetLastError());
    uint32_t bad = 0;
    CPM_HIP(hipMemcpyAsync(&bad, ws + kNaFlag, 4, hipMemcpyDeviceToHost, st));
    CPM_HIP(hipStreamSynchronize(st));
    if (bad) return fail(CPM_ERR_INVALID, "a candidate column is >= the store's %u items", b->n_items);
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

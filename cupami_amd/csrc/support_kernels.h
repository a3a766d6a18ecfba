// Device-side description of a search level and the support kernels'
// host interface (support.hip).
#pragma once

#include "cpm_internal.h"

namespace cpm {

// Items in a task prefix are stored with this stride (the longest itemset a
// search can reach is kMaxDepth + 1 items).
constexpr uint32_t kMaxDepth = 64;

// Databases whose rows are one 64-word stride (at most 4096 transactions)
// keep each counted itemset's row, the AND of its items' rows, for the next
// level: a task then reads its prefix as one row instead of d.
constexpr uint32_t kRetainWords = 64;
constexpr uint64_t kRetainBytes = 256ull << 20;   // per level set

// One database of a search.  Frequent items of every database live in one
// global index space; database i owns indexes [fbase, fbase + nf).
struct DbDev {
    const uint64_t *store;   // item rows, `stride` words apart
    uint64_t stride;         // words per row (multiple of 64)
    uint64_t words;          // ceil(n / 64)
    uint64_t n;              // transactions
    uint32_t tiles;          // tiles of a row
    uint32_t tiles_eff;      // tiles rounded up to a multiple of 8 when >= 8
    uint64_t minsup;
    int64_t maxper, end;     // periodic searches
    const int64_t *ts;       // timestamps (periodic searches)
    uint32_t fbase, nf;
    uint64_t pair_off;       // offset (uint32 words) of this database's pair table
    uint32_t pair_w32, pad;
};

// Sizes of one level, in device memory (written by the level's producers).
struct LevelSizes {
    unsigned long long n_tasks, d, n_ext, n_work;
    unsigned long long n_pat, n_child, n_cext;  // totals of planning this level
    unsigned long long n_cand;                  // entries before the subset test
    unsigned long long group;                   // entries per support work range
    unsigned long long pad;
};

// A level's arrays.  Task t: prefix items prefix[t*kMaxDepth .. +d), database
// db[t], entries ext[ext_off[t] .. ext_off[t+1]) (entry e is the candidate
// prefix + {ext[e]}).  Entries hold global frequent-item indexes (through
// fcol) or, at level 1, columns of the task's database.
struct LevelPtrs {
    uint32_t *prefix = nullptr;
    uint32_t *pid = nullptr;
    uint32_t *par = nullptr;  // parent task: its index in the previous level (levels >= 3)
    uint32_t *src = nullptr;  // the previous level's entry whose itemset is the prefix (levels >= 3)
    uint32_t *db = nullptr;
    uint32_t *ext_off = nullptr;
    uint32_t *ext = nullptr;
    uint32_t *sup = nullptr;
    int64_t *per = nullptr;
    LevelSizes *sz = nullptr;
};

// Partial periodicity over one tile of words: the occurrences' first and last
// timestamps, the largest gap between consecutive occurrences, their count.
struct PeriodPart {
    int64_t first, last, gap;
    uint32_t count, pad;
};

// Totals of the support launches, in device memory (fold_stamps).
struct SupportAcc {
    unsigned long long ticks;       // summed launch spans (wall-clock ticks)
    unsigned long long launches;
    unsigned long long bytes;       // algorithmic bitset bytes streamed
    unsigned long long candidates;  // entries counted
};

// Support work: the level's entries are cut into ranges of G consecutive
// entries (a range may span several tasks), and work item i covers range
// i / te on tile i % te of every database's rows.  te is the largest
// database's tile count, rounded up to a multiple of 8 when >= 8; as the grid
// is a multiple of 8 too, a block -- which stays on XCD blockIdx % 8 -- only
// touches tiles congruent to its XCD, so each XCD's L2 holds its own slice of
// the hot rows.
struct SupportArgs {
    LevelPtrs L;
    const DbDev *dbs = nullptr;
    const uint32_t *fcol = nullptr;  // global frequent index -> column (null: entries are columns)
    uint32_t te = 1;                 // tiles per work range (entries per range: L.sz->group)
    const uint32_t *stop = nullptr;  // skip the launch when *stop != 0
    PeriodPart *parts = nullptr;     // periodic: n_ext x parts_tiles partials
    uint32_t parts_tiles = 0;
    // per-block stamps (start, end, bytes, candidates), one set per grid
    // block; null to skip
    unsigned long long *stamps = nullptr;
    // L.par holds each task's parent: tasks with one parent share the AND
    // of their first d - 1 prefix rows
    int siblings = 0;
    // Retained rows (kRetainWords words per slot, `slots` slots each; null:
    // none).  rout[e] receives entry e's itemset row when its database's
    // stride is kRetainWords and e < slots; a task of such a database whose
    // prefix has 3 or more items reads its prefix row as rin[L.src[t]] (the
    // previous launch's rout) when L.src[t] < slots.
    const uint64_t *rin = nullptr;
    uint64_t *rout = nullptr;
    uint64_t slots = 0;
    // Per work range r (entries [r G, r G + G)), filled by launch_ranges
    // before the count: ranges[r] = its first task, ranges[rcap + r] = the
    // most tiles any of its tasks' databases has, so a work item on a tile
    // no database of its range reaches is skipped with one load.  Null (or
    // r >= rcap): the kernel finds the task by binary search and walks the
    // range's tasks.
    uint32_t *ranges = nullptr;
    uint64_t rcap = 0;
};

// Entries per work range for a level of n_ext entries: at most G, and small
// enough that the level has about `target` work items when it is short (a
// short level's tasks are small, and a block walks the tasks of its range one
// after another).
__host__ __device__ inline uint32_t level_group(uint64_t n_ext, uint32_t G, uint32_t te, uint64_t target) {
    const uint64_t g = (n_ext * te + target - 1) / target;
    return (uint32_t)(g < 1 ? 1 : (g > G ? G : g));
}

__host__ __device__ inline uint64_t level_work(uint64_t n_ext, uint32_t g, uint32_t te) {
    return (n_ext + g - 1) / g * te;
}

inline uint32_t tiles_eff_of(uint32_t tiles) { return tiles >= 8 ? (tiles + 7) / 8 * 8 : tiles; }

// sup[e] += |prefix rows AND row ext[e]| for every entry (sup zeroed by the
// producer of the level).  tile_words is 1024 or 2048.
int launch_support_count(cpm_ctx *ctx, const SupportArgs &a, uint32_t tile_words, uint32_t grid);
// support and periodicity of every entry (partials, then an in-order fold)
int launch_support_periodic(cpm_ctx *ctx, const SupportArgs &a, uint32_t grid);
// a.ranges for the level a describes (sizes read from device memory, so it
// can be replayed from a graph); nothing when a.ranges is null
int launch_ranges(cpm_ctx *ctx, const SupportArgs &a, uint32_t grid);

// child_j = prefix & ext_j (rows of `stride` words), sup[j] = popcount.
// out may be null (count only).
int launch_intersect(cpm_ctx *ctx, const uint64_t *prefix, const uint64_t *ext, uint32_t k,
                     uint64_t stride, unsigned long long *sup, uint64_t *out);

inline uint32_t tiles_per_row(uint64_t stride, uint32_t tile_words = kTileWords) {
    return (uint32_t)((stride + tile_words - 1) / tile_words);
}

// persistent grid for the support kernels on this device (a multiple of 8)
inline uint32_t support_grid(const cpm_ctx *ctx) { return (uint32_t)ctx->cu_count * 8; }


}  // namespace cpm

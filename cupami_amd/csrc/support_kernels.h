// Device-side description of a search level and the support kernels'
// host interface (support.hip).
#pragma once

#include "cpm_internal.h"

namespace cpm {

// Items in a task prefix are stored with this stride (the longest itemset a
// search can reach is kMaxDepth + 1 items).
constexpr uint32_t kMaxDepth = 64;

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
    unsigned long long pad;
};

// A level's arrays.  Task t: prefix items prefix[t*kMaxDepth .. +d), database
// db[t], entries ext[ext_off[t] .. ext_off[t+1]) (entry e is the candidate
// prefix + {ext[e]}), support work items [work_off[t], work_off[t+1]).
// Entries hold global frequent-item indexes (through fcol) or, at level 1,
// columns of the task's database.
struct LevelPtrs {
    uint32_t *prefix = nullptr;
    uint32_t *pid = nullptr;
    uint32_t *db = nullptr;
    uint32_t *ext_off = nullptr;
    uint32_t *work_off = nullptr;
    uint32_t *nwork = nullptr;
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

struct SupportArgs {
    LevelPtrs L;
    const DbDev *dbs = nullptr;
    const uint32_t *fcol = nullptr;  // global frequent index -> column (null: entries are columns)
    uint32_t G = 32;                 // entries per work item
    const uint32_t *stop = nullptr;  // skip the launch when *stop != 0
    PeriodPart *parts = nullptr;     // periodic: n_ext x parts_tiles partials
    uint32_t parts_tiles = 0;
    // per-block device clock stamps of the launch (wall_clock64 ticks):
    // clock[2*b] start, clock[2*b+1] end of block b; null to skip
    unsigned long long *clock = nullptr;
};

// Number of support work items of a task with k entries in a database with
// tiles_eff tiles per row.  When rows have 8 or more tiles, tiles_eff is a
// multiple of 8 and so is every task's work count: work item i then always
// handles a tile congruent to i mod 8, i.e. to its block's XCD.
__host__ __device__ inline uint32_t task_work(uint32_t k, uint32_t G, uint32_t tiles_eff) {
    return (k + G - 1) / G * tiles_eff;
}

inline uint32_t tiles_eff_of(uint32_t tiles) { return tiles >= 8 ? (tiles + 7) / 8 * 8 : tiles; }

// sup[e] += |prefix rows AND row ext[e]| for every entry (sup zeroed by the
// producer of the level).  tile_words is 1024 or 2048.
int launch_support_count(cpm_ctx *ctx, const SupportArgs &a, uint32_t tile_words, uint32_t grid);
// support and periodicity of every entry (partials, then an in-order fold)
int launch_support_periodic(cpm_ctx *ctx, const SupportArgs &a, uint32_t grid);

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

// The support-counting kernels' host interface (support.hip).
#pragma once

#include "cpm_internal.h"

namespace cpm {

// One level of prefix tasks.  Task t owns the prefix columns
// prefix[t*d .. t*d+d) and the extension entries ext[ext_off[t] .. ext_off[t+1]);
// entry e is the candidate prefix(t) + {ext[e]}.  The entries of a task are
// split into groups of G; grp_off[t] is the first group of task t and
// grp_off[n_tasks] the number of groups.  Columns go through fcol[] when it
// is non-null (frequent-item index -> store column).  All pointers are device
// pointers; n_tasks is a host value.
struct TaskView {
    const uint32_t *prefix = nullptr;
    uint32_t d = 0;
    const uint32_t *ext_off = nullptr;
    const uint32_t *ext = nullptr;
    const uint32_t *grp_off = nullptr;
    uint32_t n_tasks = 0;
    const uint32_t *fcol = nullptr;
    uint32_t G = 1;
};

// Partial periodicity over one tile of words: the occurrences' first and last
// timestamps, the largest gap between consecutive occurrences, their count.
struct PeriodPart {
    int64_t first, last, gap;
    uint32_t count, pad;
};

// sup[e] += |AND of the prefix rows and row ext[e]| for every entry of the
// level (sup zeroed by the caller).  n_ext bounds the number of entries.
int launch_support_count(cpm_ctx *ctx, const cpm_bits *bits, const TaskView &tv, uint64_t n_ext,
                         uint32_t *sup);

// Support and periodicity of every entry: per-tile partials into parts
// (n_ext x tiles), then an in-order fold into sup[e] and per[e].
// When accept != nullptr, accept[e] = (sup >= minsup && sup > 0 && per <= maxper)
// is also written (minsup/maxper passed through the last arguments).
int launch_support_periodic(cpm_ctx *ctx, const cpm_bits *bits, const TaskView &tv,
                            uint64_t n_ext, const int64_t *ts, int64_t end, PeriodPart *parts,
                            uint32_t *sup, int64_t *per, uint32_t *unused, int flags);

// child_j = prefix & ext_j (rows of `stride` words), sup[j] = popcount.
// out may be null (count only).
int launch_intersect(cpm_ctx *ctx, const uint64_t *prefix, const uint64_t *ext, uint32_t k,
                     uint64_t stride, unsigned long long *sup, uint64_t *out);

inline uint32_t tiles_per_row(uint64_t stride) {
    return (uint32_t)((stride + kTileWords - 1) / kTileWords);
}

}  // namespace cpm

// Device-wide scan and radix sort used by the FP-tree build (sort.hip).
#pragma once

#include "cpm_internal.h"

namespace cpm {

struct SortWs {
    DBuf keys_tmp, vals_tmp, hist, offs, part, total;
    int ensure(uint64_t n);
};

// out[i] = sum(in[0..i)), out[n] = total (out has n+1 entries).  When d_total
// is non-null it also receives the 64-bit total.
int exclusive_scan_u32(cpm_ctx *ctx, SortWs &ws, const uint32_t *in, uint32_t *out, uint64_t n,
                       uint64_t *d_total);
int exclusive_scan_u32_to_u64(cpm_ctx *ctx, SortWs &ws, const uint32_t *in, uint64_t *out,
                              uint64_t n, uint64_t *d_total);

// Stable sort of (keys[i], vals[i]) by bits [0, end_bit) of the key, in place.
int radix_sort_pairs(cpm_ctx *ctx, SortWs &ws, uint64_t *keys, uint32_t *vals, uint64_t n,
                     int end_bit);

// number of bits needed to hold max_value
int bits_for(uint64_t max_value);

}  // namespace cpm

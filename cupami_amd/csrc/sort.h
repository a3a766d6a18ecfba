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

// Sorts P (a power of two) entries of (sk, si) in LDS by (sk, si) ascending,
// the whole workgroup taking part; callers set si to the input index, which
// makes the order stable.  Ends with a barrier.
__device__ __forceinline__ void bitonic_sort_lds(uint64_t *sk, uint32_t *si, uint32_t P) {
    for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint64_t a = sk[i], b = sk[l];
                    const uint32_t ia = si[i], ib = si[l];
                    const bool gt = a > b || (a == b && ia > ib);
                    if (gt == ((i & k) == 0)) {
                        sk[i] = b;
                        sk[l] = a;
                        si[i] = ib;
                        si[l] = ia;
                    }
                }
            }
            __syncthreads();
        }
}

// Exclusive scan over a workgroup of up to 1024 lanes (16 waves): returns the
// lane's offset, *total the sum.  lds holds 16 entries.  Ends with a barrier.
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t *lds, uint32_t *total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
        if (w < wave) base += lds[w];
        all += lds[w];
    }
    *total = all;
    __syncthreads();
    return base + x - v;
}

}  // namespace cpm

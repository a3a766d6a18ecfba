// Level-2 pair supports as a dense contraction on the INT8 matrix cores.
//
// With D the N x F 0/1 matrix of the F item rows, the pair supports are
// G = D^T D: G[a][b] = |row_a & row_b|.  Each 64-bit word of a row is 64
// consecutive K elements.  A workgroup owns a 128x128 block of G (upper
// block triangle only) and one slice of the K (transaction) axis; it stages
// 16-word tiles of its 128 + 128 rows in LDS and each wave accumulates a
// 64x64 quadrant with v_mfma_i32_16x16x64_i8 (4 x 4 tiles).  The bits become int8 0/1 in
// registers: lane group g (lane / 16) takes 16 bits of each word and spreads
// them into 16 bytes with shift-and-mask; A (row items) and B (column items)
// use the same lane -> bit map, so every bit is paired with itself exactly
// once whatever order the hardware gives the K slots.  Accumulation is int32:
// exact while N < 2^31.  Slices are summed with integer atomics.
#include "cpm_internal.h"

#include <algorithm>

namespace cpm {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kBT = 128;  // block tile (items)
constexpr int kKW = 16;   // words per K tile

// 16 bits of `w` -> 16 int8 0/1 values for lane group g (0..3).
__device__ __forceinline__ v4i spread_bits(uint64_t w, uint32_t g) {
    const uint32_t half = (g & 2) ? (uint32_t)(w >> 32) : (uint32_t)w;
    const uint32_t q = (g & 1) * 4;
    v4i r;
    r.x = (int)((half >> (q + 0)) & 0x01010101u);
    r.y = (int)((half >> (q + 1)) & 0x01010101u);
    r.z = (int)((half >> (q + 2)) & 0x01010101u);
    r.w = (int)((half >> (q + 3)) & 0x01010101u);
    return r;
}

// One workgroup: a 128x128 block of G (bi <= bj) over one K slice.  Each
// wave owns a 64x64 quadrant as 4x4 MFMA tiles; the next K tile is loaded
// into registers while the current one is multiplied out of LDS.
__global__ __launch_bounds__(256) void k_gram_i8(const uint64_t *__restrict__ store,
                                                uint64_t stride, uint64_t words,
                                                const uint32_t *__restrict__ cols, uint32_t n,
                                                uint32_t nb, uint32_t slices,
                                                uint64_t slice_words, int32_t *__restrict__ out,
                                                uint32_t ld) {
    __shared__ uint64_t sA[kBT][kKW + 1];
    __shared__ uint64_t sB[kBT][kKW + 1];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t slice = blockIdx.x % slices;
    uint32_t tri = blockIdx.x / slices;
    uint32_t bi = 0;
    while (tri >= nb - bi) {
        tri -= nb - bi;
        ++bi;
    }
    const uint32_t bj = bi + tri;
    const uint64_t k_begin = (uint64_t)slice * slice_words;
    const uint64_t k_end = std::min<uint64_t>(words, k_begin + slice_words);

    // loader: lane pair (2r, 2r+1) stages row r, 8 words each
    const uint32_t lrow = tid >> 1, lw = (tid & 1) * 8;
    const uint32_t ia = bi * kBT + lrow, ib = bj * kBT + lrow;
    const uint64_t *pa = ia < n ? store + (uint64_t)cols[ia] * stride : nullptr;
    const uint64_t *pb = ib < n ? store + (uint64_t)cols[ib] * stride : nullptr;
    uint4 ra[4], rb[4];
    auto fetch = [&](uint64_t k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint64_t w = k0 + lw + 2 * q;  // even; w + 1 < stride (zero padded)
            const bool in = w < k_end;
            ra[q] = (pa && in) ? *reinterpret_cast<const uint4 *>(pa + w) : make_uint4(0, 0, 0, 0);
            rb[q] = (pb && in) ? *reinterpret_cast<const uint4 *>(pb + w) : make_uint4(0, 0, 0, 0);
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            sA[lrow][lw + 2 * q] = (uint64_t)ra[q].x | ((uint64_t)ra[q].y << 32);
            sA[lrow][lw + 2 * q + 1] = (uint64_t)ra[q].z | ((uint64_t)ra[q].w << 32);
            sB[lrow][lw + 2 * q] = (uint64_t)rb[q].x | ((uint64_t)rb[q].y << 32);
            sB[lrow][lw + 2 * q + 1] = (uint64_t)rb[q].z | ((uint64_t)rb[q].w << 32);
        }
    };

    const uint32_t wr = wave >> 1, wc = wave & 1;
    const uint32_t li = lane & 15, lg = lane >> 4;
    // on a diagonal block the lower-left quadrant lies below the diagonal
    const bool idle = (bi == bj) && (wr > wc);
    v4i acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (v4i){0, 0, 0, 0};

    if (k_begin < k_end) fetch(k_begin);
    for (uint64_t k0 = k_begin; k0 < k_end; k0 += kKW) {
        stash();
        __syncthreads();
        if (k0 + kKW < k_end) fetch(k0 + kKW);
        if (!idle) {
#pragma unroll 2
            for (int w = 0; w < kKW; ++w) {
                v4i fa[4], fb[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    fa[t] = spread_bits(sA[wr * 64 + t * 16 + li][w], lg);
                    fb[t] = spread_bits(sB[wc * 64 + t * 16 + li][w], lg);
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[a], fb[b], acc[a][b], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (idle) return;
    // D tile layout: lane holds rows 4*(lane/16) + r, column lane % 16
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t col = bj * kBT + wc * 64 + b * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t row = bi * kBT + wr * 64 + a * 16 + lg * 4 + r;
                const int v = acc[a][b][r];
                if (row < n && col < n && row <= col && v) atomicAdd(out + (uint64_t)row * ld + col, v);
            }
        }
}

}  // namespace

int launch_gram_i8(cpm_ctx *ctx, const cpm_bits *bits, const uint32_t *d_cols, uint32_t n,
                   int32_t *d_out, uint32_t ld) {
    if (n == 0) return CPM_OK;
    if (bits->n_transactions >= (1ull << 31))
        return fail(CPM_ERR_INVALID, "int32 pair counts need fewer than 2^31 transactions");
    CPM_HIP(hipMemsetAsync(d_out, 0, (size_t)n * ld * sizeof(int32_t), ctx->stream));
    if (bits->words == 0) return CPM_OK;
    const uint32_t nb = (n + kBT - 1) / kBT;
    const uint64_t tri = (uint64_t)nb * (nb + 1) / 2;
    const uint64_t want = (uint64_t)ctx->cu_count * 2;
    uint64_t slices = std::max<uint64_t>(1, (want + tri - 1) / tri);
    const uint64_t ktiles = (bits->words + kKW - 1) / kKW;
    slices = std::min<uint64_t>(slices, ktiles);
    const uint64_t slice_words = (ktiles + slices - 1) / slices * kKW;
    slices = (bits->words + slice_words - 1) / slice_words;
    const uint64_t grid = tri * slices;
    if (grid >= (1ull << 31)) return fail(CPM_ERR_INVALID, "pair block grid too large");
    Timed timer(ctx, 5);
    hipLaunchKernelGGL(k_gram_i8, dim3((unsigned)grid), dim3(256), 0, ctx->stream,
                       bits->store.as<const uint64_t>(), bits->stride, bits->words, d_cols, n, nb,
                       (uint32_t)slices, slice_words, d_out, ld);
    CPM_HIP(hipGetLastError());
    CPM_CHECK(timer.finish());
    ctx->stats[4] += 1;
    // 2 ops per multiply-add over the upper triangle including the diagonal
    ctx->stats[6] += (double)bits->n_transactions * (double)n * (double)(n + 1);
    return CPM_OK;
}

}  // namespace cpm

extern "C" int cpm_pair_gram_i8(cpm_ctx *ctx, const cpm_bits *bits, const uint32_t *cols,
                                uint32_t n, uint32_t *gram) {
    using namespace cpm;
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!bits) return fail(CPM_ERR_INVALID, "null store");
    if (n == 0) return CPM_OK;
    if (!cols || !gram) return fail(CPM_ERR_INVALID, "null argument");
    if (n > 16384) return fail(CPM_ERR_INVALID, "at most 16384 items per pair matrix");
    for (uint32_t i = 0; i < n; ++i)
        if (cols[i] >= bits->n_items)
            return fail(CPM_ERR_INVALID, "column %u >= %u items", cols[i], bits->n_items);
    DBuf dc, dout;
    CPM_CHECK(dc.ensure((size_t)n * 4));
    CPM_CHECK(dout.ensure((size_t)n * n * 4));
    CPM_HIP(hipMemcpyAsync(dc.p, cols, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    CPM_CHECK(launch_gram_i8(ctx, bits, dc.as<uint32_t>(), n, dout.as<int32_t>(), n));
    CPM_HIP(hipMemcpyAsync(gram, dout.p, (size_t)n * n * 4, hipMemcpyDeviceToHost, ctx->stream));
    CPM_HIP(hipStreamSynchronize(ctx->stream));
    return CPM_OK;
}

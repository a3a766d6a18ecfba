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
// exact while N < 2^31.  Slices are summed with integer atomics.  One launch
// serves any number of databases: the job table lists each workgroup's
// (database, block row, block column, K range).
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

// One workgroup: a 128x128 block of G (bi <= bj) over one K slice of one
// database, as listed in the job table.  Each wave owns a 64x64 quadrant as
// 4x4 MFMA tiles; the next K tile is loaded into registers while the current
// one is multiplied out of LDS.
__global__ __launch_bounds__(256) void k_gram_i8(const GramJob *__restrict__ jobs,
                                                const GramDb *__restrict__ dbs,
                                                int32_t *__restrict__ out) {
    __shared__ uint64_t sA[kBT][kKW + 1];
    __shared__ uint64_t sB[kBT][kKW + 1];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const GramJob job = jobs[blockIdx.x];
    const GramDb db = dbs[job.db];
    const gptr<uint64_t> store = to_global(db.store);
    const uint64_t stride = db.stride;
    const uint32_t n = db.n_items;
    const gptr<uint32_t> cols = to_global(db.cols);
    const uint32_t bi = job.bi, bj = job.bj;
    const uint64_t k_begin = job.k_begin, k_end = job.k_end;
    int32_t *res = out + db.out_off;
    const uint32_t ld = n;

    // loader: lane pair (2r, 2r+1) stages row r, 8 words each
    const uint32_t lrow = tid >> 1, lw = (tid & 1) * 8;
    const uint32_t ia = bi * kBT + lrow, ib = bj * kBT + lrow;
    const gptr<uint4> pa = ia < n ? (gptr<uint4>)(store + (uint64_t)cols[ia] * stride) : nullptr;
    const gptr<uint4> pb = ib < n ? (gptr<uint4>)(store + (uint64_t)cols[ib] * stride) : nullptr;
    uint4 ra[4], rb[4];
    auto fetch = [&](uint64_t k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint64_t w = k0 + lw + 2 * q;  // even; w + 1 < stride (zero padded)
            const bool in = w < k_end;
            ra[q] = make_uint4(0, 0, 0, 0);
            rb[q] = make_uint4(0, 0, 0, 0);
            if (pa && in) ra[q] = pa[w / 2];
            if (pb && in) rb[q] = pb[w / 2];
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
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = (v4i){0, 0, 0, 0};

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
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y)
                        acc[x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[x], fb[y], acc[x][y], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (idle) return;
    // D tile layout: lane holds rows 4*(lane/16) + r, column lane % 16
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const uint32_t col = bj * kBT + wc * 64 + y * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t row = bi * kBT + wr * 64 + x * 16 + lg * 4 + r;
                const int v = acc[x][y][r];
                if (row < n && col < n && row <= col && v) atomicAdd(res + (uint64_t)row * ld + col, v);
            }
        }
}

}  // namespace

int gram_plan(const cpm_ctx *ctx, const std::vector<GramDb> &dbs, std::vector<GramJob> &jobs,
              uint64_t *out_elems) {
    jobs.clear();
    uint64_t tri_total = 0, out = 0;
    for (const GramDb &d : dbs) {
        const uint64_t nb = (d.n_items + kBT - 1) / kBT;
        tri_total += nb * (nb + 1) / 2;
        out += (uint64_t)d.n_items * d.n_items;
    }
    *out_elems = out;
    // enough workgroups to fill the device: split K when there are few blocks
    const uint64_t want = (uint64_t)ctx->cu_count * 2;
    for (uint32_t i = 0; i < dbs.size(); ++i) {
        const GramDb &d = dbs[i];
        if (d.n_items == 0 || d.words == 0) continue;
        if (d.n_transactions >= (1ull << 31))
            return fail(CPM_ERR_INVALID, "int32 pair counts need fewer than 2^31 transactions");
        const uint32_t nb = (d.n_items + kBT - 1) / kBT;
        const uint64_t ktiles = (d.words + kKW - 1) / kKW;
        uint64_t slices = tri_total ? std::max<uint64_t>(1, (want + tri_total - 1) / tri_total) : 1;
        slices = std::min<uint64_t>(slices, ktiles);
        const uint64_t slice_words = (ktiles + slices - 1) / slices * kKW;
        for (uint32_t bi = 0; bi < nb; ++bi)
            for (uint32_t bj = bi; bj < nb; ++bj)
                for (uint64_t k0 = 0; k0 < d.words; k0 += slice_words)
                    jobs.push_back({i, bi, bj, 0, k0, std::min<uint64_t>(d.words, k0 + slice_words)});
    }
    if (jobs.size() >= (1ull << 31)) return fail(CPM_ERR_INVALID, "pair block grid too large");
    return CPM_OK;
}

int launch_gram_i8(cpm_ctx *ctx, const GramJob *d_jobs, uint64_t n_jobs, const GramDb *d_dbs,
                   int32_t *d_out, uint64_t out_elems, uint64_t n_ops) {
    CPM_HIP(hipMemsetAsync(d_out, 0, out_elems * sizeof(int32_t), ctx->stream));
    if (n_jobs == 0) return CPM_OK;
    Timed timer(ctx, 5);
    hipLaunchKernelGGL(k_gram_i8, dim3((unsigned)n_jobs), dim3(256), 0, ctx->stream, d_jobs, d_dbs,
                       d_out);
    CPM_HIP(hipGetLastError());
    CPM_CHECK(timer.finish());
    ctx->stats[4] += 1;
    ctx->stats[6] += (double)n_ops;
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
    DBuf dc, dout, djobs, ddb;
    CPM_CHECK(dc.ensure((size_t)n * 4));
    CPM_HIP(hipMemcpyAsync(dc.p, cols, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    std::vector<GramDb> dbs(1);
    dbs[0].store = bits->store.as<const uint64_t>();
    dbs[0].stride = bits->stride;
    dbs[0].words = bits->words;
    dbs[0].n_transactions = bits->n_transactions;
    dbs[0].cols = dc.as<uint32_t>();
    dbs[0].n_items = n;
    dbs[0].out_off = 0;
    std::vector<GramJob> jobs;
    uint64_t elems = 0;
    CPM_CHECK(gram_plan(ctx, dbs, jobs, &elems));
    CPM_CHECK(dout.ensure(elems * 4));
    CPM_CHECK(djobs.ensure(std::max<size_t>(jobs.size(), 1) * sizeof(GramJob)));
    CPM_CHECK(ddb.ensure(sizeof(GramDb)));
    if (!jobs.empty())
        CPM_HIP(hipMemcpyAsync(djobs.p, jobs.data(), jobs.size() * sizeof(GramJob),
                               hipMemcpyHostToDevice, ctx->stream));
    CPM_HIP(hipMemcpyAsync(ddb.p, dbs.data(), sizeof(GramDb), hipMemcpyHostToDevice, ctx->stream));
    const uint64_t ops = (uint64_t)bits->n_transactions * n * (uint64_t)(n + 1);
    CPM_CHECK(launch_gram_i8(ctx, djobs.as<GramJob>(), jobs.size(), ddb.as<GramDb>(),
                             dout.as<int32_t>(), elems, ops));
    CPM_HIP(hipMemcpyAsync(gram, dout.p, (size_t)n * n * 4, hipMemcpyDeviceToHost, ctx->stream));
    CPM_HIP(hipStreamSynchronize(ctx->stream));
    return CPM_OK;
}

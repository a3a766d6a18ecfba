// Level-2 pair supports as a dense contraction on the matrix cores.
//
// With D the N x F 0/1 matrix of the F item rows, the pair supports are
// G = D^T D: G[a][b] = |row_a & row_b|.  Each 64-bit word of a row is 64
// consecutive K elements.  A 512-lane workgroup owns a 256 x 256 block of G
// (upper block triangle only) and one slice of the K (transaction) axis.  It
// stages kKW-word tiles of its 2 x 256 rows in LDS -- eight lanes per row, so
// one load instruction reads eight whole 128-byte lines -- and each of its
// eight waves accumulates a 128 x 64 piece with
// v_mfma_scale_f32_32x32x64_f8f6f4 in its FP4 (E2M1) form, 4 x 2 tiles of
// 32 x 32.  (That form runs at twice the rate of the 16 x 16 x 128 one on
// MI355X: ~9.8 against ~5.2 POP/s in a loop of independent MFMAs.)  0 and 1
// are exact E2M1 values (1.0 is the nibble 0b0010) and every product is 0 or
// 1, so the FP32 accumulators hold exact integers while a slice spans fewer
// than 2^24 transactions; the planner cuts K so that it does.
//
// An MFMA step takes 32 K elements per lane: lanes 0-31 feed rows (columns)
// 0-31, lanes 32-63 the same rows with other K elements.  Over four steps a
// lane uses one 16-byte LDS read per 32-row tile (u32 8j + 4h .. +3 of its
// row's tile, h = lane / 32): step p takes bit plane p of those four u32 (the
// bits 4m + p, masked in place, see plane_fp4).  A (row items) and B (column items) use the
// same lane -> bit map, so every bit is paired with itself exactly once
// whatever order the hardware gives the K slots.  128 accumulator registers
// per wave leave room for two waves per SIMD, so one wave's barrier and LDS
// waits hide under the other's MFMAs.  Tiles are double-buffered in LDS: the
// next K tile is loaded into registers and stashed into the other buffer
// while the current one is multiplied, one barrier per tile.  On diagonal
// blocks the six pieces with work are spread 1.5 per SIMD (two waves share
// two of them, each taking every other K group, and their halves are added
// through LDS at the end).  A block of one slice writes its outputs; a block
// of many outputs cut into K slices stores each slice into a partial slot
// and a second launch adds the slots (plain stores and one pass instead of
// per-element atomics from every slice); a small matrix cut into many
// slices adds them with atomics.  One launch serves any number of databases:
// the job table lists each workgroup's (database, block row, block column,
// K range).
#include "cpm_internal.h"

#include <algorithm>

namespace cpm {

namespace {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kTX = 4, kTY = 2;             // 32 x 32 tiles per wave: rows, columns
constexpr int kBT = 256;                     // block tile (items)
constexpr int kWaveRows = 32 * kTX, kWaveCols = 32 * kTY;
constexpr int kWR = kBT / kWaveRows;         // waves along rows
constexpr int kWC = kBT / kWaveCols;         // waves along columns
constexpr int kThreads = 64 * kWR * kWC;     // 512: two waves per SIMD
constexpr int kKW = 16;                      // words per K tile: one 128-byte line per row
constexpr int kGroups = kKW / 4;             // 16-byte LDS reads (four MFMA steps) per K tile
constexpr int kLdsW = 2 * kKW + 4;           // u32 per LDS row (stride = 4 mod 32: 8-lane groups of
                                             // 16-byte reads cover all banks)
constexpr int kFmtFp4 = 4;                   // f8f6f4 operand format code of E2M1
static_assert(kKW % 2 == 0, "K tiles hold whole word pairs");

// Bit planes.  Four u32 of a row (128 K elements) feed four MFMAs, one per
// plane p: plane p takes bit 4m + p of every u32, left in place where that is
// a positive E2M1 value -- bit 0 of a nibble is 0.5, bit 1 is 1.0, bit 2 is
// 2.0 -- and bit 3 (the sign) shifted down to bit 2.  An E8M0 block scale per
// plane (2^1, 2^0, 2^-1, 2^-1) brings every set value to 1.0 in both
// operands, so each product is still 0 or 1 and a plane costs one AND (two
// for the last) per u32 instead of a shift and an AND per nibble position.
__device__ __forceinline__ v8i plane_fp4(uint4 q, int p) {
    constexpr uint32_t kMask[4] = {0x11111111u, 0x22222222u, 0x44444444u, 0x44444444u};
    const uint32_t m = kMask[p];
    if (p == 3) return (v8i){(int)((q.x >> 1) & m), (int)((q.y >> 1) & m), (int)((q.z >> 1) & m),
                             (int)((q.w >> 1) & m), 0, 0, 0, 0};
    return (v8i){(int)(q.x & m), (int)(q.y & m), (int)(q.z & m), (int)(q.w & m), 0, 0, 0, 0};
}
// E8M0 scales of the planes, the same byte in all four lanes of the register
__device__ __forceinline__ int plane_scale(int p) {
    return p == 0 ? 0x80808080 : p == 1 ? 0x7f7f7f7f : 0x7e7e7e7e;
}

__global__ __launch_bounds__(kThreads) void k_gram(const GramJob *__restrict__ jobs, const GramDb *__restrict__ dbs,
                                                   int32_t *__restrict__ out, int32_t *__restrict__ partials) {
    __shared__ uint32_t sA[2][kBT][kLdsW];  // two buffers of A and B tiles: 144 KB
    __shared__ uint32_t sB[2][kBT][kLdsW];
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar loop control
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

    // loader: lanes 4r .. 4r+3 of a wave stage the 64 bytes (kKW words) of
    // one row, so one load instruction touches 16 rows; wave w's pass q
    // covers rows 16 (4w + q) .. +15 of the 512 (A then B).  Loads are
    // unconditional: a block row past the last item reads item n - 1's row,
    // which only reaches outputs past the last item (never stored), and a
    // tile reaches past k_end only in the last slice, where the words past
    // the row's data are the store's zero padding (slices start and end on
    // kKW-word tiles, and rows are padded to a multiple of 64 words).
    constexpr int kLanesPerRow = kKW / 2;            // uint4 per row per tile
    constexpr int kRowsPerPass = 64 / kLanesPerRow;  // rows per load instruction
    constexpr int kPasses = 2 * kBT / (kRowsPerPass * (kThreads / 64));
    static_assert(kPasses * kRowsPerPass * (kThreads / 64) == 2 * kBT, "passes cover A and B");
    const uint32_t part = lane % kLanesPerRow;
    uint32_t col[kPasses];
    uint32_t *srow[kPasses];
#pragma unroll
    for (int q = 0; q < kPasses; ++q) {  // all column loads in flight together
        const uint32_t R = (wave * kPasses + q) * kRowsPerPass + lane / kLanesPerRow;  // 0 .. 511
        const bool isB = R >= (uint32_t)kBT;
        const uint32_t lr = R % kBT;
        const uint32_t item = (isB ? bj : bi) * kBT + lr;
        col[q] = cols[min(item, n - 1)];
        srow[q] = (isB ? &sB[0][lr][0] : &sA[0][lr][0]) + 4 * part;
    }
    gptr<uint4> pr[kPasses];
#pragma unroll
    for (int q = 0; q < kPasses; ++q) pr[q] = (gptr<uint4>)(store + (uint64_t)col[q] * stride);
    uint4 rr[kPasses] = {};
    auto fetch = [&](uint64_t k0) {
        const uint64_t w = k0 + 2 * part;  // even; w + 1 < stride
#pragma unroll
        for (int q = 0; q < kPasses; ++q) {
            rr[q] = make_uint4(0, 0, 0, 0);
            if (w < k_end) rr[q] = pr[q][w / 2];
        }
    };
    constexpr uint32_t kBufU32 = kBT * kLdsW;  // u32 per LDS buffer of A (or B)
    auto stash = [&](int buf) {
#pragma unroll
        for (int q = 0; q < kPasses; ++q) *reinterpret_cast<uint4 *>(srow[q] + buf * kBufU32) = rr[q];
    };

    // Wave -> (piece of the block, K groups).  Off the diagonal wave w takes
    // piece (w / kWC, w % kWC) over every K group.  On a diagonal block two of
    // the eight pieces lie wholly below the diagonal; the six others are
    // spread so that each SIMD (waves w and w + 4) carries 1.5 pieces: waves
    // 0-3 take (0,0) (0,1) (1,2) (1,3) whole, waves 4-7 take (0,2) (0,3) on
    // even K groups and (0,2) (0,3) on odd ones.  A shared piece's two halves
    // add up in the epilogue's atomics.
    uint32_t wr = wave / kWC, wc = wave % kWC;
    int j0 = 0, jstep = 1;
    if (bi == bj) {
        static_assert(kWR == 2 && kWC == 4, "diagonal wave map is for 2 x 4 waves");
        if (wave < 4) {
            wr = wave >> 1;
            wc = (wave & 1) + 2 * (wave >> 1);
        } else {
            wr = 0;
            wc = 2 + (wave & 1);
            j0 = (wave - 4) >> 1;
            jstep = 2;
        }
    }
    const uint32_t r32 = lane & 31, h = lane >> 5;
    v16f acc[kTX][kTY];
#pragma unroll
    for (int x = 0; x < kTX; ++x)
#pragma unroll
        for (int y = 0; y < kTY; ++y)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[x][y][i] = 0.f;

    const uint32_t *rowA0 = &sA[0][wr * kWaveRows + r32][4 * h];
    const uint32_t *rowB0 = &sB[0][wc * kWaveCols + r32][4 * h];
    auto compute = [&](int buf) {
        const uint32_t *rowA = rowA0 + buf * kBufU32, *rowB = rowB0 + buf * kBufU32;
#pragma unroll 1
        for (int j = j0; j < kGroups; j += jstep) {
            uint4 qa[kTX], qb[kTY];
#pragma unroll
            for (int t = 0; t < kTX; ++t) qa[t] = *reinterpret_cast<const uint4 *>(rowA + t * 32 * kLdsW + 8 * j);
#pragma unroll
            for (int t = 0; t < kTY; ++t) qb[t] = *reinterpret_cast<const uint4 *>(rowB + t * 32 * kLdsW + 8 * j);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int sc = plane_scale(c);
                v8i fb[kTY];
#pragma unroll
                for (int y = 0; y < kTY; ++y) fb[y] = plane_fp4(qb[y], c);
#pragma unroll
                for (int x = 0; x < kTX; ++x) {
                    const v8i fa = plane_fp4(qa[x], c);
#pragma unroll
                    for (int y = 0; y < kTY; ++y)
                        acc[x][y] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                            fa, fb[y], acc[x][y], kFmtFp4, kFmtFp4, 0, sc, 0, sc);
                }
            }
        }
    };
    // double-buffered tiles: tile t+1 is loaded and stashed into the other
    // buffer while tile t is multiplied; one barrier per tile
    if (k_begin < k_end) {
        fetch(k_begin);
        stash(0);
    }
    __syncthreads();
    int cur = 0;
    for (uint64_t k0 = k_begin; k0 < k_end; k0 += kKW) {
        const bool more = k0 + kKW < k_end;
        if (more) fetch(k0 + kKW);
        compute(cur);
        if (more) stash(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }
    // On a diagonal block waves 6 and 7 hold the odd K groups of the pieces
    // waves 4 and 5 hold the even ones: those halves are added through LDS
    // (the tiles are free once the K loop's last barrier is passed), so one
    // wave stores each piece.
    const bool diag = bi == bj;
    if (diag) {
        float *xch = reinterpret_cast<float *>(&sA[0][0][0]);  // 2 waves x 128 x 64 sums: 64 KB
        static_assert(2 * kTX * kTY * 16 * 64 * sizeof(float) <= sizeof(sA), "exchange fits the A tiles");
        auto at = [&](uint32_t w, int x, int y, int i) { return ((w * kTX * kTY + x * kTY + y) * 16 + i) * 64 + lane; };
        if (wave >= 6)
#pragma unroll
            for (int x = 0; x < kTX; ++x)
#pragma unroll
                for (int y = 0; y < kTY; ++y)
#pragma unroll
                    for (int i = 0; i < 16; ++i) xch[at(wave - 6, x, y, i)] = acc[x][y][i];
        __syncthreads();
        if (wave == 4 || wave == 5)
#pragma unroll
            for (int x = 0; x < kTX; ++x)
#pragma unroll
                for (int y = 0; y < kTY; ++y)
#pragma unroll
                    for (int i = 0; i < 16; ++i) acc[x][y][i] += xch[at(wave - 4, x, y, i)];
    }
    if (diag && wave >= 6) return;
    // D tile layout of the 32 x 32 form: element i of lane l is row
    // 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32.  A block of one K slice
    // writes its upper-triangle outputs and a slice of a block adding with
    // atomics adds them; a slice of a block with partial slots stores the
    // whole piece into its slot (coalesced rows), for the reduce.
    int32_t *const slot = job.slot < kGramAtomic ? partials + (uint64_t)job.slot * kBT * kBT : nullptr;
    const bool atomic = job.slot == kGramAtomic;
#pragma unroll
    for (int x = 0; x < kTX; ++x)
#pragma unroll
        for (int y = 0; y < kTY; ++y) {
            const uint32_t cl = wc * kWaveCols + y * 32 + r32;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const uint32_t rl = wr * kWaveRows + x * 32 + 8 * (i / 4) + 4 * h + (i % 4);
                const int v = (int)acc[x][y][i];
                if (slot) {
                    slot[rl * kBT + cl] = v;
                } else {
                    const uint32_t row = bi * kBT + rl, col = bj * kBT + cl;
                    if (row < n && col < n && row <= col) {
                        if (!atomic) res[(uint64_t)row * ld + col] = v;
                        else if (v) atomicAdd(res + (uint64_t)row * ld + col, v);
                    }
                }
            }
        }
}

// The upper-triangle outputs of the blocks cut into K slices: the sum of
// their slots.
__global__ __launch_bounds__(256) void k_gram_reduce(const GramBlock *__restrict__ blocks, uint64_t n_blocks,
                                                     const GramDb *__restrict__ dbs,
                                                     const int32_t *__restrict__ partials, int32_t *__restrict__ out) {
    constexpr uint64_t kE = (uint64_t)kBT * kBT;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_blocks * kE;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const GramBlock b = blocks[i / kE];
        const uint32_t e = (uint32_t)(i % kE), rl = e / kBT, cl = e % kBT;
        const uint32_t n = dbs[b.db].n_items;
        const uint32_t row = b.bi * kBT + rl, col = b.bj * kBT + cl;
        if (row >= n || col >= n || row > col) continue;
        const int32_t *p = partials + (uint64_t)b.slot0 * kE + e;
        int32_t sum = 0;
        for (uint32_t k = 0; k < b.nslots; ++k) sum += p[k * kE];
        out[dbs[b.db].out_off + (uint64_t)row * n + col] = sum;
    }
}

}  // namespace

// transactions per K slice: every partial sum is at most this, below 2^24
constexpr uint64_t kMaxSliceBits = 1ull << 23;

// waves of a diagonal block with work (the others lie below the diagonal)
constexpr int diagonal_active_waves() {
    int a = 0;
    for (int wr = 0; wr < kWR; ++wr)
        for (int wc = 0; wc < kWC; ++wc) a += !(wr * kWaveRows > wc * kWaveCols + kWaveCols - 1);
    return a;
}


int gram_plan(const cpm_ctx *ctx, const std::vector<GramDb> &dbs, GramPlan &plan) {
    std::vector<GramJob> &jobs = plan.jobs;
    jobs.clear();
    plan.blocks.clear();
    plan.slots = 0;
    // work of a block: its K tiles x the share of its waves with work
    // (a diagonal block's six pieces are spread 1.5 per SIMD, so it takes
    // 6/8 of a full block's time)
    const double diag = (double)diagonal_active_waves() / (kWR * kWC);
    double total = 0;
    uint64_t out = 0;
    for (const GramDb &d : dbs) {
        out += (uint64_t)d.n_items * d.n_items;
        if (d.n_items == 0 || d.words == 0) continue;
        const uint64_t nb = (d.n_items + kBT - 1) / kBT;
        const double kt = (double)((d.words + kKW - 1) / kKW);
        total += kt * ((double)nb * (nb - 1) / 2 + diag * nb);
    }
    plan.out_elems = out;
    // One workgroup per CU at a time (the accumulators take half the
    // register file): cut each block's K range in proportion to its work so
    // that the jobs fill one round of CUs evenly.
    const double per_job = total / (double)std::max(1, ctx->cu_count);
    for (uint32_t i = 0; i < dbs.size(); ++i) {
        const GramDb &d = dbs[i];
        if (d.n_items == 0 || d.words == 0) continue;
        if (d.n_transactions >= (1ull << 31))
            return fail(CPM_ERR_INVALID, "int32 pair counts need fewer than 2^31 transactions");
        const uint32_t nb = (d.n_items + kBT - 1) / kBT;
        const uint64_t ktiles = (d.words + kKW - 1) / kKW;
        // FP32 sums stay exact below 2^24 transactions per slice
        const uint64_t exact = (ktiles * kKW * 64 + kMaxSliceBits - 1) / kMaxSliceBits;
        for (uint32_t bi = 0; bi < nb; ++bi)
            for (uint32_t bj = bi; bj < nb; ++bj) {
                const double work = (double)ktiles * (bi == bj ? diag : 1.0);
                uint64_t slices = per_job > 0 ? (uint64_t)(work / per_job) : 1;
                slices = std::max<uint64_t>(std::min<uint64_t>(std::max<uint64_t>(slices, 1), ktiles), exact);
                const uint64_t slice_words = (ktiles + slices - 1) / slices * kKW;
                const uint64_t n_slices = (d.words + slice_words - 1) / slice_words;
                const uint64_t slot0 = plan.slots;
                // measured: slots win at 1000 items in ~28 slices, atomics at
                // ~120 items in 256 slices
                const bool slots = n_slices > 1 && d.n_items >= 2 * kBT && n_slices <= 64;
                if (slots) {
                    plan.blocks.push_back({i, bi, bj, (uint32_t)slot0, (uint32_t)n_slices, 0});
                    plan.slots += n_slices;
                }
                uint64_t s = 0;
                for (uint64_t k0 = 0; k0 < d.words; k0 += slice_words, ++s)
                    jobs.push_back({i, bi, bj, slots ? (uint32_t)(slot0 + s) : n_slices > 1 ? kGramAtomic : kGramDirect,
                                    k0, std::min<uint64_t>(d.words, k0 + slice_words)});
            }
    }
    if (jobs.size() >= (1ull << 31) || plan.slots >= kGramAtomic)
        return fail(CPM_ERR_INVALID, "pair block grid too large");
    return CPM_OK;
}

int launch_gram(cpm_ctx *ctx, const GramDev &g, const GramDb *d_dbs, int32_t *d_out, uint64_t out_elems,
                uint64_t n_ops, bool clear) {
    if (clear) CPM_HIP(hipMemsetAsync(d_out, 0, out_elems * sizeof(int32_t), ctx->stream));
    if (g.n_jobs == 0) return CPM_OK;
    if (g.n_blocks && !g.partials) return fail(CPM_ERR_INVALID, "no partial slots for sliced pair blocks");
    Timed timer(ctx, 5);  // both launches
    hipLaunchKernelGGL(k_gram, dim3((unsigned)g.n_jobs), dim3(kThreads), 0, ctx->stream, g.jobs, d_dbs,
                       d_out, g.partials);
    CPM_HIP(hipGetLastError());
    if (g.n_blocks) {
        const uint64_t items = g.n_blocks * (uint64_t)kBT * kBT;
        hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)std::min<uint64_t>((items + 255) / 256, 8192)), dim3(256),
                           0, ctx->stream, g.blocks, g.n_blocks, d_dbs, g.partials, d_out);
        CPM_HIP(hipGetLastError());
    }
    CPM_CHECK(timer.finish());
    ctx->stats[4] += 1;
    ctx->stats[6] += (double)n_ops;
    return CPM_OK;
}

}  // namespace cpm

extern "C" int cpm_pair_gram(cpm_ctx *ctx, const cpm_bits *bits, const uint32_t *cols, uint32_t n,
                             uint32_t *gram) {
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
    dbs[0].store = bits->rows();
    dbs[0].stride = bits->stride;
    dbs[0].words = bits->words;
    dbs[0].n_transactions = bits->n_transactions;
    dbs[0].cols = dc.as<uint32_t>();
    dbs[0].n_items = n;
    dbs[0].out_off = 0;
    GramPlan plan;
    CPM_CHECK(gram_plan(ctx, dbs, plan));
    const uint64_t elems = plan.out_elems;
    DBuf dblocks, dparts;
    CPM_CHECK(dout.ensure(elems * 4));
    CPM_CHECK(djobs.ensure(std::max<size_t>(plan.jobs.size(), 1) * sizeof(GramJob)));
    CPM_CHECK(dblocks.ensure(std::max<size_t>(plan.blocks.size(), 1) * sizeof(GramBlock)));
    CPM_CHECK(dparts.ensure(std::max<uint64_t>(plan.slots, 1) * kBT * kBT * 4));
    CPM_CHECK(ddb.ensure(sizeof(GramDb)));
    if (!plan.jobs.empty())
        CPM_HIP(hipMemcpyAsync(djobs.p, plan.jobs.data(), plan.jobs.size() * sizeof(GramJob),
                               hipMemcpyHostToDevice, ctx->stream));
    if (!plan.blocks.empty())
        CPM_HIP(hipMemcpyAsync(dblocks.p, plan.blocks.data(), plan.blocks.size() * sizeof(GramBlock),
                               hipMemcpyHostToDevice, ctx->stream));
    CPM_HIP(hipMemcpyAsync(ddb.p, dbs.data(), sizeof(GramDb), hipMemcpyHostToDevice, ctx->stream));
    const uint64_t ops = (uint64_t)bits->n_transactions * n * (uint64_t)(n + 1);
    GramDev g;
    g.jobs = djobs.as<GramJob>();
    g.n_jobs = plan.jobs.size();
    g.blocks = dblocks.as<GramBlock>();
    g.n_blocks = plan.blocks.size();
    g.partials = dparts.as<int32_t>();
    CPM_CHECK(launch_gram(ctx, g, ddb.as<GramDb>(), dout.as<int32_t>(), elems, ops, true));
    CPM_CHECK(copy_to_host(ctx, gram, dout.p, (size_t)n * n * 4));
    return CPM_OK;
}

// Store build straight from a basket file's arrays (CUPAMI01 / CUPAMIT1):
// offsets u64[n+1] and item ids u32[m], as the file holds them.  Replaces the
// host work of the reference's BIN path (cuApriori.py:273-276, 301-305: the
// array checks; cuApriori.py:381-396: unique ids, labels sorted as strings,
// row and column of every entry) and the COO it hands to build_bitsets.
//
//   1. k_ingest_scan_input  checks the offsets and ids, marks each kept
//                           basket, and sets bit x of a presence bitmap for
//                           every id x that occurs (LDS first when the
//                           bitmap is small)
//   2. k_ingest_totals      kept baskets summed; presence popcounts scanned
//                           -> pre[w] (ids below word w) in one workgroup for
//                           bitmaps up to 4096 words (a device-wide scan above)
//   3. host                 reads the two flags and two totals (one copy):
//                           the only synchronization
//   4. k_ingest_labels      per present id: its rank in decimal-string order
//                           (the column) from range counts over the bitmap;
//                           labels[column] = id, and a direct id -> column
//                           map for universes up to 2^26
//   5. k_ingest_rows        row starts of the kept baskets, after a scan of
//                           the kept flags (only when a basket is left out)
//   6. k_ingest_tile        one workgroup per B words of every column: its
//                           baskets' ids are read coalesced, each entry's row
//                           found by a binary search over the rows' starts in
//                           LDS, its column from the direct map (in LDS for
//                           universes up to 4096), its bit ORed into an LDS
//                           tile written once
//      (k_ingest_scatter: global atomics when the columns exceed the LDS tile)

#include "cpm_internal.h"
#include "sort.h"

#include <algorithm>

namespace cpm {
namespace {

constexpr uint32_t kLdsPresenceWords = 2048;  // ids below 131072 are marked in LDS first
constexpr uint32_t kRowsPerTileMax = 64 * 8;  // rows of one tile workgroup (B <= 8 words)
constexpr uint64_t kDirectMapMax = 1ull << 26;  // ids with a direct id -> column map (256 MB)
constexpr uint32_t kLdsMapMax = 4096;           // ... copied into LDS by the tile build
constexpr uint32_t kTileThreads = 512;          // lanes of a tile-build workgroup
constexpr size_t kKeepBytes = (size_t)64 << 20;  // larger workspace buffers are freed after a build

struct IngestWs {
    DBuf state;   // u64 totals[2] (present ids, store rows), u32 flags[2]
    DBuf pres;    // presence bitmap, (universe >> 6) + 1 words
    DBuf cnt;     // popcount of each presence word
    DBuf pre;     // ids present below each word (+ the total)
    DBuf keep;    // u32 per basket: 1 if it becomes a store row
    DBuf act;     // u64 per basket (+1): its store row (when some basket is left out)
    DBuf kept;    // kept baskets counted by each workgroup of the input scan
    DBuf rows;    // u64 row starts of the store rows (+1)
    DBuf col;     // column of the u-th present id in numeric order
    DBuf direct;  // column of each id (universes up to kDirectMapMax)
    SortWs sort;
    HBuf host;
};

void free_ingest_ws(void *p) { delete static_cast<IngestWs *>(p); }

IngestWs &ingest_ws(cpm_ctx *ctx) {
    if (!ctx->ingest_ws) {
        ctx->ingest_ws = new IngestWs();
        ctx->ingest_ws_free = free_ingest_ws;
    }
    return *static_cast<IngestWs *>(ctx->ingest_ws);
}

// Offsets non-decreasing from 0 to m (flags[0] otherwise), ids <= universe
// (flags[1] otherwise); keep[i] = basket i becomes a store row; presence bit
// x set for every id x.  A bit is tested before the atomic that sets it, so
// frequent ids cost one load after their first occurrence.  Ids are read four
// to a 16-byte load, two loads in flight per lane, when the array is 16-byte
// aligned.
__global__ __launch_bounds__(256) void k_ingest_scan_input(const uint64_t *__restrict__ off, uint64_t n,
                                                          const uint32_t *__restrict__ ids, uint64_t m,
                                                          uint32_t universe, uint64_t pres_words,
                                                          int keep_empty, uint32_t *__restrict__ keep,
                                                          uint64_t *__restrict__ kept_part,
                                                          unsigned long long *pres, unsigned int *flags) {
    __shared__ unsigned long long lp[kLdsPresenceWords];
    __shared__ uint64_t wave_kept[4];
    const bool lds = pres_words <= kLdsPresenceWords;
    const uint32_t tid = threadIdx.x;
    if (lds)
        for (uint32_t i = tid; i < pres_words; i += blockDim.x) lp[i] = 0;
    __syncthreads();
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + tid;
    bool bad_off = g == 0 && (off[0] != 0 || off[n] != m);
    bool bad_id = false;
    uint64_t kept = 0;
    for (uint64_t i = g; i < n; i += step) {
        const uint64_t a = off[i], b = off[i + 1];
        bad_off |= a > b;
        const bool k = keep_empty || b > a;
        keep[i] = k;
        kept += k;
    }
#pragma unroll
    for (int d = 32; d; d >>= 1) kept += __shfl_down(kept, d, 64);
    if ((tid & 63) == 0) wave_kept[tid >> 6] = kept;
    auto mark = [&](uint32_t x) {
        if (x > universe) {
            bad_id = true;
            return;
        }
        const unsigned long long bit = 1ull << (x & 63);
        unsigned long long *w = lds ? &lp[x >> 6] : &pres[x >> 6];
        if (!(*w & bit)) atomicOr(w, bit);
    };
    uint64_t tail = 0;
    if ((reinterpret_cast<uintptr_t>(ids) & 15) == 0) {
        const uint4 *__restrict__ ids4 = reinterpret_cast<const uint4 *>(ids);
        const uint64_t m4 = m / 4;
        for (uint64_t k = g; k < m4; k += 2 * step) {
            const bool two = k + step < m4;
            const uint4 a = ids4[k];
            const uint4 b = two ? ids4[k + step] : make_uint4(0, 0, 0, 0);
            mark(a.x);
            mark(a.y);
            mark(a.z);
            mark(a.w);
            if (two) {
                mark(b.x);
                mark(b.y);
                mark(b.z);
                mark(b.w);
            }
        }
        tail = m4 * 4;
    }
    for (uint64_t k = tail + g; k < m; k += step) mark(ids[k]);
    if (bad_off) atomicOr(flags, 1u);
    if (bad_id) atomicOr(flags + 1, 1u);
    __syncthreads();
    if (tid == 0) kept_part[blockIdx.x] = wave_kept[0] + wave_kept[1] + wave_kept[2] + wave_kept[3];
    if (lds) {
        // only bits the global word lacks: once the common ids are in, most
        // workgroups flush nothing
        __syncthreads();
        for (uint32_t i = tid; i < pres_words; i += blockDim.x) {
            const unsigned long long miss = lp[i] & ~pres[i];
            if (miss) atomicOr(&pres[i], miss);
        }
    }
}

// One workgroup: totals[1] = the kept baskets (the input scan's per-workgroup
// counts summed) and, for a bitmap of at most kSmallScanWords words, pre[w] =
// ids present below word w, pre[pw] = totals[0] = all present ids.
constexpr uint32_t kSmallScanWords = 4096;
__global__ __launch_bounds__(1024) void k_ingest_totals(const uint64_t *__restrict__ kept_part, uint32_t parts,
                                                       const uint64_t *__restrict__ pres, uint64_t pw, int scan_pres,
                                                       uint32_t *__restrict__ pre, uint64_t *__restrict__ totals) {
    __shared__ uint64_t part[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint64_t k = 0;
    for (uint32_t i = tid; i < parts; i += blockDim.x) k += kept_part[i];
#pragma unroll
    for (int d = 32; d; d >>= 1) k += __shfl_down(k, d, 64);
    if (lane == 0) part[wave] = k;
    __syncthreads();
    if (tid == 0) {
        uint64_t t = 0;
        for (int w = 0; w < 16; ++w) t += part[w];
        totals[1] = t;
    }
    if (!scan_pres) return;
    __syncthreads();
    // four consecutive words per lane
    uint32_t c[4], sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint64_t w = (uint64_t)tid * 4 + q;
        c[q] = w < pw ? (uint32_t)__popcll(pres[w]) : 0;
        sum += c[q];
    }
    uint32_t x = sum;  // inclusive scan over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) part[wave] = x;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t w = 0; w < 16; ++w) {
        if (w < wave) base += (uint32_t)part[w];
        all += (uint32_t)part[w];
    }
    uint32_t run = base + x - sum;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint64_t w = (uint64_t)tid * 4 + q;
        if (w < pw) pre[w] = run;
        run += c[q];
    }
    if (tid == 0) {
        pre[pw] = all;
        totals[0] = all;
    }
}

__global__ __launch_bounds__(256) void k_popcount_words(const uint64_t *__restrict__ w, uint64_t n,
                                                       uint32_t *__restrict__ cnt) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        cnt[i] = (uint32_t)__popcll(w[i]);
}

// Ids present below v (v may exceed the bitmap: then all of them).
__device__ __forceinline__ uint64_t present_below(const uint64_t *__restrict__ pres, const uint32_t *__restrict__ pre,
                                                  uint64_t nbits, uint64_t v) {
    if (v >= nbits) return pre[(nbits + 63) >> 6];
    const uint64_t w = v >> 6;
    return pre[w] + (uint64_t)__popcll(pres[w] & ((1ull << (v & 63)) - 1));
}

__device__ __forceinline__ bool is_present(const uint64_t *__restrict__ pres, uint64_t nbits, uint64_t v) {
    return v < nbits && ((pres[v >> 6] >> (v & 63)) & 1);
}

__device__ __forceinline__ uint64_t pow10u(int d) {
    uint64_t p = 1;
#pragma unroll
    for (int i = 0; i < 10; ++i)
        if (i < d) p *= 10;
    return p;
}

// Position of str(x) among the decimal strings of the present ids, sorted as
// strings (the reference orders labels with sorted() over str, so "10" < "2").
// A present y of m digits sorts before x (L digits) when
//   m <  L: y < the first m digits p of x, or y == p (a proper prefix);
//   m == L: y < x;
//   m >  L: y < x * 10^(m-L) (its first L digits are below x).
// Each case is a range count over the bitmap; the loop is unrolled so that
// the loads of all ten lengths are independent and issue together.
__device__ uint32_t string_rank(uint64_t x, const uint64_t *__restrict__ pres, const uint32_t *__restrict__ pre,
                                uint64_t nbits) {
    int L = 1;
    for (uint64_t p = 10; L < 10 && x >= p; p *= 10) ++L;
    uint64_t r = 0;
#pragma unroll
    for (int m = 1; m <= 10; ++m) {
        const uint64_t lo = m == 1 ? 0 : pow10u(m - 1);
        const uint64_t hi = pow10u(m);
        uint64_t end;
        if (m < L) {
            end = x / pow10u(L - m);
            r += is_present(pres, nbits, end);
        } else if (m == L) {
            end = x;
        } else {
            end = std::min(hi, x * pow10u(m - L));
        }
        end = std::max(end, lo);
        r += present_below(pres, pre, nbits, end) - present_below(pres, pre, nbits, lo);
    }
    return (uint32_t)r;
}

// One lane per present id: the u-th present id x (numeric order) is found by
// a binary search of pre[] for its word and a select within the word; its
// column c = its string rank; col[u] = c, labels[c] = x, and direct[x] = c
// when the direct map is kept.
__global__ __launch_bounds__(256) void k_ingest_labels(const uint64_t *__restrict__ pres, const uint32_t *__restrict__ pre,
                                                      uint64_t pres_words, uint64_t nbits, uint64_t n_present,
                                                      uint32_t *__restrict__ col, uint32_t *__restrict__ labels,
                                                      uint32_t *__restrict__ direct) {
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n_present;
         u += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t a = 0, b = pres_words;  // pre[a] <= u < pre[b]
        while (b - a > 1) {
            const uint64_t mid = (a + b) >> 1;
            if (pre[mid] <= u) a = mid; else b = mid;
        }
        uint64_t bits = pres[a];
        for (uint32_t k = (uint32_t)(u - pre[a]); k; --k) bits &= bits - 1;
        const uint64_t x = a * 64 + (uint64_t)__builtin_ctzll(bits);
        const uint32_t c = string_rank(x, pres, pre, nbits);
        col[u] = c;
        labels[c] = (uint32_t)x;
        if (direct) direct[x] = c;
    }
}

// Row starts of the kept baskets: rows[act[i]] = off[i]; rows[n_rows] = m.
__global__ __launch_bounds__(256) void k_ingest_rows(const uint64_t *__restrict__ off, uint64_t n,
                                                    const uint32_t *__restrict__ keep, const uint64_t *__restrict__ act,
                                                    uint64_t n_rows, uint64_t m, uint64_t *__restrict__ rows) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) rows[n_rows] = m;
    for (uint64_t i = g; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (keep[i]) rows[act[i]] = off[i];
}

// How an id finds its column: the direct map when the universe allows one,
// else its rank among the present ids (bitmap popcounts) into col[].
struct ColumnMap {
    const uint32_t *direct;
    const uint64_t *pres;
    const uint32_t *pre, *col;
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
        if (direct) return direct[x];
        const uint32_t w = x >> 6;
        return col[pre[w] + (uint32_t)__popcll(pres[w] & ((1ull << (x & 63)) - 1))];
    }
};

// Workgroup b owns words [b*B, b*B+B) of every column, i.e. store rows
// [64*b*B, 64*(b+1)*B).  Those rows' entries are one contiguous run of ids;
// lanes read it coalesced, eight entries in flight each, find each entry's
// row by a binary search over the row starts held in LDS, and OR its bit into
// the LDS tile (n_items x B words), written out once: the store needs no
// clearing and HBM sees no atomics.  With lds_map > 0 the direct map of ids
// 0..lds_map-1 is copied into LDS (16-bit columns) after the tile.
__global__ __launch_bounds__(512) void k_ingest_tile(const uint64_t *__restrict__ rows, uint64_t n_rows,
                                                    const uint32_t *__restrict__ ids, ColumnMap cm,
                                                    uint32_t lds_map, uint32_t n_items, uint32_t B,
                                                    uint64_t stride, uint64_t *__restrict__ store) {
    extern __shared__ unsigned long long tile[];
    __shared__ uint64_t rs[kRowsPerTileMax + 1];
    const uint32_t tid = threadIdx.x;
    const uint64_t w0 = (uint64_t)xcd_contiguous_block() * B;
    const uint32_t cells = n_items * B;
    uint16_t *lmap = reinterpret_cast<uint16_t *>(tile + cells);
    for (uint32_t i = tid; i < cells; i += blockDim.x) tile[i] = 0;
    for (uint32_t i = tid; i < lds_map; i += blockDim.x) lmap[i] = (uint16_t)cm.direct[i];
    const uint64_t r_lo = w0 * 64;
    const uint32_t nr = r_lo < n_rows ? (uint32_t)std::min<uint64_t>((uint64_t)B * 64, n_rows - r_lo) : 0;
    if (nr)
        for (uint32_t i = tid; i <= nr; i += blockDim.x) rs[i] = rows[r_lo + i];
    __syncthreads();
    if (nr) {
        const uint64_t lo = rs[0], hi = rs[nr];
        constexpr int kBatch = 8;
        for (uint64_t i0 = lo; i0 < hi; i0 += (uint64_t)kBatch * blockDim.x) {
            uint32_t x[kBatch];
#pragma unroll
            for (int q = 0; q < kBatch; ++q) {
                const uint64_t i = i0 + (uint64_t)q * blockDim.x + tid;
                x[q] = i < hi ? ids[i] : 0;
            }
#pragma unroll
            for (int q = 0; q < kBatch; ++q) {
                const uint64_t i = i0 + (uint64_t)q * blockDim.x + tid;
                if (i >= hi) continue;
                // the last row r in [0, nr) with rs[r] <= i
                uint32_t a = 0, b = nr;
                while (b - a > 1) {
                    const uint32_t mid = (a + b) >> 1;
                    if (rs[mid] <= i) a = mid; else b = mid;
                }
                const uint32_t c = lds_map ? lmap[x[q]] : cm(x[q]);
                atomicOr(&tile[c * B + (a >> 6)], 1ull << (a & 63));
            }
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < cells; i += blockDim.x) {
        const uint32_t c = i / B, k = i % B;
        if (w0 + k < stride) store[(uint64_t)c * stride + w0 + k] = tile[i];
    }
}

// One lane per store row, its entries ORed into the (cleared) store.
__global__ __launch_bounds__(256) void k_ingest_scatter(const uint64_t *__restrict__ rows, uint64_t n_rows,
                                                       const uint32_t *__restrict__ ids, ColumnMap cm,
                                                       uint64_t stride, unsigned long long *store) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows;
         r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t a = rows[r], b = rows[r + 1];
        for (uint64_t i = a; i < b; ++i)
            atomicOr(&store[(uint64_t)cm(ids[i]) * stride + (r >> 6)], 1ull << (r & 63));
    }
}

unsigned grid_for(cpm_ctx *ctx, uint64_t n, uint64_t per_cu = 8) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, (uint64_t)ctx->cu_count * per_cu));
}

}  // namespace

// The whole build from device arrays; the store's labels are kept with it.
int build_store_csr(cpm_ctx *ctx, const uint64_t *d_off, uint64_t n, const uint32_t *d_ids, uint64_t m,
                    uint32_t universe, int keep_empty, cpm_bits **out) {
    if (!d_off || (m && !d_ids)) return fail(CPM_ERR_INVALID, "offsets/ids are NULL");
    if (universe < 1) return fail(CPM_ERR_INVALID, "Invalid BIN file: item universe must fit in uint32");
    hipStream_t st = ctx->stream;
    IngestWs &ws = ingest_ws(ctx);
    const uint64_t nbits = (uint64_t)universe + 1;
    const uint64_t pw = (nbits + 63) / 64;
    CPM_CHECK(ws.state.ensure(32));
    CPM_CHECK(ws.pres.ensure(pw * 8));
    CPM_CHECK(ws.cnt.ensure(pw * 4));
    CPM_CHECK(ws.pre.ensure((pw + 1) * 4));
    CPM_CHECK(ws.keep.ensure(std::max<uint64_t>(n, 1) * 4));
    const unsigned g_in = grid_for(ctx, std::max(n, m));
    CPM_CHECK(ws.kept.ensure((size_t)g_in * 8));
    CPM_CHECK(ws.host.ensure(32));
    uint64_t *totals = ws.state.as<uint64_t>();
    unsigned int *flags = reinterpret_cast<unsigned int *>(totals + 2);
    uint64_t *pres = ws.pres.as<uint64_t>();
    uint32_t *pre = ws.pre.as<uint32_t>();
    uint32_t *keep = ws.keep.as<uint32_t>();

    Timed t_build(ctx, 11);
    CPM_HIP(hipMemsetAsync(totals, 0, 32, st));
    CPM_HIP(hipMemsetAsync(pres, 0, pw * 8, st));
    hipLaunchKernelGGL(k_ingest_scan_input, dim3(g_in), dim3(256), 0, st, d_off, n, d_ids, m, universe, pw,
                       keep_empty, keep, ws.kept.as<uint64_t>(), reinterpret_cast<unsigned long long *>(pres), flags);
    const bool small = pw <= kSmallScanWords;
    hipLaunchKernelGGL(k_ingest_totals, dim3(1), dim3(1024), 0, st, ws.kept.as<uint64_t>(), g_in, pres, pw,
                       (int)small, pre, totals);
    if (!small)
        hipLaunchKernelGGL(k_popcount_words, dim3(grid_for(ctx, pw)), dim3(256), 0, st, pres, pw,
                           ws.cnt.as<uint32_t>());
    CPM_HIP(hipGetLastError());
    if (!small) CPM_CHECK(exclusive_scan_u32(ctx, ws.sort, ws.cnt.as<uint32_t>(), pre, pw, totals));
    uint64_t *h = ws.host.as<uint64_t>();
    CPM_HIP(hipMemcpyAsync(h, totals, 24, hipMemcpyDeviceToHost, st));
    CPM_HIP(hipStreamSynchronize(st));
    const unsigned int *hflags = reinterpret_cast<const unsigned int *>(h + 2);
    // the reference checks the offsets before the ids (cuApriori.py:273-276)
    if (hflags[0]) return fail(CPM_ERR_INVALID, "Invalid BIN file: invalid transaction offsets");
    if (hflags[1]) return fail(CPM_ERR_INVALID, "Invalid BIN file: item ID outside configured universe");
    const uint64_t n_items = h[0], n_rows = h[1];
    if (n_items > 0xFFFFFFFFull) return fail(CPM_ERR_INVALID, "too many distinct items");
    if (n_rows >= (1ull << 40)) return fail(CPM_ERR_INVALID, "too many transactions");

    const uint32_t B = tile_words_for(n_items);
    cpm_bits *b = nullptr;
    CPM_CHECK(bits_alloc(ctx, n_rows, (uint32_t)n_items, &b, B == 0));
    int s = b->labels.ensure(std::max<uint64_t>(n_items, 1) * 4);
    if (s == CPM_OK) s = ws.col.ensure(std::max<uint64_t>(n_items, 1) * 4);
    // with no empty basket left out, the file's offsets are the row starts;
    // otherwise the kept baskets' starts are compacted (scan of keep[])
    const uint64_t *rows = d_off;
    if (s == CPM_OK && n_rows < n) {
        s = ws.rows.ensure((n_rows + 1) * 8);
        if (s == CPM_OK) s = ws.act.ensure((n + 1) * 8);
        if (s == CPM_OK) s = exclusive_scan_u32_to_u64(ctx, ws.sort, keep, ws.act.as<uint64_t>(), n, nullptr);
        if (s == CPM_OK) {
            hipLaunchKernelGGL(k_ingest_rows, dim3(grid_for(ctx, n)), dim3(256), 0, st, d_off, n, keep,
                               ws.act.as<uint64_t>(), n_rows, m, ws.rows.as<uint64_t>());
            rows = ws.rows.as<uint64_t>();
        }
    }
    const bool use_direct = nbits <= kDirectMapMax;
    if (s == CPM_OK && use_direct) s = ws.direct.ensure(nbits * 4);
    if (s == CPM_OK && n_items) {
        uint32_t *col = ws.col.as<uint32_t>();
        uint32_t *direct = use_direct ? ws.direct.as<uint32_t>() : nullptr;
        hipLaunchKernelGGL(k_ingest_labels, dim3(grid_for(ctx, n_items)), dim3(256), 0, st, pres, pre, pw, nbits,
                           n_items, col, b->labels.as<uint32_t>(), direct);
        const ColumnMap cm{direct, pres, pre, col};
        if (B) {
            const uint64_t blocks = (b->stride + B - 1) / B;
            if (blocks >= (1ull << 31)) {
                cpm_bits_destroy(b);
                return fail(CPM_ERR_INVALID, "too many transactions for one store");
            }
            const uint32_t lds_map = use_direct && nbits <= kLdsMapMax ? (uint32_t)nbits : 0;
            const size_t lds = (size_t)n_items * B * 8 + ((size_t)lds_map * 2 + 7) / 8 * 8;
            Timed t_tile(ctx, 9);
            hipLaunchKernelGGL(k_ingest_tile, dim3((unsigned)blocks), dim3(kTileThreads), lds, st, rows, n_rows, d_ids, cm,
                               lds_map, (uint32_t)n_items, B, b->stride, b->rows());
            s = t_tile.finish();
            ctx->stats[10] += 8.0 * (double)(n_rows + 1) + 4.0 * (double)m + 8.0 * (double)n_items * (double)b->stride;
        } else {
            hipLaunchKernelGGL(k_ingest_scatter, dim3(grid_for(ctx, n_rows, 32)), dim3(256), 0, st, rows, n_rows,
                               d_ids, cm, b->stride, reinterpret_cast<unsigned long long *>(b->rows()));
        }
    }
    if (s == CPM_OK) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) s = fail(CPM_ERR_RUNTIME, "store build failed: %s", hipGetErrorString(e));
    }
    if (s != CPM_OK) {
        (void)hipStreamSynchronize(st);
        cpm_bits_destroy(b);
        return s;
    }
    b->has_labels = true;
    *out = b;
    ctx->stats[8] += 1;
    s = t_build.finish();
    // a universe near 2^32 needs ~1 GB of bitmap and prefix counts: give
    // workspace buffers above kKeepBytes back (hipFree waits for this build)
    for (DBuf *d : {&ws.pres, &ws.cnt, &ws.pre, &ws.direct})
        if (d->cap > kKeepBytes) d->release();
    return s;
}

}  // namespace cpm

using namespace cpm;

extern "C" {

int cpm_bits_build_csr_device(cpm_ctx *ctx, const uint64_t *d_offsets, uint64_t n_baskets, const uint32_t *d_ids,
                              uint64_t n_ids, uint32_t universe, int keep_empty, cpm_bits **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    return build_store_csr(ctx, d_offsets, n_baskets, d_ids, n_ids, universe, keep_empty, out);
}

int cpm_bits_build_csr(cpm_ctx *ctx, const uint64_t *offsets, uint64_t n_baskets, const uint32_t *ids,
                       uint64_t n_ids, uint32_t universe, int keep_empty, cpm_bits **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!offsets || (n_ids && !ids)) return fail(CPM_ERR_INVALID, "offsets/ids are NULL");
    DBuf &d_off = ctx->scratch[0], &d_ids = ctx->scratch[1];
    CPM_CHECK(d_off.ensure((n_baskets + 1) * sizeof(uint64_t)));
    CPM_CHECK(d_ids.ensure(std::max<uint64_t>(n_ids, 1) * sizeof(uint32_t)));
    CPM_HIP(hipMemcpyAsync(d_off.p, offsets, (n_baskets + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
    if (n_ids)
        CPM_HIP(hipMemcpyAsync(d_ids.p, ids, n_ids * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    return build_store_csr(ctx, d_off.as<uint64_t>(), n_baskets, d_ids.as<uint32_t>(), n_ids, universe, keep_empty,
                           out);
}

int cpm_bits_labels(cpm_ctx *ctx, const cpm_bits *bits, uint32_t *ids) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!bits || (!ids && bits->n_items)) return fail(CPM_ERR_INVALID, "null argument");
    if (!bits->has_labels) return fail(CPM_ERR_INVALID, "the store was not built from basket arrays");
    return copy_to_host(ctx, ids, bits->labels.p, (size_t)bits->n_items * sizeof(uint32_t));
}

}  // extern "C"

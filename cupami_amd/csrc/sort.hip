// Device-wide exclusive scan and stable LSD radix sort of (u64 key, u32 value)
// pairs: the sorting and grouping primitives of the FP-tree build.
//
// Radix pass over one 8-bit digit:
//   hist    : per tile of 2048 elements, a 256-bin digit histogram in LDS,
//             stored bin-major (bin * tiles + tile)
//   scan    : device-wide exclusive scan of the histograms -> scatter bases
//   scatter : the tile is walked in element order, 256 elements at a time;
//             lanes with equal digits find each other with 8 ballots, so an
//             element's rank among equal digits follows input order (stable)
#include "sort.h"

#include <algorithm>

namespace cpm {

namespace {

constexpr uint32_t kSortBlock = 256;
constexpr uint32_t kSortIPT = 8;
constexpr uint32_t kSortTile = kSortBlock * kSortIPT;  // 2048
constexpr uint32_t kScanChunk = 2048;

// ---- scan -------------------------------------------------------------------------

__device__ __forceinline__ uint64_t block_scan_u64(uint64_t v, uint64_t *lds, uint64_t *total) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint64_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint64_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    const uint32_t nw = blockDim.x / 64;
    uint64_t base = 0, tot = 0;
    for (uint32_t w = 0; w < nw; ++w) {
        if (w < wave) base += lds[w];
        tot += lds[w];
    }
    *total = tot;
    __syncthreads();
    return base + x - v;
}

// chunk sums
__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t *__restrict__ in, uint64_t n,
                                                    uint64_t *__restrict__ part) {
    __shared__ uint64_t lds[8];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint64_t s = 0;
    for (uint32_t i = threadIdx.x; i < kScanChunk; i += blockDim.x)
        if (base + i < n) s += in[base + i];
    uint64_t tot;
    (void)block_scan_u64(s, lds, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// exclusive scan of the chunk sums in one block; part[nchunks] = total
__global__ __launch_bounds__(1024) void k_scan_partials(uint64_t *__restrict__ part, uint64_t nchunks,
                                                       uint64_t *__restrict__ total) {
    __shared__ uint64_t lds[16];
    __shared__ uint64_t carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (uint64_t i0 = 0; i0 < nchunks; i0 += blockDim.x) {
        const uint64_t i = i0 + threadIdx.x;
        const uint64_t v = i < nchunks ? part[i] : 0;
        uint64_t tot;
        const uint64_t off = block_scan_u64(v, lds, &tot);
        const uint64_t carry = carry_s;
        if (i < nchunks) part[i] = carry + off;
        __syncthreads();
        if (threadIdx.x == 0) carry_s = carry + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[nchunks] = carry_s;
        if (total) *total = carry_s;
    }
}

template <class OutT>
__global__ __launch_bounds__(256) void k_scan_down(const uint32_t *__restrict__ in, uint64_t n,
                                                  const uint64_t *__restrict__ part,
                                                  OutT *__restrict__ out) {
    __shared__ uint64_t lds[8];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    const uint32_t tid = threadIdx.x;
    // each lane owns 8 consecutive elements
    uint32_t v[8];
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t k = base + tid * 8 + i;
        v[i] = k < n ? in[k] : 0;
        s += v[i];
    }
    uint64_t tot;
    uint64_t off = part[blockIdx.x] + block_scan_u64(s, lds, &tot);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t k = base + tid * 8 + i;
        if (k < n) out[k] = (OutT)off;
        off += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && tid == 0) out[n] = (OutT)part[gridDim.x];
}

// ---- radix sort ---------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_radix_hist(const uint64_t *__restrict__ keys, uint64_t n,
                                                   uint32_t shift, uint32_t tiles,
                                                   uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[256];
    const uint32_t tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    for (uint32_t j = 0; j < kSortIPT; ++j) {
        const uint64_t i = base + (uint64_t)j * kSortBlock + tid;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)tid * tiles + blockIdx.x] = h[tid];
}

__global__ __launch_bounds__(256) void k_radix_scatter(const uint64_t *__restrict__ keys,
                                                      const uint32_t *__restrict__ vals, uint64_t n,
                                                      uint32_t shift, uint32_t tiles,
                                                      const uint32_t *__restrict__ offs,
                                                      uint64_t *__restrict__ okeys,
                                                      uint32_t *__restrict__ ovals) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcount[4][256];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    running[tid] = offs[(uint64_t)tid * tiles + blockIdx.x];
    for (uint32_t w = 0; w < 4; ++w) wcount[w][tid] = 0;
    __syncthreads();
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    for (uint32_t j = 0; j < kSortIPT; ++j) {
        const uint64_t i = base + (uint64_t)j * kSortBlock + tid;
        const bool in = i < n;
        uint64_t key = 0;
        uint32_t val = 0, digit = 0;
        if (in) {
            key = keys[i];
            val = vals[i];
            digit = (uint32_t)(key >> shift) & 255u;
        }
        // lanes of this wave with the same digit
        uint64_t match = __ballot(in);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bal = __ballot(in && ((digit >> b) & 1u));
            match &= ((digit >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(match & lt);
        if (in && rank == 0) wcount[wave][digit] = (uint32_t)__popcll(match);
        __syncthreads();
        if (in) {
            uint32_t pos = running[digit] + rank;
            for (uint32_t w = 0; w < wave; ++w) pos += wcount[w][digit];
            okeys[pos] = key;
            ovals[pos] = val;
        }
        __syncthreads();
        running[tid] += wcount[0][tid] + wcount[1][tid] + wcount[2][tid] + wcount[3][tid];
        for (uint32_t w = 0; w < 4; ++w) wcount[w][tid] = 0;
        __syncthreads();
    }
}

// ---- single-workgroup paths for small inputs -----------------------------------------
// The FP-tree build sorts and scans many small arrays (conditional trees of a
// few hundred nodes): one launch each instead of the device-wide passes.

constexpr uint32_t kSmallSort = 2048;  // pairs sorted by one workgroup
constexpr uint32_t kSmallScan = 4096;  // elements scanned by one workgroup (4 per lane)

// Sort of n <= kSmallSort pairs by (key & mask), stable: bitonic over
// (masked key, input index), a total order in which equal keys keep their
// input order.  In place.
__global__ __launch_bounds__(1024) void k_sort_small(uint64_t *keys, uint32_t *vals, uint32_t n, uint64_t mask) {
    __shared__ uint64_t sk[kSmallSort];
    __shared__ uint32_t si[kSmallSort];
    const uint32_t tid = threadIdx.x;
    uint32_t P = 1;
    while (P < n) P <<= 1;
    for (uint32_t i = tid; i < P; i += blockDim.x) {
        sk[i] = i < n ? (keys[i] & mask) : ~0ull;  // padding sorts last (index breaks ties)
        si[i] = i;
    }
    __syncthreads();
    bitonic_sort_lds(sk, si, P);
    // every lane reads its sorted pairs before any lane writes (in place)
    constexpr uint32_t kPer = kSmallSort / 1024;
    uint64_t ok[kPer];
    uint32_t ov[kPer];
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t i = tid + r * blockDim.x;
        if (i < n) {
            ok[r] = keys[si[i]];
            ov[r] = vals[si[i]];
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t i = tid + r * blockDim.x;
        if (i < n) {
            keys[i] = ok[r];
            vals[i] = ov[r];
        }
    }
}

// Exclusive scan of n <= kSmallScan values: out[n] = total (and *d_total).
template <class OutT>
__global__ __launch_bounds__(1024) void k_scan_small(const uint32_t *in, uint64_t n, OutT *out, uint64_t *d_total) {
    __shared__ uint64_t lds[16];
    const uint32_t tid = threadIdx.x;
    uint32_t v[4];
    uint64_t sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint64_t i = (uint64_t)tid * 4 + q;
        v[q] = i < n ? in[i] : 0;
        sum += v[q];
    }
    uint64_t total;
    uint64_t run = block_scan_u64(sum, lds, &total);  // its barriers also order reads before writes
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint64_t i = (uint64_t)tid * 4 + q;
        if (i < n) out[i] = (OutT)run;
        run += v[q];
    }
    if (tid == 0) {
        out[n] = (OutT)total;
        if (d_total) *d_total = total;
    }
}

__global__ void k_copy_pairs(const uint64_t *__restrict__ k, const uint32_t *__restrict__ v,
                             uint64_t n, uint64_t *__restrict__ ok, uint32_t *__restrict__ ov) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        ok[i] = k[i];
        ov[i] = v[i];
    }
}

}  // namespace

int SortWs::ensure(uint64_t n) {
    const uint64_t tiles = (n + kSortTile - 1) / kSortTile;
    CPM_CHECK(keys_tmp.ensure(std::max<uint64_t>(n, 1) * 8));
    CPM_CHECK(vals_tmp.ensure(std::max<uint64_t>(n, 1) * 4));
    CPM_CHECK(hist.ensure((256 * tiles + 1) * 4));
    CPM_CHECK(offs.ensure((256 * tiles + 1) * 4));
    const uint64_t chunks = (256 * tiles + kScanChunk - 1) / kScanChunk + 1;
    const uint64_t chunks2 = (n + kScanChunk - 1) / kScanChunk + 1;
    CPM_CHECK(part.ensure((std::max(chunks, chunks2) + 1) * 8));
    CPM_CHECK(total.ensure(16));
    return CPM_OK;
}

int exclusive_scan_u32(cpm_ctx *ctx, SortWs &ws, const uint32_t *in, uint32_t *out, uint64_t n,
                       uint64_t *d_total) {
    if (n <= kSmallScan) {
        hipLaunchKernelGGL(k_scan_small<uint32_t>, dim3(1), dim3(1024), 0, ctx->stream, in, n, out, d_total);
        CPM_HIP(hipGetLastError());
        return CPM_OK;
    }
    const uint64_t chunks = std::max<uint64_t>(1, (n + kScanChunk - 1) / kScanChunk);
    CPM_CHECK(ws.part.ensure((chunks + 1) * 8));
    uint64_t *part = ws.part.as<uint64_t>();
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)chunks), dim3(256), 0, ctx->stream, in, n, part);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, ctx->stream, part, chunks, d_total);
    hipLaunchKernelGGL(k_scan_down<uint32_t>, dim3((unsigned)chunks), dim3(256), 0, ctx->stream, in, n,
                       part, out);
    CPM_HIP(hipGetLastError());
    return CPM_OK;
}

int exclusive_scan_u32_to_u64(cpm_ctx *ctx, SortWs &ws, const uint32_t *in, uint64_t *out,
                              uint64_t n, uint64_t *d_total) {
    if (n <= kSmallScan) {
        hipLaunchKernelGGL(k_scan_small<uint64_t>, dim3(1), dim3(1024), 0, ctx->stream, in, n, out, d_total);
        CPM_HIP(hipGetLastError());
        return CPM_OK;
    }
    const uint64_t chunks = std::max<uint64_t>(1, (n + kScanChunk - 1) / kScanChunk);
    CPM_CHECK(ws.part.ensure((chunks + 1) * 8));
    uint64_t *part = ws.part.as<uint64_t>();
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)chunks), dim3(256), 0, ctx->stream, in, n, part);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, ctx->stream, part, chunks, d_total);
    hipLaunchKernelGGL(k_scan_down<uint64_t>, dim3((unsigned)chunks), dim3(256), 0, ctx->stream, in,
                       n, part, out);
    CPM_HIP(hipGetLastError());
    return CPM_OK;
}

int radix_sort_pairs(cpm_ctx *ctx, SortWs &ws, uint64_t *keys, uint32_t *vals, uint64_t n,
                     int end_bit) {
    if (n <= 1 || end_bit <= 0) return CPM_OK;
    if (n <= kSmallSort) {
        // the device-wide passes order by whole 8-bit digits: bits [0, 8 * ceil(end_bit / 8))
        const int bits = std::min(64, (end_bit + 7) / 8 * 8);
        const uint64_t mask = bits >= 64 ? ~0ull : (1ull << bits) - 1;
        hipLaunchKernelGGL(k_sort_small, dim3(1), dim3(1024), 0, ctx->stream, keys, vals, (uint32_t)n, mask);
        CPM_HIP(hipGetLastError());
        return CPM_OK;
    }
    CPM_CHECK(ws.ensure(n));
    const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
    uint64_t *ka = keys, *kb = ws.keys_tmp.as<uint64_t>();
    uint32_t *va = vals, *vb = ws.vals_tmp.as<uint32_t>();
    int passes = 0;
    for (int shift = 0; shift < end_bit; shift += 8, ++passes) {
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kSortBlock), 0, ctx->stream, ka, n,
                           (uint32_t)shift, tiles, ws.hist.as<uint32_t>());
        CPM_HIP(hipGetLastError());
        CPM_CHECK(exclusive_scan_u32(ctx, ws, ws.hist.as<uint32_t>(), ws.offs.as<uint32_t>(),
                                     (uint64_t)256 * tiles, nullptr));
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortBlock), 0, ctx->stream, ka, va,
                           n, (uint32_t)shift, tiles, ws.offs.as<uint32_t>(), kb, vb);
        CPM_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (passes & 1) {
        const uint32_t g = (uint32_t)std::min<uint64_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(k_copy_pairs, dim3(g), dim3(256), 0, ctx->stream, ka, va, n, keys, vals);
        CPM_HIP(hipGetLastError());
    }
    return CPM_OK;
}

int bits_for(uint64_t max_value) {
    int b = 0;
    while (b < 64 && (max_value >> b)) ++b;
    return b;
}

}  // namespace cpm

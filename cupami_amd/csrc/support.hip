// Support-counting kernels over the vertical store.
//
// A level of prefix tasks is cut into work items (task, group of G extension
// entries, tile of kTileWords words).  A 256-lane workgroup builds its tile of
// the task's prefix in registers (the AND of the d prefix item rows), then
// streams the group's extension rows over it: AND, popcount, wave reduction,
// one atomic per entry per workgroup.  Work items are handed out by a
// persistent loop whose blockIdx -> tile mapping keeps a tile on one XCD, so
// the slice of every row that an XCD reads stays in that XCD's L2.
//
// Reference semantics: _bitsets.py:30-53 (complete intersection,
// popcount, block reduction) and _temporal.py:28-55 (occurrence walk).
#include "support_kernels.h"

#include <algorithm>

namespace cpm {

namespace {

constexpr uint32_t kQuadsPerTile = kTileWords / 2;          // uint4 per tile
constexpr uint32_t kQuadsPerLane = kQuadsPerTile / kBlock;  // 4
constexpr uint32_t kMaxGroup = 64;                          // entries per group

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint32_t pop4(uint4 a) {
    return __popc(a.x) + __popc(a.y) + __popc(a.z) + __popc(a.w);
}

__device__ __forceinline__ uint4 and4(uint4 a, uint4 b) {
    return make_uint4(a.x & b.x, a.y & b.y, a.z & b.z, a.w & b.w);
}

// First task whose group range holds g: largest t with grp_off[t] <= g.
__device__ __forceinline__ uint32_t find_task(const uint32_t *__restrict__ grp_off, uint32_t n_tasks,
                                              uint32_t g) {
    uint32_t lo = 0, hi = n_tasks;  // answer in [lo, hi)
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (grp_off[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// Work-item schedule: XCD x = blockIdx % 8 owns tiles x, x+8, ... (when there
// are at least 8 tiles), so each XCD's L2 only ever holds its own slice of the
// rows.  Returns false when the block has no more items.
struct Schedule {
    uint32_t pool, slot, nslots, tiles_in_pool, pools;
    uint64_t items;
    __device__ Schedule(uint32_t n_tiles, uint32_t total_groups) {
        pools = (n_tiles >= 8 && (gridDim.x % 8) == 0) ? 8u : 1u;
        pool = blockIdx.x % pools;
        slot = blockIdx.x / pools;
        nslots = gridDim.x / pools;
        tiles_in_pool = (n_tiles - pool + pools - 1) / pools;
        items = (uint64_t)total_groups * tiles_in_pool;
    }
    __device__ void decode(uint64_t it, uint32_t &g, uint32_t &tile) const {
        g = (uint32_t)(it / tiles_in_pool);
        tile = pool + pools * (uint32_t)(it % tiles_in_pool);
    }
};

__device__ __forceinline__ uint32_t col_of(const uint32_t *__restrict__ fcol, uint32_t f) {
    return fcol ? fcol[f] : f;
}

// ---- frequency: support of every entry ---------------------------------------

__global__ __launch_bounds__(256) void k_support_count(const uint4 *__restrict__ store,
                                                      uint64_t stride_q, uint32_t n_tiles,
                                                      TaskView tv, uint32_t *__restrict__ sup) {
    __shared__ uint32_t part[4][kMaxGroup];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tv.n_tasks == 0) return;
    const uint32_t total_groups = tv.grp_off[tv.n_tasks];
    Schedule s(n_tiles, total_groups);
    for (uint64_t it = s.slot; it < s.items; it += s.nslots) {
        uint32_t g, tile;
        s.decode(it, g, tile);
        const uint32_t t = find_task(tv.grp_off, tv.n_tasks, g);
        const uint32_t e0 = tv.ext_off[t] + (g - tv.grp_off[t]) * tv.G;
        const uint32_t e1 = min(tv.ext_off[t + 1], e0 + tv.G);

        // prefix tile in registers: quads q = tile*1024 + r*256 + tid
        const uint64_t q0 = (uint64_t)tile * kQuadsPerTile + tid;
        uint4 P[kQuadsPerLane];
        bool valid[kQuadsPerLane];
#pragma unroll
        for (int r = 0; r < kQuadsPerLane; ++r) {
            valid[r] = q0 + (uint64_t)r * kBlock < stride_q;
            P[r] = valid[r] ? make_uint4(~0u, ~0u, ~0u, ~0u) : make_uint4(0, 0, 0, 0);
        }
        const uint32_t *pre = tv.prefix + (uint64_t)t * tv.d;
        for (uint32_t j = 0; j < tv.d; ++j) {
            const uint4 *row = store + (uint64_t)col_of(tv.fcol, pre[j]) * stride_q + q0;
#pragma unroll
            for (int r = 0; r < kQuadsPerLane; ++r)
                if (valid[r]) P[r] = and4(P[r], row[r * kBlock]);
        }

        // extension rows, two at a time to keep loads in flight
        uint32_t e = e0;
        for (; e + 1 < e1; e += 2) {
            const uint4 *ra = store + (uint64_t)col_of(tv.fcol, tv.ext[e]) * stride_q + q0;
            const uint4 *rb = store + (uint64_t)col_of(tv.fcol, tv.ext[e + 1]) * stride_q + q0;
            uint4 A[kQuadsPerLane], B[kQuadsPerLane];
#pragma unroll
            for (int r = 0; r < kQuadsPerLane; ++r) {
                A[r] = valid[r] ? ra[r * kBlock] : make_uint4(0, 0, 0, 0);
                B[r] = valid[r] ? rb[r * kBlock] : make_uint4(0, 0, 0, 0);
            }
            uint32_t ca = 0, cb = 0;
#pragma unroll
            for (int r = 0; r < kQuadsPerLane; ++r) {
                ca += pop4(and4(P[r], A[r]));
                cb += pop4(and4(P[r], B[r]));
            }
            ca = wave_sum(ca);
            cb = wave_sum(cb);
            if (lane == 0) {
                part[wave][e - e0] = ca;
                part[wave][e + 1 - e0] = cb;
            }
        }
        if (e < e1) {
            const uint4 *ra = store + (uint64_t)col_of(tv.fcol, tv.ext[e]) * stride_q + q0;
            uint32_t ca = 0;
#pragma unroll
            for (int r = 0; r < kQuadsPerLane; ++r)
                if (valid[r]) ca += pop4(and4(P[r], ra[r * kBlock]));
            ca = wave_sum(ca);
            if (lane == 0) part[wave][e - e0] = ca;
        }
        __syncthreads();
        if (tid < e1 - e0) {
            uint32_t v = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
            if (v) atomicAdd(sup + e0 + tid, v);
        }
        __syncthreads();
    }
}

// ---- periodicity ----------------------------------------------------------------

__device__ __forceinline__ void pp_empty(PeriodPart &m) {
    m.first = 0;
    m.last = 0;
    m.gap = 0;
    m.count = 0;
    m.pad = 0;
}

// a then b (a covers earlier rows)
__device__ __forceinline__ PeriodPart pp_combine(const PeriodPart &a, const PeriodPart &b) {
    if (a.count == 0) return b;
    if (b.count == 0) return a;
    PeriodPart m;
    m.first = a.first;
    m.last = b.last;
    int64_t g = a.gap > b.gap ? a.gap : b.gap;
    int64_t x = b.first - a.last;
    m.gap = g > x ? g : x;
    m.count = a.count + b.count;
    m.pad = 0;
    return m;
}

__device__ __forceinline__ PeriodPart pp_shfl_down(const PeriodPart &m, int off) {
    PeriodPart o;
    o.first = __shfl_down(m.first, off, 64);
    o.last = __shfl_down(m.last, off, 64);
    o.gap = __shfl_down(m.gap, off, 64);
    o.count = __shfl_down(m.count, off, 64);
    o.pad = 0;
    return o;
}

// In-order reduction over the 64 lanes (lane i covers earlier rows than lane
// i+1); the result is valid in lane 0.
__device__ __forceinline__ PeriodPart pp_wave_reduce(PeriodPart m, uint32_t lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        PeriodPart o = pp_shfl_down(m, off);
        if (lane + off < 64) m = pp_combine(m, o);
    }
    return m;
}

// Walk the set bits of one word of an occurrence row, in row order.
__device__ __forceinline__ void pp_scan_word(PeriodPart &m, uint64_t v, uint64_t row0,
                                             const int64_t *__restrict__ ts, uint64_t n) {
    while (v) {
        const uint64_t r = row0 + (uint64_t)__builtin_ctzll(v);
        v &= v - 1;
        if (r < n) {
            const int64_t t = ts[r];
            if (m.count == 0) {
                m.first = t;
            } else {
                const int64_t g = t - m.last;
                if (g > m.gap) m.gap = g;
            }
            m.last = t;
            ++m.count;
        }
    }
}

// Blocked lanes: lane `tid` owns words tile*2048 + tid*8 .. +8 (in row order),
// so lane order is row order for the in-order reductions.
__global__ __launch_bounds__(256) void k_support_periodic(const uint64_t *__restrict__ store,
                                                         uint64_t stride, uint32_t n_tiles,
                                                         TaskView tv,
                                                         const int64_t *__restrict__ ts,
                                                         uint64_t n,
                                                         PeriodPart *__restrict__ parts) {
    __shared__ PeriodPart wpart[4][16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tv.n_tasks == 0) return;
    const uint32_t total_groups = tv.grp_off[tv.n_tasks];
    Schedule s(n_tiles, total_groups);
    for (uint64_t it = s.slot; it < s.items; it += s.nslots) {
        uint32_t g, tile;
        s.decode(it, g, tile);
        const uint32_t t = find_task(tv.grp_off, tv.n_tasks, g);
        const uint32_t e0 = tv.ext_off[t] + (g - tv.grp_off[t]) * tv.G;
        const uint32_t e1 = min(tv.ext_off[t + 1], e0 + tv.G);
        const uint64_t w0 = (uint64_t)tile * kTileWords + (uint64_t)tid * 8;
        const bool valid = w0 < stride;  // stride is a multiple of 64 words
        uint64_t P[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) P[v] = valid ? ~0ull : 0ull;
        const uint32_t *pre = tv.prefix + (uint64_t)t * tv.d;
        if (valid) {
            for (uint32_t j = 0; j < tv.d; ++j) {
                const uint64_t *row = store + (uint64_t)col_of(tv.fcol, pre[j]) * stride + w0;
#pragma unroll
                for (int v = 0; v < 8; ++v) P[v] &= row[v];
            }
        }
        for (uint32_t eb = e0; eb < e1; eb += 16) {
            const uint32_t ee = min(e1, eb + 16);
            for (uint32_t e = eb; e < ee; ++e) {
                PeriodPart m;
                pp_empty(m);
                if (valid) {
                    const uint64_t *row = store + (uint64_t)col_of(tv.fcol, tv.ext[e]) * stride + w0;
                    uint64_t X[8];
#pragma unroll
                    for (int v = 0; v < 8; ++v) X[v] = P[v] & row[v];
#pragma unroll
                    for (int v = 0; v < 8; ++v) pp_scan_word(m, X[v], (w0 + v) * 64, ts, n);
                }
                m = pp_wave_reduce(m, lane);
                if (lane == 0) wpart[wave][e - eb] = m;
            }
            __syncthreads();
            if (tid < ee - eb) {
                PeriodPart m = pp_combine(pp_combine(wpart[0][tid], wpart[1][tid]),
                                          pp_combine(wpart[2][tid], wpart[3][tid]));
                parts[(uint64_t)(eb + tid) * n_tiles + tile] = m;
            }
            __syncthreads();
        }
    }
}

// Fold the tile partials of each entry in row order: support and the largest
// gap of [0, ts..., end] (_temporal.py:36-54: prev starts at 0, end gap last).
__global__ __launch_bounds__(256) void k_periodic_finalize(const PeriodPart *__restrict__ parts,
                                                          uint32_t n_tiles, uint64_t n_ext,
                                                          int64_t end, uint32_t *__restrict__ sup,
                                                          int64_t *__restrict__ per) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_ext;
         e += (uint64_t)gridDim.x * blockDim.x) {
        PeriodPart m;
        pp_empty(m);
        for (uint32_t c = 0; c < n_tiles; ++c) m = pp_combine(m, parts[e * n_tiles + c]);
        int64_t p;
        if (m.count == 0) {
            p = end;
        } else {
            p = m.first;
            if (m.gap > p) p = m.gap;
            if (end - m.last > p) p = end - m.last;
        }
        sup[e] = m.count;
        per[e] = p;
    }
}

// ---- operator: child rows of one prefix -----------------------------------------

__global__ __launch_bounds__(256) void k_intersect(const uint64_t *__restrict__ prefix,
                                                  const uint64_t *__restrict__ ext, uint64_t stride,
                                                  unsigned long long *__restrict__ sup,
                                                  uint64_t *__restrict__ out) {
    __shared__ uint32_t red[4];
    const uint32_t j = blockIdx.x, tid = threadIdx.x;
    const uint64_t *row = ext + (uint64_t)j * stride;
    uint64_t *dst = out ? out + (uint64_t)j * stride : nullptr;
    uint32_t c = 0;
    for (uint64_t w = tid; w < stride; w += blockDim.x) {
        const uint64_t v = prefix[w] & row[w];
        if (dst) dst[w] = v;
        c += (uint32_t)__popcll(v);
    }
    c = wave_sum(c);
    if ((tid & 63) == 0) red[tid >> 6] = c;
    __syncthreads();
    if (tid == 0) sup[j] = (unsigned long long)red[0] + red[1] + red[2] + red[3];
}

uint32_t grid_for(const cpm_ctx *ctx, uint64_t items) {
    uint64_t cap = (uint64_t)ctx->cu_count * 8;
    uint64_t g = std::min<uint64_t>(cap, std::max<uint64_t>(items, 1));
    g = (g + 7) / 8 * 8;
    return (uint32_t)g;
}

}  // namespace

int launch_support_count(cpm_ctx *ctx, const cpm_bits *bits, const TaskView &tv, uint64_t n_ext,
                         uint32_t *sup) {
    if (tv.n_tasks == 0 || n_ext == 0) return CPM_OK;
    if (tv.G == 0 || tv.G > kMaxGroup) return fail(CPM_ERR_INVALID, "group size %u", tv.G);
    const uint32_t n_tiles = tiles_per_row(bits->stride);
    const uint32_t grid = grid_for(ctx, n_ext * n_tiles);
    Timed timer(ctx, 1);
    hipLaunchKernelGGL(k_support_count, dim3(grid), dim3(kBlock), 0, ctx->stream,
                       bits->store.as<const uint4>(), bits->stride / 2, n_tiles, tv, sup);
    CPM_HIP(hipGetLastError());
    CPM_CHECK(timer.finish());
    ctx->stats[0] += 1;
    ctx->stats[3] += (double)n_ext;
    return CPM_OK;
}

int launch_support_periodic(cpm_ctx *ctx, const cpm_bits *bits, const TaskView &tv,
                            uint64_t n_ext, const int64_t *ts, int64_t end, PeriodPart *parts,
                            uint32_t *sup, int64_t *per, uint32_t *, int) {
    if (tv.n_tasks == 0 || n_ext == 0) return CPM_OK;
    if (tv.G == 0 || tv.G > kMaxGroup) return fail(CPM_ERR_INVALID, "group size %u", tv.G);
    const uint32_t n_tiles = tiles_per_row(bits->stride);
    const uint32_t grid = grid_for(ctx, n_ext * n_tiles);
    Timed timer(ctx, 1);
    hipLaunchKernelGGL(k_support_periodic, dim3(grid), dim3(kBlock), 0, ctx->stream,
                       bits->store.as<const uint64_t>(), bits->stride, n_tiles, tv, ts,
                       bits->n_transactions, parts);
    CPM_HIP(hipGetLastError());
    CPM_CHECK(timer.finish());
    const uint32_t fgrid = (uint32_t)std::min<uint64_t>((n_ext + 255) / 256, 4096);
    hipLaunchKernelGGL(k_periodic_finalize, dim3(fgrid), dim3(256), 0, ctx->stream, parts,
                       n_tiles, n_ext, end, sup, per);
    CPM_HIP(hipGetLastError());
    ctx->stats[0] += 1;
    ctx->stats[3] += (double)n_ext;
    return CPM_OK;
}

int launch_intersect(cpm_ctx *ctx, const uint64_t *prefix, const uint64_t *ext, uint32_t k,
                     uint64_t stride, unsigned long long *sup, uint64_t *out) {
    if (k == 0) return CPM_OK;
    hipLaunchKernelGGL(k_intersect, dim3(k), dim3(256), 0, ctx->stream, prefix, ext, stride, sup,
                       out);
    CPM_HIP(hipGetLastError());
    return CPM_OK;
}

}  // namespace cpm

// Support-counting kernels over the vertical stores of one or more databases.
//
// A level's entries are cut into work items (a range of G consecutive
// entries, a tile of words).  A 256-lane workgroup walks the tasks its range
// touches; for each it builds its tile of the task's prefix in registers (the
// AND of the d prefix item rows -- of only the last one when the previous task
// was a sibling, whose parent part is kept), then streams the task's
// extension rows of the range over it: AND, popcount, wave reduction, one
// atomic per entry per workgroup.  The grid is persistent and a multiple of 8 and the tiles per
// range are too (when rows have 8 or more tiles), so a block -- which stays
// on XCD blockIdx % 8 -- only ever touches tiles congruent to that XCD: each
// XCD's L2 holds its own slice of the hot rows.  Sizes are read from device
// memory, so the launch can be replayed from a graph.
//
// Reference semantics: _bitsets.py:30-53 (complete intersection, popcount,
// block reduction) and _temporal.py:28-55 (occurrence walk).
#include "support_kernels.h"

#include <algorithm>
#include <type_traits>

namespace cpm {

namespace {

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint32_t pop4(uint4 a) {
    return __popc(a.x) + __popc(a.y) + __popc(a.z) + __popc(a.w);
}

// 16 bytes at byte `off` of a row of `bytes` bytes, as a raw buffer load:
// zero when off is at or past the end.  Word 3 of the resource is the gfx9
// family's raw-access setting (32-bit data format; the format fields are
// unused by untyped loads).
constexpr int kRawBufferWord3 = 0x00020000;
__device__ __forceinline__ uint4 row_load(gptr<uint4> row, uint32_t bytes, uint32_t off) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)(const void *)row, 0, (int)bytes, kRawBufferWord3);
    const auto x = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
    return make_uint4(x[0], x[1], x[2], x[3]);
}

__device__ __forceinline__ uint4 and4(uint4 a, uint4 b) {
    return make_uint4(a.x & b.x, a.y & b.y, a.z & b.z, a.w & b.w);
}

// largest t < n with off[t] <= i (off non-decreasing, off[n] > i): the task
// holding entry i
__device__ __forceinline__ uint32_t find_task(cptr<uint32_t> off, uint32_t n, uint64_t i) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// find_task by the whole wave: a 64-ary search, one load per lane per step
// (the interval shrinks 64-fold: two steps for 4096 tasks, three for 2^18)
// instead of a binary search's chain of dependent loads.  Keeps
// off[lo] <= i < off[hi]; lane 0's pivot is lo, so the ballot is never empty.
__device__ __forceinline__ uint32_t wave_find_task(cptr<uint32_t> off, uint32_t n, uint64_t i) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t p = lo + lane * step;
        const unsigned long long le = __ballot(p < hi && off[p] <= i);
        lo += (63u - (uint32_t)__builtin_clzll(le)) * step;
        hi = min(hi, lo + step);
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
}

__device__ __forceinline__ uint32_t col_of(cptr<uint32_t> fcol, uint32_t f) {
    return fcol ? fcol[f] : f;
}

// The level's index arrays, read-only while a support kernel runs: read
// through the constant address space, so that the per-row lookups of a work
// item (task offsets, prefix and entry items, columns) are scalar loads.
struct LevelView {
    cptr<uint32_t> ext_off, dbi, prefix, ext, fcol, par, src;
    cptr<DbDev> dbs;
    uint32_t G, te, n_tasks, d, n_ext;
    int siblings;
    uint64_t total;
    const uint64_t *rin;   // retained rows (SupportArgs)
    uint64_t *rout;
    uint64_t slots;
    cptr<uint32_t> rtask, rtiles;  // per work range (SupportArgs::ranges), or null
    uint64_t rcap;
};

// false when the launch has nothing to do (the same answer in every block)
__device__ __forceinline__ bool level_view(const SupportArgs &a, LevelView &v) {
    if (a.stop && *a.stop) return false;
    v.n_tasks = (uint32_t)a.L.sz->n_tasks;
    v.total = a.L.sz->n_work;
    if (v.n_tasks == 0 || v.total == 0) return false;
    v.d = (uint32_t)a.L.sz->d;
    v.n_ext = (uint32_t)a.L.sz->n_ext;
    v.ext_off = to_const(a.L.ext_off);
    v.dbi = to_const(a.L.db);
    v.prefix = to_const(a.L.prefix);
    v.ext = to_const(a.L.ext);
    v.fcol = a.fcol ? to_const(a.fcol) : nullptr;
    v.dbs = to_const(a.dbs);
    v.G = (uint32_t)a.L.sz->group;
    v.te = a.te;
    v.siblings = a.siblings && a.L.par;
    v.par = v.siblings ? to_const(a.L.par) : nullptr;
    v.rin = v.d >= 3 && a.L.src ? a.rin : nullptr;
    v.src = v.rin ? to_const(a.L.src) : nullptr;
    v.rout = a.rout;
    v.slots = a.slots;
    v.rtask = a.ranges ? to_const(a.ranges) : nullptr;
    v.rtiles = a.ranges ? to_const(a.ranges + a.rcap) : nullptr;
    v.rcap = a.ranges ? a.rcap : 0;
    return true;
}

// The work range of item `it` reaches no database tile `tile` (nothing to
// count), from the range table; false when there is no table entry.
__device__ __forceinline__ bool range_empty(const LevelView &v, uint64_t it, uint32_t tile) {
    const uint64_t g = it / v.te;
    return g < v.rcap && tile >= v.rtiles[g];
}

// the first task of item `it`'s range: from the range table, else by binary
// search
__device__ __forceinline__ uint32_t range_task(const LevelView &v, uint64_t it, uint32_t r0) {
    const uint64_t g = it / v.te;
    return g < v.rcap ? v.rtask[g] : wave_find_task(v.ext_off, v.n_tasks, r0);
}

// Work item `it`: entries [r0, r1) on tile `tile`.
struct Range {
    uint32_t r0, r1, tile;
};

// Item `it` is the k-th of its block.  When the grid is a multiple of the
// tiles per range, item it would always land on tile it % te of the same
// block; the tile is rotated by k (by 8 k when te is a multiple of 8, which
// keeps a block on the tiles of its XCD) so that the tiles ranges of short
// rows leave empty spread over all blocks.  Each run of gridDim.x items covers
// whole ranges, so the map stays one to one.
__device__ __forceinline__ Range range_of(const LevelView &v, uint64_t it, uint64_t k) {
    const uint64_t g = it / v.te;
    Range r;
    uint64_t tile = it - g * v.te;
    if (v.te > 1 && gridDim.x % v.te == 0) tile = (tile + k * (v.te % 8 == 0 ? 8 : 1)) % v.te;
    r.tile = (uint32_t)tile;
    r.r0 = (uint32_t)(g * v.G);
    r.r1 = min(v.n_ext, r.r0 + v.G);
    return r;
}

// valid data words of tile `tile` (tw words per tile) of a database
__device__ __forceinline__ uint64_t tile_words_of(uint64_t words, uint32_t tile, uint32_t tw) {
    const uint64_t w0 = (uint64_t)tile * tw;
    return w0 >= words ? 0 : min<uint64_t>(tw, words - w0);
}

// This block's stamps (start, end, algorithmic bytes, candidates).  Plain
// stores: the launch that follows folds them (fold_stamps), so no block
// fences or contends on a shared counter here.
__device__ __forceinline__ void launch_epilogue(const SupportArgs &a, unsigned long long t_start,
                                                unsigned long long bytes, unsigned long long cands) {
    if (!a.stamps || threadIdx.x != 0) return;
    unsigned long long *s = a.stamps + 4ull * blockIdx.x;
    s[0] = t_start;
    s[1] = wall_clock64();
    s[2] = bytes;
    s[3] = cands;
}

// ---- frequency ----------------------------------------------------------------

// Short rows.  A database of at most 4096 transactions has one 64-word
// stride per row: on the tile path it fills 32 / QPL lanes of a 256-lane
// block per entry, so its deep levels (many small tasks) are walked one entry
// pair at a time by nearly idle blocks.  A work item whose tasks all belong
// to such databases, and whose prefixes are each one retained row (levels
// past 3) or at most two item rows, is counted by lane groups instead: S =
// 32 / QPL lanes per entry, 256 / S entries at a time, each group reading its
// own task's prefix.  emap (64 entries) and ebytes are the block's LDS.
// Returns false, with nothing counted, when the work item does not qualify
// (the same answer for every tile of its range).
template <int QPL>
__device__ bool short_rows(const SupportArgs &a, const LevelView &v, const Range &r, uint32_t t0, uint32_t colv,
                           uint32_t *emap, unsigned long long *ebytes, unsigned long long &bytes,
                           unsigned long long &cands) {
    constexpr uint32_t kQuads = kRetainWords / 2;  // uint4 per row
    constexpr uint32_t S = kQuads / QPL;           // lanes per entry
    constexpr uint32_t NG = 256 / S;               // entries at a time
    static_assert(S == 8 || S == 16, "group sums cover 8 or 16 lanes");
    const uint32_t tid = threadIdx.x;
    // the range's tasks, one per lane: entry -> task map, qualification,
    // algorithmic bytes (entry row, prefix rows, retained row written)
    if (tid == 0) *ebytes = 0;
    __syncthreads();
    int ok = 1;
    unsigned long long my_bytes = 0;
    const uint32_t prefix_rows = v.d == 0 ? 0u : v.d <= 2 ? v.d : 1u;
    for (uint32_t t = t0 + tid; t < v.n_tasks && v.ext_off[t] < r.r1; t += blockDim.x) {
        const uint32_t e0 = max(v.ext_off[t], r.r0), e1 = min(v.ext_off[t + 1], r.r1);
        if (e0 >= e1) continue;
        const cptr<DbDev> db = v.dbs + v.dbi[t];
        const bool kept = v.rin && v.src[t] < v.slots;
        if (db->stride != kRetainWords || (v.d > 2 && !kept)) ok = 0;
        for (uint32_t e = e0; e < e1; ++e) {
            emap[e - r.r0] = t;
            my_bytes += 8ull * db->words * (1 + prefix_rows + (a.rout && e < v.slots ? 1 : 0));
        }
    }
    if (my_bytes && r.tile == 0) atomicAdd(ebytes, my_bytes);
    if (!__syncthreads_and(ok)) return false;
    if (r.tile == 0) {
        bytes += *ebytes;
        cands += r.r1 - r.r0;
        const uint32_t g = tid / S, sl = tid % S;
        const uint32_t ne = r.r1 - r.r0;
        uint4 *const keep = reinterpret_cast<uint4 *>(a.rout);
        for (uint32_t base = 0; base < ne; base += NG) {  // the same trip count in every lane
            const uint32_t i = base + g;
            const bool live = i < ne;
            const uint32_t ii = live ? i : ne - 1;
            const uint32_t e = r.r0 + ii;
            const uint32_t t = emap[ii];
            const uint32_t col = (uint32_t)__shfl((int)colv, (int)ii, 64);
            const cptr<DbDev> db = v.dbs + v.dbi[t];
            const gptr<uint4> store = to_global(reinterpret_cast<const uint4 *>(db->store));
            const uint32_t q0 = sl * QPL;
            uint4 P[QPL];
            if (v.d == 0) {
#pragma unroll
                for (int q = 0; q < QPL; ++q) P[q] = make_uint4(~0u, ~0u, ~0u, ~0u);
            } else if (v.d <= 2) {
                const cptr<uint32_t> pre = v.prefix + (uint64_t)t * kMaxDepth;
                const gptr<uint4> r0 = store + (uint64_t)col_of(v.fcol, pre[0]) * kQuads;
                const gptr<uint4> r1 = store + (uint64_t)col_of(v.fcol, pre[v.d - 1]) * kQuads;
#pragma unroll
                for (int q = 0; q < QPL; ++q) P[q] = and4(r0[q0 + q], r1[q0 + q]);
            } else {
                const gptr<uint4> row = to_global(reinterpret_cast<const uint4 *>(v.rin)) + (uint64_t)v.src[t] * kQuads;
#pragma unroll
                for (int q = 0; q < QPL; ++q) P[q] = row[q0 + q];
            }
            const gptr<uint4> xr = store + (uint64_t)col * kQuads;
            uint32_t c = 0;
            uint4 X[QPL];
#pragma unroll
            for (int q = 0; q < QPL; ++q) {
                X[q] = and4(P[q], xr[q0 + q]);
                c += pop4(X[q]);
            }
            if (live && keep && e < v.slots) {
#pragma unroll
                for (int q = 0; q < QPL; ++q) keep[(uint64_t)e * kQuads + q0 + q] = X[q];
            }
            // sum over the group's S lanes (quad swaps, half-row mirror, and
            // for 16 lanes the row mirror)
            c += (uint32_t)__builtin_amdgcn_mov_dpp((int)c, 0xb1, 0xf, 0xf, false);
            c += (uint32_t)__builtin_amdgcn_mov_dpp((int)c, 0x4e, 0xf, 0xf, false);
            c += (uint32_t)__builtin_amdgcn_mov_dpp((int)c, 0x141, 0xf, 0xf, false);
            if (S == 16) c += (uint32_t)__builtin_amdgcn_mov_dpp((int)c, 0x140, 0xf, 0xf, false);
            if (live && sl == 0) a.L.sup[e] = c;
        }
    }
    __syncthreads();  // emap and ebytes are read before the next work item rewrites them
    return true;
}

// The register hint keeps the 1024-word kernel without retained rows at 7
// waves per SIMD (72 VGPRs; the persistent grid is 7 blocks per CU) and the
// others at 4 at most.
template <int QPL, bool KEEP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(QPL == 2 && !KEEP ? 7 : 4)))
void k_support_count(SupportArgs a) {
    LevelView v;
    if (!level_view(a, v)) return;
    uint32_t *__restrict__ sup = a.L.sup;
    __shared__ uint32_t part[16][64];
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    __shared__ uint32_t emap[64];                // short rows: entry -> task
    __shared__ unsigned long long ebytes;
    const uint32_t slot = tid >> 4;              // wave * 4 + row of 16 lanes
    const bool leader = (lane & 15) == 0;
    constexpr uint32_t kQuadsPerTile = 256 * QPL;
    // entries counted per step: row loads in flight per lane = U x QPL
    constexpr int U = QPL == 2 ? 4 : 2;
    // a row of 16 lanes counts at most 16 x 128 x QPL bits: two entries'
    // counts share one 32-bit word through the row sums
    static_assert(16 * 128 * QPL < 65536, "packed row counts");
    const unsigned long long t_start = wall_clock64();
    unsigned long long bytes = 0, cands = 0;
    // work item `it`, the k-th this block takes
    auto item = [&](uint64_t it, uint64_t k) {
        const Range r = range_of(v, it, k);
        if (range_empty(v, it, r.tile)) return;  // the same answer in every lane
        const uint32_t ne = r.r1 - r.r0;  // <= 64 (the group is at most 64)
        const uint32_t colv = lane < ne ? col_of(v.fcol, v.ext[r.r0 + lane]) : 0u;
        uint32_t t = range_task(v, it, r.r0);
        if (KEEP && short_rows<QPL>(a, v, r, t, colv, emap, &ebytes, bytes, cands)) return;
        // The AND of the first d - 1 prefix rows of the last task built here
        // (its parent's itemset), reused while the next tasks are its
        // siblings: children of one task are stored next to each other.
        uint4 PP[QPL];
        bool pp_ok = false;      // PP holds a parent's rows for this tile
        uint32_t pp_par = 0;     // that parent
        for (uint32_t e0 = r.r0; e0 < r.r1;) {
            while (v.ext_off[t + 1] <= e0) ++t;  // empty tasks
            const uint32_t e1 = min(r.r1, v.ext_off[t + 1]);
            const cptr<DbDev> db = v.dbs + v.dbi[t];
            if (r.tile == 0) cands += e1 - e0;
            if (r.tile >= db->tiles) {
                if (leader)
                    for (uint32_t e = e0; e < e1; ++e) part[slot][e - r.r0] = 0;
                e0 = e1;
                continue;
            }
            const gptr<uint4> store = to_global(reinterpret_cast<const uint4 *>(db->store));
            const uint64_t stride_q = db->stride / 2;
            const uint64_t q0 = (uint64_t)r.tile * kQuadsPerTile + tid;
            // Row loads are buffer loads: the row is a buffer resource (its
            // scalar base, num_records = its bytes) and a lane passes a 32-bit
            // byte offset, so a load costs no 64-bit address arithmetic, and
            // a lane past the row's end reads zero (an out-of-range buffer
            // load returns 0) with no branch or clamp of its own; zero prefix
            // words zero those lanes' popcounts.  (A row is under 2^32 bytes.)
            // off/valid: the clamped quad index, for the retained-row stores.
            const uint32_t row_bytes = (uint32_t)(stride_q * sizeof(uint4));
            uint64_t off[QPL];
            uint32_t boff[QPL];
            bool valid[QPL];
#pragma unroll
            for (int q = 0; q < QPL; ++q) {
                const uint64_t o = q0 + (uint64_t)q * 256;
                valid[q] = o < stride_q;
                off[q] = valid[q] ? o : stride_q - 1;
                boff[q] = (uint32_t)((valid[q] ? o : stride_q) * sizeof(uint4));
            }
            auto at = [&](gptr<uint4> row, int q) { return row_load(row, row_bytes, boff[q]); };
            auto row_of = [&](uint32_t f) { return store + (uint64_t)col_of(v.fcol, f) * stride_q; };
            // entry e's row (e in this work item's range)
            auto entry_row = [&](uint32_t e) {
                return store + (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)colv, (int)(e - r.r0)) * stride_q;
            };
            const cptr<uint32_t> pre = v.prefix + (uint64_t)t * kMaxDepth;
            // one-stride rows keep every counted itemset's row: read the
            // prefix's from the previous launch when it has one
            const bool kept_rows = KEEP && db->stride == kRetainWords;
            const uint32_t src = kept_rows && v.rin ? v.src[t] : ~0u;
            const bool from_kept = KEEP && src < v.slots;
            uint4 P[QPL];
            uint32_t rows_read = e1 - e0;
            if (from_kept) {
                const gptr<uint4> row = to_global(reinterpret_cast<const uint4 *>(v.rin)) + (uint64_t)src * stride_q;
#pragma unroll
                for (int q = 0; q < QPL; ++q) P[q] = at(row, q);
#pragma unroll
                for (int q = 0; q < QPL; ++q)
                    if (!valid[q]) P[q] = make_uint4(0, 0, 0, 0);
                rows_read += 1;
                pp_ok = false;
            } else {
            // the parent PP was built for (siblings share the database and the
            // first d - 1 prefix items); a prefix of at most one item has
            // nothing to share
            const uint32_t par = v.siblings && v.d >= 2 ? v.par[t] : ~0u;
            const bool same = pp_ok && par != ~0u && par == pp_par;
            // the last prefix row first, so its load overlaps the parent's
            uint4 last[QPL];
#pragma unroll
            for (int q = 0; q < QPL; ++q) last[q] = make_uint4(~0u, ~0u, ~0u, ~0u);
            if (v.d) {
                const gptr<uint4> row = row_of(pre[v.d - 1]);
#pragma unroll
                for (int q = 0; q < QPL; ++q) last[q] = at(row, q);
                rows_read += 1;
            }
            if (!same) {
#pragma unroll
                for (int q = 0; q < QPL; ++q)
                    PP[q] = valid[q] ? make_uint4(~0u, ~0u, ~0u, ~0u) : make_uint4(0, 0, 0, 0);
                const uint32_t np = v.d ? v.d - 1 : 0;  // parent rows, two in flight
                uint32_t j = 0;
                for (; j + 2 <= np; j += 2) {
                    const gptr<uint4> r0 = row_of(pre[j]), r1 = row_of(pre[j + 1]);
                    uint4 x[QPL], y[QPL];
#pragma unroll
                    for (int q = 0; q < QPL; ++q) {
                        x[q] = at(r0, q);
                        y[q] = at(r1, q);
                    }
#pragma unroll
                    for (int q = 0; q < QPL; ++q) PP[q] = and4(PP[q], and4(x[q], y[q]));
                }
                if (j < np) {
                    const gptr<uint4> r0 = row_of(pre[j]);
#pragma unroll
                    for (int q = 0; q < QPL; ++q) PP[q] = and4(PP[q], at(r0, q));
                }
                rows_read += np;
            }
            pp_ok = true;
            pp_par = par;
#pragma unroll
            for (int q = 0; q < QPL; ++q) P[q] = and4(PP[q], last[q]);
            }
            // this launch's itemset rows, for the next level's prefixes
            uint4 *const keep = reinterpret_cast<uint4 *>(v.rout);
            const bool put = kept_rows && v.rout;
            const uint64_t tile_w = tile_words_of(db->words, r.tile, 2 * kQuadsPerTile);
            bytes += 8ull * tile_w * rows_read;
            // NE entries from e: loads first (NE x QPL in flight per lane),
            // then the counts, packed two per row sum
            auto count = [&](uint32_t e, auto ne_c) {
                constexpr int NE = decltype(ne_c)::value;
                uint4 X[NE][QPL];
#pragma unroll
                for (int u = 0; u < NE; ++u) {
                    const gptr<uint4> row = entry_row(e + u);
#pragma unroll
                    for (int q = 0; q < QPL; ++q) X[u][q] = at(row, q);
                }
                uint32_t c[NE];
#pragma unroll
                for (int u = 0; u < NE; ++u) {
                    c[u] = 0;
#pragma unroll
                    for (int q = 0; q < QPL; ++q) {
                        X[u][q] = and4(P[q], X[u][q]);
                        c[u] += pop4(X[u][q]);
                    }
                }
                if (put) {
#pragma unroll
                    for (int u = 0; u < NE; ++u)
                        if (e + u < v.slots) {
#pragma unroll
                            for (int q = 0; q < QPL; ++q)
                                if (valid[q]) keep[(uint64_t)(e + u) * stride_q + off[q]] = X[u][q];
                            bytes += 8ull * tile_w;
                        }
                }
#pragma unroll
                for (int u = 0; u + 1 < NE; u += 2) {
                    const uint32_t s = dpp_row_sum(c[u] | (c[u + 1] << 16));
                    if (leader) {
                        part[slot][e + u - r.r0] = s & 0xffffu;
                        part[slot][e + u + 1 - r.r0] = s >> 16;
                    }
                }
                if (NE & 1) {
                    const uint32_t s = dpp_row_sum(c[NE - 1]);
                    if (leader) part[slot][e + NE - 1 - r.r0] = s;
                }
            };
            uint32_t e = e0;
            for (; e + U <= e1; e += U) count(e, std::integral_constant<int, U>{});
            if (U > 2 && e + 2 <= e1) {
                count(e, std::integral_constant<int, 2>{});
                e += 2;
            }
            if (e < e1) count(e, std::integral_constant<int, 1>{});
            e0 = e1;
        }
        __syncthreads();
        if (tid < ne) {
            uint32_t c = 0;
#pragma unroll
            for (int s = 0; s < 16; ++s) c += part[s][tid];
            if (c) atomicAdd(sup + r.r0 + tid, c);
        }
        __syncthreads();
    };
    // items are dealt statically: block b takes b, b + grid, ...
    for (uint64_t it = blockIdx.x, k = 0; it < v.total; it += gridDim.x, ++k) item(it, k);
    launch_epilogue(a, t_start, bytes, cands);
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
    const int64_t g = a.gap > b.gap ? a.gap : b.gap;
    const int64_t x = b.first - a.last;
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

// in-order reduction over the 64 lanes (lane i covers earlier rows than lane
// i+1); the result is valid in lane 0
__device__ __forceinline__ PeriodPart pp_wave_reduce(PeriodPart m, uint32_t lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const PeriodPart o = pp_shfl_down(m, off);
        if (lane + off < 64) m = pp_combine(m, o);
    }
    return m;
}

// Fold a lane's 8 words (rows row0 .. row0 + 511, in order) into m, four
// words at a time.  The timestamps of each word's first and last occurrence
// are gathered first, eight loads in flight together, which fixes first,
// last, count and every gap between words (and inside words of two
// occurrences).  A word of three or more occurrences has inner gaps, each at
// most its span (last - first: timestamps increase); its occurrences are
// walked only when that span exceeds the largest gap already known, which
// then bounds them.
__device__ __forceinline__ void pp_scan_lane(PeriodPart &m, const uint64_t (&X)[8], uint64_t row0,
                                             gptr<int64_t> ts, uint64_t n) {
#pragma unroll 1
    for (int h = 0; h < 8; h += 4) {
        uint64_t x[4];
        uint32_t c[4];
        int64_t tf[4], tl[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint64_t r = row0 + 64ull * (h + v);  // rows at and past n hold no occurrence
            const uint64_t xv = h == 0 ? X[v] : X[4 + v];
            x[v] = r >= n ? 0ull : n - r >= 64 ? xv : xv & ((1ull << (n - r)) - 1);
            c[v] = (uint32_t)__popcll(x[v]);
            const uint64_t rf = x[v] ? r + (uint64_t)__builtin_ctzll(x[v]) : 0;
            const uint64_t rl = x[v] ? r + 63 - (uint64_t)__builtin_clzll(x[v]) : 0;
            tf[v] = ts[rf];
            tl[v] = ts[rl];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (!c[v]) continue;
            if (m.count == 0) {
                m.first = tf[v];
            } else {
                const int64_t g = tf[v] - m.last;
                if (g > m.gap) m.gap = g;
            }
            if (c[v] == 2 && tl[v] - tf[v] > m.gap) m.gap = tl[v] - tf[v];
            m.last = tl[v];
            m.count += c[v];
        }
#pragma unroll 1
        for (int v = 0; v < 4; ++v) {
            if (c[v] < 3 || tl[v] - tf[v] <= m.gap) continue;
            // the word's inner gaps: its occurrences after the first, four
            // timestamp loads at a time
            uint64_t w = x[v] & (x[v] - 1);
            int64_t prev = tf[v];
            const uint64_t r = row0 + 64ull * (h + v);
            while (w) {
                uint64_t rr[4];
                bool ok[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    ok[j] = w != 0;
                    rr[j] = r + (ok[j] ? (uint64_t)__builtin_ctzll(w) : 0);
                    w &= w - 1;
                }
                int64_t t[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) t[j] = ts[rr[j]];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (!ok[j]) continue;
                    if (t[j] - prev > m.gap) m.gap = t[j] - prev;
                    prev = t[j];
                }
            }
        }
    }
}

// Blocked lanes: lane `tid` owns words tile*2048 + tid*8 .. +8, so lane order
// is row order for the in-order reductions.
__global__ __launch_bounds__(256) void k_support_periodic(SupportArgs a) {
    __shared__ PeriodPart wpart[4][16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    LevelView lv;
    if (!level_view(a, lv)) return;
    const unsigned long long t_start = wall_clock64();
    unsigned long long bytes = 0, cands = 0;
    for (uint64_t it = blockIdx.x, k = 0; it < lv.total; it += gridDim.x, ++k) {
        const Range r = range_of(lv, it, k);
        uint32_t t = wave_find_task(lv.ext_off, lv.n_tasks, r.r0);  // periodic searches are one database: no range table
        for (uint32_t e0 = r.r0; e0 < r.r1;) {
            while (lv.ext_off[t + 1] <= e0) ++t;  // empty tasks
            const uint32_t e1 = min(r.r1, lv.ext_off[t + 1]);
            const cptr<DbDev> db = lv.dbs + lv.dbi[t];
            if (r.tile == 0) cands += e1 - e0;
            if (r.tile >= db->tiles) {  // the fold reads tiles < db->tiles only
                e0 = e1;
                continue;
            }
            bytes += 8ull * tile_words_of(db->words, r.tile, kTileWords) * (lv.d + e1 - e0);
            const gptr<uint64_t> store = to_global(db->store);
            const gptr<int64_t> ts = to_global(db->ts);
            const uint64_t n = db->n;
            const uint64_t stride = db->stride;
            const uint64_t w0 = (uint64_t)r.tile * kTileWords + (uint64_t)tid * 8;
            const bool valid = w0 < stride;  // stride is a multiple of 64 words
            uint64_t P[8];
#pragma unroll
            for (int v = 0; v < 8; ++v) P[v] = valid ? ~0ull : 0ull;
            const cptr<uint32_t> pre = lv.prefix + (uint64_t)t * kMaxDepth;
            if (valid) {
                for (uint32_t j = 0; j < lv.d; ++j) {
                    const gptr<uint64_t> row = store + (uint64_t)col_of(lv.fcol, pre[j]) * stride + w0;
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
                        const gptr<uint64_t> row = store + (uint64_t)col_of(lv.fcol, lv.ext[e]) * stride + w0;
                        uint64_t X[8];
#pragma unroll
                        for (int v = 0; v < 8; ++v) X[v] = P[v] & row[v];
                        pp_scan_lane(m, X, w0 * 64, ts, n);
                    }
                    m = pp_wave_reduce(m, lane);
                    if (lane == 0) wpart[wave][e - eb] = m;
                }
                __syncthreads();
                if (tid < ee - eb) {
                    const PeriodPart m = pp_combine(pp_combine(wpart[0][tid], wpart[1][tid]),
                                                    pp_combine(wpart[2][tid], wpart[3][tid]));
                    a.parts[(uint64_t)(eb + tid) * a.parts_tiles + r.tile] = m;
                }
                __syncthreads();
            }
            e0 = e1;
        }
    }
    launch_epilogue(a, t_start, bytes, cands);
}

// Fold the tile partials of each entry in row order: support and the largest
// gap of [0, ts..., end] (_temporal.py:36-54: prev starts at 0, end gap last).
__global__ __launch_bounds__(256) void k_periodic_finalize(SupportArgs a) {
    if (a.stop && *a.stop) return;
    const uint32_t n_tasks = (uint32_t)a.L.sz->n_tasks;
    if (n_tasks == 0) return;
    const uint64_t n_ext = a.L.sz->n_ext;
    // A periodic search holds one database (tasks are in database order, so
    // the first and last tasks' agree): then no entry searches for its task.
    const bool one_db = a.L.db[0] == a.L.db[n_tasks - 1];
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_ext;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const DbDev *db = a.dbs + a.L.db[one_db ? 0u : find_task(to_const(a.L.ext_off), n_tasks, e)];
        const uint32_t tiles = db->tiles;
        const PeriodPart *pe = a.parts + e * a.parts_tiles;
        PeriodPart m;
        pp_empty(m);
        // the tiles' partials four at a time (their loads in flight
        // together), folded in tile order
        constexpr uint32_t kB = 4;
        for (uint32_t c0 = 0; c0 < tiles; c0 += kB) {
            int64_t f[kB], l[kB], g[kB];
            uint32_t n[kB];
#pragma unroll
            for (uint32_t i = 0; i < kB; ++i) {
                const PeriodPart *x = pe + (c0 + i < tiles ? c0 + i : c0);
                f[i] = x->first;
                l[i] = x->last;
                g[i] = x->gap;
                n[i] = x->count;
            }
#pragma unroll
            for (uint32_t i = 0; i < kB; ++i) {
                if (c0 + i >= tiles || n[i] == 0) continue;
                if (m.count == 0) {
                    m.first = f[i];
                    m.gap = g[i];
                } else {
                    const int64_t x = f[i] - m.last;
                    m.gap = max(max(m.gap, g[i]), x);
                }
                m.last = l[i];
                m.count += n[i];
            }
        }
        int64_t p;
        if (m.count == 0) {
            p = db->end;
        } else {
            p = m.first;
            if (m.gap > p) p = m.gap;
            if (db->end - m.last > p) p = db->end - m.last;
        }
        a.L.sup[e] = m.count;
        a.L.per[e] = p;
    }
}

// ---- range table ------------------------------------------------------------------

// One wave per work range: its first task by a 64-ary search over the task
// offsets (each step narrows the interval 64-fold with one load per lane;
// three steps for 2^18 tasks), then the tiles of its tasks' databases, 64
// tasks per load.
__global__ __launch_bounds__(256) void k_ranges(SupportArgs a) {
    LevelView v;
    if (!level_view(a, v) || !a.ranges) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t n_r = ((uint64_t)v.n_ext + v.G - 1) / v.G;
    const uint64_t n = n_r < a.rcap ? n_r : a.rcap;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t g = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < n; g += waves) {
        const uint32_t r0 = (uint32_t)(g * v.G);
        const uint32_t r1 = min(v.n_ext, r0 + v.G);
        // the largest t < n_tasks with ext_off[t] <= r0 lies in [lo, hi)
        // (ext_off[0] = 0 <= r0 < ext_off[n_tasks] = n_ext)
        uint32_t lo = 0, hi = v.n_tasks;
        while (hi - lo > 1) {
            const uint32_t step = (hi - lo + 63) / 64;
            const uint32_t p = lo + lane * step;
            const unsigned long long le = __ballot(p < hi && v.ext_off[p] <= r0);  // a prefix of lanes, lane 0 set
            const uint32_t k = 63u - (uint32_t)__builtin_clzll(le);
            lo += k * step;
            hi = min(hi, lo + step);
        }
        uint32_t tiles = 0;
        for (uint32_t base = lo;; base += 64) {
            const uint32_t t = base + lane;
            const bool in = t < v.n_tasks && v.ext_off[t] < r1;
            if (in && v.ext_off[t + 1] > v.ext_off[t]) tiles = max(tiles, v.dbs[v.dbi[t]].tiles);
            if (__ballot(in) != ~0ull) break;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tiles = max(tiles, (uint32_t)__shfl_xor((int)tiles, off, 64));
        if (lane == 0) {
            a.ranges[g] = lo;
            a.ranges[a.rcap + g] = tiles;
        }
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

}  // namespace

// Persistent grid of a support kernel, in blocks per CU (a multiple of 8 in
// all, so tile == XCD), at most `grid`.  Measured on MI355X: the 1024-word
// kernel holds 8 blocks per CU and runs best at 7 (a full grid waits for
// slots the previous launch's last waves still hold); the 2048-word kernel
// holds 4 and runs best oversubscribed at 8 (late blocks take the leftover
// work); the periodic kernel (119 VGPRs: four waves per SIMD) runs best at 4.
uint32_t persistent_grid(const cpm_ctx *ctx, uint32_t per_cu, uint32_t grid) {
    const uint32_t g = (uint32_t)ctx->cu_count * per_cu / 8 * 8;
    return std::max<uint32_t>(8, std::min(grid, g));
}

int launch_support_count(cpm_ctx *ctx, const SupportArgs &a, uint32_t tile_words, uint32_t grid) {
    if (a.te == 0) return fail(CPM_ERR_INVALID, "zero tiles per work range");
    const bool keep = a.rout != nullptr;
    if (tile_words == 1024 && keep)
        hipLaunchKernelGGL((k_support_count<2, true>), dim3(persistent_grid(ctx, 7, grid)), dim3(256), 0,
                           ctx->stream, a);
    else if (tile_words == 1024)
        hipLaunchKernelGGL((k_support_count<2, false>), dim3(persistent_grid(ctx, 7, grid)), dim3(256), 0,
                           ctx->stream, a);
    else if (tile_words == 2048 && keep)
        hipLaunchKernelGGL((k_support_count<4, true>), dim3(persistent_grid(ctx, 8, grid)), dim3(256), 0,
                           ctx->stream, a);
    else if (tile_words == 2048)
        hipLaunchKernelGGL((k_support_count<4, false>), dim3(persistent_grid(ctx, 8, grid)), dim3(256), 0,
                           ctx->stream, a);
    else
        return fail(CPM_ERR_INVALID, "tile of %u words", tile_words);
    CPM_HIP(hipGetLastError());
    return CPM_OK;
}

int launch_support_periodic(cpm_ctx *ctx, const SupportArgs &a, uint32_t grid) {
    if (a.te == 0) return fail(CPM_ERR_INVALID, "zero tiles per work range");
    hipLaunchKernelGGL(k_support_periodic, dim3(persistent_grid(ctx, 4, grid)), dim3(256), 0,
                       ctx->stream, a);
    CPM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_periodic_finalize, dim3(grid), dim3(256), 0, ctx->stream, a);
    CPM_HIP(hipGetLastError());
    return CPM_OK;
}

int launch_ranges(cpm_ctx *ctx, const SupportArgs &a, uint32_t grid) {
    if (!a.ranges || a.te == 0) return CPM_OK;
    hipLaunchKernelGGL(k_ranges, dim3(std::max<uint32_t>(1, grid)), dim3(256), 0, ctx->stream, a);
    CPM_HIP(hipGetLastError());
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

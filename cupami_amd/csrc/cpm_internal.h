// Internal definitions shared by the libcupami_hip.so translation units.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdint.h>
#include <stdarg.h>
#include <stdio.h>
#include <memory>
#include <mutex>
#include <new>
#include <utility>
#include <string>
#include <vector>

#include "../../include/cupami_hip.h"

namespace cpm {

// ---- errors -----------------------------------------------------------------

int fail(int code, const char *fmt, ...);
void clear_error();

#define CPM_HIP(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ::cpm::fail(e_ == hipErrorOutOfMemory ? CPM_ERR_NOMEM : CPM_ERR_RUNTIME, \
                               "%s failed: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

#define CPM_CHECK(expr)                \
    do {                               \
        int s_ = (expr);               \
        if (s_ != CPM_OK) return s_;   \
    } while (0)

// ---- device buffers -----------------------------------------------------------

// Grow-only device allocation owned by a context (or a handle).
struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);       // keeps contents only if no growth is needed
    void release();
    ~DBuf() { release(); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Held around device / pinned allocation and release and around every
// graph capture, process-wide (core.hip).
std::recursive_mutex &alloc_capture_mutex();

// Pinned host staging buffer.
struct HBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    void release();
    ~HBuf() { release(); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Row stride of the vertical store, in 64-bit words: rows are padded with
// zero words to a multiple of 64 words (512 bytes).
inline uint64_t row_stride(uint64_t words) { return (words + 63) / 64 * 64; }

// Words handled by one workgroup of the support kernels (256 lanes x 8 words).
constexpr uint32_t kTileWords = 2048;
constexpr uint32_t kBlock = 256;

// Words per workgroup of the LDS tile store builds: the tile (n_items x B
// words) stays within 32 KB up to 4096 items (four 512-lane workgroups fit a
// CU's LDS), one word (56 KB, under 64 KB with the kernels' own LDS) up to
// 7168; 0 above (global atomics).
inline uint32_t tile_words_for(uint64_t n_items) {
    if (n_items == 0 || n_items > 7168) return 0;
    return n_items <= 4096 ? (uint32_t)std::min<uint64_t>(8, 4096 / n_items) : 1;
}

// Workgroup order for builds whose workgroups own consecutive word ranges:
// workgroups run on XCD blockIdx % 8, and each XCD takes one contiguous run of
// the ranges, so the partial 128-byte lines of neighbouring workgroups meet in
// the same L2 before they are written back.  A bijection on [0, gridDim.x).
__device__ __forceinline__ uint32_t xcd_contiguous_block() {
    const uint32_t G = gridDim.x, x8 = blockIdx.x % 8, j = blockIdx.x / 8, per = G / 8, rem = G % 8;
    return x8 < rem ? x8 * (per + 1) + j : rem * (per + 1) + (x8 - rem) * per + j;
}

}  // namespace cpm

struct cpm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 256;
    bool timing = true;
    // [0..8) the device search (cpm_stats), [8..12) the store builds from
    // basket arrays (cpm_ingest_stats)
    double stats[12] = {};
    // event pairs of timed launches not read yet (read by cpm_stats), and
    // spare events
    struct TimedPair {
        int slot;
        hipEvent_t start, stop;
    };
    std::vector<TimedPair> ev_pending;
    std::vector<hipEvent_t> ev_pool;
    // scratch used by the operator entry points
    cpm::DBuf scratch[6];
    cpm::HBuf host_scratch;
    // search buffers (miner.hip), reused across searches
    void *search_ws = nullptr;
    void (*search_ws_free)(void *) = nullptr;
    // store build from basket arrays (ingest.hip)
    void *ingest_ws = nullptr;
    void (*ingest_ws_free)(void *) = nullptr;
    // allocation of the last destroyed store, reused by the next store build
    cpm::DBuf spare_store;
    cpm::DBuf spare_arena;   // batch-build arena kept for the next batch
    cpm::HBuf bounce;        // pinned bounce buffer of copy_to_host
    // search schedule (cpm_set_search_options)
    uint32_t opt_group = 32, opt_tile_words = 1024;
    bool opt_graph = true, opt_prune = true;
};

namespace cpm {
// Device -> host copies into any host memory (pageable included): staged
// through the context's pinned bounce buffer in chunks, each chunk copied
// out after a stream synchronization, so the data is in place on return
// whatever the runtime does with pageable destinations.
int copy_to_host(cpm_ctx *ctx, void *dst, const void *src, size_t bytes);
int copy_to_host_2d(cpm_ctx *ctx, void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                    size_t height);
// Host -> device upload of `rows` rows of `words` uint64 from any host
// memory into rows `stride` words apart, the padding zeroed: rows are padded
// on the host in the pinned bounce buffer and sent in chunks (no device
// memset, no pageable source handed to the runtime).
int upload_rows(cpm_ctx *ctx, uint64_t *dst, uint64_t stride, const uint64_t *src, uint64_t words,
                uint64_t rows);

// Device memory shared by the stores of one batch build; handed back to its
// context for the next batch when the last store goes.
struct StoreArena {
    cpm_ctx *ctx = nullptr;
    DBuf buf;
    ~StoreArena();
};
}  // namespace cpm

struct cpm_bits {
    cpm_ctx *ctx = nullptr;
    uint64_t n_transactions = 0;
    uint32_t n_items = 0;
    uint64_t words = 0;    // ceil(N/64)
    uint64_t stride = 0;   // padded words per row
    cpm::DBuf store;       // n_items * stride uint64 (own rows) ...
    std::shared_ptr<cpm::StoreArena> arena;  // ... or rows inside a batch arena
    uint64_t *view = nullptr;
    uint64_t *rows() const { return view ? view : store.as<uint64_t>(); }
    // item id of each column, for stores built from basket arrays
    cpm::DBuf labels;
    bool has_labels = false;
};

namespace cpm {
// Large host blocks are pinned and recycled through a small process-wide
// cache, so a search's result arrays are copied from the device straight
// into memory that is already mapped (fresh pages cost the kernel a zeroing
// pass: ~9 ms per 120 MB, and pinning costs more).  Blocks under 1 MiB go
// straight to malloc / free.
void *host_block_take(size_t bytes);
void host_block_give(void *p, size_t bytes);  // bytes: as passed to host_block_take
// Allocator over that cache; elements are default-initialised, so a resize
// writes nothing before the real data does.
template <class T>
struct HostCache {
    using value_type = T;
    HostCache() = default;
    template <class U>
    HostCache(const HostCache<U> &) {}
    T *allocate(size_t n) {
        void *p = host_block_take(n * sizeof(T));
        if (!p) throw std::bad_alloc();
        return static_cast<T *>(p);
    }
    void deallocate(T *p, size_t n) { host_block_give(p, n * sizeof(T)); }
    template <class U>
    void construct(U *p) noexcept {
        ::new (static_cast<void *>(p)) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
    template <class U>
    bool operator==(const HostCache<U> &) const { return true; }
    template <class U>
    bool operator!=(const HostCache<U> &) const { return false; }
};
template <class T>
using hvec = std::vector<T, HostCache<T>>;
}  // namespace cpm

struct cpm_patterns {
    uint64_t n = 0;
    bool periodic = false;
    cpm::hvec<uint32_t> parent, item, db;
    cpm::hvec<uint64_t> support;
    cpm::hvec<int64_t> period;
};

namespace cpm {

// Make the context's device current for the calling thread.
int bind(cpm_ctx *ctx);
// Records the device address of the no-allocation entry point's scratch
// array for `device` (the current device); once per device (search.hip).
int noalloc_setup(int device);
// A store of n_items rows for n_trans transactions, zeroed when `clear`.
int bits_alloc(cpm_ctx *ctx, uint64_t n_trans, uint32_t n_items, cpm_bits **out, bool clear = true);

// Event-timed launch bracket: accumulates milliseconds into stats[slot].
// Event timing of one launch: two events on the stream, read later by
// collect_timing (no synchronization here).
hipEvent_t take_event(cpm_ctx *ctx);
int collect_timing(cpm_ctx *ctx);
struct Timed {
    cpm_ctx *ctx;
    int slot;
    hipEvent_t start = nullptr;
    Timed(cpm_ctx *c, int s) : ctx(c), slot(s) {
        if (c->timing && (start = take_event(c)) != nullptr && hipEventRecord(start, c->stream) != hipSuccess) {
            ctx->ev_pool.push_back(start);
            start = nullptr;
        }
    }
    // a bracket left open by an early return gives its event back
    ~Timed() {
        if (start) ctx->ev_pool.push_back(start);
    }
    Timed(const Timed &) = delete;
    Timed &operator=(const Timed &) = delete;
    int finish() {
        if (!start) return CPM_OK;
        hipEvent_t stop = take_event(ctx);
        if (!stop) {
            ctx->ev_pool.push_back(start);
            return CPM_OK;
        }
        CPM_HIP(hipEventRecord(stop, ctx->stream));
        ctx->ev_pending.push_back({slot, start, stop});
        start = nullptr;
        // bound the backlog of a context whose statistics are never read
        if (ctx->ev_pending.size() >= 256) return collect_timing(ctx);
        return CPM_OK;
    }
};

// ---- launchers implemented in the kernel translation units --------------------

struct TaskLevel;  // miner.hip

// gram.hip ------------------------------------------------------------------
// One database of a batched pair-count launch: G[a][b] = |row(cols[a]) &
// row(cols[b])| for a <= b, written to out + out_off with leading dimension
// n_items (int32).
// Pointer into global memory.  Row pointers that a kernel reads out of a
// table (not from its arguments) are generic to the compiler, which then
// emits flat loads: those share the LDS wait counter, so every wait on them
// also drains LDS traffic.  Casting to the global address space restores
// global loads.
//
// Pointer into the constant address space: memory that no lane writes while
// the kernel runs.  Loads from it at wave-uniform addresses compile to scalar
// loads, whatever else the kernel does (a clock read or a global store
// ahead of a loop otherwise makes the compiler assume the memory may have
// changed and fall back to vector loads).
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = const __attribute__((address_space(1))) T *;
template <class T>
using cptr = const __attribute__((address_space(4))) T *;
#else
template <class T>
using gptr = const T *;  // host pass: kernels are parsed, never run
template <class T>
using cptr = const T *;
#endif

template <class T>
__device__ __forceinline__ gptr<T> to_global(const T *p) {
    return (gptr<T>)p;
}

template <class T>
__device__ __forceinline__ cptr<T> to_const(const T *p) {
    return (cptr<T>)p;
}

// ---- wave sums and scans on DPP ------------------------------------------------
// Data-parallel moves between lanes of a row of 16 inside one VALU
// instruction: no LDS round trip and no wait between steps (a __shfl is an
// LDS permute with a wait after it).  Every lane of a row the result is read
// from must be active (of the wave, for the wave-wide ones).

// Row sums: after the two quad swaps and the half-row and row mirrors every
// lane of a row of 16 holds the sum of its row.
__device__ __forceinline__ uint32_t dpp_row_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xb1, 0xf, 0xf, false);   // quad_perm [1,0,3,2]
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4e, 0xf, 0xf, false);   // quad_perm [2,3,0,1]
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xf, 0xf, false);  // row_half_mirror
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xf, 0xf, false);  // row_mirror
    return v;
}

// Wave sum, uniform across the wave (the four row sums read as scalars).
__device__ __forceinline__ uint32_t dpp_wave_sum(uint32_t v) {
    v = dpp_row_sum(v);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 0) + (uint32_t)__builtin_amdgcn_readlane((int)v, 16) +
           (uint32_t)__builtin_amdgcn_readlane((int)v, 32) + (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
}

// Inclusive scan inside each row of 16 lanes (lanes shifted in from outside
// the row read 0).
__device__ __forceinline__ uint32_t dpp_row_scan_incl(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);  // row_shr:8
    return x;
}

// Row sums of 64-bit values (each half moved, then a 64-bit add).
__device__ __forceinline__ unsigned long long dpp_row_sum64(unsigned long long v) {
    auto step = [&](auto ctrl) {
        constexpr int c = decltype(ctrl)::value;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)v, c, 0xf, 0xf, false);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(v >> 32), c, 0xf, 0xf, false);
        v += ((unsigned long long)hi << 32) | lo;
    };
    step(std::integral_constant<int, 0xb1>{});
    step(std::integral_constant<int, 0x4e>{});
    step(std::integral_constant<int, 0x141>{});
    step(std::integral_constant<int, 0x140>{});
    return v;
}

__device__ __forceinline__ unsigned long long dpp_wave_sum64(unsigned long long v) {
    v = dpp_row_sum64(v);
    unsigned long long t = 0;
#pragma unroll
    for (int r = 0; r < 64; r += 16)
        t += ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), r) << 32) |
             (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, r);
    return t;
}

// Inclusive scan over the wave: shifts by 1, 2, 4 and 8 lanes inside each row
// (lanes shifted in from outside the row read 0), then the last lane of rows
// 0 and 2 added into rows 1 and 3, then the last lane of row 1 into rows 2
// and 3 (rows outside a move's row mask take 0).
__device__ __forceinline__ uint32_t dpp_wave_scan_incl(uint32_t x) {
    x = dpp_row_scan_incl(x);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false); // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false); // row_bcast:31
    return x;
}

struct GramDb {
    const uint64_t *store;
    uint64_t stride, words, n_transactions;
    const uint32_t *cols;
    uint32_t n_items, pad;
    uint64_t out_off;
};
// One workgroup's share of the pair matrix: a 256 x 256 block of database
// db over the K (word) range [k_begin, k_end).  A block of one slice writes
// its outputs (slot = kGramDirect).  A block cut into several K slices either
// stores each slice's sums in its own partial slot (256 x 256 int32), which
// k_gram's reduce adds, or adds them into the zeroed output with atomics
// (slot = kGramAtomic): slots pay off for blocks of many outputs, atomics
// for small matrices cut into many slices (a reduce thread would sum one
// output over every slice).
constexpr uint32_t kGramDirect = 0xffffffffu;
constexpr uint32_t kGramAtomic = 0xfffffffeu;
struct GramJob {
    uint32_t db, bi, bj, slot;
    uint64_t k_begin, k_end;
};
// a block of several slices: partial slots [slot0, slot0 + nslots)
struct GramBlock {
    uint32_t db, bi, bj, slot0, nslots, pad;
};
struct GramPlan {
    std::vector<GramJob> jobs;
    std::vector<GramBlock> blocks;
    uint64_t slots = 0;      // partial slots (256 x 256 int32 each)
    uint64_t out_elems = 0;  // sum of n_items^2
};
// Work table for a batch of databases.
int gram_plan(const cpm_ctx *ctx, const std::vector<GramDb> &dbs, GramPlan &plan);
// Device copies of a plan's tables and its partial slots.
struct GramDev {
    const GramJob *jobs = nullptr;
    uint64_t n_jobs = 0;
    const GramBlock *blocks = nullptr;
    uint64_t n_blocks = 0;
    int32_t *partials = nullptr;
};
// Run the jobs into `out` (the upper triangle of each database's matrix;
// the rest is left as it is, zeroed first when `clear`); n_ops feeds the
// statistics.
int launch_gram(cpm_ctx *ctx, const GramDev &g, const GramDb *d_dbs, int32_t *d_out, uint64_t out_elems,
                uint64_t n_ops, bool clear);

}  // namespace cpm

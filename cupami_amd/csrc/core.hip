// Context, errors, device buffers, the vertical store and the operator-level
// entry points (count_supports, intersect_extensions, periodic_eval).
#include "cpm_internal.h"

#include <mutex>
#include "support_kernels.h"

#include <string.h>
#include <algorithm>

namespace cpm {

static thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

void clear_error() { g_error.clear(); }

// Serialises device/pinned allocation and release with the graph captures
// of every context in the process: an allocation call that synchronizes the
// device must not run on one thread while another thread records a stream.
std::recursive_mutex &alloc_capture_mutex() {
    static std::recursive_mutex mu;
    return mu;
}

int DBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return CPM_OK;
    std::lock_guard<std::recursive_mutex> lock(alloc_capture_mutex());
    release();
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        return fail(CPM_ERR_NOMEM, "device allocation of %zu bytes failed: %s", want,
                    hipGetErrorString(e));
    }
    cap = want;
    return CPM_OK;
}

void DBuf::release() {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> lock(alloc_capture_mutex());
    (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

int HBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return CPM_OK;
    std::lock_guard<std::recursive_mutex> lock(alloc_capture_mutex());
    release();
    size_t want = std::max<size_t>(bytes, 4096);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        return fail(CPM_ERR_NOMEM, "pinned host allocation of %zu bytes failed: %s", want,
                    hipGetErrorString(e));
    }
    cap = want;
    return CPM_OK;
}

void HBuf::release() {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> lock(alloc_capture_mutex());
    (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

int bind(cpm_ctx *ctx) {
    if (!ctx) return fail(CPM_ERR_INVALID, "null context");
    CPM_HIP(hipSetDevice(ctx->device));
    return CPM_OK;
}

// ---- store build ------------------------------------------------------------

// One lane per COO pair: OR the bit into its 64-bit word.  Pairs that fall
// outside the store raise *bad instead of writing.
__global__ __launch_bounds__(256) void k_bits_scatter(const uint64_t *__restrict__ rows,
                                                      const uint32_t *__restrict__ cols,
                                                      uint64_t nnz, uint64_t n_trans,
                                                      uint32_t n_items, uint64_t stride,
                                                      unsigned long long *store,
                                                      unsigned int *bad) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (; i < nnz; i += step) {
        uint64_t r = rows[i];
        uint32_t c = cols[i];
        if (r >= n_trans || c >= n_items) {
            atomicOr(bad, 1u);
            continue;
        }
        atomicOr(&store[(uint64_t)c * stride + (r >> 6)], 1ull << (r & 63));
    }
}

// dst row i = src row idx[i] (words of each; rows `stride` words apart in src)
__global__ __launch_bounds__(256) void k_gather_rows(const uint64_t *__restrict__ src, uint64_t stride,
                                                    uint64_t words, const uint32_t *__restrict__ idx, uint32_t n,
                                                    uint64_t *__restrict__ dst) {
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint64_t *s = src + (uint64_t)idx[i] * stride;
        uint64_t *d = dst + (uint64_t)i * words;
        for (uint64_t w = threadIdx.x; w < words; w += blockDim.x) d[w] = s[w];
    }
}

// One database of a (batched) store build.
struct BuildJob {
    const uint64_t *rows;
    const uint32_t *cols;
    uint64_t nnz, n_trans, stride;
    uint64_t *store;        // tile build only
    unsigned int *flags;    // [0] rows not sorted into the store's range, [2] a column out of range
    uint32_t n_items, B;    // B: words per workgroup of the tile build
    uint64_t block0;        // first workgroup of this database in the tile launch
};

// Store build for row-sorted pairs, any number of databases per launch:
// workgroup b of database d owns words [b*B, b*B+B) of every item row of d.
// Its pairs are one contiguous run of the input; they are ORed into an LDS
// tile (n_items x B words) and the tile is written out once, so the store
// needs no clearing and HBM sees no atomics.
//
// The kernel checks its own input instead of a separate pass over it: each
// workgroup tests that its run is non-decreasing (its last pair against the
// next run's first too) and that its rows fall in its word range, and the
// last workgroup that the runs end at nnz.  The runs of consecutive
// workgroups meet (each end is the same search), so when no workgroup raises
// flags[0] the rows are sorted and below n_trans and the store is right;
// otherwise the caller builds that store again by scatter.
// First i in [0, n) with rows[i] >= x (n if none), for non-decreasing rows,
// by one wave: each step probes 64 positions that split the interval into 65
// parts, so 10^7 rows take four dependent loads.
__device__ uint64_t wave_lower_bound(const uint64_t *__restrict__ rows, uint64_t n, uint64_t x) {
    const uint32_t lane = threadIdx.x & 63;
    uint64_t lo = 0, hi = n;  // the answer lies in [lo, hi]
    while (hi - lo > 64) {
        const uint64_t span = hi - lo;
        const bool less = rows[lo + (uint64_t)(lane + 1) * span / 65] < x;
        const uint32_t cnt = (uint32_t)__popcll(__ballot(less));  // a prefix of the lanes
        const uint64_t nlo = cnt ? lo + (uint64_t)cnt * span / 65 + 1 : lo;
        const uint64_t nhi = cnt < 64 ? lo + (uint64_t)(cnt + 1) * span / 65 : hi;
        lo = nlo;
        hi = nhi;
    }
    const uint64_t q = lo + lane;
    return lo + (uint64_t)__popcll(__ballot(q < hi && rows[q] < x));
}

__global__ __launch_bounds__(512) void k_bits_tile(const BuildJob *__restrict__ jobs, uint32_t n_jobs) {
    extern __shared__ unsigned long long tile[];
    __shared__ uint64_t range[2];
    const uint32_t blk = xcd_contiguous_block();
    uint32_t lo_j = 0, hi_j = n_jobs;  // last job with block0 <= blk
    while (hi_j - lo_j > 1) {
        const uint32_t m = (lo_j + hi_j) >> 1;
        if (jobs[m].block0 <= blk) lo_j = m; else hi_j = m;
    }
    const BuildJob &j = jobs[lo_j];
    const uint64_t *__restrict__ rows = j.rows;
    const uint32_t *__restrict__ cols = j.cols;
    const uint32_t n_items = j.n_items, B = j.B;
    const uint64_t stride = j.stride, nnz = j.nnz;
    const uint32_t tid = threadIdx.x;
    const uint64_t w0 = (blk - j.block0) * B;
    const uint32_t cells = n_items * B;
    for (uint32_t i = tid; i < cells; i += blockDim.x) tile[i] = 0;
    // [lo, hi): pairs whose row lies in this word range, the two ends found
    // by waves 0 and 1 at once
    const uint64_t r_lo = w0 * 64, r_hi = min((w0 + B) * 64, j.n_trans);
    if (tid < 128) {
        const uint64_t b = wave_lower_bound(rows, nnz, tid < 64 ? r_lo : r_hi);
        if ((tid & 63) == 0) range[tid >> 6] = b;
    }
    __syncthreads();
    const uint64_t lo = range[0], lo2 = range[1];
    const bool last_block = blk + 1 == j.block0 + (stride + B - 1) / B;
    bool unsorted = lo2 < lo || (last_block && lo2 != nnz);
    constexpr int kPairBatch = 4;  // loads in flight per lane before its LDS atomics
    const uint32_t lane = tid & 63;
    for (uint64_t i0 = lo; i0 < lo2; i0 += (uint64_t)kPairBatch * blockDim.x) {
        uint64_t r[kPairBatch], nx[kPairBatch];
        uint32_t c[kPairBatch];
        bool v[kPairBatch];
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
            const uint64_t i = i0 + (uint64_t)q * blockDim.x + tid;
            v[q] = i < lo2;
            r[q] = v[q] ? rows[i] : 0;
            c[q] = v[q] ? cols[i] : 0;
            // the next row: the next lane's, or loaded at the end of a wave or of the run
            const bool own = v[q] && (lane == 63 || i + 1 == lo2) && i + 1 < nnz;
            nx[q] = own ? rows[i + 1] : ~0ull;
        }
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
            const uint64_t i = i0 + (uint64_t)q * blockDim.x + tid;
            const uint64_t down = __shfl_down(r[q], 1, 64);
            const uint64_t next = (lane == 63 || i + 1 == lo2) ? nx[q] : down;
            if (!v[q]) continue;
            const uint64_t wr = (r[q] >> 6) - w0;
            if (r[q] > next || wr >= B) {
                unsorted = true;
                continue;
            }
            if (c[q] >= n_items) {
                atomicOr(j.flags + 2, 1u);
                continue;
            }
            atomicOr(&tile[c[q] * B + (uint32_t)wr], 1ull << (r[q] & 63));
        }
    }
    if (unsorted) atomicOr(j.flags, 1u);
    __syncthreads();
    uint64_t *__restrict__ store = j.store;
    for (uint32_t i = tid; i < cells; i += blockDim.x) {
        const uint32_t c = i / B, k = i % B;
        if (w0 + k < stride) store[(uint64_t)c * stride + w0 + k] = tile[i];
    }
}

}  // namespace cpm

using namespace cpm;

extern "C" {

int cpm_version(void) { return 10000; }

const char *cpm_last_error(void) { return cpm::g_error.c_str(); }

int cpm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int cpm_ctx_create(int device, cpm_ctx **out) {
    clear_error();
    if (!out) return fail(CPM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int n = cpm_device_count();
    if (n <= 0) return fail(CPM_ERR_NODEVICE, "no HIP device was found");
    if (device < 0 || device >= n)
        return fail(CPM_ERR_INVALID, "device %d out of range (found %d)", device, n);
    CPM_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    CPM_HIP(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(CPM_ERR_NODEVICE, "device %d is %s; this library is built for gfx950", device,
                    prop.gcnArchName);
    CPM_CHECK(noalloc_setup(device));
    cpm_ctx *ctx = new cpm_ctx();
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        cpm_ctx_destroy(ctx);
        return fail(CPM_ERR_RUNTIME, "stream/event creation failed: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return CPM_OK;
}

int cpm_ctx_destroy(cpm_ctx *ctx) {
    if (!ctx) return CPM_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &b : ctx->scratch) b.release();
    ctx->host_scratch.release();
    ctx->bounce.release();
    if (ctx->search_ws_free) ctx->search_ws_free(ctx->search_ws);
    ctx->search_ws = nullptr;
    if (ctx->ingest_ws_free) ctx->ingest_ws_free(ctx->ingest_ws);
    ctx->ingest_ws = nullptr;
    ctx->spare_store.release();
    ctx->spare_arena.release();
    for (auto &t : ctx->ev_pending) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (hipEvent_t ev : ctx->ev_pool) (void)hipEventDestroy(ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CPM_OK;
}

int cpm_ctx_synchronize(cpm_ctx *ctx) {
    CPM_CHECK(bind(ctx));
    CPM_HIP(hipStreamSynchronize(ctx->stream));
    return CPM_OK;
}

void *cpm_ctx_stream(cpm_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

}  // extern "C"

int cpm::bits_alloc(cpm_ctx *ctx, uint64_t n_trans, uint32_t n_items, cpm_bits **out, bool clear) {
    cpm_bits *b = new cpm_bits();
    b->ctx = ctx;
    b->n_transactions = n_trans;
    b->n_items = n_items;
    b->words = (n_trans + 63) / 64;
    b->stride = row_stride(b->words);
    size_t bytes = (size_t)n_items * b->stride * sizeof(uint64_t);
    if (ctx->spare_store.p && ctx->spare_store.cap >= bytes) {
        std::swap(b->store.p, ctx->spare_store.p);
        std::swap(b->store.cap, ctx->spare_store.cap);
    }
    int s = b->store.ensure(bytes ? bytes : 8);
    if (s != CPM_OK) {
        delete b;
        return s;
    }
    hipError_t e = clear ? hipMemsetAsync(b->store.p, 0, bytes ? bytes : 8, ctx->stream) : hipSuccess;
    if (e != hipSuccess) {
        delete b;
        return fail(CPM_ERR_RUNTIME, "hipMemsetAsync failed: %s", hipGetErrorString(e));
    }
    *out = b;
    return CPM_OK;
}

namespace cpm {

constexpr size_t kBounceBytes = 8u << 20;

// ---- host block cache ------------------------------------------------------
// A large block is pinned host memory (plain memory if pinning fails) with a
// 64-byte header: its capacity and whether it is pinned.  The cache keeps
// up to kHostCacheBlocks blocks and kHostCacheBytes in all (held until the
// process ends); a request takes the smallest cached block that fits and is
// at most twice its size.
namespace {
constexpr size_t kHostBlockMin = 1 << 20;
constexpr size_t kHostCacheBlocks = 16;
constexpr size_t kHostCacheBytes = (size_t)512 << 20;
constexpr size_t kHostHeader = 64;
std::mutex g_host_mu;
std::vector<std::pair<size_t, void *>> g_host_cache;  // (capacity, block)
size_t g_host_cached = 0;
}  // namespace

void *host_block_take(size_t bytes) {
    if (bytes < kHostBlockMin) return malloc(bytes ? bytes : 1);
    {
        std::lock_guard<std::mutex> lock(g_host_mu);
        size_t best = g_host_cache.size();
        for (size_t i = 0; i < g_host_cache.size(); ++i) {
            const size_t cap = g_host_cache[i].first;
            if (cap >= bytes && cap <= 2 * bytes && (best == g_host_cache.size() || cap < g_host_cache[best].first))
                best = i;
        }
        if (best < g_host_cache.size()) {
            void *b = g_host_cache[best].second;
            g_host_cached -= g_host_cache[best].first;
            g_host_cache.erase(g_host_cache.begin() + best);
            return static_cast<unsigned char *>(b) + kHostHeader;
        }
    }
    // pinned, so that results are copied from the device straight into
    // them; plain memory when pinning fails
    void *b = nullptr;
    size_t pinned = 1;
    if (hipHostMalloc(&b, bytes + kHostHeader, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        b = malloc(bytes + kHostHeader);
        pinned = 0;
    }
    if (!b) return nullptr;
    static_cast<size_t *>(b)[0] = bytes;
    static_cast<size_t *>(b)[1] = pinned;
    return static_cast<unsigned char *>(b) + kHostHeader;
}

void host_block_give(void *p, size_t bytes) {
    if (!p) return;
    if (bytes < kHostBlockMin) {  // the malloc path of host_block_take
        free(p);
        return;
    }
    void *b = static_cast<unsigned char *>(p) - kHostHeader;
    const size_t cap = *static_cast<size_t *>(b);
    {
        std::lock_guard<std::mutex> lock(g_host_mu);
        if (g_host_cache.size() < kHostCacheBlocks && g_host_cached + cap <= kHostCacheBytes) {
            g_host_cache.push_back({cap, b});
            g_host_cached += cap;
            return;
        }
    }
    if (static_cast<size_t *>(b)[1]) (void)hipHostFree(b);
    else free(b);
}

hipEvent_t take_event(cpm_ctx *ctx) {
    if (!ctx->ev_pool.empty()) {
        hipEvent_t ev = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    return hipEventCreate(&ev) == hipSuccess ? ev : nullptr;
}

int collect_timing(cpm_ctx *ctx) {
    if (ctx->ev_pending.empty()) return CPM_OK;
    CPM_HIP(hipStreamSynchronize(ctx->stream));
    for (const auto &t : ctx->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) ctx->stats[t.slot] += ms;
        ctx->ev_pool.push_back(t.start);
        ctx->ev_pool.push_back(t.stop);
    }
    ctx->ev_pending.clear();
    return CPM_OK;
}

int copy_to_host(cpm_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return CPM_OK;
    CPM_CHECK(ctx->bounce.ensure(std::min(bytes, kBounceBytes)));
    unsigned char *h = ctx->bounce.as<unsigned char>();
    for (size_t off = 0; off < bytes; off += kBounceBytes) {
        const size_t n = std::min(kBounceBytes, bytes - off);
        CPM_HIP(hipMemcpyAsync(h, static_cast<const unsigned char *>(src) + off, n, hipMemcpyDeviceToHost,
                               ctx->stream));
        CPM_HIP(hipStreamSynchronize(ctx->stream));
        memcpy(static_cast<unsigned char *>(dst) + off, h, n);
    }
    return CPM_OK;
}

int copy_to_host_2d(cpm_ctx *ctx, void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                    size_t height) {
    if (width == 0 || height == 0) return CPM_OK;
    if (width > kBounceBytes) {
        for (size_t r = 0; r < height; ++r)
            CPM_CHECK(copy_to_host(ctx, static_cast<unsigned char *>(dst) + r * dpitch,
                                   static_cast<const unsigned char *>(src) + r * spitch, width));
        return CPM_OK;
    }
    const size_t per = kBounceBytes / width;
    CPM_CHECK(ctx->bounce.ensure(std::min(height, per) * width));
    unsigned char *h = ctx->bounce.as<unsigned char>();
    for (size_t r0 = 0; r0 < height; r0 += per) {
        const size_t rows = std::min(per, height - r0);
        CPM_HIP(hipMemcpy2DAsync(h, width, static_cast<const unsigned char *>(src) + r0 * spitch, spitch, width,
                                 rows, hipMemcpyDeviceToHost, ctx->stream));
        CPM_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t r = 0; r < rows; ++r)
            memcpy(static_cast<unsigned char *>(dst) + (r0 + r) * dpitch, h + r * width, width);
    }
    return CPM_OK;
}

int upload_rows(cpm_ctx *ctx, uint64_t *dst, uint64_t stride, const uint64_t *src, uint64_t words,
                uint64_t rows) {
    if (rows == 0 || stride == 0) return CPM_OK;
    const size_t row_bytes = (size_t)stride * 8;
    const uint64_t per = std::max<uint64_t>(1, kBounceBytes / row_bytes);
    CPM_CHECK(ctx->bounce.ensure(std::min<uint64_t>(rows, per) * row_bytes));
    uint64_t *h = ctx->bounce.as<uint64_t>();
    for (uint64_t r0 = 0; r0 < rows; r0 += per) {
        const uint64_t n = std::min(per, rows - r0);
        for (uint64_t r = 0; r < n; ++r) {
            memcpy(h + r * stride, src + (r0 + r) * words, (size_t)words * 8);
            memset(h + r * stride + words, 0, (size_t)(stride - words) * 8);
        }
        CPM_HIP(hipMemcpyAsync(dst + r0 * stride, h, (size_t)n * row_bytes, hipMemcpyHostToDevice, ctx->stream));
        CPM_HIP(hipStreamSynchronize(ctx->stream));  // the bounce buffer is refilled next
    }
    return CPM_OK;
}

StoreArena::~StoreArena() {
    // the stores of the batch are gone (their destroy synchronized the
    // stream): keep the larger allocation for the next batch
    if (ctx && buf.cap > ctx->spare_arena.cap) {
        std::swap(buf.p, ctx->spare_arena.p);
        std::swap(buf.cap, ctx->spare_arena.cap);
    }
}

namespace {

struct BuildIn {
    const uint64_t *rows;
    const uint32_t *cols;
    uint64_t nnz, n_trans;
    uint32_t n_items;
};

// Build n stores from device COO.  Every store with few enough items takes
// the tile build, which checks its own input; the flags of the whole batch are
// read back in one copy, so a batch costs one host synchronization whatever
// its size (two when a store's rows turn out unsorted and it is built again
// by scatter).  batch = true puts all stores in one arena.
int build_stores(cpm_ctx *ctx, const BuildIn *in, uint32_t n, bool batch, cpm_bits **out) {
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = nullptr;
        if (in[i].nnz && (!in[i].rows || !in[i].cols)) return fail(CPM_ERR_INVALID, "rows/cols are NULL");
        if (in[i].n_trans >= (1ull << 40)) return fail(CPM_ERR_INVALID, "too many transactions");
    }
    if (n == 0) return CPM_OK;
    hipStream_t st = ctx->stream;
    // pinned host area: flags (3 per store) then the job table; device: the same
    const size_t flag_bytes = ((size_t)3 * n * sizeof(unsigned int) + 255) / 256 * 256;
    const size_t job_bytes = (size_t)n * sizeof(BuildJob);
    CPM_CHECK(ctx->scratch[5].ensure(flag_bytes + job_bytes));
    CPM_CHECK(ctx->host_scratch.ensure(flag_bytes + job_bytes));
    unsigned int *flags = ctx->scratch[5].as<unsigned int>();
    BuildJob *d_jobs = reinterpret_cast<BuildJob *>(ctx->scratch[5].as<unsigned char>() + flag_bytes);
    unsigned int *hflags = ctx->host_scratch.as<unsigned int>();
    BuildJob *h_jobs = reinterpret_cast<BuildJob *>(ctx->host_scratch.as<unsigned char>() + flag_bytes);

    // 1. which stores take the tile build: all with pairs and few enough
    // items for the LDS tile; the tile kernel checks the rows are sorted, and
    // a store whose rows are not is built again by scatter (step 4)
    std::vector<char> tiled(n, 0);
    for (uint32_t i = 0; i < n; ++i) tiled[i] = in[i].nnz && tile_words_for(in[i].n_items);
    CPM_HIP(hipMemsetAsync(flags, 0, (size_t)3 * n * sizeof(unsigned int), st));
    // 2. allocation
    std::shared_ptr<StoreArena> arena;
    std::vector<uint64_t> offset(n, 0);
    if (batch) {
        uint64_t total = 0;
        for (uint32_t i = 0; i < n; ++i) {
            offset[i] = total;
            const uint64_t words = (in[i].n_trans + 63) / 64;
            total += std::max<uint64_t>((uint64_t)in[i].n_items * row_stride(words) * 8, 256);
        }
        arena = std::make_shared<StoreArena>();
        arena->ctx = ctx;
        if (ctx->spare_arena.p && ctx->spare_arena.cap >= total) {
            std::swap(arena->buf.p, ctx->spare_arena.p);
            std::swap(arena->buf.cap, ctx->spare_arena.cap);
        } else {
            // a new arena, with headroom so that batches of varying size
            // rarely allocate again (the old one is freed)
            ctx->spare_arena.release();
            total += total / 2;
        }
        CPM_CHECK(arena->buf.ensure(total));
    }
    int status = CPM_OK;
    for (uint32_t i = 0; i < n && status == CPM_OK; ++i) {
        cpm_bits *b = nullptr;
        if (batch) {
            b = new cpm_bits();
            b->ctx = ctx;
            b->n_transactions = in[i].n_trans;
            b->n_items = in[i].n_items;
            b->words = (in[i].n_trans + 63) / 64;
            b->stride = row_stride(b->words);
            b->arena = arena;
            b->view = reinterpret_cast<uint64_t *>(arena->buf.as<unsigned char>() + offset[i]);
            const size_t bytes = (size_t)b->n_items * b->stride * 8;
            if (!tiled[i] && bytes && hipMemsetAsync(b->view, 0, bytes, st) != hipSuccess)
                status = fail(CPM_ERR_RUNTIME, "hipMemsetAsync failed");
        } else {
            status = bits_alloc(ctx, in[i].n_trans, in[i].n_items, &b, !tiled[i]);
        }
        if (!b) break;
        out[i] = b;
    }
    // 3. the tile builds of every sorted store in one launch; scatter builds
    // (unsorted rows, or too many items for the LDS tile) one per store
    bool scattered = false;  // a scatter build ran: its range flag must be read
    if (status == CPM_OK) {
        uint32_t nt = 0;
        uint64_t blocks = 0;
        size_t lds = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (!tiled[i]) continue;
            BuildJob &j = h_jobs[nt++];
            memset(&j, 0, sizeof(j));
            j.rows = in[i].rows;
            j.cols = in[i].cols;
            j.nnz = in[i].nnz;
            j.n_trans = in[i].n_trans;
            j.stride = out[i]->stride;
            j.store = out[i]->rows();
            j.flags = flags + 3 * i;
            j.n_items = in[i].n_items;
            j.B = tile_words_for(in[i].n_items);
            j.block0 = blocks;
            blocks += (j.stride + j.B - 1) / j.B;
            lds = std::max(lds, (size_t)j.n_items * j.B * 8);
        }
        if (nt && blocks >= (1ull << 31)) status = fail(CPM_ERR_INVALID, "store batch too large");
        if (nt && status == CPM_OK) {
            if (hipMemcpyAsync(d_jobs, h_jobs, (size_t)nt * sizeof(BuildJob), hipMemcpyHostToDevice, st) != hipSuccess)
                status = fail(CPM_ERR_RUNTIME, "job table upload failed");
            else
                hipLaunchKernelGGL(k_bits_tile, dim3((unsigned)blocks), dim3(512), lds, st, d_jobs, nt);
        }
        for (uint32_t i = 0; i < n && status == CPM_OK; ++i) {
            if (tiled[i] || !in[i].nnz) continue;
            const uint64_t g = std::min<uint64_t>((in[i].nnz + 255) / 256, (uint64_t)ctx->cu_count * 32);
            hipLaunchKernelGGL(k_bits_scatter, dim3((unsigned)g), dim3(256), 0, st, in[i].rows, in[i].cols,
                               in[i].nnz, in[i].n_trans, in[i].n_items, out[i]->stride,
                               reinterpret_cast<unsigned long long *>(out[i]->rows()), flags + 3 * i + 2);
            scattered = true;
        }
    }
    hipError_t e = hipGetLastError();
    if (status == CPM_OK && e != hipSuccess)
        status = fail(CPM_ERR_RUNTIME, "store build failed: %s", hipGetErrorString(e));
    // 4. the flags of every store, one read-back: unsorted rows are built
    // again by scatter (one more read-back for its range check)
    auto read_flags = [&]() {
        hipError_t r = hipMemcpyAsync(hflags, flags, (size_t)3 * n * sizeof(unsigned int), hipMemcpyDeviceToHost, st);
        if (r == hipSuccess) r = hipStreamSynchronize(st);
        if (r != hipSuccess) status = fail(CPM_ERR_RUNTIME, "store build failed: %s", hipGetErrorString(r));
    };
    bool any_tiled = false;
    for (uint32_t i = 0; i < n; ++i) any_tiled = any_tiled || tiled[i];
    if (status == CPM_OK && (any_tiled || scattered)) read_flags();
    bool rebuilt = false;
    for (uint32_t i = 0; i < n && status == CPM_OK; ++i) {
        if (!tiled[i] || !hflags[3 * i]) continue;
        const size_t bytes = (size_t)out[i]->n_items * out[i]->stride * 8;
        if (hipMemsetAsync(out[i]->rows(), 0, bytes, st) != hipSuccess) {
            status = fail(CPM_ERR_RUNTIME, "hipMemsetAsync failed");
            break;
        }
        hflags[3 * i + 2] = 0;
        if (hipMemsetAsync(flags + 3 * i + 2, 0, sizeof(unsigned int), st) != hipSuccess) {
            status = fail(CPM_ERR_RUNTIME, "hipMemsetAsync failed");
            break;
        }
        const uint64_t g = std::min<uint64_t>((in[i].nnz + 255) / 256, (uint64_t)ctx->cu_count * 32);
        hipLaunchKernelGGL(k_bits_scatter, dim3((unsigned)g), dim3(256), 0, st, in[i].rows, in[i].cols,
                           in[i].nnz, in[i].n_trans, in[i].n_items, out[i]->stride,
                           reinterpret_cast<unsigned long long *>(out[i]->rows()), flags + 3 * i + 2);
        rebuilt = true;
    }
    if (status == CPM_OK && rebuilt) {
        e = hipGetLastError();
        if (e != hipSuccess) status = fail(CPM_ERR_RUNTIME, "store build failed: %s", hipGetErrorString(e));
        else read_flags();
    }
    for (uint32_t i = 0; i < n && status == CPM_OK; ++i)
        if (in[i].nnz && hflags[3 * i + 2])
            status = fail(CPM_ERR_INVALID, "a (row, column) pair lies outside %llu x %u",
                          (unsigned long long)in[i].n_trans, in[i].n_items);
    if (status != CPM_OK) {
        (void)hipStreamSynchronize(st);
        for (uint32_t i = 0; i < n; ++i) {
            cpm_bits_destroy(out[i]);
            out[i] = nullptr;
        }
    }
    return status;
}

}  // namespace
}  // namespace cpm

extern "C" {

int cpm_bits_build_device(cpm_ctx *ctx, const uint64_t *d_rows, const uint32_t *d_cols,
                          uint64_t nnz, uint64_t n_trans, uint32_t n_items, cpm_bits **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    const BuildIn in{d_rows, d_cols, nnz, n_trans, n_items};
    return build_stores(ctx, &in, 1, false, out);
}

int cpm_bits_build_device_batch(cpm_ctx *ctx, uint32_t n_db, const uint64_t *const *d_rows,
                                const uint32_t *const *d_cols, const uint64_t *nnz, const uint64_t *n_trans,
                                const uint32_t *n_items, cpm_bits **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "out is NULL");
    if (n_db && (!d_rows || !d_cols || !nnz || !n_trans || !n_items)) return fail(CPM_ERR_INVALID, "null argument");
    if (n_db > 65535) return fail(CPM_ERR_INVALID, "at most 65535 databases per batch build");
    std::vector<BuildIn> in(n_db);
    for (uint32_t i = 0; i < n_db; ++i) in[i] = {d_rows[i], d_cols[i], nnz[i], n_trans[i], n_items[i]};
    return build_stores(ctx, in.data(), n_db, true, out);
}

int cpm_bits_build_batch(cpm_ctx *ctx, uint32_t n_db, const uint64_t *const *rows, const uint32_t *const *cols,
                         const uint64_t *nnz, const uint64_t *n_trans, const uint32_t *n_items, cpm_bits **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "out is NULL");
    if (n_db && (!rows || !cols || !nnz || !n_trans || !n_items)) return fail(CPM_ERR_INVALID, "null argument");
    if (n_db > 65535) return fail(CPM_ERR_INVALID, "at most 65535 databases per batch build");
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_db; ++i) {
        if (nnz[i] && (!rows[i] || !cols[i])) return fail(CPM_ERR_INVALID, "rows/cols are NULL");
        total += nnz[i];
    }
    // every database's pairs in one device staging area (rows, then columns)
    DBuf &dr = ctx->scratch[0], &dc = ctx->scratch[1];
    CPM_CHECK(dr.ensure(std::max<uint64_t>(total, 1) * sizeof(uint64_t)));
    CPM_CHECK(dc.ensure(std::max<uint64_t>(total, 1) * sizeof(uint32_t)));
    std::vector<BuildIn> in(n_db);
    uint64_t off = 0;
    for (uint32_t i = 0; i < n_db; ++i) {
        if (nnz[i]) {
            CPM_HIP(hipMemcpyAsync(dr.as<uint64_t>() + off, rows[i], nnz[i] * sizeof(uint64_t),
                                   hipMemcpyHostToDevice, ctx->stream));
            CPM_HIP(hipMemcpyAsync(dc.as<uint32_t>() + off, cols[i], nnz[i] * sizeof(uint32_t),
                                   hipMemcpyHostToDevice, ctx->stream));
        }
        in[i] = {dr.as<uint64_t>() + off, dc.as<uint32_t>() + off, nnz[i], n_trans[i], n_items[i]};
        off += nnz[i];
    }
    return build_stores(ctx, in.data(), n_db, true, out);
}

int cpm_bits_build(cpm_ctx *ctx, const uint64_t *rows, const uint32_t *cols, uint64_t nnz,
                   uint64_t n_trans, uint32_t n_items, cpm_bits **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (nnz && (!rows || !cols)) return fail(CPM_ERR_INVALID, "rows/cols are NULL");
    DBuf &dr = ctx->scratch[0], &dc = ctx->scratch[1];
    CPM_CHECK(dr.ensure(nnz * sizeof(uint64_t)));
    CPM_CHECK(dc.ensure(nnz * sizeof(uint32_t)));
    if (nnz) {
        CPM_HIP(hipMemcpyAsync(dr.p, rows, nnz * sizeof(uint64_t), hipMemcpyHostToDevice,
                               ctx->stream));
        CPM_HIP(hipMemcpyAsync(dc.p, cols, nnz * sizeof(uint32_t), hipMemcpyHostToDevice,
                               ctx->stream));
    }
    return cpm_bits_build_device(ctx, dr.as<uint64_t>(), dc.as<uint32_t>(), nnz, n_trans,
                                 n_items, out);
}

int cpm_bits_destroy(cpm_bits *b) {
    if (!b) return CPM_OK;
    if (b->arena && b->arena.use_count() > 1) {
        delete b;  // a view into a batch arena that other stores still hold
        return CPM_OK;
    }
    if (b->ctx) {
        cpm_ctx *ctx = b->ctx;
        (void)hipSetDevice(ctx->device);
        // queued work may still read this memory before it is reused
        (void)hipStreamSynchronize(ctx->stream);
        // keep the larger allocation for the next build on this context
        if (b->store.cap > ctx->spare_store.cap) {
            std::swap(b->store.p, ctx->spare_store.p);
            std::swap(b->store.cap, ctx->spare_store.cap);
        }
    }
    delete b;
    return CPM_OK;
}

int cpm_bits_shape(const cpm_bits *b, uint64_t *n_trans, uint32_t *n_items, uint64_t *words,
                   uint64_t *stride) {
    if (!b) return fail(CPM_ERR_INVALID, "null store");
    if (n_trans) *n_trans = b->n_transactions;
    if (n_items) *n_items = b->n_items;
    if (words) *words = b->words;
    if (stride) *stride = b->stride;
    return CPM_OK;
}

int cpm_bits_download(cpm_ctx *ctx, const cpm_bits *b, uint64_t *host) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!b || !host) return fail(CPM_ERR_INVALID, "null argument");
    if (b->words && b->n_items)
        CPM_CHECK(copy_to_host_2d(ctx, host, b->words * 8, b->rows(), b->stride * 8, b->words * 8, b->n_items));
    CPM_HIP(hipStreamSynchronize(ctx->stream));
    return CPM_OK;
}

// ---- operators ----------------------------------------------------------------

int cpm_intersect_extensions(cpm_ctx *ctx, const uint64_t *prefix, const uint64_t *ext, uint32_t k,
                             uint64_t words, uint64_t keep_min, uint64_t *supports,
                             uint64_t *kept_rows, uint32_t *kept_index, uint32_t *n_kept) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (n_kept) *n_kept = 0;
    if (k == 0) return CPM_OK;
    if (!supports || (words && (!prefix || !ext))) return fail(CPM_ERR_INVALID, "null argument");
    const uint64_t stride = row_stride(words);
    if (words) {
        DBuf &dpre = ctx->scratch[0], &dext = ctx->scratch[1], &dout = ctx->scratch[2],
             &dsup = ctx->scratch[3];
        CPM_CHECK(dpre.ensure(stride * 8));
        CPM_CHECK(dext.ensure((size_t)k * stride * 8));
        CPM_CHECK(dsup.ensure((size_t)k * 8));
        if (kept_rows) CPM_CHECK(dout.ensure((size_t)k * stride * 8));
        CPM_CHECK(upload_rows(ctx, dpre.as<uint64_t>(), stride, prefix, words, 1));
        CPM_CHECK(upload_rows(ctx, dext.as<uint64_t>(), stride, ext, words, k));
        CPM_CHECK(launch_intersect(ctx, dpre.as<uint64_t>(), dext.as<uint64_t>(), k, stride,
                                   dsup.as<unsigned long long>(),
                                   kept_rows ? dout.as<uint64_t>() : nullptr));
        CPM_CHECK(copy_to_host(ctx, supports, dsup.p, (size_t)k * 8));
    } else {
        for (uint32_t j = 0; j < k; ++j) supports[j] = 0;
    }
    // keep step (cuEclat.py:234-242): children at or above keep_min, in order
    std::vector<uint32_t> keep;
    for (uint32_t j = 0; j < k; ++j)
        if (supports[j] >= keep_min) keep.push_back(j);
    const uint32_t nk = (uint32_t)keep.size();
    if (kept_index && nk) memcpy(kept_index, keep.data(), (size_t)nk * 4);
    if (kept_rows && words && nk) {
        // gather the kept rows on the device, then one copy out
        DBuf &didx = ctx->scratch[4], &dgat = ctx->scratch[1];  // the extension rows are consumed
        CPM_CHECK(didx.ensure((size_t)nk * 4));
        CPM_CHECK(dgat.ensure((size_t)nk * words * 8));
        CPM_CHECK(ctx->host_scratch.ensure((size_t)nk * 4));
        memcpy(ctx->host_scratch.p, keep.data(), (size_t)nk * 4);
        CPM_HIP(hipMemcpyAsync(didx.p, ctx->host_scratch.p, (size_t)nk * 4, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_gather_rows, dim3(std::min<uint32_t>(nk, 65535)), dim3(256), 0, ctx->stream,
                           ctx->scratch[2].as<uint64_t>(), stride, words, didx.as<uint32_t>(), nk,
                           dgat.as<uint64_t>());
        CPM_HIP(hipGetLastError());
        CPM_CHECK(copy_to_host(ctx, kept_rows, dgat.p, (size_t)nk * words * 8));
    }
    CPM_HIP(hipStreamSynchronize(ctx->stream));
    if (n_kept) *n_kept = nk;
    return CPM_OK;
}

// ---- patterns -----------------------------------------------------------------

int cpm_patterns_count(const cpm_patterns *p, uint64_t *n) {
    if (!p || !n) return fail(CPM_ERR_INVALID, "null argument");
    *n = p->n;
    return CPM_OK;
}

int cpm_patterns_download(const cpm_patterns *p, uint32_t *parent, uint32_t *item,
                          uint64_t *support, int64_t *period) {
    if (!p) return fail(CPM_ERR_INVALID, "null patterns");
    if (p->n == 0) return CPM_OK;
    if (parent) memcpy(parent, p->parent.data(), p->n * 4);
    if (item) memcpy(item, p->item.data(), p->n * 4);
    if (support) memcpy(support, p->support.data(), p->n * 8);
    if (period) {
        if (p->periodic) memcpy(period, p->period.data(), p->n * 8);
        else memset(period, 0, p->n * 8);
    }
    return CPM_OK;
}

int cpm_patterns_view(const cpm_patterns *p, const uint32_t **parent, const uint32_t **item,
                      const uint64_t **support, const int64_t **period, const uint32_t **db) {
    if (!p) return fail(CPM_ERR_INVALID, "null patterns");
    const bool any = p->n != 0;
    if (parent) *parent = any ? p->parent.data() : nullptr;
    if (item) *item = any ? p->item.data() : nullptr;
    if (support) *support = any ? p->support.data() : nullptr;
    if (period) *period = any && p->periodic ? p->period.data() : nullptr;
    if (db) *db = any && p->db.size() == p->n ? p->db.data() : nullptr;
    return CPM_OK;
}

int cpm_patterns_destroy(cpm_patterns *p) {
    delete p;
    return CPM_OK;
}

// ---- measurement ----------------------------------------------------------------

int cpm_stats(cpm_ctx *ctx, double *out8) {
    if (!ctx || !out8) return fail(CPM_ERR_INVALID, "null argument");
    CPM_CHECK(bind(ctx));
    CPM_CHECK(collect_timing(ctx));
    memcpy(out8, ctx->stats, 8 * sizeof(double));
    return CPM_OK;
}

int cpm_ingest_stats(cpm_ctx *ctx, double *out4) {
    if (!ctx || !out4) return fail(CPM_ERR_INVALID, "null argument");
    CPM_CHECK(bind(ctx));
    CPM_CHECK(collect_timing(ctx));
    memcpy(out4, ctx->stats + 8, 4 * sizeof(double));
    return CPM_OK;
}

int cpm_stats_reset(cpm_ctx *ctx) {
    if (!ctx) return fail(CPM_ERR_INVALID, "null context");
    CPM_CHECK(bind(ctx));
    CPM_CHECK(collect_timing(ctx));  // launches before the reset count for nothing
    memset(ctx->stats, 0, sizeof(ctx->stats));
    return CPM_OK;
}

int cpm_set_timing(cpm_ctx *ctx, int on) {
    if (!ctx) return fail(CPM_ERR_INVALID, "null context");
    ctx->timing = on != 0;
    return CPM_OK;
}

}  // extern "C"

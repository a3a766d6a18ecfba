// Delimited text of decimal item ids, parsed on the device into basket
// arrays (offsets u64[baskets + 1], ids u32[entries]) for the device store
// build (cpm_bits_build_csr_device).
//
// The host reader this replaces (cuApriori.py:202-230, 428-450) splits lines
// with csv.reader, strips every field and drops empty fields and repeats.
// Here every line is a basket, every delimiter (the separator or '\n') ends a
// field, and each field is stripped of ASCII whitespace; a non-empty field
// must be a canonical decimal id (no sign, no leading zero, <= 2^32 - 1),
// whose label is its own text.  Anything the csv module would read
// differently -- a quote, a byte >= 0x80, a NUL, a '\r' that is neither
// followed by '\n' nor the last byte -- and any field that is not such an id
// make the parse return CPM_ERR_LAYOUT, and the caller reads the text on the
// host.  Repeats within a basket and empty baskets are left to the store
// build, which sets one bit per (basket, id) and leaves empty baskets out.
//
// Three launches: per 64-byte lane range, count line ends and ids (each
// field is parsed by the lane holding its closing delimiter, scanning back
// at most kMaxField bytes for its start); scan the per-block counts; write
// each id and each line's offset.  One host read between the first two and
// the last (the totals size the arrays), one after (non-empty baskets).
#include "cpm_internal.h"

struct cpm_baskets {
    cpm_ctx *ctx = nullptr;
    cpm::DBuf offsets, ids;
    uint64_t n_baskets = 0, n_ids = 0, n_nonempty = 0;
    uint32_t max_id = 0;
};

namespace cpm {

namespace {

constexpr uint32_t kLaneBytes = 64;
constexpr uint32_t kTextBlock = 256;
constexpr uint64_t kBlockBytes = (uint64_t)kLaneBytes * kTextBlock;
constexpr uint32_t kMaxField = 64;   // bytes of a field, whitespace included

enum TextFlag : uint32_t { kBadByte = 1, kBadField = 2, kLongField = 4 };

struct TextCtl {
    unsigned long long lines, entries, nonempty;
    unsigned int flags, max_id;
};

__device__ __forceinline__ bool ascii_space(uint8_t c) {
    return c == ' ' || c == '\t' || c == '\v' || c == '\f' || c == '\r' || (c >= 0x1c && c <= 0x1f);
}

// Bytes of the text as a block sees them: its own kBlockBytes, the
// kMaxField bytes before them and the byte after them, staged in LDS.
struct TextView {
    const uint8_t *lds;
    uint64_t base;   // text index of lds[0]
    __device__ __forceinline__ uint8_t operator[](uint64_t i) const { return lds[i - base]; }
};

// The field closed by the delimiter at `e` (or by the end of the text):
// returns 0 (empty), 1 (an id, in *value) or a TextFlag.
__device__ __forceinline__ uint32_t parse_field(const TextView &t, uint64_t e, uint8_t sep, uint32_t *value) {
    uint64_t s = 0;
    bool found = false;
    for (uint64_t k = 1; k <= kMaxField && k <= e; ++k) {
        const uint8_t c = t[e - k];
        if (c == sep || c == '\n') {
            s = e - k + 1;
            found = true;
            break;
        }
    }
    if (!found && e > kMaxField) return kLongField;
    while (s < e && ascii_space(t[s])) ++s;
    uint64_t f = e;
    while (f > s && ascii_space(t[f - 1])) --f;
    if (s == f) return 0;
    const uint64_t len = f - s;
    if (len > 10 || (t[s] == '0' && len > 1)) return kBadField;
    uint64_t v = 0;
    for (uint64_t i = s; i < f; ++i) {
        const uint8_t c = t[i];
        if (c < '0' || c > '9') return kBadField;
        v = v * 10 + (c - '0');
    }
    if (v > 0xffffffffull) return kBadField;
    *value = (uint32_t)v;
    return 1;
}

// Walks bytes [lo, hi) of one lane; on_field(value) for every id, on_line()
// for every '\n' (after the id of the field it closes).  Returns TextFlags.
template <class F, class L>
__device__ __forceinline__ uint32_t walk(const TextView &t, uint64_t n, uint8_t sep, uint64_t lo, uint64_t hi,
                                         F on_field, L on_line) {
    uint32_t flags = 0;
    for (uint64_t i = lo; i < hi; ++i) {
        const uint8_t c = t[i];
        if (c == '"' || c >= 0x80 || c == 0 || (c == '\r' && i + 1 < n && t[i + 1] != '\n')) flags |= kBadByte;
        if (c == sep || c == '\n') {
            uint32_t v = 0;
            const uint32_t r = parse_field(t, i, sep, &v);
            if (r == 1) on_field(v);
            else flags |= r;
            if (c == '\n') on_line();
        }
    }
    // the last line may end at the end of the text instead of a '\n'
    if (hi == n && n > 0 && t[n - 1] != '\n') {
        uint32_t v = 0;
        const uint32_t r = parse_field(t, n, sep, &v);
        if (r == 1) on_field(v);
        else flags |= r;
    }
    return flags;
}

// Stage this block's bytes, the kMaxField before them and the one after.
constexpr uint32_t kStageBytes = (uint32_t)kBlockBytes + kMaxField + 4;
__device__ __forceinline__ TextView stage_block(const uint8_t *__restrict__ g, uint64_t n, uint8_t *lds) {
    const uint64_t b0 = (uint64_t)blockIdx.x * kBlockBytes;
    const uint64_t base = b0 >= kMaxField ? b0 - kMaxField : 0;
    const uint64_t end = min(n, b0 + kBlockBytes + 1);
    for (uint64_t i = base + threadIdx.x; i < end; i += blockDim.x) lds[i - base] = g[i];
    __syncthreads();
    return TextView{lds, base};
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const unsigned long long s = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(kTextBlock) void k_text_count(const uint8_t *__restrict__ g, uint64_t n, uint8_t sep,
                                                           unsigned long long *part_lines,
                                                           unsigned long long *part_entries, TextCtl *ctl) {
    __shared__ unsigned long long red[4];
    __shared__ uint8_t stage[kStageBytes];
    const TextView t = stage_block(g, n, stage);
    const uint64_t lo = min(n, (uint64_t)blockIdx.x * kBlockBytes + (uint64_t)threadIdx.x * kLaneBytes);
    const uint64_t hi = min(n, lo + kLaneBytes);
    unsigned long long lines = 0, entries = 0;
    uint32_t vmax = 0;
    const uint32_t flags = walk(t, n, sep, lo, hi, [&](uint32_t v) { ++entries; vmax = max(vmax, v); },
                                [&]() { ++lines; });
    if (flags) atomicOr(&ctl->flags, flags);
    if (entries) atomicMax(&ctl->max_id, vmax);
    const unsigned long long L = block_sum_u64(lines, red);
    const unsigned long long E = block_sum_u64(entries, red);
    if (threadIdx.x == 0) {
        part_lines[blockIdx.x] = L;
        part_entries[blockIdx.x] = E;
    }
}

// One workgroup: exclusive scan of the per-block counts, totals into ctl.
__global__ __launch_bounds__(1024) void k_text_scan(unsigned long long *part_lines, unsigned long long *part_entries,
                                                    uint32_t nb, TextCtl *ctl) {
    __shared__ unsigned long long sl[1024], se[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t lo = min(nb, tid * per), hi = min(nb, lo + per);
    unsigned long long l = 0, e = 0;
    for (uint32_t b = lo; b < hi; ++b) {
        l += part_lines[b];
        e += part_entries[b];
    }
    sl[tid] = l;
    se[tid] = e;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const unsigned long long a = tid >= off ? sl[tid - off] : 0, b = tid >= off ? se[tid - off] : 0;
        __syncthreads();
        sl[tid] += a;
        se[tid] += b;
        __syncthreads();
    }
    unsigned long long bl = sl[tid] - l, be = se[tid] - e;
    for (uint32_t b = lo; b < hi; ++b) {
        const unsigned long long x = part_lines[b], y = part_entries[b];
        part_lines[b] = bl;
        part_entries[b] = be;
        bl += x;
        be += y;
    }
    if (tid == 1023) {
        ctl->lines = sl[1023];
        ctl->entries = se[1023];
    }
}

__device__ __forceinline__ unsigned long long block_excl_u64(unsigned long long v, unsigned long long *red) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) red[wave] = x;
    __syncthreads();
    unsigned long long base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += red[w];
    __syncthreads();
    return base + x - v;
}

__global__ __launch_bounds__(kTextBlock) void k_text_emit(const uint8_t *__restrict__ g, uint64_t n, uint8_t sep,
                                                          const unsigned long long *base_lines,
                                                          const unsigned long long *base_entries,
                                                          uint64_t *__restrict__ offsets, uint32_t *__restrict__ ids) {
    __shared__ unsigned long long red[4];
    __shared__ uint8_t stage[kStageBytes];
    const TextView t = stage_block(g, n, stage);
    const uint64_t lo = min(n, (uint64_t)blockIdx.x * kBlockBytes + (uint64_t)threadIdx.x * kLaneBytes);
    const uint64_t hi = min(n, lo + kLaneBytes);
    // this lane's counts again, for its place within the block
    unsigned long long lines = 0, entries = 0;
    (void)walk(t, n, sep, lo, hi, [&](uint32_t) { ++entries; }, [&]() { ++lines; });
    unsigned long long line = base_lines[blockIdx.x] + block_excl_u64(lines, red);
    unsigned long long entry = base_entries[blockIdx.x] + block_excl_u64(entries, red);
    (void)walk(t, n, sep, lo, hi, [&](uint32_t v) { ids[entry++] = v; }, [&]() { offsets[++line] = entry; });
}

__global__ void k_text_nonempty(const uint64_t *__restrict__ offsets, uint64_t nb, TextCtl *ctl) {
    unsigned long long c = 0;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * blockDim.x)
        c += offsets[b + 1] > offsets[b] ? 1 : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&ctl->nonempty, c);
}

}  // namespace

}  // namespace cpm

using namespace cpm;

extern "C" {

int cpm_baskets_parse_text(cpm_ctx *ctx, const char *text, uint64_t n_bytes, char sep, cpm_baskets **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (n_bytes && !text) return fail(CPM_ERR_INVALID, "text is NULL");
    const uint8_t s = (uint8_t)sep;
    if (s == 0 || s >= 0x80 || s == '"' || s == '\n' || s == '\r')
        return fail(CPM_ERR_LAYOUT, "separator 0x%02x is read on the host", s);
    const uint64_t nb64 = (n_bytes + kBlockBytes - 1) / kBlockBytes;
    if (nb64 > 0x7fffffffull) return fail(CPM_ERR_LAYOUT, "text of %llu bytes is read on the host",
                                         (unsigned long long)n_bytes);
    const uint32_t nb = (uint32_t)nb64;
    std::unique_ptr<cpm_baskets> b(new cpm_baskets());
    b->ctx = ctx;
    hipStream_t st = ctx->stream;
    DBuf dtext, parts, dctl;
    CPM_CHECK(dtext.ensure(std::max<uint64_t>(n_bytes, 1)));
    CPM_CHECK(parts.ensure((size_t)2 * std::max<uint32_t>(nb, 1) * 8));
    CPM_CHECK(dctl.ensure(sizeof(TextCtl)));
    if (n_bytes) CPM_HIP(hipMemcpyAsync(dtext.p, text, n_bytes, hipMemcpyHostToDevice, st));
    const TextCtl zero{};
    CPM_HIP(hipMemcpyAsync(dctl.p, &zero, sizeof(zero), hipMemcpyHostToDevice, st));
    unsigned long long *pl = parts.as<unsigned long long>(), *pe = pl + std::max<uint32_t>(nb, 1);
    const uint8_t *t = dtext.as<uint8_t>();
    TextCtl *ctl = dctl.as<TextCtl>();
    if (nb) {
        hipLaunchKernelGGL(k_text_count, dim3(nb), dim3(kTextBlock), 0, st, t, n_bytes, s, pl, pe, ctl);
        CPM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_text_scan, dim3(1), dim3(1024), 0, st, pl, pe, nb, ctl);
        CPM_HIP(hipGetLastError());
    }
    TextCtl h{};
    CPM_CHECK(copy_to_host(ctx, &h, ctl, sizeof(h)));
    if (h.flags)
        return fail(CPM_ERR_LAYOUT, "text holds %s: read on the host",
                    (h.flags & kBadByte) ? "quotes, non-ASCII bytes or lone carriage returns"
                    : (h.flags & kLongField) ? "a field longer than 64 bytes"
                                             : "fields that are not canonical decimal ids");
    // a last line without '\n' is a basket too
    const bool open_last = n_bytes > 0 && text[n_bytes - 1] != '\n';
    b->n_baskets = h.lines + (open_last ? 1 : 0);
    b->n_ids = h.entries;
    b->max_id = h.max_id;
    CPM_CHECK(b->offsets.ensure((b->n_baskets + 1) * 8));
    CPM_CHECK(b->ids.ensure(std::max<uint64_t>(b->n_ids, 1) * 4));
    uint64_t *offs = b->offsets.as<uint64_t>();
    const uint64_t head = 0, tail = b->n_ids;
    CPM_HIP(hipMemcpyAsync(offs, &head, 8, hipMemcpyHostToDevice, st));
    if (nb) {
        hipLaunchKernelGGL(k_text_emit, dim3(nb), dim3(kTextBlock), 0, st, t, n_bytes, s, pl, pe, offs,
                           b->ids.as<uint32_t>());
        CPM_HIP(hipGetLastError());
    }
    if (open_last) CPM_HIP(hipMemcpyAsync(offs + b->n_baskets, &tail, 8, hipMemcpyHostToDevice, st));
    if (b->n_baskets) {
        hipLaunchKernelGGL(k_text_nonempty, dim3((unsigned)std::min<uint64_t>((b->n_baskets + 255) / 256, 2048)),
                           dim3(256), 0, st, offs, b->n_baskets, ctl);
        CPM_HIP(hipGetLastError());
    }
    CPM_CHECK(copy_to_host(ctx, &h, ctl, sizeof(h)));
    b->n_nonempty = h.nonempty;
    *out = b.release();
    return CPM_OK;
}

int cpm_baskets_info(const cpm_baskets *b, uint64_t *n_baskets, uint64_t *n_ids, uint64_t *n_nonempty,
                     uint32_t *max_id) {
    if (!b) return fail(CPM_ERR_INVALID, "null baskets");
    if (n_baskets) *n_baskets = b->n_baskets;
    if (n_ids) *n_ids = b->n_ids;
    if (n_nonempty) *n_nonempty = b->n_nonempty;
    if (max_id) *max_id = b->max_id;
    return CPM_OK;
}

int cpm_baskets_download(cpm_ctx *ctx, const cpm_baskets *b, uint64_t *offsets, uint32_t *ids) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!b) return fail(CPM_ERR_INVALID, "null baskets");
    if (offsets) CPM_CHECK(copy_to_host(ctx, offsets, b->offsets.p, (b->n_baskets + 1) * 8));
    if (ids && b->n_ids) CPM_CHECK(copy_to_host(ctx, ids, b->ids.p, b->n_ids * 4));
    return CPM_OK;
}

int cpm_bits_build_baskets(cpm_ctx *ctx, const cpm_baskets *b, int keep_empty, cpm_bits **out) {
    if (!b) return fail(CPM_ERR_INVALID, "null baskets");
    return cpm_bits_build_csr_device(ctx, b->offsets.as<uint64_t>(), b->n_baskets, b->ids.as<uint32_t>(), b->n_ids,
                                     std::max<uint32_t>(b->max_id, 1), keep_empty, out);
}

int cpm_baskets_destroy(cpm_baskets *b) {
    if (!b) return CPM_OK;
    if (b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
    }
    delete b;
    return CPM_OK;
}

}  // extern "C"

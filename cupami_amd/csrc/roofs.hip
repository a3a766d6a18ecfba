// Measured roofs of this device, for placing the path's kernels against the
// level of the memory hierarchy (or the matrix cores) that bounds them.
//
// k_roof_stream reads two equal streams of a buffer (its halves: the prefix
// rows and the extension rows of the support kernel) and ANDs and popcounts
// them, 16 bytes per lane per load -- the support kernel's inner loop with
// nothing else.  With `xcd_sliced`, XCD x (blocks b with b % 8 == x) reads only
// its own eighth of each stream, as the support kernel's tile == XCD mapping
// does, so a buffer of 8 x (an L2's share) stays L2-resident, one of a few
// hundred MB stays in the Infinity Cache, and a multi-GB one streams from
// HBM.  k_roof_mfma issues independent FP4 MFMAs from registers (the Gram's
// v_mfma_scale_f32_16x16x128_f8f6f4 form, and the 32x32x64 one): the loop rate
// of the matrix cores.
#include "cpm_internal.h"

namespace cpm {

namespace {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t pop_and(uint4 a, uint4 b) {
    return __popc(a.x & b.x) + __popc(a.y & b.y) + __popc(a.z & b.z) + __popc(a.w & b.w);
}

__global__ __launch_bounds__(256) void k_roof_stream(const uint4 *__restrict__ buf, uint64_t half, uint32_t reps,
                                                     int xcd_sliced, unsigned long long *sink) {
    const uint32_t tid = threadIdx.x;
    uint64_t lo = 0, hi = half, start, step;
    if (xcd_sliced) {
        const uint32_t x = blockIdx.x % 8, j = blockIdx.x / 8, nb = gridDim.x / 8;
        const uint64_t slice = half / 8;
        lo = x * slice;
        hi = lo + slice;
        start = lo + (uint64_t)j * 256 + tid;
        step = (uint64_t)nb * 256;
    } else {
        start = (uint64_t)blockIdx.x * 256 + tid;
        step = (uint64_t)gridDim.x * 256;
    }
    const uint4 *A = buf, *B = buf + half;
    uint32_t acc = 0;
    for (uint32_t r = 0; r < reps; ++r) {
        uint64_t i = start;
        // four independent load pairs in flight per lane
        for (; i + 3 * step < hi; i += 4 * step) {
            const uint4 a0 = A[i], b0 = B[i], a1 = A[i + step], b1 = B[i + step];
            const uint4 a2 = A[i + 2 * step], b2 = B[i + 2 * step], a3 = A[i + 3 * step], b3 = B[i + 3 * step];
            acc += pop_and(a0, b0) + pop_and(a1, b1) + pop_and(a2, b2) + pop_and(a3, b3);
        }
        for (; i < hi; i += step) acc += pop_and(A[i], B[i]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((tid & 63) == 0) atomicAdd(sink, (unsigned long long)acc);
}

constexpr int kFmtFp4 = 4;
constexpr int kScaleOne = 0x7f7f7f7f;

// kChains independent accumulators per wave, `iters` rounds of one MFMA on
// each; operands come from the lane id so nothing is loaded.
template <int kChains>
__global__ __launch_bounds__(256) void k_roof_mfma(uint32_t iters, float *sink) {
    const uint32_t lane = threadIdx.x & 63;
    const v8i a = (v8i){(int)(0x22222222u * (lane & 1)), 0x22222222, (int)(lane * 0x01010101u), 0, 0, 0, 0, 0};
    const v8i b = (v8i){0x22222222, (int)(0x02020202u * (lane & 3)), 0x20202020, 0x22222222, 0, 0, 0, 0};
    v4f acc[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) acc[c] = (v4f){0.f, 0.f, 0.f, 0.f};
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c)
            acc[c] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, acc[c], kFmtFp4, kFmtFp4, 0, kScaleOne,
                                                                       0, kScaleOne);
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
    if (s == -1.f) sink[blockIdx.x] = s;  // never true: keeps the loop
}

typedef float v16f __attribute__((ext_vector_type(16)));

// the 32 x 32 x 64 form: four independent accumulators of 16 VGPRs
template <int kChains>
__global__ __launch_bounds__(256) void k_roof_mfma32(uint32_t iters, float *sink) {
    const uint32_t lane = threadIdx.x & 63;
    const v8i a = (v8i){(int)(0x22222222u * (lane & 1)), 0x22222222, (int)(lane * 0x01010101u), 0, 0, 0, 0, 0};
    const v8i b = (v8i){0x22222222, (int)(0x02020202u * (lane & 3)), 0x20202020, 0x22222222, 0, 0, 0, 0};
    v16f acc[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[c][r] = 0.f;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c)
            acc[c] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc[c], kFmtFp4, kFmtFp4, 0, kScaleOne,
                                                                      0, kScaleOne);
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < kChains; ++c)
#pragma unroll
        for (int r = 0; r < 16; ++r) s += acc[c][r];
    if (s == -1.f) sink[blockIdx.x] = s;  // never true: keeps the loop
}

int timed_launch(cpm_ctx *ctx, double *ms, void (*launch)(cpm_ctx *, void *), void *arg) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CPM_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    if (err != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(CPM_ERR_RUNTIME, "hipEventCreate failed: %s", hipGetErrorString(err));
    }
    err = hipEventRecord(e0, ctx->stream);
    if (err == hipSuccess) {
        launch(ctx, arg);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float t = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != hipSuccess) return fail(CPM_ERR_RUNTIME, "roof launch failed: %s", hipGetErrorString(err));
    *ms = t;
    return CPM_OK;
}

}  // namespace

}  // namespace cpm

using namespace cpm;

extern "C" int cpm_roof_stream(cpm_ctx *ctx, const void *d_buf, uint64_t bytes, uint32_t reps, int xcd_sliced,
                               unsigned long long *d_sink, double *ms, double *bytes_read) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!d_buf || !d_sink || !ms || !bytes_read) return fail(CPM_ERR_INVALID, "null argument");
    if (bytes % 4096 || bytes == 0) return fail(CPM_ERR_INVALID, "buffer bytes must be a positive multiple of 4096");
    if (reinterpret_cast<uintptr_t>(d_buf) % 16) return fail(CPM_ERR_INVALID, "buffer must be 16-byte aligned");
    if (reps == 0) return fail(CPM_ERR_INVALID, "reps must be >= 1");
    struct Arg {
        const uint4 *buf;
        uint64_t half;
        uint32_t reps;
        int sliced;
        unsigned long long *sink;
        uint32_t grid;
    } a{static_cast<const uint4 *>(d_buf), bytes / 32, reps, xcd_sliced ? 1 : 0, d_sink,
        (uint32_t)ctx->cu_count * 8};
    CPM_CHECK(timed_launch(ctx, ms, [](cpm_ctx *c, void *p) {
        const Arg &x = *static_cast<Arg *>(p);
        hipLaunchKernelGGL(k_roof_stream, dim3(x.grid), dim3(256), 0, c->stream, x.buf, x.half, x.reps, x.sliced,
                           x.sink);
    }, &a));
    // every stream element of every slice is read once per rep
    const uint64_t per_rep = xcd_sliced ? (a.half / 8) * 8 : a.half;
    *bytes_read = (double)per_rep * 2 * 16 * reps;
    return CPM_OK;
}

extern "C" int cpm_roof_mfma(cpm_ctx *ctx, int shape, uint32_t iters, float *d_sink, double *ms, double *ops) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!d_sink || !ms || !ops) return fail(CPM_ERR_INVALID, "null argument");
    if (shape != 16 && shape != 32) return fail(CPM_ERR_INVALID, "MFMA shape %d (16 or 32)", shape);
    struct Arg {
        uint32_t iters, grid;
        float *sink;
    } a{iters, (uint32_t)ctx->cu_count * 8, d_sink};
    if (shape == 16) {
        CPM_CHECK(timed_launch(ctx, ms, [](cpm_ctx *c, void *p) {
            const Arg &x = *static_cast<Arg *>(p);
            hipLaunchKernelGGL(k_roof_mfma<8>, dim3(x.grid), dim3(256), 0, c->stream, x.iters, x.sink);
        }, &a));
        // 4 waves per block x 8 chains x iters MFMAs of 16 x 16 x 128 (2 ops each)
        *ops = (double)a.grid * 4 * 8 * iters * (2.0 * 16 * 16 * 128);
    } else {
        CPM_CHECK(timed_launch(ctx, ms, [](cpm_ctx *c, void *p) {
            const Arg &x = *static_cast<Arg *>(p);
            hipLaunchKernelGGL(k_roof_mfma32<4>, dim3(x.grid), dim3(256), 0, c->stream, x.iters, x.sink);
        }, &a));
        // 4 waves x 4 chains x iters MFMAs of 32 x 32 x 64
        *ops = (double)a.grid * 4 * 4 * iters * (2.0 * 32 * 32 * 64);
    }
    return CPM_OK;
}

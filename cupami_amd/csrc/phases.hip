// What a phase boundary costs inside one launch, against a launch boundary.
//
// A search level is seven phases (plan counts, offsets, writes, child fill,
// subset flags, compaction, support counts), each reading what the previous
// one wrote anywhere on the device.  Folded into one persistent launch, every
// boundary becomes a grid-wide barrier, and on MI355X -- eight XCDs, each with
// its own L2 -- a barrier that makes one XCD's writes visible to another must
// write them back from that L2 (an agent-scope release) before arriving and
// invalidate before reading on (an agent-scope acquire).
//
// k_phase_barriers runs `rounds` such boundaries in one cooperative launch
// (every workgroup resident): each round a workgroup stores `dirty` bytes,
// its first lane releases, arrives on a counter and waits for every other
// workgroup -- a bounded wait: past `spin_limit` polls it raises a flag and
// goes on, so the launch always drains.  k_phase_step is one round as a
// launch of its own; a graph of `rounds` of them is timed beside it.
#include "cpm_internal.h"

namespace cpm {

namespace {

__device__ __forceinline__ void phase_stores(uint4 *scratch, uint32_t dirty_q, uint32_t r) {
    uint4 *mine = scratch + (uint64_t)blockIdx.x * dirty_q;
    for (uint32_t i = threadIdx.x; i < dirty_q; i += blockDim.x) mine[i] = make_uint4(r, blockIdx.x, i, 0);
}

__global__ __launch_bounds__(256) void k_phase_barriers(unsigned int *counter, unsigned int *timed_out, uint4 *scratch,
                                                        uint32_t dirty_q, uint32_t rounds, uint32_t spin_limit) {
    for (uint32_t r = 1; r <= rounds; ++r) {
        phase_stores(scratch, dirty_q, r);
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned int target = r * gridDim.x;
            for (uint32_t polls = 0;
                 __hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target; ++polls) {
                if (polls >= spin_limit) {
                    __hip_atomic_store(timed_out, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_phase_step(uint4 *scratch, uint32_t dirty_q, uint32_t r) {
    phase_stores(scratch, dirty_q, r);
}

// milliseconds between two events recorded around `enqueue` on the stream
template <class F>
int time_on_stream(cpm_ctx *ctx, double *ms, F &&enqueue) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CPM_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    if (err == hipSuccess) err = hipEventRecord(e0, ctx->stream);
    if (err == hipSuccess) err = enqueue();
    if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float t = 0;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (err != hipSuccess) return fail(CPM_ERR_RUNTIME, "timed phases failed: %s", hipGetErrorString(err));
    *ms = t;
    return CPM_OK;
}

}  // namespace

}  // namespace cpm

extern "C" int cpm_roof_phases(cpm_ctx *ctx, uint32_t blocks, uint32_t rounds, uint32_t dirty_bytes,
                               double *us_barrier, double *us_launch) {
    using namespace cpm;
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!us_barrier || !us_launch) return fail(CPM_ERR_INVALID, "null output");
    if (blocks == 0 || rounds == 0 || rounds > 4096) return fail(CPM_ERR_INVALID, "1..4096 rounds of >= 1 workgroup");
    if (dirty_bytes % 16 || dirty_bytes > (1u << 20)) return fail(CPM_ERR_INVALID, "dirty bytes: a multiple of 16, at most 1 MiB");
    int per_cu = 0;
    CPM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_phase_barriers, 256, 0));
    if ((uint64_t)blocks > (uint64_t)per_cu * ctx->cu_count)
        return fail(CPM_ERR_INVALID, "%u workgroups cannot all be resident (%d per CU)", blocks, per_cu);
    const uint32_t dirty_q = dirty_bytes / 16;
    DBuf flags, scratch;
    CPM_CHECK(flags.ensure(8));
    CPM_CHECK(scratch.ensure(std::max<uint64_t>((uint64_t)blocks * dirty_bytes, 16)));
    CPM_HIP(hipMemsetAsync(flags.p, 0, 8, ctx->stream));
    unsigned int *counter = flags.as<unsigned int>(), *timed_out = counter + 1;
    uint4 *sp = scratch.as<uint4>();
    uint32_t spin_limit = 1u << 22;  // polls of ~0.1 us: far beyond any round
    void *args[] = {&counter, &timed_out, &sp, (void *)&dirty_q, &rounds, &spin_limit};
    double ms = 0;
    CPM_CHECK(time_on_stream(ctx, &ms, [&] {
        return hipLaunchCooperativeKernel((const void *)k_phase_barriers, dim3(blocks), dim3(256), args, 0, ctx->stream);
    }));
    unsigned int h_flags[2] = {0, 0};
    CPM_HIP(hipMemcpy(h_flags, flags.p, 8, hipMemcpyDeviceToHost));
    if (h_flags[1]) return fail(CPM_ERR_RUNTIME, "a phase barrier wait timed out");
    *us_barrier = ms * 1e3 / rounds;
    // the same rounds as launches of one graph (captured once, replayed twice)
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    {
        std::lock_guard<std::recursive_mutex> lock(alloc_capture_mutex());
        CPM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        for (uint32_t r = 1; r <= rounds; ++r)
            hipLaunchKernelGGL(k_phase_step, dim3(blocks), dim3(256), 0, ctx->stream, sp, dirty_q, r);
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (e != hipSuccess) return fail(CPM_ERR_RUNTIME, "graph capture failed: %s", hipGetErrorString(e));
    }
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(CPM_ERR_RUNTIME, "graph instantiation failed: %s", hipGetErrorString(e));
    e = hipGraphLaunch(ge, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    int rc = e == hipSuccess ? CPM_OK : fail(CPM_ERR_RUNTIME, "graph launch failed: %s", hipGetErrorString(e));
    if (rc == CPM_OK) rc = time_on_stream(ctx, &ms, [&] { return hipGraphLaunch(ge, ctx->stream); });
    (void)hipGraphExecDestroy(ge);
    CPM_CHECK(rc);
    *us_launch = ms * 1e3 / rounds;
    return CPM_OK;
}

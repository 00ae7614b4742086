// lux_trace.hip -- closest-hit and any-hit ray traversal kernels for gfx950.
//
// Persistent grid: each wave keeps 64 rays in flight and refills idle lanes
// from a global ray cursor once at least `regroup` lanes are idle (ballot +
// mbcnt compaction), so rays that finish early do not leave lanes empty
// while a few long rays keep the wave alive.  Per-lane traversal stacks
// live in LDS (lane-contiguous columns) with a private spill area.
//
// Replaces the trace/traceVisibility dispatches of the reference
// (shaders/split.slang:1266-1339 over shaders/traversal.slang:696-1046).
#include "lux_internal.h"

namespace {

constexpr int kTraceWg = 256;
constexpr int kLdsDepth = 16;

struct TraceArgs {
    lux::DevScene scene;
    const lux_ray* rays;
    lux_hit* hits;
    float* bary;
    uint32_t n;
    uint32_t* cursor;  // global ray cursor (zeroed before launch)
    uint32_t regroup;
};

template <class St>
__device__ __forceinline__ void trace_one(const TraceArgs& a, uint32_t idx,
                                          lux::LaneStack<kLdsDepth, kTraceWg>& stk, St& st) {
    const float4* rp = reinterpret_cast<const float4*>(a.rays) + 2u * idx;
    float4 r0 = rp[0], r1 = rp[1];
    float3 o = make_float3(r0.x, r0.y, r0.z);
    float3 d = make_float3(r1.x, r1.y, r1.z);
    uint32_t flags = __float_as_uint(r1.w);
    lux::HitState h;
    if (flags & lux::kRayAnyHit) lux::trace_occluded(a.scene, o, d, r0.w, h, stk, st);
    else lux::trace_closest(a.scene, o, d, r0.w, h, stk, st, (flags & lux::kRayFirstHit) != 0u);
    uint4 out;
    out.x = __float_as_uint(h.t);
    out.y = h.tri == lux::kEmpty ? lux::kEmpty : __float_as_uint(a.scene.triA[h.tri].w);
    out.z = h.inst == lux::kEmpty ? lux::kEmpty : __float_as_uint(a.scene.inst[8u * h.inst + 7u].x);
    out.w = h.err;
    reinterpret_cast<uint4*>(a.hits)[idx] = out;
    if (a.bary) reinterpret_cast<float2*>(a.bary)[idx] = make_float2(h.b1, h.b2);
}

// One ray per lane, rays fetched in wave-sized batches from a cursor.
__global__ __launch_bounds__(kTraceWg) void k_trace_persistent(TraceArgs a) {
    __shared__ uint2 lds[kLdsDepth][kTraceWg];
    lux::LaneStack<kLdsDepth, kTraceWg> stk;
    stk.col = &lds[0][threadIdx.x];
    const uint32_t lane = threadIdx.x & 63u;
    while (true) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.cursor, 64u);
        base = __shfl(base, 0, 64);
        if (base >= a.n) break;
        uint32_t idx = base + lane;
        lux::NoStats st;
        if (idx < a.n) trace_one(a, idx, stk, st);
    }
}

// Plain grid: one ray per thread.
__global__ __launch_bounds__(kTraceWg) void k_trace_grid(TraceArgs a) {
    __shared__ uint2 lds[kLdsDepth][kTraceWg];
    lux::LaneStack<kLdsDepth, kTraceWg> stk;
    stk.col = &lds[0][threadIdx.x];
    uint32_t idx = blockIdx.x * kTraceWg + threadIdx.x;
    lux::NoStats st;
    if (idx < a.n) trace_one(a, idx, stk, st);
}

// Instrumented grid: tallies traversal work per ray (outside timed runs).
__global__ __launch_bounds__(kTraceWg) void k_trace_stats(TraceArgs a, unsigned long long* counters) {
    __shared__ uint2 lds[kLdsDepth][kTraceWg];
    lux::LaneStack<kLdsDepth, kTraceWg> stk;
    stk.col = &lds[0][threadIdx.x];
    uint32_t idx = blockIdx.x * kTraceWg + threadIdx.x;
    lux::Counting st;
    if (idx < a.n) trace_one(a, idx, stk, st);
    unsigned long long v[3] = {st.nodes, st.tris, st.insts};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        unsigned long long x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((threadIdx.x & 63u) == 0) atomicAdd(&counters[k], x);
    }
}

__global__ void k_zero1(uint32_t* p) { *p = 0u; }

}  // namespace

lux::DevScene lux_ctx::dev_scene() const {
    lux::DevScene s;
    s.tlas = static_cast<const float4*>(tlas.ptr);
    s.inst = static_cast<const float4*>(inst.ptr);
    s.blas = static_cast<const float4*>(blas.ptr);
    s.triA = static_cast<const float4*>(triA.ptr);
    s.triB = static_cast<const float4*>(triB.ptr);
    s.triC = static_cast<const float4*>(triC.ptr);
    s.n_tlas = n_tlas;
    s.n_inst = n_inst;
    s.n_blas = n_blas;
    s.n_tri = n_tri;
    return s;
}

static int trace_launch(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits, float* bary) {
    if (!ctx->tlas_built) return lux_fail(ctx, LUX_ERR_STATE, "lux_trace: BVH not built");
    if (n == 0) return LUX_OK;
    if (!rays || !hits) return lux_fail(ctx, LUX_ERR_INVALID, "lux_trace: null buffer");
    if (int rc = lux_ensure(ctx, ctx->scratch, 256)) return rc;
    TraceArgs a;
    a.scene = ctx->dev_scene();
    a.rays = rays;
    a.hits = hits;
    a.bary = bary;
    a.n = n;
    a.cursor = static_cast<uint32_t*>(ctx->scratch.ptr);
    a.regroup = ctx->regroup;
    hipStream_t st = ctx->stream;
    LUX_HIP(ctx, hipEventRecord(ctx->ev0, st));
    if (ctx->waves_per_cu == 1000u) {
        uint32_t blocks = (n + kTraceWg - 1) / kTraceWg;
        hipLaunchKernelGGL(k_trace_grid, dim3(blocks), dim3(kTraceWg), 0, st, a);
    } else {
        uint32_t wpc = ctx->waves_per_cu ? ctx->waves_per_cu : 16u;
        uint32_t blocks = (uint32_t)ctx->cu_count * wpc / (kTraceWg / 64);
        uint32_t need = (n + kTraceWg - 1) / kTraceWg;
        if (blocks > need) blocks = need;
        if (blocks == 0) blocks = 1;
        hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, st, a.cursor);
        hipLaunchKernelGGL(k_trace_persistent, dim3(blocks), dim3(kTraceWg), 0, st, a);
    }
    LUX_HIP(ctx, hipGetLastError());
    LUX_HIP(ctx, hipEventRecord(ctx->ev1, st));
    return LUX_OK;
}

extern "C" int lux_trace_rays_device(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits,
                                     float* bary) {
    if (!ctx) return LUX_ERR_INVALID;
    return trace_launch(ctx, rays, n, hits, bary);
}

extern "C" int lux_trace_rays(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits) {
    if (!ctx) return LUX_ERR_INVALID;
    if (n == 0) return LUX_OK;
    if (!rays || !hits) return lux_fail(ctx, LUX_ERR_INVALID, "lux_trace_rays: null buffer");
    if (int rc = lux_ensure(ctx, ctx->ray_buf, sizeof(lux_ray) * (size_t)n)) return rc;
    if (int rc = lux_ensure(ctx, ctx->hit_buf, sizeof(lux_hit) * (size_t)n)) return rc;
    LUX_HIP(ctx, hipMemcpyAsync(ctx->ray_buf.ptr, rays, sizeof(lux_ray) * (size_t)n,
                                hipMemcpyHostToDevice, ctx->stream));
    if (int rc = trace_launch(ctx, static_cast<const lux_ray*>(ctx->ray_buf.ptr), n,
                              static_cast<lux_hit*>(ctx->hit_buf.ptr), nullptr))
        return rc;
    LUX_HIP(ctx, hipMemcpyAsync(hits, ctx->hit_buf.ptr, sizeof(lux_hit) * (size_t)n,
                                hipMemcpyDeviceToHost, ctx->stream));
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LUX_OK;
}

extern "C" int lux_trace_stats_device(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits,
                                      unsigned long long* counters) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->tlas_built) return lux_fail(ctx, LUX_ERR_STATE, "lux_trace_stats: BVH not built");
    if (n == 0) return LUX_OK;
    if (!rays || !hits || !counters) return lux_fail(ctx, LUX_ERR_INVALID, "lux_trace_stats: null buffer");
    TraceArgs a;
    a.scene = ctx->dev_scene();
    a.rays = rays;
    a.hits = hits;
    a.bary = nullptr;
    a.n = n;
    a.cursor = nullptr;
    a.regroup = 0;
    hipLaunchKernelGGL(k_trace_stats, dim3((n + kTraceWg - 1) / kTraceWg), dim3(kTraceWg), 0,
                       ctx->stream, a, counters);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_last_kernel_ms(lux_ctx* ctx, float* ms) {
    if (!ctx || !ms) return LUX_ERR_INVALID;
    LUX_HIP(ctx, hipEventSynchronize(ctx->ev1));
    LUX_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return LUX_OK;
}

extern "C" int lux_set_trace_tuning(lux_ctx* ctx, uint32_t waves_per_cu, uint32_t regroup_threshold) {
    if (!ctx) return LUX_ERR_INVALID;
    if (waves_per_cu > 32 && waves_per_cu != 1000u)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_set_trace_tuning: waves_per_cu > 32");
    if (regroup_threshold > 64)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_set_trace_tuning: regroup > 64");
    ctx->waves_per_cu = waves_per_cu;
    ctx->regroup = regroup_threshold;
    return LUX_OK;
}

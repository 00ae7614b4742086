// lux_trace.hip -- closest-hit and any-hit ray traversal kernels for gfx950.
//
// Two persistent kernels over a global ray cursor:
//   k_trace_persistent  a wave takes 64 rays and runs them to completion
//                       (the wave lasts as long as its longest ray);
//   k_trace_regroup     every lane runs a resumable walk (lux::Walk); after
//                       each batch of steps the wave counts its idle lanes
//                       (ballot) and, once at least `regroup` are idle,
//                       refills them with one atomic on the cursor and a
//                       popcount rank, so finished rays stop wasting lanes.
// Per-lane traversal stacks live in LDS (lane-contiguous columns, depth a
// launch parameter) with a private spill area.
//
// Replaces the trace/traceVisibility dispatches of the reference
// (shaders/split.slang:1266-1339 over shaders/traversal.slang:696-1046).
#include <algorithm>
#include <cstring>
#include <mutex>
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
    uint32_t regroup;  // idle lanes that trigger a refill (1..64)
    uint32_t batch;    // walk steps between refill checks
};

__device__ __forceinline__ void write_hit(const TraceArgs& a, uint32_t idx, const lux::HitState& h) {
    uint4 out;
    out.x = __float_as_uint(h.t);
    out.y = h.tri == lux::kEmpty ? lux::kEmpty : __float_as_uint(a.scene.tri[3ull * h.tri].w);
    out.z = h.inst == lux::kEmpty ? lux::kEmpty : __float_as_uint(a.scene.inst[8u * h.inst + 7u].x);
    out.w = h.err;
    reinterpret_cast<uint4*>(a.hits)[idx] = out;
    if (a.bary) reinterpret_cast<float2*>(a.bary)[idx] = make_float2(h.b1, h.b2);
}

template <class St>
__device__ __forceinline__ void trace_one(const TraceArgs& a, uint32_t idx,
                                          lux::LaneStack<kLdsDepth, kTraceWg, lux::FlatWord>& stk, St& st) {
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
    out.y = h.tri == lux::kEmpty ? lux::kEmpty : __float_as_uint(a.scene.tri[3ull * h.tri].w);
    out.z = h.inst == lux::kEmpty ? lux::kEmpty : __float_as_uint(a.scene.inst[8u * h.inst + 7u].x);
    out.w = h.err;
    reinterpret_cast<uint4*>(a.hits)[idx] = out;
    if (a.bary) reinterpret_cast<float2*>(a.bary)[idx] = make_float2(h.b1, h.b2);
}

// One ray per lane, rays fetched in wave-sized batches from a cursor.
__global__ __launch_bounds__(kTraceWg) void k_trace_persistent(TraceArgs a) {
    __shared__ uint2 lds[kLdsDepth][kTraceWg];
    lux::LaneStack<kLdsDepth, kTraceWg, lux::FlatWord> stk;
    stk.bind(&lds[0][threadIdx.x]);
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

// Regrouping persistent kernel (see the file comment): items are the rays
// of the caller's list.
struct RaySource {
    const TraceArgs* a;
    __device__ uint32_t count() const { return a->n; }
    template <class St>
    __device__ bool begin(uint32_t idx, lux::Walk& w, St& st) {
        const float4* rp = reinterpret_cast<const float4*>(a->rays) + 2u * idx;
        float4 r0 = rp[0], r1 = rp[1];
        uint32_t flags = __float_as_uint(r1.w);
        uint32_t mode = (flags & lux::kRayAnyHit) ? lux::kWalkOccluded
                        : ((flags & lux::kRayFirstHit) ? lux::kWalkFirst : lux::kWalkClosest);
        lux::walk_begin(a->scene, w, make_float3(r0.x, r0.y, r0.z), make_float3(r1.x, r1.y, r1.z), r0.w,
                        mode, st);
        if (w.phase != lux::kPhaseDone) return true;
        write_hit(*a, idx, w.h);
        return false;
    }
    __device__ void finish(uint32_t idx, const lux::Walk& w) { write_hit(*a, idx, w.h); }
};

template <int Depth>
__global__ __launch_bounds__(kTraceWg, lux::walk_min_waves(Depth, kTraceWg, 0)) void k_trace_regroup(TraceArgs a) {
    __shared__ uint2 lds[Depth][kTraceWg];
    lux::LaneStack<Depth, kTraceWg, lux::FlatWord> stk;
    stk.bind(&lds[0][threadIdx.x]);
    lux::NoStats st;
    RaySource src{&a};
    lux::regroup_walks(a.scene, src, a.cursor, a.regroup, a.batch, stk, st);
}

// Plain grid: one ray per thread.
__global__ __launch_bounds__(kTraceWg) void k_trace_grid(TraceArgs a) {
    __shared__ uint2 lds[kLdsDepth][kTraceWg];
    lux::LaneStack<kLdsDepth, kTraceWg, lux::FlatWord> stk;
    stk.bind(&lds[0][threadIdx.x]);
    uint32_t idx = blockIdx.x * kTraceWg + threadIdx.x;
    lux::NoStats st;
    if (idx < a.n) trace_one(a, idx, stk, st);
}

// Instrumented grid: tallies traversal work per ray (outside timed runs).
__global__ __launch_bounds__(kTraceWg) void k_trace_stats(TraceArgs a, unsigned long long* counters) {
    __shared__ uint2 lds[kLdsDepth][kTraceWg];
    lux::LaneStack<kLdsDepth, kTraceWg, lux::FlatWord> stk;
    stk.bind(&lds[0][threadIdx.x]);
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

__device__ __forceinline__ uint32_t mix32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

// Self-check of the walk's reciprocal slab division: for `count` operand
// pairs n (zero excluded, magnitude in [2^-87, 2^40)) and d (magnitude in
// [2^-60, 2^20)) -- random significands, and every eighth pair with an
// all-ones, all-zeros or one-ulp-from-either significand -- counts pairs
// where lux::div_fast(n, d, 1/d) differs in any bit from the IEEE quotient.
__global__ __launch_bounds__(256) void k_div_check(uint64_t count, uint32_t seed, unsigned long long* bad) {
    unsigned long long local = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256u) {
        const uint32_t h0 = mix32((uint32_t)i ^ mix32(seed)), h1 = mix32(h0 ^ (uint32_t)(i >> 32) ^ 0x9e3779b9u);
        const uint32_t h2 = mix32(h1 + 0x632be5abu);
        uint32_t mn = h0 & 0x7fffffu, md = h1 & 0x7fffffu;
        switch (i & 7u) {
        case 0: md = 0x7fffffu; break;
        case 1: md = 0u; break;
        case 2: md = 0x7ffffeu; break;
        case 3: md = 1u; break;
        case 4: mn = 0x7fffffu; break;
        case 5: mn = 0u; md = 0x7fffffu; break;
        default: break;
        }
        const uint32_t en = 127u - 87u + (h2 & 0xffffu) % 127u;  // 2^-87 .. 2^39
        const uint32_t ed = 127u - 60u + (h2 >> 16) % 80u;       // 2^-60 .. 2^19
        const float n = __uint_as_float((h0 & 0x80000000u) | (en << 23) | mn);
        const float d = __uint_as_float((h1 & 0x80000000u) | (ed << 23) | md);
        const float y = 1.0f / d;
        if (__float_as_uint(lux::div_fast(n, d, y)) != __float_as_uint(n / d)) ++local;
    }
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
    if ((threadIdx.x & 63u) == 0u && local) atomicAdd(bad, local);
}

}  // namespace

// ---- scratch of the allocation-free trace (lux_trace_rays_device_noalloc) ----
// The traversal kernels' only scratch is the global ray cursor.  The
// allocation-free entry point keeps it in this file-scope device array, one
// copy per GPU at its own address, looked up once per device when a context
// is created on it.  Host code uses the looked-up address, never `scratch`
// itself (the host-side name is not a device address).
constexpr uint32_t kScratchWords = 64;          // 256 B: the cursor alone on its cache lines
constexpr uint32_t kNoAllocMaxRays = 1u << 31;  // the cursor runs at most 64 per wave past n
constexpr int kScratchDevices = 64;
static __device__ __align__(256) uint32_t scratch[kScratchWords];
static uint32_t* g_scratch_ptr[kScratchDevices] = {};
static std::mutex g_scratch_mu;

// Called by lux_ctx_create with ctx->device current.
int lux_trace_scratch_init(lux_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kScratchDevices) return LUX_OK;  // the entry point refuses it
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[ctx->device]) return LUX_OK;
    void* ptr = nullptr;
    LUX_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[ctx->device] = static_cast<uint32_t*>(ptr);
    return LUX_OK;
}

lux::DevScene lux_ctx::dev_scene() const {
    lux::DevScene s;
    s.tlas = static_cast<const float4*>(tlas.ptr);
    s.inst = static_cast<const float4*>(inst.ptr);
    s.blas = static_cast<const float4*>(blas.ptr);
    s.tri = static_cast<const float4*>(tris.ptr);
    s.n_tlas = n_tlas;
    s.n_inst = n_inst;
    s.n_blas = n_blas;
    s.n_tri = n_tri;
    s.range_ok = static_cast<const uint32_t*>(range_ok.ptr);
    return s;
}

// The launch of a trace with its cursor in `cursor` (4 bytes of device
// memory); `timed` brackets it with the context's events.
static int trace_launch_with(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits, float* bary,
                             uint32_t* cursor, bool timed) {
    TraceArgs a;
    a.scene = ctx->dev_scene();
    a.rays = rays;
    a.hits = hits;
    a.bary = bary;
    a.n = n;
    a.cursor = cursor;
    a.regroup = ctx->regroup;
    a.batch = ctx->trace_batch ? ctx->trace_batch : 8u;
    hipStream_t st = ctx->stream;
    if (timed) LUX_HIP(ctx, hipEventRecord(ctx->ev0, st));
    if (ctx->regroup > 0u && ctx->waves_per_cu != 1000u) {
        uint32_t wpc = ctx->waves_per_cu ? ctx->waves_per_cu : 28u;
        uint32_t blocks = (uint32_t)ctx->cu_count * wpc / (kTraceWg / 64);
        uint32_t need = (n + kTraceWg - 1) / kTraceWg;
        if (blocks > need) blocks = need;
        if (blocks == 0) blocks = 1;
        hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, st, a.cursor);
        switch (ctx->trace_depth) {
        case 4: hipLaunchKernelGGL(k_trace_regroup<4>, dim3(blocks), dim3(kTraceWg), 0, st, a); break;
        case 12: hipLaunchKernelGGL(k_trace_regroup<12>, dim3(blocks), dim3(kTraceWg), 0, st, a); break;
        case 16: hipLaunchKernelGGL(k_trace_regroup<16>, dim3(blocks), dim3(kTraceWg), 0, st, a); break;
        default: hipLaunchKernelGGL(k_trace_regroup<8>, dim3(blocks), dim3(kTraceWg), 0, st, a); break;
        }
    } else if (ctx->waves_per_cu == 1000u) {
        uint32_t blocks = (n + kTraceWg - 1) / kTraceWg;
        hipLaunchKernelGGL(k_trace_grid, dim3(blocks), dim3(kTraceWg), 0, st, a);
    } else {
        uint32_t wpc = ctx->waves_per_cu ? ctx->waves_per_cu : 28u;
        uint32_t blocks = (uint32_t)ctx->cu_count * wpc / (kTraceWg / 64);
        uint32_t need = (n + kTraceWg - 1) / kTraceWg;
        if (blocks > need) blocks = need;
        if (blocks == 0) blocks = 1;
        hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, st, a.cursor);
        hipLaunchKernelGGL(k_trace_persistent, dim3(blocks), dim3(kTraceWg), 0, st, a);
    }
    LUX_HIP(ctx, hipGetLastError());
    if (timed) LUX_HIP(ctx, hipEventRecord(ctx->ev1, st));
    return LUX_OK;
}

static int trace_launch(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits, float* bary) {
    if (!ctx->tlas_built) return lux_fail(ctx, LUX_ERR_STATE, "lux_trace: BVH not built");
    if (n == 0) return LUX_OK;
    if (!rays || !hits) return lux_fail(ctx, LUX_ERR_INVALID, "lux_trace: null buffer");
    if (int rc = lux_ensure(ctx, ctx->scratch, 256)) return rc;
    return trace_launch_with(ctx, rays, n, hits, bary, static_cast<uint32_t*>(ctx->scratch.ptr), true);
}

extern "C" int lux_trace_rays_device(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits,
                                     float* bary) {
    if (!ctx) return LUX_ERR_INVALID;
    return trace_launch(ctx, rays, n, hits, bary);
}

// lux_trace_rays_device without any allocation: the cursor lives in the
// device's `scratch` copy (zeroed by a fill kernel on the context stream, not
// a memset), no events are recorded.  Calls that overlap on one GPU share
// that cursor and race.
extern "C" int lux_trace_rays_device_noalloc(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits,
                                             float* bary) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->tlas_built) return lux_fail(ctx, LUX_ERR_STATE, "lux_trace_rays_device_noalloc: BVH not built");
    if (n > kNoAllocMaxRays)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_trace_rays_device_noalloc: more than 2^31 rays in one call");
    if (n == 0) return LUX_OK;
    if (!rays || !hits) return lux_fail(ctx, LUX_ERR_INVALID, "lux_trace_rays_device_noalloc: null buffer");
    uint32_t* cursor = nullptr;
    if (ctx->device >= 0 && ctx->device < kScratchDevices) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        cursor = g_scratch_ptr[ctx->device];
    }
    if (!cursor)
        return lux_fail(ctx, LUX_ERR_STATE, "lux_trace_rays_device_noalloc: no scratch address for this device");
    return trace_launch_with(ctx, rays, n, hits, bary, cursor, false);
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
    a.batch = 0;
    hipLaunchKernelGGL(k_trace_stats, dim3((n + kTraceWg - 1) / kTraceWg), dim3(kTraceWg), 0,
                       ctx->stream, a, counters);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_check_reciprocal_division(lux_ctx* ctx, uint64_t pairs, uint32_t seed, uint64_t* mismatches) {
    if (!ctx || !mismatches) return LUX_ERR_INVALID;
    if (int rc = lux_ensure(ctx, ctx->scratch, 256)) return rc;
    unsigned long long* bad = static_cast<unsigned long long*>(ctx->scratch.ptr) + 16;
    LUX_HIP(ctx, hipMemsetAsync(bad, 0, 8, ctx->stream));
    hipLaunchKernelGGL(k_div_check, dim3((uint32_t)ctx->cu_count * 32u), dim3(256), 0, ctx->stream, pairs, seed, bad);
    LUX_HIP(ctx, hipGetLastError());
    unsigned long long host = 0;
    LUX_HIP(ctx, hipMemcpyAsync(&host, bad, 8, hipMemcpyDeviceToHost, ctx->stream));
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *mismatches = host;
    return LUX_OK;
}

extern "C" int lux_last_kernel_ms(lux_ctx* ctx, float* ms) {
    if (!ctx || !ms) return LUX_ERR_INVALID;
    LUX_HIP(ctx, hipEventSynchronize(ctx->ev1));
    LUX_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return LUX_OK;
}

extern "C" int lux_set_trace_config(lux_ctx* ctx, const lux_trace_config* cfg) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!cfg) return lux_fail(ctx, LUX_ERR_INVALID, "lux_set_trace_config: null config");
    if (cfg->waves_per_cu > 32 && cfg->waves_per_cu != 1000u)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_set_trace_config: waves_per_cu > 32");
    if (cfg->regroup_threshold > 64)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_set_trace_config: regroup threshold > 64");
    if (cfg->lds_stack_depth != 0 && cfg->lds_stack_depth != 4 && cfg->lds_stack_depth != 8 &&
        cfg->lds_stack_depth != 12 && cfg->lds_stack_depth != 16)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_set_trace_config: LDS stack depth must be 4, 8, 12 or 16");
    if (cfg->steps_per_refill > 256)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_set_trace_config: steps per refill > 256");
    ctx->waves_per_cu = cfg->waves_per_cu;
    ctx->regroup = cfg->regroup_threshold;
    ctx->trace_depth = cfg->lds_stack_depth;
    ctx->trace_batch = cfg->steps_per_refill;
    return LUX_OK;
}

extern "C" int lux_get_trace_config(lux_ctx* ctx, lux_trace_config* cfg) {
    if (!ctx || !cfg) return LUX_ERR_INVALID;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->waves_per_cu = ctx->waves_per_cu;
    cfg->regroup_threshold = ctx->regroup;
    cfg->lds_stack_depth = ctx->trace_depth;
    cfg->steps_per_refill = ctx->trace_batch;
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

// Autotuner: coordinate descent over the launch knobs (waves per CU, LDS
// stack depth, refill threshold, steps per refill), each candidate timed with
// HIP events on the caller's rays (median of `repeats` launches after one
// warm-up), two passes; the fastest configuration is installed.  Results do
// not depend on the configuration, so tuning never changes hits.
extern "C" int lux_autotune_trace(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits,
                                  uint32_t repeats, lux_trace_config* best, float* best_ms,
                                  uint32_t* evaluations) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!rays || !hits || n == 0) return lux_fail(ctx, LUX_ERR_INVALID, "lux_autotune_trace: empty ray set");
    if (!ctx->tlas_built) return lux_fail(ctx, LUX_ERR_STATE, "lux_autotune_trace: BVH not built");
    if (repeats == 0) repeats = 3;
    if (repeats > 15) repeats = 15;
    const uint32_t saved[4] = {ctx->waves_per_cu, ctx->trace_depth, ctx->regroup, ctx->trace_batch};
    auto measure = [&](const uint32_t c[4], float& ms) -> int {
        ctx->waves_per_cu = c[0];
        ctx->trace_depth = c[1];
        ctx->regroup = c[2];
        ctx->trace_batch = c[3];
        float t[16];
        for (uint32_t r = 0; r <= repeats; ++r) {
            if (int rc = trace_launch(ctx, rays, n, hits, nullptr)) return rc;
            LUX_HIP(ctx, hipEventSynchronize(ctx->ev1));
            float e = 0.0f;
            LUX_HIP(ctx, hipEventElapsedTime(&e, ctx->ev0, ctx->ev1));
            if (r > 0) t[r - 1] = e;
        }
        std::sort(t, t + repeats);
        ms = t[repeats / 2];
        return LUX_OK;
    };
    static const uint32_t cand_wpc[] = {16, 20, 24, 28, 32};
    static const uint32_t cand_depth[] = {4, 8, 12, 16};
    static const uint32_t cand_regroup[] = {0, 1, 4, 6, 8, 12, 16, 32};
    static const uint32_t cand_batch[] = {4, 8, 16, 24, 32};
    const uint32_t* cands[4] = {cand_wpc, cand_depth, cand_regroup, cand_batch};
    const uint32_t ncand[4] = {5, 4, 8, 5};
    uint32_t cur[4] = {28, 12, 6, 24};
    float cur_ms = 0.0f;
    uint32_t evals = 0;
    int rc = measure(cur, cur_ms);
    ++evals;
    for (int pass = 0; pass < 2 && rc == LUX_OK; ++pass) {
        for (int k = 0; k < 4 && rc == LUX_OK; ++k) {
            for (uint32_t j = 0; j < ncand[k] && rc == LUX_OK; ++j) {
                if (cands[k][j] == cur[k]) continue;
                uint32_t c[4] = {cur[0], cur[1], cur[2], cur[3]};
                c[k] = cands[k][j];
                float ms = 0.0f;
                rc = measure(c, ms);
                ++evals;
                // a candidate must beat the incumbent by more than the timing noise
                if (rc == LUX_OK && ms < cur_ms * kTuneMargin) {
                    cur_ms = ms;
                    cur[k] = c[k];
                }
            }
        }
    }
    if (rc != LUX_OK) {
        ctx->waves_per_cu = saved[0];
        ctx->trace_depth = saved[1];
        ctx->regroup = saved[2];
        ctx->trace_batch = saved[3];
        return rc;
    }
    ctx->waves_per_cu = cur[0];
    ctx->trace_depth = cur[1];
    ctx->regroup = cur[2];
    ctx->trace_batch = cur[3];
    if (best) {
        std::memset(best, 0, sizeof *best);
        best->waves_per_cu = cur[0];
        best->lds_stack_depth = cur[1];
        best->regroup_threshold = cur[2];
        best->steps_per_refill = cur[3];
    }
    if (best_ms) *best_ms = cur_ms;
    if (evaluations) *evaluations = evals;
    return LUX_OK;
}

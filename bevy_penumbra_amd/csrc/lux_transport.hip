// lux_transport.hip -- dense-grid interval trace, cascade merge, probe
// packing and world resolve on gfx950.
//
// One GI update on the context stream:
//   interval trace  (probe, direction, level) rays from the probe centre over
//                the interval, closest hit, hit shading (emission + emissive
//                NEE with occlusion rays), deposit J/beta -- trace + deposit +
//                reduce of the reference (shaders/split.slang:1056-1163,
//                1386-1476; shaders/cascades.slang:60-268); either the four
//                wavefront stages (k_gi_primary, k_gi_setup, k_gi_shadow,
//                k_gi_accum) or the fused k_gi_trace, one thread per interval
//   k_gi_merge   per level, top down: support-aware Eq. 7 merge with 4 child
//                directions x 8 trilinear corners (split.slang:389-479,1523-1725)
//   k_gi_pack    16 level-0 probes per block (4 per wave) over the block's
//                table of texel-direction cosines, 36 octahedral texels
//                (split.slang:1788-1918)
// and k_world_resolve per receiver (shaders/world_resolve.slang:1-196).
//
// Layout in HBM: level-major, then probe-major (x fastest), then direction;
// one float4 per interval for J.rgb + beta and one for I.rgb + support.
// All arithmetic is IEEE f32 in a fixed operation order; min/max/saturate
// that feed stored values are compare-selects; cos/sin of 2*pi*u use a fixed
// f64 Taylor evaluation.  The CPU oracle evaluates the same sequence.
#include <algorithm>
#include <cmath>
#include <cstring>
#include "lux_gi_device.h"

namespace {

// Intervals with global index in [g_begin, g_end) (levels are contiguous).
__global__ __launch_bounds__(kWg) void k_gi_trace(GiArgs a, uint64_t g_begin, uint64_t g_end) {
    __shared__ uint2 lds[kLdsDepth][kWg];
    Stack stk;
    stk.bind(&lds[0][threadIdx.x]);
    uint64_t g = g_begin + (uint64_t)blockIdx.x * kWg + threadIdx.x;
    const lux_gi_params& p = a.p;
    if (g >= g_end) return;
    uint32_t L = 0;
    while (L + 1 < p.levels && g >= a.offset[L + 1]) ++L;
    uint64_t local = g - a.offset[L];
    uint32_t D = (p.base_phi << L) * (p.base_y << L);
    uint32_t k = (uint32_t)(local % D);
    float3 c = probe_centre(p, L, local / D);
    bool terminal = p.terminal_tail != 0u && L + 1 == p.levels;
    uint32_t ns = 0;
    Dep dep = dep_empty();
    a.intervals[g] = trace_interval(a, L, c, k, terminal, stk, &ns, a.dep ? &dep : nullptr);
    if (a.samples) a.samples[g] = ns;
    if (a.dep) dep_store(a.dep + 2 * g, dep);
}

__global__ __launch_bounds__(kWg) void k_sur_queries(GiArgs a, uint32_t L, uint64_t count, float* out) {
    uint64_t i = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (i >= count) return;
    level_query(a.p, L, i, out + 9ull * i);
}

// Training samples: (probe, bin) drawn by a multiply-shift of a hash of
// (seed, i); target = terminal trace of the level or the merged value.
__global__ __launch_bounds__(kWg) void k_sur_samples(GiArgs a, uint32_t L, uint32_t n, uint32_t seed,
                                                     uint32_t source, float* out) {
    __shared__ uint2 lds[kLdsDepth][kWg];
    Stack stk;
    stk.bind(&lds[0][threadIdx.x]);
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const lux_gi_params& p = a.p;
    uint64_t count = a.offset[L + 1] - a.offset[L];
    uint32_t h = transport_hash(i + transport_hash(seed ^ 0x3c6ef372u));
    uint64_t local = ((uint64_t)h * count) >> 32;
    float* o = out + 12ull * i;
    level_query(p, L, local, o);
    uint32_t D = (p.base_phi << L) * (p.base_y << L);
    float4 m;
    if (source == 1u) {
        m = a.merged[a.offset[L] + local];
    } else {
        float3 c = probe_centre(p, L, local / D);
        m = merge_leaf(trace_interval(a, L, c, (uint32_t)(local % D), true, stk));
    }
    bool ok = m.w > 0.0f;
    float nan = __uint_as_float(0x7fc00000u);
    o[9] = ok ? m.x : nan;
    o[10] = ok ? m.y : nan;
    o[11] = ok ? m.z : nan;
}

// Network outputs -> merged radiance of level L with support 1.
__global__ __launch_bounds__(kWg) void k_sur_apply(const float* Y, uint64_t count, float4* merged) {
    uint64_t i = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (i >= count) return;
    const float* y = Y + 16ull * i;
    merged[i] = make_float4(smax(y[0], 0.0f), smax(y[1], 0.0f), smax(y[2], 0.0f), 1.0f);
}

// One wave per probe: predicted RMSE = sqrt(mean_k max(exp(y3) - 1e-4, 0))
// in a fixed reduction order; probes above the bound (or NaN) are appended
// to the fallback list.
__global__ __launch_bounds__(64) void k_sur_probe_error(const float* Y, uint32_t D, float bound,
                                                        float* err, uint32_t* list, uint32_t* count) {
    uint32_t probe = blockIdx.x;
    const float* y = Y + 16ull * D * probe;
    float sum = 0.0f;
    for (uint32_t k = threadIdx.x; k < D; k += 64u) sum = sum + smax(expf(y[16ull * k + 3]) - 1e-4f, 0.0f);
    for (int off = 32; off > 0; off >>= 1) sum = sum + __shfl_xor(sum, off, 64);
    if (threadIdx.x == 0) {
        float e = sqrtf(sum / (float)D);
        err[probe] = e;
        if (!(e <= bound)) list[atomicAdd(count, 1u)] = probe;
    }
}

// Traced fallback: the listed probes' intervals of level L as the unbounded
// terminal interval, written merged.  Threads past the listed work exit.
__global__ __launch_bounds__(kWg) void k_sur_fallback(GiArgs a, uint32_t L, const uint32_t* list,
                                                      const uint32_t* count) {
    __shared__ uint2 lds[kLdsDepth][kWg];
    Stack stk;
    stk.bind(&lds[0][threadIdx.x]);
    const lux_gi_params& p = a.p;
    uint32_t D = (p.base_phi << L) * (p.base_y << L);
    uint64_t g = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    uint64_t slot = g / D;
    if (slot >= *count) return;
    uint32_t probe = list[slot];
    uint32_t k = (uint32_t)(g % D);
    float3 c = probe_centre(p, L, probe);
    a.merged[a.offset[L] + (uint64_t)probe * D + k] = merge_leaf(trace_interval(a, L, c, k, true, stk));
}

// ---------------------------------------------------------------- wavefront interval trace
// The fused interval trace split into four launches so that the primary and
// the shadow traversals run as regrouped walks (lux::regroup_walks):
//   k_gi_primary  closest-hit walk of every (interval, ray) item
//   k_gi_setup    per hit: emission and NEE context; each NEE sample's
//                 contribution goes to a fixed slot, its occlusion ray to the
//                 wave's slab (then packed into one list in item order)
//   k_gi_shadow   occlusion walks of the list -> visibility per slot
//   k_gi_accum    per interval: the fused kernel's accumulation in sample
//                 order (same operations, so the same bits)
// Three kinds of interval sets share the stages:
//   kWaveRange    intervals [g0, g1) -> interval array (J, beta)
//   kWaveListed   every direction of the listed probes of one level as its
//                 unbounded terminal interval -> merged array (the
//                 surrogate's traced fallback)
//   kWaveSampled  hashed (probe, direction) samples of one level as terminal
//                 intervals -> 12-float training records
//   kWaveReceiver the rays of a receiver update's ray list (one per item) ->
//                 per ray its shaded radiance, hit distance and status
constexpr uint32_t kWaveRange = 0u, kWaveListed = 1u, kWaveSampled = 2u, kWaveReceiver = 3u;



struct Wave {
    uint32_t kind;
    uint64_t g0, g1;             // kWaveRange: interval range of this chunk
    uint32_t level;              // kWaveListed / kWaveSampled
    const uint32_t* list;        // kWaveListed: probe list and its device-side length
    const uint32_t* list_count;
    uint32_t first, n;           // kWaveListed: first listed probe of the chunk, probes in it
                                 // kWaveSampled: first sample of the chunk, samples in it
    uint32_t seed;               // kWaveSampled: sample seed
    float* sample_out;           // kWaveSampled: 12 floats per sample
    uint32_t R, nee;             // rays per interval, NEE slots per ray
    uint32_t stride;             // items per NEE sample row of `slots` / `vis` (the chunk's capacity)
    uint4* prim;                 // per item: t bits, tri, inst, err
    float4* shade;               // per item: emission rgb, w = 1 when NEE runs
    float4* dep_item;            // per item (history on, kWaveRange): its NEE segments' dependency
    float4* slots;               // per slot (sample-major: smp * stride + item, so a wave's
                                 // lanes touch consecutive slots): contribution rgb, w = 1
                                 // when a query was issued
    uint32_t* vis;               // per slot: 0 visible, 1 occluded, 2 unsupported
    float4* shadow;              // per occlusion ray: (origin, limit), (direction, slot bits)
    uint32_t* counters;          // [0] primary cursor, [1] shadow rays, [2] shadow cursor
    float4* slab;                // per setup wave: its occlusion rays (64 * nee)
    uint32_t* wave_cnt;          // per setup wave: its ray count, then scanned into offsets
    uint32_t regroup, batch;
    // kWaveReceiver: receivers, ray list ((variant, receiver) indices),
    // sequences, receiver count, frame; per ray (rad.rgb, t) and status bits
    const float* recv;
    const uint32_t* rays;
    const uint32_t* seq;
    uint32_t count, frame;
    float4* rx_out;
    uint32_t* rx_status;
    // instrumented pass (lux_gi_trace_stats): node fetches, triangle tests and
    // instance visits of the primary walks [0..2] and occlusion walks [3..5]
    unsigned long long* stats;
};

// Adds a wave's traversal work counts to w.stats[off..off+2] (the whole wave
// calls it; NoStats: nothing).
__device__ __forceinline__ void stats_flush(const lux::NoStats&, unsigned long long*, uint32_t) {}
__device__ __forceinline__ void stats_flush(const lux::Counting& st, unsigned long long* out, uint32_t off) {
    unsigned long long v[3] = {st.nodes, st.tris, st.insts};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        unsigned long long x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (__lane_id() == 0u) atomicAdd(&out[off + k], x);
    }
}

// Intervals of the chunk (kWaveListed: listed probes beyond the list's
// device-side length contribute none).
__device__ __forceinline__ uint32_t wave_intervals(const GiArgs& a, const Wave& w) {
    if (w.kind == kWaveRange) return (uint32_t)(w.g1 - w.g0);
    if (w.kind == kWaveSampled || w.kind == kWaveReceiver) return w.n;
    uint32_t total = *w.list_count;
    uint32_t probes = total > w.first ? total - w.first : 0u;
    if (probes > w.n) probes = w.n;
    return probes * ((a.p.base_phi << w.level) * (a.p.base_y << w.level));
}

// Global interval index of the chunk's j-th interval and whether it is traced
// as the unbounded terminal interval of its level.
__device__ __forceinline__ uint64_t wave_interval(const GiArgs& a, const Wave& w, uint32_t j, bool& force) {
    force = w.kind != kWaveRange;
    if (w.kind == kWaveRange) return w.g0 + j;
    uint64_t count = a.offset[w.level + 1] - a.offset[w.level];
    if (w.kind == kWaveSampled) {
        uint32_t h = transport_hash((w.first + j) + transport_hash(w.seed ^ 0x3c6ef372u));
        return a.offset[w.level] + (((uint64_t)h * count) >> 32);
    }
    uint32_t D = (a.p.base_phi << w.level) * (a.p.base_y << w.level);
    uint32_t probe = w.list[w.first + j / D];
    return a.offset[w.level] + (uint64_t)probe * D + (j % D);
}

__device__ __forceinline__ void wave_item_ray(const GiArgs& a, const Wave& w, uint32_t i, float3& o, float3& d,
                                              float& lim, bool& terminal) {
    if (w.kind == kWaveReceiver) {
        const uint32_t j = w.rays[w.first + i];
        float3 pos;
        // a listed ray always has a surface (its receiver was seeded)
        rx_ray(w.recv, w.count, j, w.seq[j], w.frame, a.p.normal_bias, pos, o, d);
        lim = 1e20f;
        terminal = true;
        return;
    }
    bool force;
    uint64_t g = wave_interval(a, w, i / w.R, force);
    item_ray(a, g, i % w.R, force, o, d, lim, terminal);
}

struct PrimarySource {
    const GiArgs* a;
    const Wave* w;
    uint32_t n;
    uint32_t tally;  // queries admitted under an unbounded budget
    __device__ uint32_t count() const { return n; }
    template <class St>
    __device__ bool begin(uint32_t i, lux::Walk& wk, St& st) {
        float3 o, d;
        float lim;
        bool term;
        wave_item_ray(*a, *w, i, o, d, lim, term);
        if (!admit_query_tallied(*a, tally)) {  // denied: no sample, as an unsupported traversal
            w->prim[i] = make_uint4(0u, lux::kEmpty, lux::kEmpty, lux::kErrUnsupported);
            return false;
        }
        lux::walk_begin(a->scene, wk, o, d, lim, lux::kWalkClosest, st);
        if (wk.phase != lux::kPhaseDone) return true;
        finish(i, wk);
        return false;
    }
    __device__ void finish(uint32_t i, const lux::Walk& wk) {
        w->prim[i] = make_uint4(__float_as_uint(wk.h.t), wk.h.tri, wk.h.inst, wk.h.err);
    }
};

template <int Depth, class St = lux::NoStats>
__global__ __launch_bounds__(kWg, lux::walk_min_waves(Depth, kWg, 0)) void k_gi_primary(GiArgs a, Wave w) {
    __shared__ uint2 lds[Depth][kWg];
    lux::LaneStack<Depth, kWg> stk;
    stk.bind(&lds[0][threadIdx.x]);
    St st;
    PrimarySource src{&a, &w, wave_intervals(a, w) * w.R, 0u};
    lux::regroup_walks(a.scene, src, &w.counters[0], w.regroup, w.batch, stk, st);
    query_tally_flush(a, src.tally);
    stats_flush(st, w.stats, 0u);
}

__global__ __launch_bounds__(kWg) void k_gi_setup(GiArgs a, Wave w) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    const uint32_t gw = blockIdx.x * (kWg / 64u) + (threadIdx.x >> 6);
    const bool live = i < wave_intervals(a, w) * w.R;
    uint4 pr = live ? w.prim[i] : make_uint4(0u, lux::kEmpty, lux::kEmpty, lux::kErrUnsupported);
    float4 sh = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t running = 0;  // rays of this wave so far (the same in every lane of a ballot)
    if ((pr.w & lux::kErrUnsupported) == 0u && pr.y != lux::kEmpty) {
        float3 o, d;
        float lim;
        bool term;
        wave_item_ray(a, w, i, o, d, lim, term);
        lux::HitState h;
        h.t = __uint_as_float(pr.x);
        h.tri = pr.y;
        h.inst = pr.z;
        h.err = pr.w;
        h.b1 = h.b2 = 0.0f;
        float3 rad;
        NeeCtx c;
        bool nee = shade_begin(a, o, d, h, rad, c);
        sh = make_float4(rad.x, rad.y, rad.z, nee ? 1.0f : 0.0f);
        if (nee) {
            const uint32_t lane = __lane_id();
            Dep nd = dep_empty();
            for (uint32_t smp = 0; smp < c.count; ++smp) {
                float3 origin = make_float3(0.0f, 0.0f, 0.0f), dir = origin, t3 = origin;
                float sl = 0.0f;
                bool q = nee_sample(a, c, smp, origin, dir, sl, t3);
                if (q && w.dep_item) dep_ray(nd, a.scene, origin, dir, sl, false, 0.0f);
                uint32_t slot = smp * w.stride + i;
                w.slots[slot] = make_float4(t3.x, t3.y, t3.z, q ? 1.0f : 0.0f);
                uint64_t mask = __ballot(q);
                if (q) {
                    const uint64_t at = (uint64_t)gw * 64u * w.nee + running +
                                        (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
                    w.slab[2ull * at] = make_float4(origin.x, origin.y, origin.z, sl);
                    w.slab[2ull * at + 1] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(slot));
                }
                running += (uint32_t)__popcll(mask);
            }
            if (w.dep_item) dep_store(w.dep_item + 2ull * i, nd);
        }
    }
    if (live) w.shade[i] = sh;
    const uint32_t total = wave_max_u32(running);
    if (__lane_id() == 0u) w.wave_cnt[gw] = total;
}

struct ShadowSource {
    const GiArgs* a;
    const Wave* w;
    uint32_t n;
    uint32_t tally;  // queries admitted under an unbounded budget
    __device__ uint32_t count() const { return n; }
    template <class St>
    __device__ bool begin(uint32_t i, lux::Walk& wk, St& st) {
        float4 r0 = w->shadow[2ull * i], r1 = w->shadow[2ull * i + 1];
        if (!admit_query_tallied(*a, tally)) {
            w->vis[__float_as_uint(r1.w)] = 2u;
            return false;
        }
        lux::walk_begin(a->scene, wk, make_float3(r0.x, r0.y, r0.z), make_float3(r1.x, r1.y, r1.z), r0.w,
                        lux::kWalkOccluded, st);
        if (wk.phase != lux::kPhaseDone) return true;
        finish(i, wk);
        return false;
    }
    __device__ void finish(uint32_t i, const lux::Walk& wk) {
        uint32_t slot = __float_as_uint(w->shadow[2ull * i + 1].w);
        w->vis[slot] = (wk.h.err & lux::kHitOccluded) ? 1u : ((wk.h.err & lux::kErrUnsupported) ? 2u : 0u);
    }
};

template <int Depth, class St = lux::NoStats>
__global__ __launch_bounds__(kWg, lux::walk_min_waves(Depth, kWg, 0)) void k_gi_shadow(GiArgs a, Wave w) {
    __shared__ uint2 lds[Depth][kWg];
    lux::LaneStack<Depth, kWg> stk;
    stk.bind(&lds[0][threadIdx.x]);
    St st;
    ShadowSource src{&a, &w, w.counters[1], 0u};
    lux::regroup_walks(a.scene, src, &w.counters[2], w.regroup, w.batch, stk, st);
    query_tally_flush(a, src.tally);
    stats_flush(st, w.stats, 3u);
}

__global__ __launch_bounds__(kWg) void k_gi_accum(GiArgs a, Wave w) {
    uint32_t j = blockIdx.x * kWg + threadIdx.x;
    if (j >= wave_intervals(a, w)) return;
    bool force;
    uint64_t g = wave_interval(a, w, j, force);
    float3 o, d;
    float lim;
    bool terminal;
    item_ray(a, g, 0u, force, o, d, lim, terminal);
    uint32_t count = a.p.emissive_samples / 2u;
    if (count < 1u) count = 1u;
    float3 jsum = make_float3(0.0f, 0.0f, 0.0f);
    float bsum = 0.0f;
    uint32_t n = 0;
    const bool track = a.dep != nullptr && w.kind == kWaveRange;
    Dep dep = dep_empty();
    for (uint32_t r = 0; r < w.R; ++r) {
        uint32_t i = j * w.R + r;
        uint4 pr = w.prim[i];
        if (pr.w & lux::kErrUnsupported) continue;
        bool nee_ran = false;
        float3 rad = make_float3(0.0f, 0.0f, 0.0f);
        float beta;
        if (pr.y != lux::kEmpty) {
            float4 sh = w.shade[i];
            float3 rr = make_float3(sh.x, sh.y, sh.z);
            if (sh.w != 0.0f) {
                nee_ran = true;
                float3 sum = make_float3(0.0f, 0.0f, 0.0f);
                bool ok = true;
                for (uint32_t smp = 0; smp < count; ++smp) {
                    uint32_t slot = smp * w.stride + i;
                    float4 s4 = w.slots[slot];
                    if (s4.w == 0.0f) continue;
                    uint32_t v = w.vis[slot];
                    if (v == 1u) continue;
                    if (v == 2u) {
                        ok = false;
                        break;
                    }
                    sum.x = sum.x + s4.x;
                    sum.y = sum.y + s4.y;
                    sum.z = sum.z + s4.z;
                }
                if (!ok) continue;
                float fc = (float)count;
                rr = make_float3(rr.x + sum.x / fc, rr.y + sum.y / fc, rr.z + sum.z / fc);
            }
            rad = shade_end(rr);
            beta = 0.0f;
        } else {
            beta = terminal ? 0.0f : 1.0f;
        }
        if (track) {
            // the ray's segment, then its NEE occlusion segments
            float3 ro, rdir;
            float rlim;
            bool rterm;
            item_ray(a, g, r, force, ro, rdir, rlim, rterm);
            Dep rd = dep_empty();
            dep_ray(rd, a.scene, ro, rdir, rlim, pr.y != lux::kEmpty, __uint_as_float(pr.x));
            if (nee_ran) dep_union(rd, dep_load(w.dep_item + 2ull * i));
            dep_union(dep, rd);
        }
        jsum = make_float3(jsum.x + rad.x, jsum.y + rad.y, jsum.z + rad.z);
        bsum = bsum + beta;
        n++;
    }
    float4 out;
    if (n == 0) {
        out = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    } else {
        float fn = (float)n;
        out = make_float4(jsum.x / fn, jsum.y / fn, jsum.z / fn, bsum / fn);
    }
    if (w.kind == kWaveRange) {
        a.intervals[g] = out;
        if (a.samples) a.samples[g] = n;
        if (track) dep_store(a.dep + 2 * g, dep);
    } else if (w.kind == kWaveListed) {
        a.merged[g] = merge_leaf(out);
    } else {
        float* rec = w.sample_out + 12ull * (w.first + j);
        level_query(a.p, w.level, g - a.offset[w.level], rec);
        float4 m = merge_leaf(out);
        bool ok = m.w > 0.0f;
        float nan = __uint_as_float(0x7fc00000u);
        rec[9] = ok ? m.x : nan;
        rec[10] = ok ? m.y : nan;
        rec[11] = ok ? m.z : nan;
    }
}

// Receiver rays: the fused shade()'s sum of visible NEE contributions in
// sample order, per ray (rad.rgb, t) and status (1 hit, 2 unsupported).
__global__ __launch_bounds__(kWg) void k_rx_wave_finish(GiArgs a, Wave w) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= w.n) return;
    uint32_t count = a.p.emissive_samples / 2u;
    if (count < 1u) count = 1u;
    const uint4 pr = w.prim[i];
    float3 rad = make_float3(0.0f, 0.0f, 0.0f);
    uint32_t status = 0;
    if (pr.w & lux::kErrUnsupported) {
        status = 2u;
    } else if (pr.y != lux::kEmpty) {
        status = 1u;
        const float4 sh = w.shade[i];
        float3 rr = make_float3(sh.x, sh.y, sh.z);
        if (sh.w != 0.0f) {
            float3 sum = make_float3(0.0f, 0.0f, 0.0f);
            for (uint32_t smp = 0; smp < count; ++smp) {
                const uint32_t slot = smp * w.stride + i;
                const float4 s4 = w.slots[slot];
                if (s4.w == 0.0f) continue;
                const uint32_t v = w.vis[slot];
                if (v == 1u) continue;
                if (v == 2u) {
                    status = 2u;
                    break;
                }
                sum.x = sum.x + s4.x;
                sum.y = sum.y + s4.y;
                sum.z = sum.z + s4.z;
            }
            const float fc = (float)count;
            rr = make_float3(rr.x + sum.x / fc, rr.y + sum.y / fc, rr.z + sum.z / fc);
        }
        if (status == 1u) rad = shade_end(rr);
    }
    w.rx_out[w.first + i] = make_float4(rad.x, rad.y, rad.z, __uint_as_float(pr.x));
    w.rx_status[w.first + i] = status;
}

// Temporal blend of intervals [0, count) (levels are contiguous).
__global__ __launch_bounds__(kWg) void k_gi_history(GiArgs a, uint64_t count) {
    uint64_t g = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    bool kept = false;
    if (g < count) {
        uint32_t L = 0;
        while (L + 1 < a.p.levels && g >= a.offset[L + 1]) ++L;
        kept = history_blend(a, g, L, g - a.offset[L]);
    }
    const uint64_t m = __ballot(kept);
    if (__lane_id() == 0u && m) atomicAdd(&a.hist_state[2], (uint32_t)__popcll(m));
}

// After a history pass: the history is valid and the dirty boxes consumed.
__global__ void k_gi_history_done(uint32_t* state) {
    if (threadIdx.x == 0) {
        state[0] = 1u;
        state[1] = 0u;
        state[3] = 0u;
    }
}

__global__ __launch_bounds__(kWg) void k_gi_merge(GiArgs a, uint32_t L) {
    uint64_t local = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (local >= a.offset[L + 1] - a.offset[L]) return;
    merge_interval(a, L, local);
}

// One 64-lane wave per level-0 probe (packProbes, split.slang:1788-1918).
__global__ __launch_bounds__(256) void k_gi_pack(GiArgs a, uint32_t probes) {
    extern __shared__ float pack_lds[];
    const uint32_t D = a.p.base_phi * a.p.base_y;
    float* ctab = pack_lds;
    float* nom = ctab + 36u * D;
    float3* dirs = reinterpret_cast<float3*>(nom + 36u);
    float4* smp = reinterpret_cast<float4*>(pack_lds + pack_smp_offset(D));
    pack_table(a.p, ctab, nom, dirs);
    const uint32_t wv = threadIdx.x >> 6;
    for (uint32_t j = 0; j < kPackBlockProbes / 4u; ++j) {
        const uint32_t probe = blockIdx.x * kPackBlockProbes + wv * (kPackBlockProbes / 4u) + j;
        if (probe < probes) pack_probe_table(a, probe, ctab, nom, smp + wv * D);
    }
}


__device__ __forceinline__ void wrap_texel(int32_t& x, int32_t& y) {
    int32_t px = x, py = y;
    if (px < 0) { int32_t nx = -px - 1, ny = 5 - py; px = nx; py = ny; }
    else if (px >= 6) { int32_t nx = 11 - px, ny = 5 - py; px = nx; py = ny; }
    if (py < 0) { int32_t nx = 5 - px, ny = -py - 1; px = nx; py = ny; }
    else if (py >= 6) { int32_t nx = 5 - px, ny = 11 - py; px = nx; py = ny; }
    x = px;
    y = py;
}

// Receiver eligibility (finite normal with |n|^2 > 0.01) and its unit normal.
__device__ __forceinline__ bool resolve_normal(const float* r, float& nx, float& ny, float& nz) {
    float x = r[3], y = r[4], z = r[5];
    float nn = (x * x + y * y) + z * z;
    if (!(lux::isfinite_bits(x) && lux::isfinite_bits(y) && lux::isfinite_bits(z)) || !(nn > 0.01f)) return false;
    float nl = sqrtf(nn);
    nx = x / nl;
    ny = y / nl;
    nz = z / nl;
    return true;
}

// irradianceLookup (world_resolve.slang:25-40): the octahedral footprint of
// the unit normal -- 4 texels and their bilinear weights.
__device__ __forceinline__ void resolve_lookup(float nx, float ny, float nz, uint32_t tex[4], float bw[4]) {
    float sabs = (fabsf(nx) + fabsf(ny)) + fabsf(nz);
    float ex = nx / sabs, ey = ny / sabs, ez = nz / sabs;
    if (ez < 0.0f) {
        float tx = (1.0f - fabsf(ey)) * (ex >= 0.0f ? 1.0f : -1.0f);
        float ty = (1.0f - fabsf(ex)) * (ey >= 0.0f ? 1.0f : -1.0f);
        ex = tx;
        ey = ty;
    }
    float gx = (ex * 0.5f + 0.5f) * 6.0f - 0.5f, gy = (ey * 0.5f + 0.5f) * 6.0f - 0.5f;
    float flx = floorf(gx), fly = floorf(gy);
    float fx = gx - flx, fy = gy - fly;
    float cx = 1.0f - fx, cy = 1.0f - fy;
    bw[0] = cx * cy;
    bw[1] = fx * cy;
    bw[2] = cx * fy;
    bw[3] = fx * fy;
    int32_t bxi = (int32_t)flx, byi = (int32_t)fly;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int32_t tx = bxi + (c & 1), ty = byi + (c >> 1);
        wrap_texel(tx, ty);
        tex[c] = (uint32_t)(ty * 6 + tx);
    }
}

// One dense probe grid of the resolve (a LOD of the pyramid).
struct ResolveGrid {
    const float4* probes;  // 36 packed texels per probe
    float origin[3];
    float spacing;
    uint32_t dims[3];
};

// diffuseAtLod + diffuseProbePage (world_resolve.slang:52-170): 8 probes
// trilinear x 4 texels bilinear, support-normalised; probes outside the grid
// are missing (they still count in the potential weight).
__device__ void diffuse_at(const ResolveGrid& g, const float* pos, const uint32_t tex[4], const float bw[4],
                           float3& val, float& cov) {
    float rel[3] = {pos[0] - g.origin[0], pos[1] - g.origin[1], pos[2] - g.origin[2]};
    float fr[3];
    int32_t base[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        float gq = rel[q] / g.spacing - 0.5f;
        float fl = floorf(gq);
        base[q] = (int32_t)fl;
        fr[q] = gq - fl;
    }
    const uint32_t dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    float3 sum = make_float3(0.0f, 0.0f, 0.0f);
    float observed = 0.0f, potential = 0.0f;
    for (uint32_t corner = 0; corner < 8u; ++corner) {
        uint32_t bx = corner & 1u, by = (corner >> 1) & 1u, bz = corner >> 2;
        float ax = bx == 0u ? 1.0f - fr[0] : fr[0];
        float ay = by == 0u ? 1.0f - fr[1] : fr[1];
        float az = bz == 0u ? 1.0f - fr[2] : fr[2];
        float w = (ax * ay) * az;
        if (!(w > 0.0f)) continue;
        potential = potential + w;
        int64_t qx = (int64_t)base[0] + bx, qy = (int64_t)base[1] + by, qz = (int64_t)base[2] + bz;
        if (qx < 0 || qy < 0 || qz < 0 || qx >= dx || qy >= dy || qz >= dz) continue;
        uint64_t probe = ((uint64_t)qz * dy + (uint64_t)qy) * dx + (uint64_t)qx;
        float3 psum = make_float3(0.0f, 0.0f, 0.0f);
        float pobs = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (!(bw[c] > 0.0f)) continue;
            float4 smp = g.probes[36ull * probe + tex[c]];
            if (!(smp.w > 0.0f) || !(lux::isfinite_bits(smp.x) && lux::isfinite_bits(smp.y) &&
                                     lux::isfinite_bits(smp.z) && lux::isfinite_bits(smp.w)))
                continue;
            float sw = bw[c] * sat(smp.w);
            psum = make_float3(psum.x + smp.x * sw, psum.y + smp.y * sw, psum.z + smp.z * sw);
            pobs = pobs + sw;
        }
        float3 pv = make_float3(0.0f, 0.0f, 0.0f);
        if (pobs > 0.0f) pv = make_float3(psum.x / pobs, psum.y / pobs, psum.z / pobs);
        float support = sat(pobs);
        if (!(support > 0.0f)) continue;
        float sw = w * support;
        sum = make_float3(sum.x + pv.x * sw, sum.y + pv.y * sw, sum.z + pv.z * sw);
        observed = observed + sw;
    }
    cov = 0.0f;
    val = make_float3(0.0f, 0.0f, 0.0f);
    if (potential > 0.0f) {
        cov = sat(observed / potential);
        if (observed > 0.0f) val = make_float3(sum.x / observed, sum.y / observed, sum.z / observed);
    }
}

// resolveLodSource (world_resolve.slang:172-182) and the raw alpha code
// (src/lib.rs:320-335): 0 not eligible, -1 unsupported, 1 + coverage / 4.
__device__ __forceinline__ float4 resolve_combine(float3 lv, float lc, float3 hv, float hc, float blend) {
    float lw = lc * (1.0f - blend), hw = hc * blend, tw = lw + hw;
    float coverage = sat(tw);
    float3 res = make_float3(0.0f, 0.0f, 0.0f);
    if (tw > 0.0f)
        res = make_float3((lv.x * lw + hv.x * hw) / tw, (lv.y * lw + hv.y * hw) / tw, (lv.z * lw + hv.z * hw) / tw);
    bool supported = coverage > 0.0f && lux::isfinite_bits(res.x) && lux::isfinite_bits(res.y) &&
                     lux::isfinite_bits(res.z);
    return make_float4(res.x, res.y, res.z, supported ? 1.0f + 0.25f * coverage : -1.0f);
}

__device__ __forceinline__ ResolveGrid grid_of(const GiArgs& a) {
    ResolveGrid g;
    g.probes = a.probes;
    g.origin[0] = a.p.origin[0];
    g.origin[1] = a.p.origin[1];
    g.origin[2] = a.p.origin[2];
    g.spacing = a.p.spacing;
    g.dims[0] = a.p.dims[0];
    g.dims[1] = a.p.dims[1];
    g.dims[2] = a.p.dims[2];
    return g;
}

// worldResolve at LOD 0 (world_resolve.slang:1-196,281-313): the LOD pair
// with blend 0 (no high LOD, its coverage 1).
__global__ __launch_bounds__(kWg) void k_world_resolve(GiArgs a, const float* recv, uint32_t n,
                                                       float4* out) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const float* r = recv + 6ull * i;
    float nx, ny, nz;
    if (!resolve_normal(r, nx, ny, nz)) {
        out[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    uint32_t tex[4];
    float bw[4];
    resolve_lookup(nx, ny, nz, tex, bw);
    ResolveGrid g = grid_of(a);
    float3 val;
    float cov;
    diffuse_at(g, r, tex, bw, val, cov);
    out[i] = resolve_combine(val, cov, make_float3(0.0f, 0.0f, 0.0f), 1.0f, 0.0f);
}

constexpr uint32_t kMaxResolveLods = 8;
struct ResolveLods {
    ResolveGrid g[kMaxResolveLods];
    uint32_t count;
};

// resolveSplitDiffuse (world_resolve.slang:184-219) over a pyramid of dense
// grids, grid k at LOD k; LODs at or past `count` have no probes.
__global__ __launch_bounds__(kWg) void k_world_resolve_lods(ResolveLods lods, lux_lod_view view, const float* recv,
                                                            uint32_t n, float4* out) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const float* r = recv + 6ull * i;
    float nx, ny, nz;
    if (!resolve_normal(r, nx, ny, nz)) {
        out[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    uint32_t tex[4];
    float bw[4];
    resolve_lookup(nx, ny, nz, tex, bw);
    uint32_t low, high;
    float blend;
    lod_select(view, r, low, high, blend);
    float3 lv = make_float3(0.0f, 0.0f, 0.0f), hv = lv;
    float lc = 0.0f, hc = 1.0f;
    if (low < lods.count) diffuse_at(lods.g[low], r, tex, bw, lv, lc);
    if (high != low) {
        hc = 0.0f;
        if (high < lods.count) diffuse_at(lods.g[high], r, tex, bw, hv, hc);
    }
    out[i] = resolve_combine(lv, lc, hv, hc, blend);
}

struct Camera {
    float o[3], f[3], r[3], u[3];
};

// Receivers of a pinhole camera (the depth/normal prepass the reference
// resolves against): one closest-hit ray per pixel, row-major with y down;
// receiver = hit position and the triangle's geometric normal turned toward
// the camera, or a zero normal (not eligible) on a miss or an unsupported
// traversal.
__global__ __launch_bounds__(kWg) void k_camera_receivers(lux::DevScene scene, Camera cam, uint32_t w,
                                                          uint32_t h, float* out) {
    __shared__ uint2 lds[kLdsDepth][kWg];
    Stack stk;
    stk.bind(&lds[0][threadIdx.x]);
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= w * h) return;
    uint32_t x = i % w, y = i / w;
    float u = ((float)x + 0.5f) / (float)w * 2.0f - 1.0f;
    float v = 1.0f - ((float)y + 0.5f) / (float)h * 2.0f;
    float3 d = make_float3((cam.f[0] + u * cam.r[0]) + v * cam.u[0], (cam.f[1] + u * cam.r[1]) + v * cam.u[1],
                           (cam.f[2] + u * cam.r[2]) + v * cam.u[2]);
    float len = sqrtf(dot3(d, d));
    d = make_float3(d.x / len, d.y / len, d.z / len);
    float3 o = make_float3(cam.o[0], cam.o[1], cam.o[2]);
    lux::HitState hs;
    lux::NoStats ns;
    // IEEE slab tests: with the reciprocal path the kernel needs 73 VGPRs
    // (six waves per SIMD) and the coherent camera rays ran 3 % slower
    lux::trace_closest(scene, o, d, 1e20f, hs, stk, ns, false, false);
    float* r = out + 6ull * i;
    if (hs.tri == lux::kEmpty || (hs.err & lux::kErrUnsupported)) {
        r[0] = o.x; r[1] = o.y; r[2] = o.z;
        r[3] = 0.0f; r[4] = 0.0f; r[5] = 0.0f;
        return;
    }
    const float4* rec = scene.inst + 8u * hs.inst;
    uint32_t iflags = __float_as_uint(rec[6].x);
    const float4* tp = scene.tri + 3ull * hs.tri;
    float4 A = tp[0], B = tp[1], C = tp[2];
    float3 n = geometric_normal(rec, make_float3(A.x, A.y, A.z), make_float3(B.x, B.y, B.z),
                                make_float3(C.x, C.y, C.z), iflags);
    if (dot3(n, d) > 0.0f) n = make_float3(-n.x, -n.y, -n.z);
    r[0] = o.x + d.x * hs.t; r[1] = o.y + d.y * hs.t; r[2] = o.z + d.z * hs.t;
    r[3] = n.x; r[4] = n.y; r[5] = n.z;
}

inline uint32_t level_dim_host(uint32_t d, uint32_t level) {
    for (uint32_t l = 0; l < level; ++l) d = (d + 1u) / 2u;
    return d;
}

// build_emissive_lights (src/scene.rs:2311-2372) over the uploaded scene, in
// caller instance order and mesh triangle order.
std::vector<Light> build_lights(lux_ctx* ctx) {
    std::vector<Light> out;
    std::vector<double> w;
    uint32_t nm = (uint32_t)ctx->host_materials.size();
    if (nm == 0) return out;
    // A triangle whose material's luminance is not positive has zero (or no
    // finite) power under every transform and never enters the table, so
    // each frame transforms only the emissive triangles: per mesh, those of
    // its own materials (cached until the scene changes); an instance with a
    // material of its own emits from all of its triangles or none.
    auto luminance = [&](uint32_t mt) {
        const lux_material& m = ctx->host_materials[mt < nm ? mt : nm - 1u];
        float em[3];
        for (int c = 0; c < 3; ++c) em[c] = m.emission[c] > 0.0f ? m.emission[c] : 0.0f;
        return (em[0] * 0.2126f + em[1] * 0.7152f) + em[2] * 0.0722f;
    };
    if (ctx->mesh_emissive.size() != ctx->host_pos.size()) {
        ctx->mesh_emissive.assign(ctx->host_pos.size(), {});
        for (size_t m = 0; m < ctx->host_pos.size(); ++m) {
            const std::vector<uint32_t>& mat = ctx->host_mat[m];
            const uint32_t nt = (uint32_t)(ctx->host_pos[m].size() / 9);
            for (uint32_t t = 0; t < nt; ++t)
                if (luminance(t < mat.size() ? mat[t] : 0u) > 0.0f) ctx->mesh_emissive[m].push_back(t);
        }
    }
    for (uint32_t k = 0; k < (uint32_t)ctx->instances.size(); ++k) {
        const lux_instance_desc& in = ctx->instances[k];
        if (in.mesh >= ctx->host_pos.size()) continue;
        const float* f = in.world_from_local;
        uint32_t cull = in.flags & 3u;
        const std::vector<float>& pos = ctx->host_pos[in.mesh];
        const std::vector<uint32_t>& mat = ctx->host_mat[in.mesh];
        uint32_t nt = (uint32_t)(pos.size() / 9);
        const bool inst_mat = k < ctx->inst_material.size();
        if (inst_mat && !(luminance(ctx->inst_material[k]) > 0.0f)) continue;
        const std::vector<uint32_t>& emissive = ctx->mesh_emissive[in.mesh];
        const uint32_t count = inst_mat ? nt : (uint32_t)emissive.size();
        for (uint32_t e = 0; e < count; ++e) {
            const uint32_t t = inst_mat ? e : emissive[e];
            uint32_t mt = inst_mat ? ctx->inst_material[k] : mat[t];
            uint32_t mi = mt < nm ? mt : nm - 1u;
            const lux_material& m = ctx->host_materials[mi];
            float em[3];
            for (int c = 0; c < 3; ++c) em[c] = m.emission[c] > 0.0f ? m.emission[c] : 0.0f;
            const float* p = pos.data() + 9ull * t;
            float v[3][3];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) {
                    const float* row = f + 4 * r;
                    v[c][r] = ((row[0] * p[3 * c] + row[1] * p[3 * c + 1]) + row[2] * p[3 * c + 2]) + row[3];
                }
            float e1[3], e2[3];
            for (int r = 0; r < 3; ++r) {
                e1[r] = v[1][r] - v[0][r];
                e2[r] = v[2][r] - v[0][r];
            }
            float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                           e1[0] * e2[1] - e1[1] * e2[0]};
            float len = std::sqrt((cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2]);
            float area = len * 0.5f;
            float lum = (em[0] * 0.2126f + em[1] * 0.7152f) + em[2] * 0.0722f;
            double power = (double)area * (double)lum * 1.0;
            if (!std::isfinite(power) || power <= 0.0) continue;
            Light L;
            std::memset(&L, 0, sizeof L);
            L.position = make_float4(v[0][0], v[0][1], v[0][2], cull == 0 ? -2.0f : -1.0f);
            L.color = make_float4(em[0], em[1], em[2], area);
            float inv = 1.0f / len;
            float sgn = cull == 2 ? -1.0f : 1.0f;
            L.direction = make_float4((cr[0] * inv) * sgn, (cr[1] * inv) * sgn, (cr[2] * inv) * sgn, 0.0f);
            L.edge1 = make_float4(e1[0], e1[1], e1[2], 0.0f);
            L.edge2 = make_float4(e2[0], e2[1], e2[2], 0.0f);
            L.source = make_uint4(k, t, 0u, 0u);
            out.push_back(L);
            w.push_back(power);
        }
    }
    uint32_t n = (uint32_t)out.size();
    if (n) {
        const uint32_t steps = 1u << 24;
        double total = 0.0;
        for (double x : w) total += x;
        double cum = 0.0;
        uint32_t prev = 0;
        for (uint32_t i = 0; i < n; ++i) {
            cum += w[i];
            uint32_t boundary;
            if (i + 1 == n) {
                boundary = steps;
            } else {
                double ideal_d = std::round(cum / total * (double)steps);
                uint32_t ideal = ideal_d < 0.0 ? 0u : (ideal_d > 4294967295.0 ? 0xffffffffu : (uint32_t)ideal_d);
                uint32_t lo = prev + 1, hi = steps - (n - i - 1);
                boundary = ideal < lo ? lo : (ideal > hi ? hi : ideal);
            }
            out[i].direction.w = (float)boundary / (float)steps;
            out[i].edge1.w = (float)(boundary - prev) / (float)steps;
            prev = boundary;
        }
    }
    return out;
}


// Emitter table of the current instance transforms (build_emissive_lights
// runs again after the instances moved, as the reference re-packs them).
int upload_lights(lux_ctx* ctx) {
    std::vector<Light> lights = build_lights(ctx);
    ctx->gi_lights = (uint32_t)lights.size();
    if (int rc = lux_ensure(ctx, ctx->gi_light_buf, sizeof(Light) * (lights.size() + 1))) return rc;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!lights.empty())
        LUX_HIP(ctx, hipMemcpy(ctx->gi_light_buf.ptr, lights.data(), sizeof(Light) * lights.size(),
                               hipMemcpyHostToDevice));
    ctx->gi_lights_dirty = false;
    return LUX_OK;
}

int gi_check(lux_ctx* ctx) {
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    if (!ctx->tlas_built) return lux_fail(ctx, LUX_ERR_STATE, "GI: BVH not built");
    if (ctx->gi_lights_dirty)
        if (int rc = upload_lights(ctx)) return rc;
    return LUX_OK;
}

}  // namespace

int lux_gi_check_internal(lux_ctx* ctx) { return gi_check(ctx); }

namespace {
int history_alloc(lux_ctx* ctx);
}

extern "C" int lux_gi_configure(lux_ctx* ctx, const lux_gi_params* params) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!params) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_configure: null params");
    const lux_gi_params& p = *params;
    if (p.levels < 1 || p.levels > 8)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_configure: levels must be 1..8");
    if (!p.dims[0] || !p.dims[1] || !p.dims[2] || !p.base_phi || !p.base_y)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_configure: empty grid or direction set");
    if ((uint64_t)p.base_phi * p.base_y > kMaxPackDirections)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_configure: more than 256 level-0 directions");
    if (!(p.spacing > 0.0f) || !std::isfinite(p.spacing) || !(p.interval_length > 0.0f))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_configure: spacing/interval must be positive");
    if (p.interval_mode > 1u) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_configure: interval mode");
    if ((uint64_t)(p.base_phi << (p.levels - 1)) * (p.base_y << (p.levels - 1)) > (1ull << 31))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_configure: too many directions");
    uint64_t off = 0;
    for (uint32_t L = 0; L <= p.levels; ++L) {
        ctx->gi_offset[L] = off;
        if (L == p.levels) break;
        uint64_t P = (uint64_t)level_dim_host(p.dims[0], L) * level_dim_host(p.dims[1], L) *
                     level_dim_host(p.dims[2], L);
        uint64_t D = (uint64_t)(p.base_phi << L) * (p.base_y << L);
        off += P * D;
    }
    for (uint32_t L = p.levels + 1; L < 17; ++L) ctx->gi_offset[L] = off;
    if (off >= (1ull << 32)) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_configure: too many intervals");
    ctx->gi_total = off;
    ctx->gi_probes0 = (uint32_t)((uint64_t)p.dims[0] * p.dims[1] * p.dims[2]);
    if (int rc = lux_ensure(ctx, ctx->gi_intervals, 16ull * off)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_merged, 16ull * off)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_probes, 16ull * 36 * ctx->gi_probes0)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_counters, 16)) return rc;
    if (int rc = lux_device_zero(ctx, ctx->gi_counters.ptr, 8)) return rc;
    if (int rc = upload_lights(ctx)) return rc;
    ctx->gi = p;
    ctx->gi_ready = true;
    ctx->gi_sur_probes = 0;
    // a new grid starts without history
    return history_alloc(ctx);
}

extern "C" int lux_gi_set_frame(lux_ctx* ctx, uint32_t frame, uint32_t seed) {
    if (!ctx) return LUX_ERR_INVALID;
    ctx->gi.frame = frame;
    ctx->gi.seed = seed;
    return LUX_OK;
}

namespace {
// Runs the four wavefront stages over a set of intervals in chunks whose
// scratch stays under about 6 GiB.  `proto` carries the kind and its inputs;
// `total` is the number of intervals (kWaveRange), listed probes
// (kWaveListed, an upper bound: the device-side list length decides) or
// samples (kWaveSampled); `per_unit` intervals per unit (D for listed
// probes, else 1).
int run_wave(lux_ctx* ctx, Wave proto, uint64_t total, uint64_t per_unit) {
    const lux_gi_params& p = ctx->gi;
    const bool rx = proto.kind == kWaveReceiver;
    const uint32_t R = rx ? 1u : (p.rays_per_interval ? p.rays_per_interval : 1u);
    const uint32_t nee = p.emissive_samples / 2u > 1u ? p.emissive_samples / 2u : 1u;
    const uint64_t per_item = 32ull + (uint64_t)nee * (16ull + 4ull + 64ull);
    uint64_t chunk_items = (6ull << 30) / per_item;
    if (chunk_items < 4096) chunk_items = 4096;
    uint64_t chunk_units = chunk_items / (R * per_unit);
    if (chunk_units < 1) chunk_units = 1;
    if (chunk_units > total) chunk_units = total;
    uint64_t items = chunk_units * per_unit * R;
    if (items * nee >= (1ull << 31)) return lux_fail(ctx, LUX_ERR_INVALID, "wavefront trace: too many NEE slots");
    if (int rc = lux_ensure(ctx, ctx->gi_w_prim, 16ull * items)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_w_shade, 16ull * items)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_w_slots, 16ull * items * nee)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_w_vis, 4ull * items * nee)) return rc;
    const bool track = ctx->gi_hist_on && proto.kind == kWaveRange;
    if (track)
        if (int rc = lux_ensure(ctx, ctx->gi_w_dep, 32ull * items)) return rc;
    // occlusion rays: per setup wave a slab of 64 * nee, then the packed list
    const uint64_t setup_waves = ((items + kWg - 1) / kWg) * (kWg / 64u);
    if (int rc = lux_ensure(ctx, ctx->gi_w_shadow, 32ull * setup_waves * 64u * nee * 2u)) return rc;
    const size_t counter_words = 4u + (setup_waves + 1u) + lux_scan_sums_words((uint32_t)setup_waves + 1u);
    if (int rc = lux_ensure(ctx, ctx->gi_w_counters, 4ull * counter_words)) return rc;
    GiArgs a = gi_args(ctx);
    if (rx) a.admit_limit = 0;  // a receiver update is bounded by its fair window
    Wave w = proto;
    w.R = R;
    w.nee = nee;
    w.stride = (uint32_t)items;
    w.prim = static_cast<uint4*>(ctx->gi_w_prim.ptr);
    w.shade = static_cast<float4*>(ctx->gi_w_shade.ptr);
    w.slots = static_cast<float4*>(ctx->gi_w_slots.ptr);
    w.vis = static_cast<uint32_t*>(ctx->gi_w_vis.ptr);
    w.dep_item = track ? static_cast<float4*>(ctx->gi_w_dep.ptr) : nullptr;
    w.shadow = static_cast<float4*>(ctx->gi_w_shadow.ptr);
    w.counters = static_cast<uint32_t*>(ctx->gi_w_counters.ptr);
    w.wave_cnt = w.counters + 4;
    uint32_t* scan_sums = w.wave_cnt + setup_waves + 1u;
    w.slab = w.shadow + 2ull * setup_waves * 64u * nee;
    w.regroup = ctx->gi_regroup ? ctx->gi_regroup : 8u;
    w.batch = ctx->gi_batch ? ctx->gi_batch : 8u;
    const uint32_t wpc = ctx->gi_wpc ? ctx->gi_wpc : 24u;
    const uint32_t pblocks = (uint32_t)ctx->cu_count * wpc / (kWg / 64);
    hipStream_t st = ctx->stream;
    for (uint64_t u0 = 0; u0 < total; u0 += chunk_units) {
        uint64_t units = total - u0 < chunk_units ? total - u0 : chunk_units;
        if (w.kind == kWaveRange) {
            w.g0 = proto.g0 + u0;
            w.g1 = w.g0 + units;
        } else {
            w.first = (uint32_t)u0;
            w.n = (uint32_t)units;
        }
        uint32_t max_intervals = (uint32_t)(units * per_unit);
        uint32_t max_items = max_intervals * R;
        if (int rc = lux_device_zero(ctx, w.counters, 16)) return rc;
        uint32_t need = (max_items + kWg - 1) / kWg;
        uint32_t pb = pblocks < need ? pblocks : need;
        const bool timing = ctx->gi_stage_timing && !w.stats;
        if (timing) LUX_HIP(ctx, hipEventRecord(ctx->ev_stage[0], st));
        if (w.stats) {
            hipLaunchKernelGGL((k_gi_primary<8, lux::Counting>), dim3(pb), dim3(kWg), 0, st, a, w);
        } else {
            switch (ctx->gi_depth) {
            case 4: hipLaunchKernelGGL(k_gi_primary<4>, dim3(pb), dim3(kWg), 0, st, a, w); break;
            case 12: hipLaunchKernelGGL(k_gi_primary<12>, dim3(pb), dim3(kWg), 0, st, a, w); break;
            case 16: hipLaunchKernelGGL(k_gi_primary<16>, dim3(pb), dim3(kWg), 0, st, a, w); break;
            default: hipLaunchKernelGGL(k_gi_primary<8>, dim3(pb), dim3(kWg), 0, st, a, w); break;
            }
        }
        if (timing) LUX_HIP(ctx, hipEventRecord(ctx->ev_stage[1], st));
        hipLaunchKernelGGL(k_gi_setup, dim3(need), dim3(kWg), 0, st, a, w);
        const uint32_t nw = need * (kWg / 64u);
        hipLaunchKernelGGL(k_fill_u32_one, dim3(1), dim3(1), 0, st, w.wave_cnt + nw, 0u);
        if (int rc = lux_scan_exclusive_on(ctx, st, w.wave_cnt, nw + 1u, scan_sums, &w.counters[1])) return rc;
        hipLaunchKernelGGL(k_shadow_compact, dim3((nw + 3u) / 4u), dim3(256), 0, st, w.slab, w.wave_cnt, nw,
                           64u * w.nee, w.shadow);
        if (timing) LUX_HIP(ctx, hipEventRecord(ctx->ev_stage[2], st));
        if (w.stats) {
            hipLaunchKernelGGL((k_gi_shadow<8, lux::Counting>), dim3(pblocks), dim3(kWg), 0, st, a, w);
        } else {
            switch (ctx->gi_depth) {
            case 4: hipLaunchKernelGGL(k_gi_shadow<4>, dim3(pblocks), dim3(kWg), 0, st, a, w); break;
            case 12: hipLaunchKernelGGL(k_gi_shadow<12>, dim3(pblocks), dim3(kWg), 0, st, a, w); break;
            case 16: hipLaunchKernelGGL(k_gi_shadow<16>, dim3(pblocks), dim3(kWg), 0, st, a, w); break;
            default: hipLaunchKernelGGL(k_gi_shadow<8>, dim3(pblocks), dim3(kWg), 0, st, a, w); break;
            }
        }
        if (timing) LUX_HIP(ctx, hipEventRecord(ctx->ev_stage[3], st));
        if (rx) hipLaunchKernelGGL(k_rx_wave_finish, dim3((max_intervals + kWg - 1) / kWg), dim3(kWg), 0, st, a, w);
        else hipLaunchKernelGGL(k_gi_accum, dim3((max_intervals + kWg - 1) / kWg), dim3(kWg), 0, st, a, w);
        LUX_HIP(ctx, hipGetLastError());
        if (timing) {
            LUX_HIP(ctx, hipEventRecord(ctx->ev_stage[4], st));
            LUX_HIP(ctx, hipEventSynchronize(ctx->ev_stage[4]));
            for (int k = 0; k < 4; ++k) {
                float ms = 0.0f;
                LUX_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_stage[k], ctx->ev_stage[k + 1]));
                ctx->gi_stage_ms[k] += ms;
            }
        }
        if (w.stats) {
            // occlusion rays of the chunk (the walks' count), accumulated on the host
            uint32_t rays = 0;
            LUX_HIP(ctx, hipMemcpyAsync(&rays, &w.counters[1], 4, hipMemcpyDeviceToHost, st));
            LUX_HIP(ctx, hipStreamSynchronize(st));
            ctx->gi_stats_rays += rays;
        }
    }
    return LUX_OK;
}

Wave wave_proto(uint32_t kind) {
    Wave w;
    std::memset(&w, 0, sizeof w);
    w.kind = kind;
    return w;
}

}  // namespace

// Receiver rays through the wavefront stages (regrouped primary and
// occlusion walks): per ray of the list (rad.rgb, t) and status.
int lux_rx_wave_trace(lux_ctx* ctx, const float* recv, uint32_t count, const uint32_t* rays, uint32_t n_rays,
                      const uint32_t* seq, uint32_t frame, float4* out, uint32_t* status) {
    if (n_rays == 0) return LUX_OK;
    Wave w = wave_proto(kWaveReceiver);
    w.recv = recv;
    w.rays = rays;
    w.seq = seq;
    w.count = count;
    w.frame = frame;
    w.rx_out = out;
    w.rx_status = status;
    return run_wave(ctx, w, n_rays, 1);
}

namespace {
int trace_wavefront(lux_ctx* ctx, uint64_t g0, uint64_t g1) {
    Wave w = wave_proto(kWaveRange);
    w.g0 = g0;
    w.g1 = g1;
    return run_wave(ctx, w, g1 - g0, 1);
}
}  // namespace

extern "C" int lux_gi_set_trace_config(lux_ctx* ctx, const lux_trace_config* cfg) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!cfg) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_set_trace_config: null config");
    if (cfg->waves_per_cu > 32)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_set_trace_config: waves_per_cu > 32");
    if (cfg->regroup_threshold > 64)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_set_trace_config: regroup threshold > 64");
    if (cfg->lds_stack_depth != 0 && cfg->lds_stack_depth != 4 && cfg->lds_stack_depth != 8 &&
        cfg->lds_stack_depth != 12 && cfg->lds_stack_depth != 16)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_set_trace_config: LDS stack depth must be 4, 8, 12 or 16");
    if (cfg->steps_per_refill > 256)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_set_trace_config: steps per refill > 256");
    ctx->gi_wpc = cfg->waves_per_cu;
    ctx->gi_depth = cfg->lds_stack_depth;
    ctx->gi_regroup = cfg->regroup_threshold;
    ctx->gi_batch = cfg->steps_per_refill;
    return LUX_OK;
}

extern "C" int lux_gi_get_trace_config(lux_ctx* ctx, lux_trace_config* cfg) {
    if (!ctx || !cfg) return LUX_ERR_INVALID;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->waves_per_cu = ctx->gi_wpc;
    cfg->lds_stack_depth = ctx->gi_depth;
    cfg->regroup_threshold = ctx->gi_regroup;
    cfg->steps_per_refill = ctx->gi_batch;
    return LUX_OK;
}

// Autotuner of the wavefront stages: coordinate descent as for ray lists,
// each candidate timed over a full interval trace of the configured grid
// (HIP events, median of `repeats` after one warm-up).  Overwrites the
// interval array.
extern "C" int lux_autotune_gi(lux_ctx* ctx, uint32_t repeats, lux_trace_config* best, float* best_ms,
                               uint32_t* evaluations) {
    if (!ctx) return LUX_ERR_INVALID;
    if (int rc = gi_check(ctx)) return rc;
    if (repeats == 0) repeats = 2;
    if (repeats > 15) repeats = 15;
    const uint32_t saved[4] = {ctx->gi_wpc, ctx->gi_depth, ctx->gi_regroup, ctx->gi_batch};
    const uint32_t saved_mode = ctx->gi_trace_mode;
    ctx->gi_trace_mode = 0u;
    auto measure = [&](const uint32_t c[4], float& ms) -> int {
        ctx->gi_wpc = c[0];
        ctx->gi_depth = c[1];
        ctx->gi_regroup = c[2];
        ctx->gi_batch = c[3];
        float t[16];
        for (uint32_t r = 0; r <= repeats; ++r) {
            LUX_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
            if (int rc = lux_gi_trace_levels(ctx, 0, ctx->gi.levels - 1)) return rc;
            LUX_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
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
    static const uint32_t cand_regroup[] = {4, 8, 12, 16, 32, 48};
    static const uint32_t cand_batch[] = {4, 8, 16, 24, 32};
    const uint32_t* cands[4] = {cand_wpc, cand_depth, cand_regroup, cand_batch};
    const uint32_t ncand[4] = {5, 4, 6, 5};
    uint32_t cur[4] = {24, 12, 12, 24};
    float cur_ms = 0.0f;
    uint32_t evals = 1;
    int rc = measure(cur, cur_ms);
    for (int pass = 0; pass < 2 && rc == LUX_OK; ++pass)
        for (int k = 0; k < 4 && rc == LUX_OK; ++k)
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
    ctx->gi_trace_mode = saved_mode;
    const uint32_t* fin = rc == LUX_OK ? cur : saved;
    ctx->gi_wpc = fin[0];
    ctx->gi_depth = fin[1];
    ctx->gi_regroup = fin[2];
    ctx->gi_batch = fin[3];
    if (rc != LUX_OK) return rc;
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

extern "C" int lux_gi_set_trace_mode(lux_ctx* ctx, uint32_t mode) {
    if (!ctx) return LUX_ERR_INVALID;
    if (mode > 1u) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_set_trace_mode: mode is 0 or 1");
    ctx->gi_trace_mode = mode;
    return LUX_OK;
}

extern "C" int lux_gi_trace_levels(lux_ctx* ctx, uint32_t first, uint32_t last) {
    if (!ctx) return LUX_ERR_INVALID;
    if (int rc = gi_check(ctx)) return rc;
    if (first > last || last >= ctx->gi.levels)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_trace_levels: level range outside the grid");
    uint64_t g0 = ctx->gi_offset[first], g1 = ctx->gi_offset[last + 1];
    for (float& t : ctx->gi_stage_ms) t = 0.0f;
    if (g1 == g0) return LUX_OK;
    if (ctx->gi_trace_mode == 0u) return trace_wavefront(ctx, g0, g1);
    GiArgs a = gi_args(ctx);
    uint32_t blocks = (uint32_t)((g1 - g0 + kWg - 1) / kWg);
    hipLaunchKernelGGL(k_gi_trace, dim3(blocks), dim3(kWg), 0, ctx->stream, a, g0, g1);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_gi_set_stage_timing(lux_ctx* ctx, int on) {
    if (!ctx) return LUX_ERR_INVALID;
    ctx->gi_stage_timing = on != 0;
    return LUX_OK;
}

extern "C" int lux_gi_stage_times(lux_ctx* ctx, float* ms) {
    if (!ctx || !ms) return LUX_ERR_INVALID;
    for (int k = 0; k < 4; ++k) ms[k] = ctx->gi_stage_ms[k];
    return LUX_OK;
}

extern "C" int lux_gi_trace_stats(lux_ctx* ctx, uint64_t* counters) {
    if (!ctx || !counters) return LUX_ERR_INVALID;
    if (int rc = gi_check(ctx)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_stats_buf, 64)) return rc;
    LUX_HIP(ctx, hipMemsetAsync(ctx->gi_stats_buf.ptr, 0, 48, ctx->stream));
    if (int rc = lux_device_zero(ctx, ctx->gi_counters.ptr, 8)) return rc;
    ctx->gi_stats_rays = 0;
    Wave w = wave_proto(kWaveRange);
    w.g0 = 0;
    w.g1 = ctx->gi_total;
    w.stats = static_cast<unsigned long long*>(ctx->gi_stats_buf.ptr);
    if (int rc = run_wave(ctx, w, ctx->gi_total, 1)) return rc;
    unsigned long long v[6] = {};
    LUX_HIP(ctx, hipMemcpyAsync(v, ctx->gi_stats_buf.ptr, 48, hipMemcpyDeviceToHost, ctx->stream));
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint64_t R = ctx->gi.rays_per_interval ? ctx->gi.rays_per_interval : 1u;
    counters[0] = ctx->gi_total * R;
    counters[1] = v[0];
    counters[2] = v[1];
    counters[3] = v[2];
    counters[4] = ctx->gi_stats_rays;
    counters[5] = v[3];
    counters[6] = v[4];
    counters[7] = v[5];
    return LUX_OK;
}

extern "C" int lux_gi_trace(lux_ctx* ctx) {
    if (!ctx) return LUX_ERR_INVALID;
    if (int rc = gi_check(ctx)) return rc;
    return lux_gi_trace_levels(ctx, 0, ctx->gi.levels - 1);
}

namespace {
int merge_levels(lux_ctx* ctx, int top, hipStream_t st = nullptr) {
    if (!st) st = ctx->stream;
    GiArgs a = gi_args(ctx);
    for (int L = top; L >= 0; --L) {
        uint64_t count = ctx->gi_offset[L + 1] - ctx->gi_offset[L];
        uint32_t blocks = (uint32_t)((count + kWg - 1) / kWg);
        hipLaunchKernelGGL(k_gi_merge, dim3(blocks), dim3(kWg), 0, st, a, (uint32_t)L);
    }
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}
}  // namespace

extern "C" int lux_gi_merge(lux_ctx* ctx) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    return merge_levels(ctx, (int)ctx->gi.levels - 1);
}

extern "C" int lux_gi_pack(lux_ctx* ctx) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    GiArgs a = gi_args(ctx);
    hipLaunchKernelGGL(k_gi_pack, dim3((ctx->gi_probes0 + kPackBlockProbes - 1) / kPackBlockProbes), dim3(256),
                       pack_lds_bytes(ctx->gi.base_phi * ctx->gi.base_y), ctx->stream, a, ctx->gi_probes0);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_gi_reset_query_counters(lux_ctx* ctx) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    return lux_device_zero(ctx, ctx->gi_counters.ptr, 8);
}

extern "C" int lux_gi_query_counters(lux_ctx* ctx, uint32_t* attempts, uint32_t* admitted) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    uint32_t c[2] = {0, 0};
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(c, ctx->gi_counters.ptr, 8, hipMemcpyDeviceToHost));
    if (attempts) *attempts = c[0];
    if (admitted) *admitted = c[1];
    return LUX_OK;
}

namespace {
// Temporal history pass over the first `count` intervals (the traced ones).
int history_pass(lux_ctx* ctx, uint64_t count, hipStream_t st = nullptr) {
    if (!ctx->gi_hist_on || count == 0) return LUX_OK;
    if (!st) st = ctx->stream;
    GiArgs a = gi_args(ctx);
    hipLaunchKernelGGL(k_fill_u32_one, dim3(1), dim3(1), 0, st, a.hist_state + 2, 0u);
    hipLaunchKernelGGL(k_gi_history, dim3((uint32_t)((count + kWg - 1) / kWg)), dim3(kWg), 0, st, a, count);
    hipLaunchKernelGGL(k_gi_history_done, dim3(1), dim3(64), 0, st, a.hist_state);
    LUX_HIP(ctx, hipGetLastError());
    ctx->gi_dirty_n = 0;
    ctx->gi_hist_valid = true;
    return LUX_OK;
}

int history_alloc(lux_ctx* ctx) {
    ctx->gi_hist_valid = false;
    if (!ctx->gi_hist_on || !ctx->gi_ready) return LUX_OK;
    if (int rc = lux_ensure(ctx, ctx->gi_samples, 4ull * ctx->gi_total)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_history, 64ull * ctx->gi_total)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_dep, 32ull * ctx->gi_total)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_hist_state, 32ull + 32ull * kMaxDirtyBoxes)) return rc;
    ctx->gi_dirty_n = 0;
    uint32_t st[8] = {};
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(ctx->gi_hist_state.ptr, st, 32, hipMemcpyHostToDevice));
    return LUX_OK;
}
}  // namespace

extern "C" int lux_gi_set_history(lux_ctx* ctx, const lux_gi_history_params* hp) {
    if (!ctx) return LUX_ERR_INVALID;
    if (hp && !(std::isfinite(hp->radiance_blend) && std::isfinite(hp->transmittance_blend)))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_set_history: blends must be finite");
    ctx->gi_hist_on = hp && (hp->radiance_blend > 0.0f || hp->transmittance_blend > 0.0f);
    if (hp) ctx->gi_hist = *hp;
    else std::memset(&ctx->gi_hist, 0, sizeof ctx->gi_hist);
    return history_alloc(ctx);
}

extern "C" int lux_gi_reset_history(lux_ctx* ctx) {
    if (!ctx) return LUX_ERR_INVALID;
    ctx->gi_hist_valid = false;
    if (!ctx->gi_hist_on || !ctx->gi_hist_state.ptr) return LUX_OK;
    ctx->gi_dirty_n = 0;
    return lux_device_zero(ctx, ctx->gi_hist_state.ptr, 8);
}

// Appends dirty boxes (6 floats each: lo, hi) for the next history pass;
// empty boxes are skipped; past the capacity the whole history is reset.
int lux_gi_invalidate_boxes(lux_ctx* ctx, const float* boxes, uint32_t count) {
    if (!ctx->gi_hist_on || !ctx->gi_hist_state.ptr) return LUX_OK;
    std::vector<float> add;
    for (uint32_t b = 0; b < count; ++b) {
        const float* lo = boxes + 6ull * b;
        const float* hi = lo + 3;
        if (!(lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2])) continue;
        add.insert(add.end(), {lo[0], lo[1], lo[2], 0.0f, hi[0], hi[1], hi[2], 0.0f});
    }
    const uint32_t n_add = (uint32_t)(add.size() / 8);
    if (n_add == 0) return LUX_OK;
    if (ctx->gi_dirty_n + n_add > kMaxDirtyBoxes) return lux_gi_reset_history(ctx);
    const uint32_t n = ctx->gi_dirty_n + n_add;
    char* st = static_cast<char*>(ctx->gi_hist_state.ptr);
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(st + 32 + 32ull * ctx->gi_dirty_n, add.data(), 32ull * n_add, hipMemcpyHostToDevice));
    LUX_HIP(ctx, hipMemcpy(st + 4, &n, 4, hipMemcpyHostToDevice));
    ctx->gi_dirty_n = n;
    return LUX_OK;
}

// The scene's root box grew: the next history pass drops measured-empty
// dependencies (state word 3, cleared after the pass).
int lux_gi_root_grew(lux_ctx* ctx) {
    if (!ctx->gi_hist_on || !ctx->gi_hist_state.ptr) return LUX_OK;
    const uint32_t one = 1u;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(static_cast<char*>(ctx->gi_hist_state.ptr) + 12, &one, 4, hipMemcpyHostToDevice));
    return LUX_OK;
}

extern "C" int lux_gi_invalidate_history(lux_ctx* ctx, const float* lo, const float* hi) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!lo || !hi) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_invalidate_history: null box");
    const float box[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
    return lux_gi_invalidate_boxes(ctx, box, 1);
}

extern "C" int lux_gi_history_reused(lux_ctx* ctx, uint32_t* reused) {
    if (!ctx || !reused) return LUX_ERR_INVALID;
    *reused = 0;
    if (!ctx->gi_hist_on || !ctx->gi_hist_state.ptr) return LUX_OK;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(reused, static_cast<char*>(ctx->gi_hist_state.ptr) + 8, 4, hipMemcpyDeviceToHost));
    return LUX_OK;
}

extern "C" int lux_gi_history_arrays(lux_ctx* ctx, float** history, uint32_t** samples) {
    if (!ctx) return LUX_ERR_INVALID;
    if (history) *history = ctx->gi_hist_on ? static_cast<float*>(ctx->gi_history.ptr) : nullptr;
    if (samples) *samples = ctx->gi_hist_on ? static_cast<uint32_t*>(ctx->gi_samples.ptr) : nullptr;
    return LUX_OK;
}

extern "C" int lux_gi_update(lux_ctx* ctx, float* ms) {
    if (!ctx) return LUX_ERR_INVALID;
    if (int rc = gi_check(ctx)) return rc;
    LUX_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (int rc = lux_device_zero(ctx, ctx->gi_counters.ptr, 8)) return rc;
    if (int rc = lux_gi_trace(ctx)) return rc;
    if (int rc = history_pass(ctx, ctx->gi_total)) return rc;
    if (int rc = lux_gi_merge(ctx)) return rc;
    if (int rc = lux_gi_pack(ctx)) return rc;
    LUX_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (ms) {
        LUX_HIP(ctx, hipEventSynchronize(ctx->ev1));
        LUX_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    }
    return LUX_OK;
}

// History blend (when on), merge and pack of a context's intervals on a
// given stream: the end of an update whose intervals another pass wrote.
namespace {
__global__ void k_zero_f4(float4* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
}  // namespace

int lux_gi_dep_unknown(lux_ctx* ctx, hipStream_t st) {
    if (!ctx->gi_hist_on || !ctx->gi_dep.ptr || ctx->gi_total == 0) return LUX_OK;
    const uint64_t n = 2ull * ctx->gi_total;  // known flag 0: unknown
    hipLaunchKernelGGL(k_zero_f4, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                       static_cast<float4*>(ctx->gi_dep.ptr), n);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

int lux_gi_finish_internal(lux_ctx* ctx, hipStream_t st) {
    // receiver deposits record no segments: their dependency is unknown
    if (int rc = lux_gi_dep_unknown(ctx, st)) return rc;
    if (int rc = history_pass(ctx, ctx->gi_total, st)) return rc;
    if (int rc = merge_levels(ctx, (int)ctx->gi.levels - 1, st)) return rc;
    GiArgs a = gi_args(ctx);
    hipLaunchKernelGGL(k_gi_pack, dim3((ctx->gi_probes0 + kPackBlockProbes - 1) / kPackBlockProbes), dim3(256),
                       pack_lds_bytes(ctx->gi.base_phi * ctx->gi.base_y), st, a, ctx->gi_probes0);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_gi_sizes(lux_ctx* ctx, uint64_t* intervals, uint32_t* probes, uint32_t* lights) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    if (intervals) *intervals = ctx->gi_total;
    if (probes) *probes = ctx->gi_probes0;
    if (lights) *lights = ctx->gi_lights;
    return LUX_OK;
}

extern "C" int lux_gi_download(lux_ctx* ctx, float* intervals, float* merged, float* probes,
                               float* lights) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (intervals)
        LUX_HIP(ctx, hipMemcpy(intervals, ctx->gi_intervals.ptr, 16ull * ctx->gi_total, hipMemcpyDeviceToHost));
    if (merged)
        LUX_HIP(ctx, hipMemcpy(merged, ctx->gi_merged.ptr, 16ull * ctx->gi_total, hipMemcpyDeviceToHost));
    if (probes)
        LUX_HIP(ctx, hipMemcpy(probes, ctx->gi_probes.ptr, 16ull * 36 * ctx->gi_probes0, hipMemcpyDeviceToHost));
    if (lights && ctx->gi_lights)
        LUX_HIP(ctx, hipMemcpy(lights, ctx->gi_light_buf.ptr, sizeof(Light) * ctx->gi_lights, hipMemcpyDeviceToHost));
    return LUX_OK;
}

extern "C" int lux_gi_upload_intervals(lux_ctx* ctx, const float* intervals) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    if (!intervals) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_upload_intervals: null");
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(ctx->gi_intervals.ptr, intervals, 16ull * ctx->gi_total, hipMemcpyHostToDevice));
    return LUX_OK;
}

extern "C" int lux_gi_device_arrays(lux_ctx* ctx, float** intervals, float** merged, float** probes) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    if (intervals) *intervals = static_cast<float*>(ctx->gi_intervals.ptr);
    if (merged) *merged = static_cast<float*>(ctx->gi_merged.ptr);
    if (probes) *probes = static_cast<float*>(ctx->gi_probes.ptr);
    return LUX_OK;
}

extern "C" int lux_world_resolve_device(lux_ctx* ctx, const float* receivers, uint32_t n, float* out) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    if (n == 0) return LUX_OK;
    if (!receivers || !out) return lux_fail(ctx, LUX_ERR_INVALID, "lux_world_resolve_device: null buffer");
    GiArgs a = gi_args(ctx);
    hipLaunchKernelGGL(k_world_resolve, dim3((n + kWg - 1) / kWg), dim3(kWg), 0, ctx->stream, a,
                       receivers, n, reinterpret_cast<float4*>(out));
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_world_resolve_lods(lux_ctx* const* lods, uint32_t lod_count, const lux_lod_view* view,
                                      const float* receivers, uint32_t n, float* out) {
    if (!lods || lod_count == 0 || !lods[0]) return LUX_ERR_INVALID;
    lux_ctx* ctx = lods[0];
    if (lod_count > kMaxResolveLods)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_world_resolve_lods: at most 8 LOD grids");
    if (!view) return lux_fail(ctx, LUX_ERR_INVALID, "lux_world_resolve_lods: null view");
    ResolveLods g{};
    g.count = lod_count;
    for (uint32_t k = 0; k < lod_count; ++k) {
        lux_ctx* c = lods[k];
        if (!c || c->device != ctx->device)
            return lux_fail(ctx, LUX_ERR_INVALID, "lux_world_resolve_lods: LOD grid of another device");
        if (!c->gi_ready)
            return lux_fail(ctx, LUX_ERR_STATE, "lux_world_resolve_lods: LOD grid not configured");
        g.g[k].probes = static_cast<const float4*>(c->gi_probes.ptr);
        for (int q = 0; q < 3; ++q) {
            g.g[k].origin[q] = c->gi.origin[q];
            g.g[k].dims[q] = c->gi.dims[q];
        }
        g.g[k].spacing = c->gi.spacing;
        // the other grids' packs run on their own streams
        if (k > 0) LUX_HIP(ctx, hipStreamSynchronize(c->stream));
    }
    if (n == 0) return LUX_OK;
    if (!receivers || !out) return lux_fail(ctx, LUX_ERR_INVALID, "lux_world_resolve_lods: null buffer");
    hipLaunchKernelGGL(k_world_resolve_lods, dim3((n + kWg - 1) / kWg), dim3(kWg), 0, ctx->stream, g, *view,
                       receivers, n, reinterpret_cast<float4*>(out));
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

// ------------------------------------------------------------------ surrogate
namespace {
uint64_t level_count(const lux_ctx* ctx, uint32_t L) { return ctx->gi_offset[L + 1] - ctx->gi_offset[L]; }
uint32_t level_directions(const lux_gi_params& p, uint32_t L) { return (p.base_phi << L) * (p.base_y << L); }
}  // namespace

extern "C" int lux_gi_surrogate_queries(lux_ctx* ctx, uint32_t level, float* queries, uint64_t* count) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    if (level >= ctx->gi.levels) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_surrogate_queries: level");
    uint64_t n = level_count(ctx, level);
    if (count) *count = n;
    if (!queries) return LUX_OK;
    GiArgs a = gi_args(ctx);
    hipLaunchKernelGGL(k_sur_queries, dim3((uint32_t)((n + kWg - 1) / kWg)), dim3(kWg), 0, ctx->stream, a,
                       level, n, queries);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_gi_surrogate_samples(lux_ctx* ctx, uint32_t level, uint32_t n, uint32_t sample_seed,
                                        uint32_t source, float* samples) {
    if (!ctx) return LUX_ERR_INVALID;
    if (source > 1u) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_surrogate_samples: source");
    if (source == 0u) {
        if (int rc = gi_check(ctx)) return rc;
    } else if (!ctx->gi_ready) {
        return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    }
    if (level >= ctx->gi.levels) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_surrogate_samples: level");
    if (n == 0) return LUX_OK;
    if (!samples) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_surrogate_samples: null buffer");
    if (source == 0u && ctx->gi_trace_mode == 0u) {
        Wave w = wave_proto(kWaveSampled);
        w.level = level;
        w.seed = sample_seed;
        w.sample_out = samples;
        return run_wave(ctx, w, n, 1);
    }
    GiArgs a = gi_args(ctx);
    hipLaunchKernelGGL(k_sur_samples, dim3((n + kWg - 1) / kWg), dim3(kWg), 0, ctx->stream, a, level, n,
                       sample_seed, source, samples);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

namespace {
__global__ void k_sur_count_zero(uint32_t* c) {
    if (threadIdx.x == 0) *c = 0u;
}
}  // namespace

// Words of the traced per-probe fallback of level `level`: the probe list,
// its count and the per-probe errors (2 * probes + 1).
uint64_t lux_sur_fallback_words(const lux_ctx* ctx, uint32_t level) {
    return 2ull * (level_count(ctx, level) / level_directions(ctx->gi, level)) + 1ull;
}

// The traced per-probe fallback on a stream (no allocation, no host wait, so
// a captured graph may hold it; `words` of lux_sur_fallback_words are the
// caller's): per-probe predicted RMSE from the network
// outputs Y of the level's queries, then the probes above the bound traced as
// the level's unbounded terminal intervals and written merged (the fused form
// of lux_gi_update_surrogate's fallback; the same bits).
int lux_sur_fallback_on(lux_ctx* ctx, hipStream_t st, uint32_t level, const float* Y, float bound,
                        uint32_t* words) {
    const uint64_t count = level_count(ctx, level);
    const uint32_t D = level_directions(ctx->gi, level);
    const uint64_t probes = count / D;
    uint32_t* list = words;
    uint32_t* fcount = list + probes;
    float* err = reinterpret_cast<float*>(fcount + 1);
    hipLaunchKernelGGL(k_sur_count_zero, dim3(1), dim3(64), 0, st, fcount);
    hipLaunchKernelGGL(k_sur_probe_error, dim3((uint32_t)probes), dim3(64), 0, st, Y, D, bound, err, list, fcount);
    GiArgs a = gi_args(ctx);
    hipLaunchKernelGGL(k_sur_fallback, dim3((uint32_t)((count + kWg - 1) / kWg)), dim3(kWg), 0, st, a, level, list,
                       static_cast<const uint32_t*>(fcount));
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_gi_update_surrogate(lux_ctx* ctx, lux_surrogate* s, uint32_t level, float error_bound,
                                       lux_gi_surrogate_stats* stats) {
    if (!ctx || !s) return LUX_ERR_INVALID;
    if (lux_surrogate_context(s) != ctx)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_update_surrogate: surrogate of another context");
    if (int rc = gi_check(ctx)) return rc;
    const lux_gi_params& p = ctx->gi;
    if (level < 1 || level >= p.levels)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_update_surrogate: level must be in 1..levels-1");
    uint64_t count = level_count(ctx, level);
    uint32_t D = level_directions(p, level);
    uint64_t probes = count / D;
    if (count >= (1ull << 31)) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_update_surrogate: level too large");
    if (int rc = lux_ensure(ctx, ctx->gi_sur_queries, 36ull * count)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_sur_err, 4ull * probes)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_sur_list, 4ull * (probes + 1))) return rc;
    hipStream_t st = ctx->stream;
    GiArgs a = gi_args(ctx);
    float4* merged = a.merged + ctx->gi_offset[level];
    uint32_t* list = static_cast<uint32_t*>(ctx->gi_sur_list.ptr);
    uint32_t* fcount = list + probes;
    float* queries = static_cast<float*>(ctx->gi_sur_queries.ptr);

    LUX_HIP(ctx, hipEventRecord(ctx->ev_phase[0], st));
    if (int rc = lux_device_zero(ctx, ctx->gi_counters.ptr, 8)) return rc;
    if (int rc = lux_gi_trace_levels(ctx, 0, level - 1)) return rc;
    LUX_HIP(ctx, hipEventRecord(ctx->ev_phase[1], st));
    uint32_t qblocks = (uint32_t)((count + kWg - 1) / kWg);
    hipLaunchKernelGGL(k_sur_queries, dim3(qblocks), dim3(kWg), 0, st, a, level, count, queries);
    const float* Y = nullptr;
    if (int rc = lux_surrogate_forward_internal(s, queries, (uint32_t)count, &Y)) return rc;
    hipLaunchKernelGGL(k_sur_apply, dim3(qblocks), dim3(kWg), 0, st, Y, count, merged);
    if (int rc = lux_device_zero(ctx, fcount, 4)) return rc;
    hipLaunchKernelGGL(k_sur_probe_error, dim3((uint32_t)probes), dim3(64), 0, st, Y, D, error_bound,
                       static_cast<float*>(ctx->gi_sur_err.ptr), list, fcount);
    LUX_HIP(ctx, hipGetLastError());
    LUX_HIP(ctx, hipEventRecord(ctx->ev_phase[2], st));
    if (ctx->gi_trace_mode == 0u) {
        Wave w = wave_proto(kWaveListed);
        w.level = level;
        w.list = list;
        w.list_count = fcount;
        if (int rc = run_wave(ctx, w, probes, D)) return rc;
    } else {
        hipLaunchKernelGGL(k_sur_fallback, dim3(qblocks), dim3(kWg), 0, st, a, level, list,
                           static_cast<const uint32_t*>(fcount));
        LUX_HIP(ctx, hipGetLastError());
    }
    if (int rc = history_pass(ctx, ctx->gi_offset[level])) return rc;
    LUX_HIP(ctx, hipEventRecord(ctx->ev_phase[3], st));
    if (int rc = merge_levels(ctx, (int)level - 1)) return rc;
    if (int rc = lux_gi_pack(ctx)) return rc;
    LUX_HIP(ctx, hipEventRecord(ctx->ev_phase[4], st));
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        uint32_t nf = 0;
        LUX_HIP(ctx, hipMemcpyAsync(&nf, fcount, 4, hipMemcpyDeviceToHost, st));
        LUX_HIP(ctx, hipStreamSynchronize(st));
        stats->level = level;
        stats->probes = (uint32_t)probes;
        stats->fallback_probes = nf;
        stats->queries = (uint32_t)count;
        LUX_HIP(ctx, hipEventElapsedTime(&stats->ms_total, ctx->ev_phase[0], ctx->ev_phase[4]));
        LUX_HIP(ctx, hipEventElapsedTime(&stats->ms_trace, ctx->ev_phase[0], ctx->ev_phase[1]));
        LUX_HIP(ctx, hipEventElapsedTime(&stats->ms_infer, ctx->ev_phase[1], ctx->ev_phase[2]));
        LUX_HIP(ctx, hipEventElapsedTime(&stats->ms_fallback, ctx->ev_phase[2], ctx->ev_phase[3]));
        LUX_HIP(ctx, hipEventElapsedTime(&stats->ms_merge_pack, ctx->ev_phase[3], ctx->ev_phase[4]));
    }
    ctx->gi_sur_probes = (uint32_t)probes;
    return LUX_OK;
}

extern "C" int lux_gi_surrogate_probe_error(lux_ctx* ctx, float* out, uint32_t count) {
    if (!ctx) return LUX_ERR_INVALID;
    if (ctx->gi_sur_probes == 0) return lux_fail(ctx, LUX_ERR_STATE, "no surrogate update has run");
    if (!out || count != ctx->gi_sur_probes)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_surrogate_probe_error: count must equal the level's probes");
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(out, ctx->gi_sur_err.ptr, 4ull * count, hipMemcpyDeviceToHost));
    return LUX_OK;
}

extern "C" int lux_camera_receivers(lux_ctx* ctx, const float* camera, uint32_t width, uint32_t height,
                                    float* receivers) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!camera || !receivers) return lux_fail(ctx, LUX_ERR_INVALID, "lux_camera_receivers: null buffer");
    if (!ctx->tlas_built) return lux_fail(ctx, LUX_ERR_STATE, "lux_camera_receivers: BVH not built");
    uint64_t n = (uint64_t)width * height;
    if (n == 0) return LUX_OK;
    if (n >= (1ull << 31)) return lux_fail(ctx, LUX_ERR_INVALID, "lux_camera_receivers: image too large");
    Camera cam;
    std::memcpy(&cam, camera, sizeof cam);
    hipLaunchKernelGGL(k_camera_receivers, dim3((uint32_t)((n + kWg - 1) / kWg)), dim3(kWg), 0, ctx->stream,
                       ctx->dev_scene(), cam, width, height, receivers);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

// lux_gi_device.h -- device-side transport code shared by the interval
// trace, merge, pack and resolve kernels (lux_transport.hip) and the batched
// multi-scene kernels (lux_batch.hip).  Internal linkage in each unit.
//
// All arithmetic is IEEE f32 in a fixed operation order; min/max/saturate
// that feed stored values are compare-selects; cos/sin of 2*pi*u use a fixed
// f64 Taylor evaluation.  The CPU oracle evaluates the same sequence.
#pragma once
#include "lux_internal.h"

namespace {

constexpr int kWg = 256;
// LDS stack slots per lane of the one-ray-per-thread kernels (camera
// receivers, surrogate samples and fallback, the fused traces): 16 KB per
// workgroup, so LDS leaves room for eight waves per SIMD (the 1080p camera
// pass: 1.93 against 2.32 ms with 16 slots, five waves; same receivers).
constexpr int kLdsDepth = 8;
constexpr float kPi = 3.14159265358979323846f;

struct Light {
    float4 position, color, direction, edge1, edge2;
    uint4 source;
};
static_assert(sizeof(Light) == 96, "light record");

struct GiArgs {
    lux::DevScene scene;
    const Light* lights;
    uint32_t n_lights;
    const lux_material* mats;
    uint32_t n_mats;
    lux_gi_params p;
    uint64_t offset[17];
    float4* intervals;
    float4* merged;
    float4* probes;
    uint32_t* counters;     // [0] attempts, [1] admitted
    uint32_t admit_limit;   // 0 = unbounded
    // temporal history of raw intervals (null when off): supported rays per
    // interval of this update; per interval (J.rgb, radiance mass) and
    // (beta, transmittance mass, last sampled frame bits, 0); state words
    // [0] history valid, [1] dirty boxes, [2] intervals whose history the last
    // pass reused, [3] the root box grew (measured-empty dependencies are
    // dropped), [4..7] unused, then 8 floats per dirty box (lo, pad, hi, pad)
    // from word 8
    uint32_t* samples;
    float4* history;
    uint32_t* hist_state;
    float hist_rb, hist_tb;
    uint32_t hist_frames, hist_persistent;
    // this update's dependency per interval (history on): (lo.xyz, known),
    // (hi.xyz, has bounds) -- the union of the world segments its supported
    // rays and their NEE occlusion rays traversed
    float4* dep;
};

constexpr uint32_t kMaxDirtyBoxes = 1024;

__device__ bool admit_query(const GiArgs& a);

// With an unbounded budget every query is admitted: count it in a per-lane
// tally instead of the shared counters, and add each wave's tally once
// (query_tally_flush, called by the whole wave) -- the same totals.
__device__ __forceinline__ bool admit_query_tallied(const GiArgs& a, uint32_t& tally) {
    if (a.admit_limit == 0u) {
        ++tally;
        return true;
    }
    return admit_query(a);
}

__device__ __forceinline__ void query_tally_flush(const GiArgs& a, uint32_t tally) {
    uint32_t s = tally;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (__lane_id() == 0u && s) {
        atomicAdd(&a.counters[0], s);
        atomicAdd(&a.counters[1], s);
    }
}

// Ticketed per-update traversal budget (admitTransportQuery,
// shaders/traversal.slang:20-38), aggregated per wave: the active lanes take
// consecutive tickets with one atomic, lanes past the limit are denied and
// their tickets returned, so the admitted count never exceeds the limit and
// ends exact.  Attempts are always counted.
__device__ bool admit_query(const GiArgs& a) {
    uint64_t mask = __ballot(1);
    uint32_t lane = __lane_id();
    uint32_t leader = (uint32_t)(__ffsll((unsigned long long)mask) - 1);
    uint32_t count = (uint32_t)__popcll(mask);
    uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    uint32_t base = 0;
    if (lane == leader) {
        atomicAdd(&a.counters[0], count);
        base = atomicAdd(&a.counters[1], count);
    }
    if (a.admit_limit == 0u) return true;
    base = __shfl(base, (int)leader, 64);
    bool ok = (uint64_t)base + rank < (uint64_t)a.admit_limit;
    uint64_t denied = __ballot(!ok);
    if (lane == leader && denied) atomicSub(&a.counters[1], (uint32_t)__popcll(denied));
    return ok;
}

__device__ __forceinline__ float smax(float a, float b) { return b > a ? b : a; }
__device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float sat(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }
__device__ __forceinline__ float fracf(float x) { return x - floorf(x); }
__device__ __forceinline__ float pow2i(int e) { return __uint_as_float((uint32_t)(127 + e) << 23); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

// ---- dependencies (DependencyContext, shaders/common.slang:138-242) ----
// The union of world-space segments that established a value: known /
// unknown, and measured-empty versus bounded.
struct Dep {
    float3 lo, hi;
    bool known, has;
};
__device__ __forceinline__ Dep dep_empty() {
    Dep d;
    d.lo = d.hi = make_float3(0.0f, 0.0f, 0.0f);
    d.known = true;
    d.has = false;
    return d;
}
__device__ __forceinline__ Dep dep_unknown() {
    Dep d = dep_empty();
    d.known = false;
    return d;
}
__device__ __forceinline__ bool finite3(float3 v) {
    return lux::isfinite_bits(v.x) && lux::isfinite_bits(v.y) && lux::isfinite_bits(v.z);
}
// dependencyIncludeBounds
__device__ __forceinline__ void dep_include_bounds(Dep& d, float3 mn, float3 mx) {
    if (!d.known) return;
    if (!finite3(mn) || !finite3(mx) || mn.x > mx.x || mn.y > mx.y || mn.z > mx.z) {
        d = dep_unknown();
        return;
    }
    if (!d.has) {
        d.lo = mn;
        d.hi = mx;
        d.has = true;
    } else {
        d.lo = make_float3(smin(d.lo.x, mn.x), smin(d.lo.y, mn.y), smin(d.lo.z, mn.z));
        d.hi = make_float3(smax(d.hi.x, mx.x), smax(d.hi.y, mx.y), smax(d.hi.z, mx.z));
    }
}
// dependencyIncludeSegment with no cone padding: endpoints widened by the
// float round-off slack max(scale * 1e-6, 2^-23), scale = max(1, |coords|).
__device__ __forceinline__ void dep_include_segment(Dep& d, float3 s, float3 e) {
    if (!finite3(s) || !finite3(e)) {
        d = dep_unknown();
        return;
    }
    float scale = smax(1.0f, smax(smax(fabsf(s.x), fabsf(s.y)), fabsf(s.z)));
    scale = smax(scale, smax(smax(fabsf(e.x), fabsf(e.y)), fabsf(e.z)));
    const float r = smax(scale * 1e-6f, 1.1920929e-7f);
    dep_include_bounds(d, make_float3(smin(s.x, e.x) - r, smin(s.y, e.y) - r, smin(s.z, e.z) - r),
                       make_float3(smax(s.x, e.x) + r, smax(s.y, e.y) + r, smax(s.z, e.z) + r));
}
// dependencyUnion
__device__ __forceinline__ void dep_union(Dep& d, const Dep& o) {
    if (!d.known || !o.known) {
        d = dep_unknown();
        return;
    }
    if (o.has) dep_include_bounds(d, o.lo, o.hi);
}
// The segment a ray traversed (traceRayTracked / traceOccludedTracked,
// shaders/traversal.slang:696-810,1097-1150): its part inside the scene's root
// box, from max(root entry, 0) to the hit (closest-hit rays that hit) or the
// root exit, which is clipped to the ray's limit; a ray missing the root
// traversed nothing (measured empty).
__device__ __forceinline__ void dep_ray(Dep& d, const lux::DevScene& s, float3 o, float3 dir, float limit, bool hit,
                                        float t) {
    if (s.n_inst == 0) return;
    const lux::Node root = lux::load_node(s.tlas, s.n_tlas, 0);
    float e, x;
    if (!lux::box_interval(o, dir, root.mn, root.mx, limit, e, x)) return;
    const float s0 = smax(e, 0.0f);
    const float s1 = hit ? smin(t, x) : x;
    dep_include_segment(d, make_float3(o.x + dir.x * s0, o.y + dir.y * s0, o.z + dir.z * s0),
                        make_float3(o.x + dir.x * s1, o.y + dir.y * s1, o.z + dir.z * s1));
}
__device__ __forceinline__ void dep_store(float4* out, const Dep& d) {
    out[0] = make_float4(d.lo.x, d.lo.y, d.lo.z, d.known ? 1.0f : 0.0f);
    out[1] = make_float4(d.hi.x, d.hi.y, d.hi.z, d.has ? 1.0f : 0.0f);
}
// storedDependency: w > 0.5 flags; a malformed bounded set is unknown.
__device__ __forceinline__ Dep dep_load(const float4* in) {
    const float4 a = in[0], b = in[1];
    Dep d;
    d.lo = make_float3(a.x, a.y, a.z);
    d.hi = make_float3(b.x, b.y, b.z);
    d.known = a.w > 0.5f;
    d.has = b.w > 0.5f;
    if (d.known && d.has && (!finite3(d.lo) || !finite3(d.hi) || d.lo.x > d.hi.x || d.lo.y > d.hi.y ||
                             d.lo.z > d.hi.z))
        d.known = false;
    return d;
}

__device__ __forceinline__ uint32_t level_dim(uint32_t d, uint32_t level) {
    for (uint32_t l = 0; l < level; ++l) d = (d + 1u) / 2u;
    return d;
}

__device__ __forceinline__ uint32_t transport_hash(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

__device__ __forceinline__ float2 r2(uint32_t index) {
    uint32_t px = index * 3242174889u, py = index * 2447445414u;
    return make_float2((float)(px >> 8) / 16777216.0f, (float)(py >> 8) / 16777216.0f);
}

__device__ __forceinline__ uint32_t hash_key(int32_t x, int32_t y, int32_t z, int32_t w,
                                             uint32_t buckets) {
    uint32_t h = (uint32_t)x * 0x8da6b343u ^ (uint32_t)y * 0xd8163841u ^
                 (uint32_t)z * 0xcb1ab31fu ^ (uint32_t)w * 0x165667b1u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    return h & (buckets - 1u);
}

// cos(2 pi u), sin(2 pi u): exact quadrant split, Taylor series in f64.
__device__ __forceinline__ void cos_sin_2pi(float u, float& c, float& s) {
    double x = (double)u * 4.0;
    double q = floor(x);
    double r = x - q;
    int qi = ((int)q) & 3;
    double t = r * 1.5707963267948966;
    double t2 = t * t;
    double sn = 1.0 / 51090942171709440000.0;
    sn = sn * t2 * -1.0 + 1.0 / 121645100408832000.0;
    sn = sn * t2 * -1.0 + 1.0 / 355687428096000.0;
    sn = sn * t2 * -1.0 + 1.0 / 1307674368000.0;
    sn = sn * t2 * -1.0 + 1.0 / 6227020800.0;
    sn = sn * t2 * -1.0 + 1.0 / 39916800.0;
    sn = sn * t2 * -1.0 + 1.0 / 362880.0;
    sn = sn * t2 * -1.0 + 1.0 / 5040.0;
    sn = sn * t2 * -1.0 + 1.0 / 120.0;
    sn = sn * t2 * -1.0 + 1.0 / 6.0;
    sn = sn * t2 * -1.0 + 1.0;
    sn = sn * t;
    double cs = 1.0 / 2432902008176640000.0;
    cs = cs * t2 * -1.0 + 1.0 / 6402373705728000.0;
    cs = cs * t2 * -1.0 + 1.0 / 20922789888000.0;
    cs = cs * t2 * -1.0 + 1.0 / 87178291200.0;
    cs = cs * t2 * -1.0 + 1.0 / 479001600.0;
    cs = cs * t2 * -1.0 + 1.0 / 3628800.0;
    cs = cs * t2 * -1.0 + 1.0 / 40320.0;
    cs = cs * t2 * -1.0 + 1.0 / 720.0;
    cs = cs * t2 * -1.0 + 1.0 / 24.0;
    cs = cs * t2 * -1.0 + 1.0 / 2.0;
    cs = cs * t2 * -1.0 + 1.0;
    double oc, os;
    if (qi == 0) { oc = cs; os = sn; }
    else if (qi == 1) { oc = -sn; os = cs; }
    else if (qi == 2) { oc = -cs; os = -sn; }
    else { oc = sn; os = -cs; }
    c = (float)oc;
    s = (float)os;
}

// Equal-area bin direction with an in-bin offset (binDirection generalised).
__device__ __forceinline__ float3 bin_direction(uint32_t index, uint32_t nphi, uint32_t ny,
                                                float ox, float oy) {
    float ux = ((float)(index % nphi) + ox) / (float)nphi;
    float uy = ((float)(index / nphi) + oy) / (float)ny;
    float c, s;
    cos_sin_2pi(ux, c, s);
    float y = uy * 2.0f - 1.0f;
    float q = 1.0f - y * y;
    float r = sqrtf(q > 0.0f ? q : 0.0f);
    return make_float3(r * c, y, r * s);
}

__device__ __forceinline__ float2 ray_offset_uv(uint32_t seed, uint32_t r) {
    float2 j = r2(transport_hash(seed ^ 0x5a17c9e3u));
    float2 q = r2(r);
    return make_float2(fracf(0.5f + j.x + q.x), fracf(0.5f + j.y + q.y));
}

__device__ __forceinline__ float offset1(float pos, float n, float bias) {
    float origin = pos + n * bias;
    int32_t step = (int32_t)(n * 256.0f);
    uint32_t moved = __float_as_uint(origin) + (uint32_t)(origin < 0.0f ? -step : step);
    float integer_offset = __uint_as_float(moved);
    float shifted = fabsf(origin) < 0.03125f ? origin + n / 65536.0f : integer_offset;
    return lux::isfinite_bits(shifted) ? shifted : pos;
}

// rayOffset (common.slang:345-359).
__device__ __forceinline__ float3 ray_offset(float3 p, float3 n, float bias) {
    return make_float3(offset1(p.x, n.x, bias), offset1(p.y, n.y, bias), offset1(p.z, n.z, bias));
}

__device__ __forceinline__ float3 unit_vector(float3 v, float3 fb) {
    float sc = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z)));
    if (sc == 0.0f || !lux::isfinite_bits(sc)) return fb;
    float div = fmaxf(sc * 0.25f, 1.17549435e-38f);
    float3 w = make_float3(v.x / div, v.y / div, v.z / div);
    float len = sqrtf(dot3(w, w));
    return make_float3(w.x / len, w.y / len, w.z / len);
}

// triangleGeometricNormal through the inverse rows (traversal.slang:248-251).
__device__ float3 geometric_normal(const float4* rec, float3 a, float3 b, float3 c, uint32_t flags) {
    float3 e1 = make_float3(b.x - a.x, b.y - a.y, b.z - a.z);
    float3 e2 = make_float3(c.x - a.x, c.y - a.y, c.z - a.z);
    float3 cr = make_float3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z,
                            e1.x * e2.y - e1.y * e2.x);
    float4 r0 = rec[0], r1 = rec[1], r2v = rec[2];
    float m0 = fmaxf(fabsf(r0.x), fmaxf(fabsf(r1.x), fabsf(r2v.x)));
    float m1 = fmaxf(fabsf(r0.y), fmaxf(fabsf(r1.y), fabsf(r2v.y)));
    float m2 = fmaxf(fabsf(r0.z), fmaxf(fabsf(r1.z), fabsf(r2v.z)));
    float sc = fmaxf(m0, fmaxf(m1, m2));
    float3 tv = make_float3(0.0f, 0.0f, 0.0f);
    if (sc != 0.0f) {
        sc = fmaxf(sc * 0.25f, 1.17549435e-38f);
        float3 a0 = make_float3(r0.x / sc, r0.y / sc, r0.z / sc);
        float3 a1 = make_float3(r1.x / sc, r1.y / sc, r1.z / sc);
        float3 a2 = make_float3(r2v.x / sc, r2v.y / sc, r2v.z / sc);
        float3 u = unit_vector(cr, make_float3(0.0f, 0.0f, 0.0f));
        tv = make_float3((a0.x * u.x + a1.x * u.y) + a2.x * u.z, (a0.y * u.x + a1.y * u.y) + a2.y * u.z,
                         (a0.z * u.x + a1.z * u.y) + a2.z * u.z);
    }
    float3 n = unit_vector(tv, make_float3(0.0f, 1.0f, 0.0f));
    float sg = (flags & 4u) != 0u ? -1.0f : 1.0f;
    return make_float3(sg * n.x, sg * n.y, sg * n.z);
}

__device__ __forceinline__ float pow5(float x) {
    float x2 = x * x;
    float x4 = x2 * x2;
    return x4 * x;
}

// disneyDiffuseMultiplier (common.slang:299-314).
__device__ float disney(float3 n, float3 l, float3 v, float pr) {
    float ndl = sat(dot3(n, l));
    float ndv = sat(dot3(n, v));
    float3 hs = make_float3(l.x + v.x, l.y + v.y, l.z + v.z);
    float hl2 = dot3(hs, hs);
    if (ndl <= 0.0f || ndv <= 0.0f || hl2 <= 1e-20f) return 0.0f;
    float rs = 1.0f / sqrtf(hl2);
    float3 h = make_float3(hs.x * rs, hs.y * rs, hs.z * rs);
    float rough = smax(sat(pr), 0.089f);
    rough = rough * rough;
    float ldh = sat(dot3(l, h));
    float f90 = 0.5f + ((2.0f * rough) * ldh) * ldh;
    float ls = 1.0f + (f90 - 1.0f) * pow5(1.0f - ndl);
    float vs = 1.0f + (f90 - 1.0f) * pow5(1.0f - ndv);
    return ls * vs;
}

using Stack = lux::LaneStack<kLdsDepth, kWg>;

// Hit shading (cascades.slang:60-134,196-268) for untextured opaque
// materials, black environment, no punctual lights, in three pieces shared by
// the fused and the wavefront interval traces:
//   shade_begin   emission of the hit and the NEE context (whether NEE runs)
//   nee_sample    one emissive sample: false when it is skipped before any
//                 visibility query, else the occlusion ray and the
//                 contribution that counts when the ray is unoccluded
//   shade         the fused loop: queries in sample order, unsupported
//                 visibility makes the whole sample unsupported
struct NeeCtx {
    float3 pos, gn, nrm, view;
    float rough, bias, phase;
    float2 rot;
    uint32_t count, frame;
    lux_material mat;
};

// Returns true when NEE runs for this hit; rad receives the emission.
__device__ bool shade_begin(const GiArgs& a, float3 o, float3 d, const lux::HitState& h, float3& rad,
                            NeeCtx& c) {
    const lux::DevScene& s = a.scene;
    const float4* rec = s.inst + 8u * h.inst;
    uint32_t iflags = __float_as_uint(rec[6].x);
    const float4* tp = s.tri + 3ull * h.tri;
    float4 A = tp[0], B = tp[1], C = tp[2];
    // a nonzero instance material (scenes installed from scene words) wins
    const uint32_t mo = __float_as_uint(rec[7].z);
    uint32_t mi = mo ? mo - 1u : __float_as_uint(B.w);
    rad = make_float3(0.0f, 0.0f, 0.0f);
    if (a.n_mats == 0) return false;
    if (mi >= a.n_mats) mi = a.n_mats - 1u;
    c.mat = a.mats[mi];
    rad = make_float3(c.mat.emission[0], c.mat.emission[1], c.mat.emission[2]);
    if ((iflags & 16u) != 0u) return false;
    float3 pos = make_float3(o.x + d.x * h.t, o.y + d.y * h.t, o.z + d.z * h.t);
    float3 gn = geometric_normal(rec, make_float3(A.x, A.y, A.z), make_float3(B.x, B.y, B.z),
                                 make_float3(C.x, C.y, C.z), iflags);
    if (dot3(gn, d) > 0.0f) gn = make_float3(-gn.x, -gn.y, -gn.z);
    float3 nrm = gn;
    float3 md = make_float3(-d.x, -d.y, -d.z);
    if (dot3(nrm, md) <= 0.0f || dot3(nrm, gn) <= 0.0f || a.n_lights == 0) return false;
    float vl = sqrtf(dot3(md, md));
    c.view = make_float3(md.x / vl, md.y / vl, md.z / vl);
    c.rough = lux::isfinite_bits(c.mat.roughness) ? c.mat.roughness : 1.0f;
    c.count = a.p.emissive_samples / 2u;
    if (c.count < 1u) c.count = 1u;
    c.bias = a.p.normal_bias;
    float sp = a.p.spacing * 0.05f;
    uint32_t seed = hash_key((int32_t)floorf(pos.x / sp), (int32_t)floorf(pos.y / sp),
                             (int32_t)floorf(pos.z / sp), 0, 16777216u);
    c.rot = r2(seed);
    uint32_t sel_seed = hash_key((int32_t)floorf(pos.x / a.p.spacing), (int32_t)floorf(pos.y / a.p.spacing),
                                 (int32_t)floorf(pos.z / a.p.spacing), 0, 65536u);
    c.frame = a.p.frame ^ 0xE11E0001u;
    c.phase = fracf(r2(c.frame).x + r2(sel_seed).x);
    c.pos = pos;
    c.gn = gn;
    c.nrm = nrm;
    return true;
}

__device__ bool nee_sample(const GiArgs& a, const NeeCtx& c, uint32_t smp, float3& origin, float3& dir,
                           float& lim, float3& term) {
    float2 q = r2(smp + c.frame * c.count + 1u);
    float2 uv = make_float2(fracf(q.x + c.rot.x), fracf(q.y + c.rot.y));
    float selection = fracf(((float)smp + 0.5f) / (float)c.count + c.phase);
    uint32_t lo = 0, hi = a.n_lights - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (a.lights[mid].direction.w < selection) lo = mid + 1u;
        else hi = mid;
    }
    const Light L = a.lights[lo];
    const float3 pos = c.pos;
    float3 lp = make_float3(L.position.x, L.position.y, L.position.z);
    float3 ldir = make_float3(L.direction.x, L.direction.y, L.direction.z);
    float3 rel = make_float3(pos.x - lp.x, pos.y - lp.y, pos.z - lp.z);
    if (L.edge1.w <= 0.0f || (L.position.w != -2.0f && dot3(ldir, rel) <= 0.0f)) return false;
    float root = sqrtf(uv.x);
    float wa = root * (1.0f - uv.y);
    float wb = root * uv.y;
    float3 target = make_float3((lp.x + L.edge1.x * wa) + L.edge2.x * wb,
                                (lp.y + L.edge1.y * wa) + L.edge2.y * wb,
                                (lp.z + L.edge1.z * wa) + L.edge2.z * wb);
    float3 tl = make_float3(target.x - pos.x, target.y - pos.y, target.z - pos.z);
    float ldist = sqrtf(dot3(tl, tl));
    if (ldist <= c.bias * 0.25f) return false;
    float3 Ld = make_float3(tl.x / ldist, tl.y / ldist, tl.z / ldist);
    float cosine = smax(0.0f, dot3(c.nrm, Ld));
    float ec = dot3(ldir, make_float3(-Ld.x, -Ld.y, -Ld.z));
    if (L.position.w == -2.0f) ec = fabsf(ec);
    float ecos = smax(0.0f, ec);
    if (cosine <= 0.0f || ecos <= 0.0f || dot3(c.gn, Ld) <= 0.0f) return false;
    float3 em = make_float3(L.color.x, L.color.y, L.color.z);
    if (!(em.x > 0.0f || em.y > 0.0f || em.z > 0.0f)) return false;
    float pdf = ((L.edge1.w * ldist) * ldist) / (L.color.w * ecos);
    float plane = fabsf(dot3(ldir, rel));
    float ob = smin(c.bias, plane * 0.25f);
    origin = ray_offset(pos, c.gn, ob);
    float3 delta = make_float3(target.x - origin.x, target.y - origin.y, target.z - origin.z);
    float dist = sqrtf(dot3(delta, delta));
    if (dist <= c.bias * 0.25f) return false;
    dir = make_float3(delta.x / dist, delta.y / dist, delta.z / dist);
    lim = dist - smin(c.bias * 0.25f, dist * 0.001f);
    float dm = disney(c.nrm, Ld, c.view, c.rough) / kPi;
    float pm = smax(pdf, 1e-20f);
    term = make_float3(em.x * (((c.mat.albedo[0] * dm) * cosine) / pm),
                       em.y * (((c.mat.albedo[1] * dm) * cosine) / pm),
                       em.z * (((c.mat.albedo[2] * dm) * cosine) / pm));
    return true;
}

// Emission + NEE sum -> the sample's radiance (non-finite -> 0, negative
// clamped), the end of the reference's outgoing-radiance evaluation.
__device__ __forceinline__ float3 shade_end(float3 rad) {
    if (!(lux::isfinite_bits(rad.x) && lux::isfinite_bits(rad.y) && lux::isfinite_bits(rad.z)))
        return make_float3(0.0f, 0.0f, 0.0f);
    return make_float3(smax(rad.x, 0.0f), smax(rad.y, 0.0f), smax(rad.z, 0.0f));
}

__device__ bool shade(const GiArgs& a, float3 o, float3 d, const lux::HitState& h, Stack& stk,
                      float3& out, Dep* dep = nullptr) {
    float3 rad;
    NeeCtx c;
    if (shade_begin(a, o, d, h, rad, c)) {
        float3 sum = make_float3(0.0f, 0.0f, 0.0f);
        for (uint32_t smp = 0; smp < c.count; ++smp) {
            float3 origin, dir, term;
            float lim;
            if (!nee_sample(a, c, smp, origin, dir, lim, term)) continue;
            if (!admit_query(a)) return false;
            lux::HitState vh;
            lux::NoStats ns;
            lux::trace_occluded(a.scene, origin, dir, lim, vh, stk, ns);
            if (dep) dep_ray(*dep, a.scene, origin, dir, lim, false, 0.0f);
            if (vh.err & lux::kHitOccluded) continue;
            if (vh.err & lux::kErrUnsupported) return false;
            sum.x = sum.x + term.x;
            sum.y = sum.y + term.y;
            sum.z = sum.z + term.z;
        }
        float fc = (float)c.count;
        rad = make_float3(rad.x + sum.x / fc, rad.y + sum.y / fc, rad.z + sum.z / fc);
    }
    out = shade_end(rad);
    return true;
}

__device__ __forceinline__ float interval_end(const lux_gi_params& p, uint32_t level) {
    if (p.interval_mode == 1u) return p.interval_length * (float)((2u << level) - 1u);
    return p.interval_length * pow2i((int)(2u * level));
}

// Probe centre of level L for a probe index (x fastest).
__device__ __forceinline__ float3 probe_centre(const lux_gi_params& p, uint32_t L, uint64_t probe) {
    uint32_t dx = level_dim(p.dims[0], L), dy = level_dim(p.dims[1], L);
    uint32_t ix = (uint32_t)(probe % dx);
    uint32_t iy = (uint32_t)((probe / dx) % dy);
    uint32_t iz = (uint32_t)(probe / ((uint64_t)dx * dy));
    float sl = p.spacing * pow2i((int)L);
    return make_float3(p.origin[0] + ((float)ix + 0.5f) * sl, p.origin[1] + ((float)iy + 0.5f) * sl,
                       p.origin[2] + ((float)iz + 0.5f) * sl);
}

// One interval: rays_per_interval jittered rays of bin k from the probe
// centre c over [start, end) of level L (or [start, inf) when terminal):
// J = mean shaded radiance, beta = mean transmittance (1 on a miss of a
// bounded interval), beta = -1 when no ray was supported.
__device__ float4 trace_interval(const GiArgs& a, uint32_t L, float3 c, uint32_t k, bool terminal,
                                 Stack& stk, uint32_t* n_out = nullptr, Dep* dep = nullptr) {
    const lux_gi_params& p = a.p;
    uint32_t nphi = p.base_phi << L, ny = p.base_y << L;
    float end = interval_end(p, L);
    float start = L == 0 ? 0.0f : interval_end(p, L - 1);
    float limit = terminal ? 1e20f : end - start;
    float3 jsum = make_float3(0.0f, 0.0f, 0.0f);
    float bsum = 0.0f;
    uint32_t n = 0;
    uint32_t R = p.rays_per_interval ? p.rays_per_interval : 1u;
    for (uint32_t r = 0; r < R; ++r) {
        float2 ouv = ray_offset_uv(p.seed, r);
        float3 d = bin_direction(k, nphi, ny, ouv.x, ouv.y);
        float3 o = make_float3(c.x + d.x * start, c.y + d.y * start, c.z + d.z * start);
        if (!admit_query(a)) continue;  // denied: no sample, as an unsupported traversal
        lux::HitState h;
        lux::NoStats ns;
        lux::trace_closest(a.scene, o, d, limit, h, stk, ns, false);
        if (h.err & lux::kErrUnsupported) continue;
        float3 rad = make_float3(0.0f, 0.0f, 0.0f);
        float beta;
        // the ray's segment, then its NEE occlusion segments (a sample that
        // is dropped contributes no dependency)
        Dep rd = dep_empty();
        if (dep) dep_ray(rd, a.scene, o, d, limit, h.tri != lux::kEmpty, h.t);
        if (h.tri != lux::kEmpty) {
            if (!shade(a, o, d, h, stk, rad, dep ? &rd : nullptr)) continue;
            beta = 0.0f;
        } else {
            beta = terminal ? 0.0f : 1.0f;
        }
        if (dep) dep_union(*dep, rd);
        jsum = make_float3(jsum.x + rad.x, jsum.y + rad.y, jsum.z + rad.z);
        bsum = bsum + beta;
        n++;
    }
    if (n_out) *n_out = n;
    if (n == 0) return make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    float fn = (float)n;
    return make_float4(jsum.x / fn, jsum.y / fn, jsum.z / fn, bsum / fn);
}

// The merge of one interval with no child contribution (its level is the
// last one traced): Eq. 7 with an empty child sum.
__device__ __forceinline__ float4 merge_leaf(float4 iv) {
    float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (iv.w >= 0.0f) {
        float tr = sat(iv.w);
        float support = 1.0f - tr;
        if (support > 0.0f)
            out = make_float4(iv.x / support, iv.y / support, iv.z / support, sat(support));
    }
    return out;
}

// Surrogate query of (probe, bin k) at level L: probe centre, the frame's
// first jittered ray direction of the bin, normal = that direction.
__device__ __forceinline__ void level_query(const lux_gi_params& p, uint32_t L, uint64_t local,
                                            float* o) {
    uint32_t nphi = p.base_phi << L, ny = p.base_y << L;
    uint32_t D = nphi * ny;
    float3 c = probe_centre(p, L, local / D);
    float2 ouv = ray_offset_uv(p.seed, 0u);
    float3 d = bin_direction((uint32_t)(local % D), nphi, ny, ouv.x, ouv.y);
    o[0] = c.x; o[1] = c.y; o[2] = c.z;
    o[3] = d.x; o[4] = d.y; o[5] = d.z;
    o[6] = d.x; o[7] = d.y; o[8] = d.z;
}

// The ray of item (interval g, ray r): exactly trace_interval's construction.
__device__ __forceinline__ void item_ray(const GiArgs& a, uint64_t g, uint32_t r, bool force_terminal,
                                         float3& o, float3& d, float& limit, bool& terminal) {
    const lux_gi_params& p = a.p;
    uint32_t L = 0;
    while (L + 1 < p.levels && g >= a.offset[L + 1]) ++L;
    uint64_t local = g - a.offset[L];
    uint32_t nphi = p.base_phi << L, ny = p.base_y << L;
    uint32_t D = nphi * ny;
    uint32_t k = (uint32_t)(local % D);
    float3 c = probe_centre(p, L, local / D);
    float end = interval_end(p, L);
    float start = L == 0 ? 0.0f : interval_end(p, L - 1);
    terminal = force_terminal || (p.terminal_tail != 0u && L + 1 == p.levels);
    limit = terminal ? 1e20f : end - start;
    float2 ouv = ray_offset_uv(p.seed, r);
    d = bin_direction(k, nphi, ny, ouv.x, ouv.y);
    o = make_float3(c.x + d.x * start, c.y + d.y * start, c.z + d.z * start);
}

// Support-weighted trilinear child lookup at level Lc.
__device__ float4 interp_child(const GiArgs& a, uint32_t Lc, uint32_t dir, float3 rel) {
    const lux_gi_params& p = a.p;
    uint32_t dx = level_dim(p.dims[0], Lc), dy = level_dim(p.dims[1], Lc), dz = level_dim(p.dims[2], Lc);
    uint32_t D = (p.base_phi << Lc) * (p.base_y << Lc);
    float sc = p.spacing * pow2i((int)Lc);
    float gx = rel.x / sc - 0.5f, gy = rel.y / sc - 0.5f, gz = rel.z / sc - 0.5f;
    float flx = floorf(gx), fly = floorf(gy), flz = floorf(gz);
    int32_t bx0 = (int32_t)flx, by0 = (int32_t)fly, bz0 = (int32_t)flz;
    float fx = gx - flx, fy = gy - fly, fz = gz - flz;
    const float4* mc = a.merged + a.offset[Lc];
    float3 sum = make_float3(0.0f, 0.0f, 0.0f);
    float total = 0.0f, potential = 0.0f;
    for (uint32_t corner = 0; corner < 8u; ++corner) {
        uint32_t bx = corner & 1u, by = (corner >> 1) & 1u, bz = corner >> 2;
        float ax = bx == 0u ? 1.0f - fx : fx;
        float ay = by == 0u ? 1.0f - fy : fy;
        float az = bz == 0u ? 1.0f - fz : fz;
        float w = (ax * ay) * az;
        if (!(w > 0.0f)) continue;
        potential = potential + w;
        int64_t cx = (int64_t)bx0 + bx, cy = (int64_t)by0 + by, cz = (int64_t)bz0 + bz;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= dx || cy >= dy || cz >= dz) continue;
        uint64_t probe = ((uint64_t)cz * dy + (uint64_t)cy) * dx + (uint64_t)cx;
        float4 smp = mc[probe * D + dir];
        if (!(smp.w > 0.0f) || !(lux::isfinite_bits(smp.x) && lux::isfinite_bits(smp.y) &&
                                 lux::isfinite_bits(smp.z) && lux::isfinite_bits(smp.w)))
            continue;
        float sw = w * sat(smp.w);
        sum = make_float3(sum.x + smp.x * sw, sum.y + smp.y * sw, sum.z + smp.z * sw);
        total = total + sw;
    }
    if (total > 0.0f)
        return make_float4(sum.x / total, sum.y / total, sum.z / total,
                           sat(total / smax(potential, 1e-8f)));
    return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Eq. 7 merge of one interval of level L (mergeEntry, split.slang:1523-1725).
__device__ void merge_interval(const GiArgs& a, uint32_t L, uint64_t local) {
    const lux_gi_params& p = a.p;
    uint32_t dx = level_dim(p.dims[0], L), dy = level_dim(p.dims[1], L);
    uint32_t nphi = p.base_phi << L;
    uint32_t D = nphi * (p.base_y << L);
    uint32_t k = (uint32_t)(local % D);
    uint64_t probe = local / D;
    uint32_t ix = (uint32_t)(probe % dx), iy = (uint32_t)((probe / dx) % dy);
    uint32_t iz = (uint32_t)(probe / ((uint64_t)dx * dy));
    float sl = p.spacing * pow2i((int)L);
    float3 rel = make_float3(((float)ix + 0.5f) * sl, ((float)iy + 0.5f) * sl, ((float)iz + 0.5f) * sl);
    const uint64_t g = a.offset[L] + local;
    float4 iv = a.intervals[g];
    float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // own radiance needs both masses; children need transmittance only
    // (beginMergeGroup, split.slang:1533-1545); without history both masses
    // are 1 for a sampled interval (beta >= 0) and 0 otherwise
    bool has;
    float tr;
    if (a.history) {
        float rm = a.history[4 * g].w, tm = a.history[4 * g + 1].y;
        tr = tm > 0.0f ? sat(iv.w) : 0.0f;
        has = rm > 0.0f && tm > 0.0f;
    } else {
        has = iv.w >= 0.0f;
        tr = has ? sat(iv.w) : 0.0f;
    }
    if (has || tr > 0.0f) {
        float support = has ? 1.0f - tr : 0.0f;
        float3 rad = make_float3(0.0f, 0.0f, 0.0f);
        if (support > 0.0f) rad = make_float3(iv.x, iv.y, iv.z);
        float3 csum = make_float3(0.0f, 0.0f, 0.0f);
        float csup = 0.0f;
        if (tr > 0.0f && L + 1 < p.levels) {
            uint32_t x = k % nphi, y = k / nphi;
            for (uint32_t child = 0; child < 4u; ++child) {
                uint32_t fx = x * 2u + (child & 1u), fy = y * 2u + (child >> 1);
                uint32_t fd = fx + fy * (2u * nphi);
                float4 v = interp_child(a, L + 1, fd, rel);
                if (!(v.w > 0.0f)) continue;
                csum = make_float3(csum.x + v.x * v.w, csum.y + v.y * v.w, csum.z + v.z * v.w);
                csup = csup + v.w;
            }
        }
        if (csup > 0.0f) {
            rad = make_float3(rad.x + (tr * csum.x) * 0.25f, rad.y + (tr * csum.y) * 0.25f,
                              rad.z + (tr * csum.z) * 0.25f);
            support = support + (tr * csup) * 0.25f;
        }
        if (support > 0.0f)
            out = make_float4(rad.x / support, rad.y / support, rad.z / support, sat(support));
    }
    a.merged[g] = out;
}

// x^n for a small integer n by repeated multiplication (pow(x, n) of the
// reference for integer sample counts; exact for n = 1).
__device__ __forceinline__ float powi(float x, uint32_t n) {
    float r = 1.0f;
    for (uint32_t i = 0; i < n; ++i) r = r * x;
    return r;
}

__device__ __forceinline__ float valid_mass(float m) { return (lux::isfinite_bits(m) && m >= 0.0f) ? m : 0.0f; }

// Temporal blend of interval g (level L, index `local` in the level) with
// its history: mergeEntry's history part (split.slang:1624-1705) and
// mergeRawRadianceGroup (:1478-1521) for the one radiance source.  History
// is reused when valid, alive (persistent world intervals, or sampled within
// hist_frames frames; intervalHistoryAlive, common.slang:85-89) and the
// interval's dependency box meets no dirty box.  The dependency box bounds
// every segment the interval's rays and their NEE occlusion rays traverse:
// the box of probe centre +- interval end joined with the emitter bounds
// (unbounded for a terminal interval).  Writes the blended (J, beta) in
// place and the new history.
__device__ bool history_blend(const GiArgs& a, uint64_t g, uint32_t L, uint64_t local) {
    const lux_gi_params& p = a.p;
    const uint32_t* st = a.hist_state;
    float4 iv = a.intervals[g];
    uint32_t n = a.samples[g];
    float4 h0 = a.history[4 * g], h1 = a.history[4 * g + 1];
    uint32_t last = __float_as_uint(h1.z);
    bool reuse = st[0] != 0u && a.hist_frames > 0u;
    if (reuse) reuse = a.hist_persistent != 0u || p.frame - last <= a.hist_frames;
    uint32_t nd = st[1] < kMaxDirtyBoxes ? st[1] : kMaxDirtyBoxes;
    const Dep old = dep_load(a.history + 4 * g + 2);
    // a measured-empty dependency holds only while the scene's root box did not grow
    if (reuse && st[3] != 0u && old.known && !old.has) reuse = false;
    if (reuse && nd > 0u) {
        // dependencyIntersectsDirty (split.slang:13-25) against each dirty box:
        // an unknown dependency meets every region, a measured-empty one none
        if (!old.known) reuse = false;
        for (uint32_t b = 0; b < nd && reuse && old.has; ++b) {
            const float* box = reinterpret_cast<const float*>(st + 8) + 8 * b;
            if (old.lo.x <= box[4] && old.hi.x >= box[0] && old.lo.y <= box[5] && old.hi.y >= box[1] &&
                old.lo.z <= box[6] && old.hi.z >= box[2])
                reuse = false;
        }
    }
    const float rbr = a.hist_rb, tbr = a.hist_tb;
    const float rb = smin(sat(rbr), 0.999f), tb = smin(sat(tbr), 0.999f);
    const bool reuse_r = reuse && rbr > 0.0f, reuse_t = reuse && tbr > 0.0f;
    const float rlim = 1.0f / (1.0f - rb), tlim = 1.0f / (1.0f - tb);
    const float om = valid_mass(h0.w), otm = valid_mass(h1.y);
    float3 J = make_float3(iv.x, iv.y, iv.z);
    float beta = iv.w, mass, ts;
    if (n > 0u) {
        float ret = powi(rb, n);
        mass = (1.0f - ret) * rlim;
        float o = smin(om, rlim) * ret;
        if (reuse_r && o > 0.0f) {
            float t = mass + o;
            J = make_float3((J.x * mass + h0.x * o) / t, (J.y * mass + h0.y * o) / t, (J.z * mass + h0.z * o) / t);
            mass = mass + o;
        }
        float tret = powi(tb, n);
        ts = (1.0f - tret) * tlim;
        float ot = smin(otm, tlim) * tret;
        if (reuse_t && ot > 0.0f) {
            beta = (beta * ts + h1.x * ot) / (ts + ot);
            ts = ts + ot;
        }
        last = p.frame;
    } else {
        if (reuse_r && om > 0.0f) {
            J = make_float3(h0.x, h0.y, h0.z);
            mass = om;
        } else {
            J = make_float3(0.0f, 0.0f, 0.0f);
            mass = 0.0f;
        }
        if (reuse_t && otm > 0.0f) {
            beta = h1.x;
            ts = otm;
        } else {
            beta = 1.0f;
            ts = 0.0f;
        }
    }
    a.intervals[g] = make_float4(J.x, J.y, J.z, beta);
    a.history[4 * g] = make_float4(J.x, J.y, J.z, mass);
    a.history[4 * g + 1] = make_float4(beta, ts, __uint_as_float(last), 0.0f);
    // the retained value depends on this update's segments and, when history
    // was reused, on everything the history depended on
    Dep cur = dep_load(a.dep + 2 * g);
    if (reuse) {
        Dep u = old;
        dep_union(u, cur);
        cur = u;
    }
    dep_store(a.history + 4 * g + 2, cur);
    return reuse && (om > 0.0f || otm > 0.0f);  // history contributed
}

// decodeNormal (common.slang:370-375).
__device__ __forceinline__ float3 decode_normal(float px, float py) {
    float x = px * 2.0f - 1.0f, y = py * 2.0f - 1.0f;
    float z = 1.0f - fabsf(x) - fabsf(y);
    float nz = -z;
    float t = nz < 0.0f ? 0.0f : (nz > 1.0f ? 1.0f : nz);
    x = x + (x >= 0.0f ? -t : t);
    y = y + (y >= 0.0f ? -t : t);
    float l = sqrtf((x * x + y * y) + z * z);
    return make_float3(x / l, y / l, z / l);
}

constexpr uint32_t kMaxPackDirections = 256;

// 16 probes per 256-lane block, 4 per wave, sharing the block's cosine table.
// Dynamic LDS (floats): the table (36 * D), the nominal sums (36), the
// directions (3 * D), then from a 16-byte boundary four staging arrays.
constexpr uint32_t kPackBlockProbes = 16;
__host__ __device__ constexpr uint32_t pack_smp_offset(uint32_t D) { return (40u * D + 36u + 3u) & ~3u; }
inline size_t pack_lds_bytes(uint32_t D) { return 4ull * (pack_smp_offset(D) + 16ull * D); }

// packProbes over many probes of one direction set: the per-texel cosines
// max(n . dir, 0) and their sum (the nominal weight) do not depend on the
// probe, so a block computes them once (in the same order, so the same
// bits) into `ctab` (direction-major, D * 36: the lanes of a wave read
// consecutive words) and `nom` (36); each wave then packs its probes
// with `smp` (D) as its staging array.  Lanes 0..35 are the texels.
__device__ void pack_table(const lux_gi_params& p, float* ctab, float* nom, float3* dirs) {
    const uint32_t D = p.base_phi * p.base_y;
    for (uint32_t k = threadIdx.x; k < D; k += blockDim.x) dirs[k] = bin_direction(k, p.base_phi, p.base_y, 0.5f, 0.5f);
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < 36u; t += blockDim.x) {
        float3 n = decode_normal(((float)(t % 6u) + 0.5f) / 6.0f, ((float)(t / 6u) + 0.5f) / 6.0f);
        float nominal = 0.0f;
        for (uint32_t k = 0; k < D; ++k) {
            float c = smax(dot3(n, dirs[k]), 0.0f);
            ctab[k * 36u + t] = c;
            nominal = nominal + c;
        }
        nom[t] = nominal;
    }
    __syncthreads();
}

__device__ void pack_probe_table(const GiArgs& a, uint32_t probe, const float* ctab, const float* nom, float4* smp) {
    const lux_gi_params& p = a.p;
    const uint32_t D = p.base_phi * p.base_y;
    const uint32_t lane = threadIdx.x & 63u;
    const float4* m = a.merged + (uint64_t)probe * D;
    for (uint32_t k = lane; k < D; k += 64u) {
        float4 v = m[k];
        bool ok = v.w > 0.0f && lux::isfinite_bits(v.x) && lux::isfinite_bits(v.y) &&
                  lux::isfinite_bits(v.z) && lux::isfinite_bits(v.w);
        smp[k] = ok ? v : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < 36u) {
        const float* c = ctab + lane;
        float3 sum = make_float3(0.0f, 0.0f, 0.0f);
        float observed = 0.0f;
        for (uint32_t k = 0; k < D; ++k) {
            const float ck = c[k * 36u];
            if (ck <= 0.0f) continue;
            float4 v = smp[k];
            if (!(v.w > 0.0f)) continue;
            float w = ck * sat(v.w);
            sum = make_float3(sum.x + v.x * w, sum.y + v.y * w, sum.z + v.z * w);
            observed = observed + w;
        }
        const float nominal = nom[lane];
        float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (nominal > 0.0f && observed > 0.0f) {
            float sc = nominal * (4.0f / (float)D);
            out = make_float4((sum.x / observed) * sc, (sum.y / observed) * sc, (sum.z / observed) * sc,
                              sat(observed / nominal));
        }
        a.probes[36ull * probe + lane] = out;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// splitLodBlendForParameters (shaders/common.slang:431-461).  floor(log2(x))
// of x >= 1 comes from the exponent field (exact); exp2 of an integer is a
// power of two; max(lodScale, 1) maps NaN to 1 (IEEE maxNum).
__device__ void lod_select(const lux_lod_view& v, const float* pos, uint32_t& low, uint32_t& high,
                           float& blend) {
    const float* pp = v.clip_row3;
    bool perspective = lux::isfinite_bits(pp[0]) && lux::isfinite_bits(pp[1]) && lux::isfinite_bits(pp[2]) &&
                       ((pp[0] * pp[0] + pp[1] * pp[1]) + pp[2] * pp[2]) > 1e-12f;
    const float* vp = v.clip_row1;
    float focal = 0.5f * (float)v.viewport_height * sqrtf((vp[0] * vp[0] + vp[1] * vp[1]) + vp[2] * vp[2]);
    uint32_t target = v.target_pixels > 1u ? v.target_pixels : 1u;
    float ts = v.probe_spacing;
    float scale;
    if (perspective) {
        float ox = fabsf(pos[0] - v.camera_position[0]);
        float oy = fabsf(pos[1] - v.camera_position[1]);
        float oz = fabsf(pos[2] - v.camera_position[2]);
        float m = oz > oy ? oz : oy;
        float radius = m > ox ? m : ox;
        float bd = focal * ts / (float)target;
        float base = ts > bd ? ts : bd;
        scale = radius / base;
    } else {
        float d = focal * ts;
        float den = 1e-20f > d ? 1e-20f : d;
        scale = (float)target / den;
    }
    float m = scale >= 1.0f ? scale : 1.0f;
    uint32_t e = ((__float_as_uint(m) >> 23) & 0xffu) - 127u;
    uint32_t lo = e < 24u ? e : 24u;
    low = lo;
    high = lo;
    blend = 0.0f;
    if (lo < 24u) {
        float next = pow2i((int)lo + 1);
        float start = 0.9f * next;
        if (scale >= start && scale < next) {
            high = lo + 1u;
            blend = sat((scale - start) / (next - start));
        }
    }
}

// Eligible receiver (finite position and normal, |n|^2 > 0.01), unit normal.
__device__ __forceinline__ bool rx_surface(const float* r, float3& pos, float3& n) {
    float x = r[3], y = r[4], z = r[5];
    float nn = (x * x + y * y) + z * z;
    if (!(lux::isfinite_bits(r[0]) && lux::isfinite_bits(r[1]) && lux::isfinite_bits(r[2]))) return false;
    if (!(lux::isfinite_bits(x) && lux::isfinite_bits(y) && lux::isfinite_bits(z)) || !(nn > 0.01f)) return false;
    float nl = sqrtf(nn);
    pos = make_float3(r[0], r[1], r[2]);
    n = make_float3(x / nl, y / nl, z / nl);
    return true;
}

// The ray of a selected (variant, receiver) j = variant * count + receiver
// (traceTransport, split.slang:1135-1150): uv = frac(0.5 + r2(sequence + 1) +
// jitter), jitter = r2(transportHash(frame ^ 0x5a17c9e3)), the decoded
// direction turned into the receiver's hemisphere, origin rayOffset from the
// surface.  False when the receiver has no surface.
__device__ __forceinline__ bool rx_ray(const float* recv, uint32_t count, uint32_t j, uint32_t sq, uint32_t frame,
                                       float bias, float3& pos, float3& o, float3& d) {
    const uint32_t smp = j < count ? j : j - count;
    float3 n;
    if (!rx_surface(recv + 6ull * smp, pos, n)) return false;
    const float2 jit = r2(transport_hash(frame ^ 0x5a17c9e3u));
    const float2 q = r2(sq + 1u);
    const float ux = fracf((0.5f + q.x) + jit.x), uy = fracf((0.5f + q.y) + jit.y);
    float c, s;
    cos_sin_2pi(ux, c, s);
    const float y = uy * 2.0f - 1.0f;
    const float qq = 1.0f - y * y;
    const float r = sqrtf(qq > 0.0f ? qq : 0.0f);
    d = make_float3(r * c, y, r * s);
    if (dot3(d, n) < 0.0f) d = make_float3(-d.x, -d.y, -d.z);
    o = ray_offset(pos, n, bias);
    return true;
}

// Occlusion rays of the NEE setup: each setup wave writes its rays to its
// own slab (wave_id * 64 * nee rays, ranked by ballots, no atomics) and its
// count; an exclusive scan over the counts and k_shadow_compact pack them
// into one list in item order for the occlusion walks.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

__global__ void k_fill_u32_one(uint32_t* p, uint32_t v) { *p = v; }

// One wave per setup wave: its rays [0, count) from the slab to the list.
// The slab and the list are distinct buffers (restrict): the copies of one
// wave's rays need not wait for each other's stores.
__global__ __launch_bounds__(256) void k_shadow_compact(const float4* __restrict__ slab,
                                                        const uint32_t* __restrict__ off, uint32_t n_waves,
                                                        uint32_t slab_rays, float4* __restrict__ list) {
    const uint32_t gw = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (gw >= n_waves) return;
    const uint32_t base = off[gw], cnt = off[gw + 1] - base;
    const float4* __restrict__ src = slab + 2ull * gw * slab_rays;
#pragma unroll 4
    for (uint32_t j = __lane_id(); j < cnt; j += 64u) {
        list[2ull * (base + j)] = src[2ull * j];
        list[2ull * (base + j) + 1] = src[2ull * j + 1];
    }
}

// Kernel arguments of a configured context (host).
inline GiArgs gi_args(lux_ctx* ctx) {
    GiArgs a;
    a.scene = ctx->dev_scene();
    a.lights = static_cast<const Light*>(ctx->gi_light_buf.ptr);
    a.n_lights = ctx->gi_lights;
    a.mats = static_cast<const lux_material*>(ctx->materials.ptr);
    a.n_mats = ctx->n_materials;
    a.p = ctx->gi;
    for (int i = 0; i < 17; ++i) a.offset[i] = ctx->gi_offset[i];
    a.intervals = static_cast<float4*>(ctx->gi_intervals.ptr);
    a.merged = static_cast<float4*>(ctx->gi_merged.ptr);
    a.probes = static_cast<float4*>(ctx->gi_probes.ptr);
    a.counters = static_cast<uint32_t*>(ctx->gi_counters.ptr);
    a.admit_limit = ctx->gi.query_limit == 0u ? 0u : (ctx->gi.query_limit / 2u > 1u ? ctx->gi.query_limit / 2u : 1u);
    const bool hist = ctx->gi_hist_on;
    a.samples = hist ? static_cast<uint32_t*>(ctx->gi_samples.ptr) : nullptr;
    a.history = hist ? static_cast<float4*>(ctx->gi_history.ptr) : nullptr;
    a.hist_state = hist ? static_cast<uint32_t*>(ctx->gi_hist_state.ptr) : nullptr;
    a.dep = hist ? static_cast<float4*>(ctx->gi_dep.ptr) : nullptr;
    a.hist_rb = ctx->gi_hist.radiance_blend;
    a.hist_tb = ctx->gi_hist.transmittance_blend;
    a.hist_frames = ctx->gi_hist.history_frames;
    a.hist_persistent = ctx->gi_hist.persistent;
    return a;
}

}  // namespace

// lux_transport.hip -- dense-grid interval trace, cascade merge, probe
// packing and world resolve on gfx950.
//
// One GI update = three kernels on the context stream:
//   k_gi_trace   one thread per interval (probe, direction, level), all levels
//                in one launch: ray from the probe centre over the interval,
//                closest hit, hit shading (emission + emissive NEE with
//                occlusion rays), deposit J/beta -- trace + deposit + reduce of
//                the reference (shaders/split.slang:1056-1163,1386-1476;
//                shaders/cascades.slang:60-268)
//   k_gi_merge   per level, top down: support-aware Eq. 7 merge with 4 child
//                directions x 8 trilinear corners (split.slang:389-479,1523-1725)
//   k_gi_pack    one wave per level-0 probe: directions staged in LDS, 36
//                octahedral texels (split.slang:1788-1918)
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
#include "lux_internal.h"

namespace {

constexpr int kWg = 256;
constexpr int kLdsDepth = 16;
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
};

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
    float4 A = s.triA[h.tri], B = s.triB[h.tri], C = s.triC[h.tri];
    uint32_t mi = __float_as_uint(B.w);
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
                      float3& out) {
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
                                 Stack& stk) {
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
        if (h.tri != lux::kEmpty) {
            if (!shade(a, o, d, h, stk, rad)) continue;
            beta = 0.0f;
        } else {
            beta = terminal ? 0.0f : 1.0f;
        }
        jsum = make_float3(jsum.x + rad.x, jsum.y + rad.y, jsum.z + rad.z);
        bsum = bsum + beta;
        n++;
    }
    if (n == 0) return make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    float fn = (float)n;
    return make_float4(jsum.x / fn, jsum.y / fn, jsum.z / fn, bsum / fn);
}

// Intervals with global index in [g_begin, g_end) (levels are contiguous).
__global__ __launch_bounds__(kWg) void k_gi_trace(GiArgs a, uint64_t g_begin, uint64_t g_end) {
    __shared__ uint2 lds[kLdsDepth][kWg];
    Stack stk;
    stk.col = &lds[0][threadIdx.x];
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
    a.intervals[g] = trace_interval(a, L, c, k, terminal, stk);
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
    stk.col = &lds[0][threadIdx.x];
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
    stk.col = &lds[0][threadIdx.x];
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
//                 contribution goes to a fixed slot, its occlusion ray to a
//                 compacted list (wave-aggregated append)
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
constexpr uint32_t kWaveRange = 0u, kWaveListed = 1u, kWaveSampled = 2u;

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
    uint4* prim;                 // per item: t bits, tri, inst, err
    float4* shade;               // per item: emission rgb, w = 1 when NEE runs
    float4* slots;               // per slot: contribution rgb, w = 1 when a query was issued
    uint32_t* vis;               // per slot: 0 visible, 1 occluded, 2 unsupported
    float4* shadow;              // per occlusion ray: (origin, limit), (direction, slot bits)
    uint32_t* counters;          // [0] primary cursor, [1] shadow rays, [2] shadow cursor
    uint32_t regroup, batch;
};

// Intervals of the chunk (kWaveListed: listed probes beyond the list's
// device-side length contribute none).
__device__ __forceinline__ uint32_t wave_intervals(const GiArgs& a, const Wave& w) {
    if (w.kind == kWaveRange) return (uint32_t)(w.g1 - w.g0);
    if (w.kind == kWaveSampled) return w.n;
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

__device__ __forceinline__ void wave_item_ray(const GiArgs& a, const Wave& w, uint32_t i, float3& o, float3& d,
                                              float& lim, bool& terminal) {
    bool force;
    uint64_t g = wave_interval(a, w, i / w.R, force);
    item_ray(a, g, i % w.R, force, o, d, lim, terminal);
}

struct PrimarySource {
    const GiArgs* a;
    const Wave* w;
    uint32_t n;
    __device__ uint32_t count() const { return n; }
    template <class St>
    __device__ bool begin(uint32_t i, lux::Walk& wk, St& st) {
        float3 o, d;
        float lim;
        bool term;
        wave_item_ray(*a, *w, i, o, d, lim, term);
        if (!admit_query(*a)) {  // denied: no sample, as an unsupported traversal
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

template <int Depth>
__global__ __launch_bounds__(kWg) void k_gi_primary(GiArgs a, Wave w) {
    __shared__ uint2 lds[Depth][kWg];
    lux::LaneStack<Depth, kWg> stk;
    stk.col = &lds[0][threadIdx.x];
    lux::NoStats st;
    PrimarySource src{&a, &w, wave_intervals(a, w) * w.R};
    lux::regroup_walks(a.scene, src, &w.counters[0], w.regroup, w.batch, stk, st);
}

__global__ __launch_bounds__(kWg) void k_gi_setup(GiArgs a, Wave w) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= wave_intervals(a, w) * w.R) return;
    uint4 pr = w.prim[i];
    float4 sh = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
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
            for (uint32_t smp = 0; smp < c.count; ++smp) {
                float3 origin = make_float3(0.0f, 0.0f, 0.0f), dir = origin, t3 = origin;
                float sl = 0.0f;
                bool q = nee_sample(a, c, smp, origin, dir, sl, t3);
                uint32_t slot = i * w.nee + smp;
                w.slots[slot] = make_float4(t3.x, t3.y, t3.z, q ? 1.0f : 0.0f);
                uint64_t mask = __ballot(q);
                if (q) {
                    uint32_t leader = (uint32_t)(__ffsll((unsigned long long)mask) - 1);
                    uint32_t base = 0;
                    if (lane == leader) base = atomicAdd(&w.counters[1], (uint32_t)__popcll(mask));
                    base = __shfl(base, (int)leader, 64);
                    uint32_t at = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
                    w.shadow[2ull * at] = make_float4(origin.x, origin.y, origin.z, sl);
                    w.shadow[2ull * at + 1] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(slot));
                }
            }
        }
    }
    w.shade[i] = sh;
}

struct ShadowSource {
    const GiArgs* a;
    const Wave* w;
    uint32_t n;
    __device__ uint32_t count() const { return n; }
    template <class St>
    __device__ bool begin(uint32_t i, lux::Walk& wk, St& st) {
        float4 r0 = w->shadow[2ull * i], r1 = w->shadow[2ull * i + 1];
        if (!admit_query(*a)) {
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

template <int Depth>
__global__ __launch_bounds__(kWg) void k_gi_shadow(GiArgs a, Wave w) {
    __shared__ uint2 lds[Depth][kWg];
    lux::LaneStack<Depth, kWg> stk;
    stk.col = &lds[0][threadIdx.x];
    lux::NoStats st;
    ShadowSource src{&a, &w, w.counters[1]};
    lux::regroup_walks(a.scene, src, &w.counters[2], w.regroup, w.batch, stk, st);
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
    for (uint32_t r = 0; r < w.R; ++r) {
        uint32_t i = j * w.R + r;
        uint4 pr = w.prim[i];
        if (pr.w & lux::kErrUnsupported) continue;
        float3 rad = make_float3(0.0f, 0.0f, 0.0f);
        float beta;
        if (pr.y != lux::kEmpty) {
            float4 sh = w.shade[i];
            float3 rr = make_float3(sh.x, sh.y, sh.z);
            if (sh.w != 0.0f) {
                float3 sum = make_float3(0.0f, 0.0f, 0.0f);
                bool ok = true;
                for (uint32_t smp = 0; smp < count; ++smp) {
                    uint32_t slot = i * w.nee + smp;
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

__global__ __launch_bounds__(kWg) void k_gi_merge(GiArgs a, uint32_t L) {
    const lux_gi_params& p = a.p;
    uint64_t local = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    uint64_t count = a.offset[L + 1] - a.offset[L];
    if (local >= count) return;
    uint32_t dx = level_dim(p.dims[0], L), dy = level_dim(p.dims[1], L);
    uint32_t nphi = p.base_phi << L;
    uint32_t D = nphi * (p.base_y << L);
    uint32_t k = (uint32_t)(local % D);
    uint64_t probe = local / D;
    uint32_t ix = (uint32_t)(probe % dx), iy = (uint32_t)((probe / dx) % dy);
    uint32_t iz = (uint32_t)(probe / ((uint64_t)dx * dy));
    float sl = p.spacing * pow2i((int)L);
    float3 rel = make_float3(((float)ix + 0.5f) * sl, ((float)iy + 0.5f) * sl, ((float)iz + 0.5f) * sl);
    float4 iv = a.intervals[a.offset[L] + local];
    float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (iv.w >= 0.0f) {
        float tr = sat(iv.w);
        float support = 1.0f - tr;
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
    a.merged[a.offset[L] + local] = out;
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

// One 64-lane wave per level-0 probe (packProbes, split.slang:1788-1918).
__global__ __launch_bounds__(64) void k_gi_pack(GiArgs a) {
    __shared__ float4 smp[kMaxPackDirections];
    __shared__ float3 dirs[kMaxPackDirections];
    const lux_gi_params& p = a.p;
    uint32_t probe = blockIdx.x;
    uint32_t D = p.base_phi * p.base_y;
    const float4* m = a.merged + (uint64_t)probe * D;
    for (uint32_t k = threadIdx.x; k < D; k += 64u) {
        float4 v = m[k];
        bool ok = v.w > 0.0f && lux::isfinite_bits(v.x) && lux::isfinite_bits(v.y) &&
                  lux::isfinite_bits(v.z) && lux::isfinite_bits(v.w);
        smp[k] = ok ? v : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        dirs[k] = bin_direction(k, p.base_phi, p.base_y, 0.5f, 0.5f);
    }
    __syncthreads();
    uint32_t t = threadIdx.x;
    if (t >= 36u) return;
    float3 n = decode_normal(((float)(t % 6u) + 0.5f) / 6.0f, ((float)(t / 6u) + 0.5f) / 6.0f);
    float3 sum = make_float3(0.0f, 0.0f, 0.0f);
    float nominal = 0.0f, observed = 0.0f;
    for (uint32_t k = 0; k < D; ++k) {
        float c = smax(dot3(n, dirs[k]), 0.0f);
        nominal = nominal + c;
        if (c <= 0.0f) continue;
        float4 v = smp[k];
        if (!(v.w > 0.0f)) continue;
        float w = c * sat(v.w);
        sum = make_float3(sum.x + v.x * w, sum.y + v.y * w, sum.z + v.z * w);
        observed = observed + w;
    }
    float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (nominal > 0.0f && observed > 0.0f) {
        float sc = nominal * (4.0f / (float)D);
        out = make_float4((sum.x / observed) * sc, (sum.y / observed) * sc, (sum.z / observed) * sc,
                          sat(observed / nominal));
    }
    a.probes[36ull * probe + t] = out;
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

// worldResolve at LOD 0 (world_resolve.slang:1-196,281-313).
__global__ __launch_bounds__(kWg) void k_world_resolve(GiArgs a, const float* recv, uint32_t n,
                                                       float4* out) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const lux_gi_params& p = a.p;
    const float* r = recv + 6ull * i;
    float nx = r[3], ny = r[4], nz = r[5];
    float nn = (nx * nx + ny * ny) + nz * nz;
    if (!(lux::isfinite_bits(nx) && lux::isfinite_bits(ny) && lux::isfinite_bits(nz)) || !(nn > 0.01f)) {
        out[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    float nl = sqrtf(nn);
    nx = nx / nl;
    ny = ny / nl;
    nz = nz / nl;
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
    float bw[4] = {cx * cy, fx * cy, cx * fy, fx * fy};
    int32_t bxi = (int32_t)flx, byi = (int32_t)fly;
    uint32_t tex[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int32_t tx = bxi + (c & 1), ty = byi + (c >> 1);
        wrap_texel(tx, ty);
        tex[c] = (uint32_t)(ty * 6 + tx);
    }
    uint32_t dx = p.dims[0], dy = p.dims[1], dz = p.dims[2];
    float rel[3] = {r[0] - p.origin[0], r[1] - p.origin[1], r[2] - p.origin[2]};
    float fr[3];
    int32_t base[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        float g = rel[q] / p.spacing - 0.5f;
        float fl = floorf(g);
        base[q] = (int32_t)fl;
        fr[q] = g - fl;
    }
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
            float4 smp = a.probes[36ull * probe + tex[c]];
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
    float cov = 0.0f;
    float3 val = make_float3(0.0f, 0.0f, 0.0f);
    if (potential > 0.0f) {
        cov = sat(observed / potential);
        if (observed > 0.0f) val = make_float3(sum.x / observed, sum.y / observed, sum.z / observed);
    }
    float lw = cov * 1.0f, hw = 1.0f * 0.0f, tw = lw + hw;
    float coverage = sat(tw);
    float3 res = make_float3(0.0f, 0.0f, 0.0f);
    if (tw > 0.0f)
        res = make_float3((val.x * lw + 0.0f * hw) / tw, (val.y * lw + 0.0f * hw) / tw,
                          (val.z * lw + 0.0f * hw) / tw);
    bool supported = coverage > 0.0f && lux::isfinite_bits(res.x) && lux::isfinite_bits(res.y) &&
                     lux::isfinite_bits(res.z);
    out[i] = make_float4(res.x, res.y, res.z, supported ? 1.0f + 0.25f * coverage : -1.0f);
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
    stk.col = &lds[0][threadIdx.x];
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
    lux::trace_closest(scene, o, d, 1e20f, hs, stk, ns, false);
    float* r = out + 6ull * i;
    if (hs.tri == lux::kEmpty || (hs.err & lux::kErrUnsupported)) {
        r[0] = o.x; r[1] = o.y; r[2] = o.z;
        r[3] = 0.0f; r[4] = 0.0f; r[5] = 0.0f;
        return;
    }
    const float4* rec = scene.inst + 8u * hs.inst;
    uint32_t iflags = __float_as_uint(rec[6].x);
    float4 A = scene.triA[hs.tri], B = scene.triB[hs.tri], C = scene.triC[hs.tri];
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
std::vector<Light> build_lights(const lux_ctx* ctx) {
    std::vector<Light> out;
    std::vector<double> w;
    uint32_t nm = (uint32_t)ctx->host_materials.size();
    if (nm == 0) return out;
    for (uint32_t k = 0; k < (uint32_t)ctx->instances.size(); ++k) {
        const lux_instance_desc& in = ctx->instances[k];
        if (in.mesh >= ctx->host_pos.size()) continue;
        const float* f = in.world_from_local;
        uint32_t cull = in.flags & 3u;
        const std::vector<float>& pos = ctx->host_pos[in.mesh];
        const std::vector<uint32_t>& mat = ctx->host_mat[in.mesh];
        uint32_t nt = (uint32_t)(pos.size() / 9);
        for (uint32_t t = 0; t < nt; ++t) {
            uint32_t mi = mat[t] < nm ? mat[t] : nm - 1u;
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

GiArgs gi_args(lux_ctx* ctx) {
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
    return a;
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
    return LUX_OK;
}

extern "C" int lux_gi_set_frame(lux_ctx* ctx, uint32_t frame, uint32_t seed) {
    if (!ctx) return LUX_ERR_INVALID;
    ctx->gi.frame = frame;
    ctx->gi.seed = seed;
    return LUX_OK;
}

namespace {
// Runs the four wavefront stages over a set of intervals in chunks whose
// scratch stays under about 2 GiB.  `proto` carries the kind and its inputs;
// `total` is the number of intervals (kWaveRange), listed probes
// (kWaveListed, an upper bound: the device-side list length decides) or
// samples (kWaveSampled); `per_unit` intervals per unit (D for listed
// probes, else 1).
int run_wave(lux_ctx* ctx, Wave proto, uint64_t total, uint64_t per_unit) {
    const lux_gi_params& p = ctx->gi;
    const uint32_t R = p.rays_per_interval ? p.rays_per_interval : 1u;
    const uint32_t nee = p.emissive_samples / 2u > 1u ? p.emissive_samples / 2u : 1u;
    const uint64_t per_item = 32ull + (uint64_t)nee * (16ull + 4ull + 32ull);
    uint64_t chunk_items = (2ull << 30) / per_item;
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
    if (int rc = lux_ensure(ctx, ctx->gi_w_shadow, 32ull * items * nee)) return rc;
    if (int rc = lux_ensure(ctx, ctx->gi_w_counters, 16)) return rc;
    GiArgs a = gi_args(ctx);
    Wave w = proto;
    w.R = R;
    w.nee = nee;
    w.prim = static_cast<uint4*>(ctx->gi_w_prim.ptr);
    w.shade = static_cast<float4*>(ctx->gi_w_shade.ptr);
    w.slots = static_cast<float4*>(ctx->gi_w_slots.ptr);
    w.vis = static_cast<uint32_t*>(ctx->gi_w_vis.ptr);
    w.shadow = static_cast<float4*>(ctx->gi_w_shadow.ptr);
    w.counters = static_cast<uint32_t*>(ctx->gi_w_counters.ptr);
    w.regroup = ctx->gi_regroup ? ctx->gi_regroup : 8u;
    w.batch = ctx->gi_batch ? ctx->gi_batch : 8u;
    const uint32_t wpc = ctx->gi_wpc ? ctx->gi_wpc : 20u;
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
        switch (ctx->gi_depth) {
        case 4: hipLaunchKernelGGL(k_gi_primary<4>, dim3(pb), dim3(kWg), 0, st, a, w); break;
        case 16: hipLaunchKernelGGL(k_gi_primary<16>, dim3(pb), dim3(kWg), 0, st, a, w); break;
        default: hipLaunchKernelGGL(k_gi_primary<8>, dim3(pb), dim3(kWg), 0, st, a, w); break;
        }
        hipLaunchKernelGGL(k_gi_setup, dim3(need), dim3(kWg), 0, st, a, w);
        switch (ctx->gi_depth) {
        case 4: hipLaunchKernelGGL(k_gi_shadow<4>, dim3(pblocks), dim3(kWg), 0, st, a, w); break;
        case 16: hipLaunchKernelGGL(k_gi_shadow<16>, dim3(pblocks), dim3(kWg), 0, st, a, w); break;
        default: hipLaunchKernelGGL(k_gi_shadow<8>, dim3(pblocks), dim3(kWg), 0, st, a, w); break;
        }
        hipLaunchKernelGGL(k_gi_accum, dim3((max_intervals + kWg - 1) / kWg), dim3(kWg), 0, st, a, w);
        LUX_HIP(ctx, hipGetLastError());
    }
    return LUX_OK;
}

Wave wave_proto(uint32_t kind) {
    Wave w;
    std::memset(&w, 0, sizeof w);
    w.kind = kind;
    return w;
}

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
        cfg->lds_stack_depth != 16)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_set_trace_config: LDS stack depth must be 4, 8 or 16");
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
    static const uint32_t cand_wpc[] = {16, 20, 24, 32};
    static const uint32_t cand_depth[] = {4, 8, 16};
    static const uint32_t cand_regroup[] = {4, 8, 16, 32, 48};
    static const uint32_t cand_batch[] = {4, 8, 16};
    const uint32_t* cands[4] = {cand_wpc, cand_depth, cand_regroup, cand_batch};
    const uint32_t ncand[4] = {4, 3, 5, 3};
    uint32_t cur[4] = {20, 8, 8, 8};
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
                if (rc == LUX_OK && ms < cur_ms) {
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
    if (g1 == g0) return LUX_OK;
    if (ctx->gi_trace_mode == 0u) return trace_wavefront(ctx, g0, g1);
    GiArgs a = gi_args(ctx);
    uint32_t blocks = (uint32_t)((g1 - g0 + kWg - 1) / kWg);
    hipLaunchKernelGGL(k_gi_trace, dim3(blocks), dim3(kWg), 0, ctx->stream, a, g0, g1);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_gi_trace(lux_ctx* ctx) {
    if (!ctx) return LUX_ERR_INVALID;
    if (int rc = gi_check(ctx)) return rc;
    return lux_gi_trace_levels(ctx, 0, ctx->gi.levels - 1);
}

namespace {
int merge_levels(lux_ctx* ctx, int top) {
    GiArgs a = gi_args(ctx);
    for (int L = top; L >= 0; --L) {
        uint64_t count = ctx->gi_offset[L + 1] - ctx->gi_offset[L];
        uint32_t blocks = (uint32_t)((count + kWg - 1) / kWg);
        hipLaunchKernelGGL(k_gi_merge, dim3(blocks), dim3(kWg), 0, ctx->stream, a, (uint32_t)L);
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
    hipLaunchKernelGGL(k_gi_pack, dim3(ctx->gi_probes0), dim3(64), 0, ctx->stream, a);
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

extern "C" int lux_gi_update(lux_ctx* ctx, float* ms) {
    if (!ctx) return LUX_ERR_INVALID;
    if (int rc = gi_check(ctx)) return rc;
    LUX_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (int rc = lux_device_zero(ctx, ctx->gi_counters.ptr, 8)) return rc;
    if (int rc = lux_gi_trace(ctx)) return rc;
    if (int rc = lux_gi_merge(ctx)) return rc;
    if (int rc = lux_gi_pack(ctx)) return rc;
    LUX_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (ms) {
        LUX_HIP(ctx, hipEventSynchronize(ctx->ev1));
        LUX_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    }
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

// lux_device.h -- device-side geometry and two-level BVH traversal for gfx950.
//
// Semantics follow the reference's software traversal operation for
// operation (shaders/traversal.slang:349-443 box/triangle tests, :474-566
// traceInstance, :696-811 traceRayTracked, :826-1046 occlusion).  All
// arithmetic is IEEE f32 with contraction disabled at compile time
// (-ffp-contract=off) and correctly rounded division, so results are
// bit-identical to a scalar CPU evaluation of the same expressions.
//
// Memory layout in HBM (structure of arrays, 16-byte loads):
//   nodes      float4 pair per node: (min.xyz, first) (max.xyz, count)
//   triangles  three float4 arrays A, B, C in BLAS leaf order;
//              A.w = triangle index within its mesh, B.w = material index
//   instances  8 float4 per instance in TLAS order: inv rows[3], fwd rows[3],
//              (flags, blas_root, tri_first, tri_count), (orig, mesh, 0, 0)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lux {

constexpr uint32_t kEmpty = 0xffffffffu;
constexpr uint32_t kMaxSteps = 4096u;
constexpr uint32_t kMaxLeafTriangles = 256u;
constexpr uint32_t kMaxLeafInstances = 4096u;
constexpr uint32_t kStackCap = 49u;
constexpr uint32_t kErrBlasOverflow = 1u;
constexpr uint32_t kErrTlasOverflow = 2u;
constexpr uint32_t kErrUnsupported = 4u;
constexpr uint32_t kHitOccluded = 0x100u;
constexpr uint32_t kHitRoot = 0x200u;
constexpr uint32_t kRayAnyHit = 1u;    // traceOccluded semantics
constexpr uint32_t kRayFirstHit = 2u;  // traceRay(anyHit = true) semantics

struct DevScene {
    const float4* tlas;
    const float4* inst;
    const float4* blas;
    const float4* tri;  // 3 per triangle, interleaved: A (w = index in mesh), B (w = material), C
    uint32_t n_tlas, n_inst, n_blas, n_tri;
    // [0] every BLAS node bound, [1] every TLAS node bound is zero or of
    // magnitude in [2^-64, 2^39) (written after each build; null: unknown).
    // With such bounds the resumable walk divides by reciprocals (div_fast).
    const uint32_t* range_ok;
};

struct Node {
    float3 mn;
    uint32_t first;
    float3 mx;
    uint32_t count;
};

__device__ __forceinline__ Node load_node(const float4* base, uint32_t n, uint32_t idx) {
    // Clamped read: a malformed index reads the last node instead of faulting
    // (the reference clamps scene word reads, traversal.slang:114-118).
    uint32_t i = idx < n ? idx : n - 1u;
    float4 a = base[2u * i];
    float4 b = base[2u * i + 1u];
    Node r;
    r.mn = make_float3(a.x, a.y, a.z);
    r.first = __float_as_uint(a.w);
    r.mx = make_float3(b.x, b.y, b.z);
    r.count = __float_as_uint(b.w);
    return r;
}

__device__ __forceinline__ bool isfinite_bits(float f) {
    return (__float_as_uint(f) & 0x7f800000u) != 0x7f800000u;
}

// boxRayInterval (traversal.slang:349-369).
__device__ __forceinline__ bool box_interval(float3 o, float3 d, float3 mn, float3 mx, float limit,
                                             float& entry, float& exit_) {
    float near_ = 0.0f, far_ = limit;
    const float os[3] = {o.x, o.y, o.z}, ds[3] = {d.x, d.y, d.z};
    const float lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
    bool ok = true;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        if (ds[axis] == 0.0f) {
            if (os[axis] < lo[axis] || os[axis] > hi[axis]) ok = false;
        } else {
            float a = (lo[axis] - os[axis]) / ds[axis];
            float b = (hi[axis] - os[axis]) / ds[axis];
            float ax = fmaxf(a, b);
            near_ = fmaxf(near_, fminf(a, b));
            float pad = fabsf(ax) * 3.5762791e-7f;
            far_ = fminf(far_, ax + pad);
        }
    }
    if (!ok) {
        entry = 0.0f;
        exit_ = 0.0f;
        return false;
    }
    entry = near_;
    exit_ = far_;
    return near_ <= far_;
}

__device__ __forceinline__ float box_entry(float3 o, float3 d, const Node& n, float limit) {
    float e, x;
    return box_interval(o, d, n.mn, n.mx, limit, e, x) ? e : -1.0f;
}

// ---- the slab test's divisions from a per-ray reciprocal ----
// For y = RN(1/d), q0 = RN(n*y) lies within one ulp of n/d, the residual
// n - d*q0 is exact, and RN(q0 + r*y) = RN(n/d) (Markstein's theorem for
// round-to-nearest), provided nothing underflows or overflows.  That holds
// when the coordinates (node bounds, ray origin) are zero or of magnitude in
// [2^-64, 2^39) and every direction component has magnitude in
// [2^-60, 2^20): then n = bound - origin is zero or in [2^-87, 2^40), the
// quotient in [2^-107, 2^100) and the residual on a grid no finer than
// 2^-133.  For n = +-0 the quotient's zero sign may differ from the IEEE
// quotient's; the slab test only compares such values, so its decisions,
// entry distances' order and the hits are the same.
__device__ __forceinline__ bool coord_in_range(float v) {
    const uint32_t m = __float_as_uint(v) & 0x7fffffffu, e = m >> 23;
    return m == 0u || (e >= 127u - 64u && e < 127u + 39u);
}
__device__ __forceinline__ bool dir_in_range(float v) {
    const uint32_t e = (__float_as_uint(v) >> 23) & 255u;
    return e >= 127u - 60u && e < 127u + 20u;
}
__device__ __forceinline__ bool ray_in_range(float3 o, float3 d) {
    return coord_in_range(o.x) && coord_in_range(o.y) && coord_in_range(o.z) && dir_in_range(d.x) &&
           dir_in_range(d.y) && dir_in_range(d.z);
}
__device__ __forceinline__ float div_fast(float n, float d, float y) {
    const float q0 = n * y;
    const float r = __builtin_fmaf(-q0, d, n);
    return __builtin_fmaf(r, y, q0);
}
__device__ __forceinline__ float3 reciprocal(float3 d) {
    return make_float3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
}

// box_entry for a ray with nonzero, in-range direction components (the zero
// branch of boxRayInterval cannot occur) and in-range coordinates; y = 1/d.
__device__ __forceinline__ float box_entry_fast(float3 o, float3 d, float3 y, const Node& n, float limit) {
    float near_ = 0.0f, far_ = limit;
    const float os[3] = {o.x, o.y, o.z}, ds[3] = {d.x, d.y, d.z}, ys[3] = {y.x, y.y, y.z};
    const float lo[3] = {n.mn.x, n.mn.y, n.mn.z}, hi[3] = {n.mx.x, n.mx.y, n.mx.z};
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        float a = div_fast(lo[axis] - os[axis], ds[axis], ys[axis]);
        float b = div_fast(hi[axis] - os[axis], ds[axis], ys[axis]);
        float ax = fmaxf(a, b);
        near_ = fmaxf(near_, fminf(a, b));
        float pad = fabsf(ax) * 3.5762791e-7f;
        far_ = fminf(far_, ax + pad);
    }
    return near_ <= far_ ? near_ : -1.0f;
}

// Whether the resumable and nested walks may divide by reciprocals: every
// BLAS and TLAS node bound is in div_fast's range (range_ok, written after
// each build; null: unknown).
__device__ __forceinline__ bool scene_in_range(const DevScene& s) {
    return s.range_ok != nullptr && s.range_ok[0] != 0u && s.range_ok[1] != 0u;
}

// orientedEdge (traversal.slang:379-392).  The reference orders the two
// points (the larger x, then the larger y, second), takes the edge of the
// ordered pair and negates it when it swapped them.  The swapped products
// are the same products (RN(b.x a.y) = RN(a.y b.x)), and RN(-v) = -RN(v), so
// the negated edge of the swapped pair is a.x b.y - a.y b.x bit for bit and
// its zeroing test is the same; only a zeroed edge keeps the swap as its
// sign, -0.
__device__ __forceinline__ float oriented_edge(float ax, float ay, float bx, float by) {
    const bool reverse = ax > bx || (ax == bx && ay > by);
    const float pa = ax * by;
    const float pb = ay * bx;
    const float edge = pa - pb;
    const float err = (fabsf(pa) + fabsf(pb)) * 2.3841858e-7f;
    return fabsf(edge) <= err ? (reverse ? -0.0f : 0.0f) : edge;
}

struct TriRay {
    int kx, ky, kz;
    float sx, sy;
    bool valid;
};

__device__ __forceinline__ float comp(float3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// triangleRay (traversal.slang:399-413).  With y = 1/d for a direction in
// div_fast's range (every component nonzero), the shear constants divide by
// the reciprocal of d[kz] (the same quotients).
__device__ __forceinline__ TriRay make_tri_ray(float3 d, const float3* y = nullptr) {
    TriRay r;
    r.kx = r.ky = 0;
    r.sx = r.sy = 0.0f;
    float m0 = fabsf(d.x), m1 = fabsf(d.y), m2 = fabsf(d.z);
    r.kz = m0 > m1 ? (m0 > m2 ? 0 : 2) : (m1 > m2 ? 1 : 2);
    float dz = comp(d, r.kz);
    r.valid = dz != 0.0f;
    if (!r.valid) return r;
    r.kx = (r.kz + 1) % 3;
    r.ky = (r.kx + 1) % 3;
    if (dz < 0.0f) {
        int s = r.kx;
        r.kx = r.ky;
        r.ky = s;
    }
    if (y) {
        const float ydz = comp(*y, r.kz);
        r.sx = div_fast(comp(d, r.kx), dz, ydz);
        r.sy = div_fast(comp(d, r.ky), dz, ydz);
    } else {
        r.sx = comp(d, r.kx) / dz;
        r.sy = comp(d, r.ky) / dz;
    }
    return r;
}

// The slab tests and triangle shear of one ray of the nested traversals:
// by the reciprocals of its direction (div_fast, the same quotients) when the
// scene and the ray are in range, else by IEEE division.
struct SlabRay {
    float3 o, d, y;
    bool fast;
    __device__ __forceinline__ SlabRay(const DevScene& s, float3 o_, float3 d_, bool allow = true)
        : o(o_), d(d_), y(make_float3(0.0f, 0.0f, 0.0f)),
          fast(allow && scene_in_range(s) && ray_in_range(o_, d_)) {
        if (fast) y = reciprocal(d_);
    }
    // boxRayInterval's entry, or -1 when the box is missed
    __device__ __forceinline__ float entry(const Node& n, float limit) const {
        return fast ? box_entry_fast(o, d, y, n, limit) : box_entry(o, d, n, limit);
    }
    __device__ __forceinline__ TriRay tri_ray() const { return make_tri_ray(d, fast ? &y : nullptr); }
};

// intersectTriangle (traversal.slang:415-443).  Returns (t, w1, w2), t = -1 on miss.
__device__ __forceinline__ float3 intersect_tri(float3 o, float3 d, const TriRay& r, float3 ta,
                                                float3 tb, float3 tc) {
    float3 miss = make_float3(-1.0f, 0.0f, 0.0f);
    if (!r.valid) return miss;
    float3 A = make_float3(ta.x - o.x, ta.y - o.y, ta.z - o.z);
    float3 B = make_float3(tb.x - o.x, tb.y - o.y, tb.z - o.z);
    float3 C = make_float3(tc.x - o.x, tc.y - o.y, tc.z - o.z);
    float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    float pax = comp(A, r.kx) - r.sx * Akz, pay = comp(A, r.ky) - r.sy * Akz;
    float pbx = comp(B, r.kx) - r.sx * Bkz, pby = comp(B, r.ky) - r.sy * Bkz;
    float pcx = comp(C, r.kx) - r.sx * Ckz, pcy = comp(C, r.ky) - r.sy * Ckz;
    float largest = fmaxf(fmaxf(fmaxf(fabsf(pax), fabsf(pay)), fmaxf(fabsf(pbx), fabsf(pby))),
                          fmaxf(fabsf(pcx), fabsf(pcy)));
    if (largest == 0.0f || !isfinite_bits(largest)) return miss;
    uint32_t exponent = (__float_as_uint(largest) >> 23) & 255u;
    uint32_t ce = exponent < 1u ? 1u : (exponent > 253u ? 253u : exponent);
    float scale = __uint_as_float((254u - ce) << 23);
    pax *= scale; pay *= scale; pbx *= scale; pby *= scale; pcx *= scale; pcy *= scale;
    float e0 = oriented_edge(pbx, pby, pcx, pcy);
    float e1 = oriented_edge(pcx, pcy, pax, pay);
    float e2 = oriented_edge(pax, pay, pbx, pby);
    bool anyneg = e0 < 0.0f || e1 < 0.0f || e2 < 0.0f;
    bool anypos = e0 > 0.0f || e1 > 0.0f || e2 > 0.0f;
    if (anyneg && anypos) return miss;
    float det = (e0 + e1) + e2;
    if (det == 0.0f || !isfinite_bits(det)) return miss;
    float w0 = e0 / det, w1 = e1 / det, w2 = e2 / det;
    float dist = ((w0 * Akz + w1 * Bkz) + w2 * Ckz) / comp(d, r.kz);
    if (!isfinite_bits(dist)) return miss;
    return make_float3(dist, w1, w2);
}

__device__ __forceinline__ bool culled(float3 a, float3 b, float3 c, float3 d, uint32_t flags) {
    if ((flags & 3u) == 0u) return false;
    float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
    float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
    float cx = e1y * e2z - e1z * e2y;
    float cy = e1z * e2x - e1x * e2z;
    float cz = e1x * e2y - e1y * e2x;
    float facing = ((cx * d.x + cy * d.y) + cz * d.z) * ((flags & 4u) != 0u ? -1.0f : 1.0f);
    return ((flags & 1u) != 0u && facing >= 0.0f) || ((flags & 2u) != 0u && facing <= 0.0f);
}

__device__ __forceinline__ float3 xform_point(float4 r0, float4 r1, float4 r2, float3 p) {
    return make_float3(((r0.x * p.x + r0.y * p.y) + r0.z * p.z) + r0.w * 1.0f,
                       ((r1.x * p.x + r1.y * p.y) + r1.z * p.z) + r1.w * 1.0f,
                       ((r2.x * p.x + r2.y * p.y) + r2.z * p.z) + r2.w * 1.0f);
}
__device__ __forceinline__ float3 xform_dir(float4 r0, float4 r1, float4 r2, float3 v) {
    return make_float3(((r0.x * v.x + r0.y * v.y) + r0.z * v.z) + r0.w * 0.0f,
                       ((r1.x * v.x + r1.y * v.y) + r1.z * v.z) + r1.w * 0.0f,
                       ((r2.x * v.x + r2.y * v.y) + r2.z * v.z) + r2.w * 0.0f);
}

// Traversal work counters: NoStats compiles away; Counting tallies the node
// fetches, triangle tests and instance visits that define the algorithmic
// bytes of a ray (the oracle counts the same events).
struct NoStats {
    __device__ __forceinline__ void node(uint32_t) {}
    __device__ __forceinline__ void tri() {}
    __device__ __forceinline__ void inst() {}
};
struct Counting {
    uint32_t nodes = 0, tris = 0, insts = 0;
    __device__ __forceinline__ void node(uint32_t k) { nodes += k; }
    __device__ __forceinline__ void tri() { tris += 1; }
    __device__ __forceinline__ void inst() { insts += 1; }
};

struct HitState {
    float t;
    uint32_t tri, inst, err;
    float b1, b2;
};

// Waves per SIMD the persistent walk kernels are compiled for: the most of
// 6, 5, 4 whose workgroups (wg lanes, each with `depth` 8-byte LDS stack
// slots, plus `extra_lds` bytes) fit a CU's 160 KB of LDS.  At six the
// register budget is 80 VGPRs (512 per SIMD lane / 6, in granules of 8); the
// walks then spill some state to scratch, and still run faster with the
// sixth wave (traversal 550 against 522 Mrays/s at 92 VGPRs, five waves).
constexpr int walk_min_waves(int depth, int wg, int extra_lds) {
    return 6 * (depth * 8 * wg + extra_lds) <= 160 * 1024   ? 6
           : 5 * (depth * 8 * wg + extra_lds) <= 160 * 1024 ? 5
                                                             : 4;
}

// Per-lane traversal stack: the first LdsDepth slots live in LDS
// (lane-contiguous columns, so a wave's access is one conflict-free row),
// deeper slots spill to a private array.  TLAS entries occupy the bottom of
// the stack and the BLAS walk of the current instance stacks on top of them,
// so the reference's two 49-entry stacks fit in one 98-slot stack.
// Word is the type the column pointer points at.  LdsWord types it as LDS
// (address space 3) and packs a slot into one 64-bit word; FlatWord keeps a
// generic uint2 pointer, for which the compiler merges the LDS and
// private-memory paths into one pointer select and flat accesses.  Both forms
// compute the same; which one runs faster depends on the kernel around it
// (the GI and batch traces gain 3-5 % from LdsWord, the plain ray trace loses
// 4 % at five waves per SIMD and 1.4 % at six; a generic 64-bit word loses it
// 8 %), so each kernel names its own.
typedef __attribute__((address_space(3))) unsigned long long LdsWord;
typedef uint2 FlatWord;
__device__ __forceinline__ void stack_store(FlatWord* p, uint2 v) { *p = v; }
__device__ __forceinline__ uint2 stack_load(const FlatWord* p) { return *p; }
__device__ __forceinline__ void stack_store(LdsWord* p, uint2 v) {
    *p = ((unsigned long long)v.y << 32) | v.x;
}
__device__ __forceinline__ uint2 stack_load(const LdsWord* p) {
    const unsigned long long w = *p;
    return make_uint2((uint32_t)w, (uint32_t)(w >> 32));
}
template <int LdsDepth, int Wg, class Word = LdsWord>
struct LaneStack {
    Word* col;  // &lds[0][lane]
    uint2 spill[2 * kStackCap - LdsDepth > 0 ? 2 * kStackCap - LdsDepth : 1];
    __device__ __forceinline__ void bind(uint2* lds_column) { col = (Word*)lds_column; }
    __device__ __forceinline__ void put(uint32_t slot, uint2 v) {
        if (slot < (uint32_t)LdsDepth) stack_store(col + slot * Wg, v);
        else spill[slot - LdsDepth] = v;
    }
    __device__ __forceinline__ uint2 get(uint32_t slot) {
        if (slot < (uint32_t)LdsDepth) return stack_load(col + slot * Wg);
        return spill[slot - LdsDepth];
    }
};

// traceInstance, closest hit (traversal.slang:474-566).  Stack slots
// [base, base+49) belong to this walk.
template <class Stack, class St>
__device__ __forceinline__ void trace_instance(const DevScene& s, float3 o, float3 d, uint32_t ii,
                                               HitState& h, Stack& stk, uint32_t base, St& st,
                                               bool any_hit, bool by_reciprocal = true) {
    const float4* rec = s.inst + 8u * ii;
    float4 i0 = rec[0], i1 = rec[1], i2 = rec[2];
    uint4 g = *reinterpret_cast<const uint4*>(rec + 6);
    uint32_t flags = g.x, root = g.y, gfirst = g.z, gcount = g.w;
    float3 lo = xform_point(i0, i1, i2, o);
    float3 ld = xform_dir(i0, i1, i2, d);
    st.inst();
    const SlabRay sr(s, lo, ld, by_reciprocal);
    uint32_t count = 0;
    Node node = load_node(s.blas, s.n_blas, root);
    st.node(1);
    bool ready = sr.entry(node, h.t) >= 0.0f;
    TriRay ray;
    ray.valid = false;
    ray.kx = ray.ky = ray.kz = 0;
    ray.sx = ray.sy = 0.0f;
    if (ready) ray = sr.tri_ray();
    uint32_t steps = 0;
    while (ready || count != 0) {
        if (++steps > kMaxSteps) {
            h.err |= kErrUnsupported;
            break;
        }
        if (!ready) {
            --count;
            uint2 ent = stk.get(base + count);
            if (__uint_as_float(ent.y) > h.t) continue;
            node = load_node(s.blas, s.n_blas, ent.x);
            st.node(1);
        }
        ready = false;
        if (node.count == 0) {
            if (count + 2 > kStackCap) {
                h.err |= kErrBlasOverflow;
                node.first = gfirst;
                node.count = gcount;
                count = 0;
            } else {
                uint32_t li = node.first, ri = li + 1u;
                Node ln = load_node(s.blas, s.n_blas, li);
                Node rn = load_node(s.blas, s.n_blas, ri);
                st.node(2);
                float le = sr.entry(ln, h.t);
                float re = sr.entry(rn, h.t);
                if (le >= 0.0f && re >= 0.0f) {
                    if (le <= re) {
                        stk.put(base + count, make_uint2(ri, __float_as_uint(re)));
                        node = ln;
                    } else {
                        stk.put(base + count, make_uint2(li, __float_as_uint(le)));
                        node = rn;
                    }
                    ++count;
                    ready = true;
                } else if (le >= 0.0f) {
                    node = ln;
                    ready = true;
                } else if (re >= 0.0f) {
                    node = rn;
                    ready = true;
                }
                continue;
            }
        }
        uint32_t tbegin = node.first > gfirst ? node.first : gfirst;
        uint32_t tlimit = gfirst + gcount;
        uint32_t req_end = node.first + node.count;
        uint32_t tend = req_end < tlimit ? req_end : tlimit;
        bool truncated = node.first < gfirst || req_end < node.first || req_end > tlimit ||
                         node.count > kMaxLeafTriangles;
        uint32_t bend = tbegin + kMaxLeafTriangles;
        if (bend > tend) bend = tend;
        truncated = truncated || bend < tend;
        if (truncated) h.err |= kErrUnsupported;
        for (uint32_t i = tbegin; i < bend; ++i) {
            uint32_t ti = i < s.n_tri ? i : s.n_tri - 1u;
            const float4* tp = s.tri + 3ull * ti;
            float4 A = tp[0], B = tp[1], C = tp[2];
            st.tri();
            float3 a = make_float3(A.x, A.y, A.z), b = make_float3(B.x, B.y, B.z),
                   c = make_float3(C.x, C.y, C.z);
            if (culled(a, b, c, ld, flags)) continue;
            float3 r3 = intersect_tri(lo, ld, ray, a, b, c);
            if (r3.x > 0.000025f && r3.x < h.t) {
                h.t = r3.x;
                h.tri = i;
                h.inst = ii;
                h.b1 = r3.y;
                h.b2 = r3.z;
            }
            if (any_hit && h.tri != kEmpty) return;
        }
    }
}

// traceRayTracked without dependency or cone terms (traversal.slang:696-811).
// by_reciprocal = false keeps every slab test an IEEE division (same results;
// a caller passing the constant compiles the reciprocal path out).
template <class Stack, class St>
__device__ __forceinline__ void trace_closest(const DevScene& s, float3 o, float3 d, float tmax,
                                              HitState& h, Stack& stk, St& st, bool any_hit,
                                              bool by_reciprocal = true) {
    h.t = tmax;
    h.tri = kEmpty;
    h.inst = kEmpty;
    h.err = 0;
    h.b1 = h.b2 = 0.0f;
    if (s.n_inst == 0) return;
    uint32_t count = 0;
    const SlabRay sr(s, o, d, by_reciprocal);
    Node node = load_node(s.tlas, s.n_tlas, 0);
    st.node(1);
    // the root's interval, when met, starts at or after 0
    if (sr.entry(node, h.t) < 0.0f) return;
    bool ready = true;
    uint32_t steps = 0, ninst = s.n_inst;
    while (ready || count != 0) {
        if (++steps > kMaxSteps) {
            h.err |= kErrUnsupported;
            break;
        }
        if (!ready) {
            --count;
            uint2 ent = stk.get(count);
            if (__uint_as_float(ent.y) > h.t) continue;
            node = load_node(s.tlas, s.n_tlas, ent.x);
            st.node(1);
        }
        ready = false;
        if (node.count == 0) {
            if (count + 2 > kStackCap) {
                h.err |= kErrTlasOverflow;
                node.first = 0;
                node.count = ninst;
                count = 0;
            } else {
                uint32_t li = node.first, ri = li + 1u;
                Node ln = load_node(s.tlas, s.n_tlas, li);
                Node rn = load_node(s.tlas, s.n_tlas, ri);
                st.node(2);
                float le = sr.entry(ln, h.t);
                float rre = sr.entry(rn, h.t);
                if (le >= 0.0f && rre >= 0.0f) {
                    if (le <= rre) {
                        stk.put(count, make_uint2(ri, __float_as_uint(rre)));
                        node = ln;
                    } else {
                        stk.put(count, make_uint2(li, __float_as_uint(le)));
                        node = rn;
                    }
                    ++count;
                    ready = true;
                } else if (le >= 0.0f) {
                    node = ln;
                    ready = true;
                } else if (rre >= 0.0f) {
                    node = rn;
                    ready = true;
                }
                continue;
            }
        }
        uint32_t req_end = node.first + node.count;
        bool truncated = req_end < node.first || node.first > ninst || req_end > ninst ||
                         node.count > kMaxLeafInstances;
        uint32_t ib = node.first < ninst ? node.first : ninst;
        uint32_t ie = req_end < ninst ? req_end : ninst;
        uint32_t be = ib + kMaxLeafInstances;
        if (be > ie) be = ie;
        truncated = truncated || be < ie;
        if (truncated) h.err |= kErrUnsupported;
        for (uint32_t i = ib; i < be; ++i) {
            trace_instance(s, o, d, i, h, stk, count, st, any_hit, by_reciprocal);
            if (any_hit && h.tri != kEmpty) {
                // an accepted hit proves an any-hit query (traversal.slang:786-792)
                h.err &= ~kErrUnsupported;
                count = 0;
                break;
            }
        }
    }
}

// occludedInstance (traversal.slang:845-932).
template <class Stack, class St>
__device__ __forceinline__ bool occluded_instance(const DevScene& s, float3 o, float3 d,
                                                  uint32_t ii, float limit, bool& supported,
                                                  HitState& h, Stack& stk, uint32_t base, St& st) {
    supported = true;
    const float4* rec = s.inst + 8u * ii;
    float4 i0 = rec[0], i1 = rec[1], i2 = rec[2];
    uint4 g = *reinterpret_cast<const uint4*>(rec + 6);
    uint32_t flags = g.x, root = g.y, gfirst = g.z, gcount = g.w;
    float3 lo = xform_point(i0, i1, i2, o);
    float3 ld = xform_dir(i0, i1, i2, d);
    st.inst();
    const SlabRay sr(s, lo, ld);
    uint32_t count = 0;
    Node node = load_node(s.blas, s.n_blas, root);
    st.node(1);
    bool ready = sr.entry(node, limit) >= 0.0f;
    TriRay ray;
    ray.valid = false;
    ray.kx = ray.ky = ray.kz = 0;
    ray.sx = ray.sy = 0.0f;
    if (ready) ray = sr.tri_ray();
    uint32_t steps = 0;
    bool incomplete = false;
    while (ready || count != 0) {
        if (++steps > kMaxSteps) {
            supported = false;
            return false;
        }
        if (!ready) {
            --count;
            node = load_node(s.blas, s.n_blas, stk.get(base + count).x);
            st.node(1);
        }
        ready = false;
        if (node.count == 0) {
            if (count + 2 > kStackCap) {
                node.first = gfirst;
                node.count = gcount;
                count = 0;
            } else {
                uint32_t li = node.first, ri = li + 1u;
                Node ln = load_node(s.blas, s.n_blas, li);
                Node rn = load_node(s.blas, s.n_blas, ri);
                st.node(2);
                float le = sr.entry(ln, limit);
                float re = sr.entry(rn, limit);
                if (le >= 0.0f && re >= 0.0f) {
                    if (le <= re) {
                        stk.put(base + count, make_uint2(ri, 0u));
                        node = ln;
                    } else {
                        stk.put(base + count, make_uint2(li, 0u));
                        node = rn;
                    }
                    ++count;
                    ready = true;
                } else if (le >= 0.0f) {
                    node = ln;
                    ready = true;
                } else if (re >= 0.0f) {
                    node = rn;
                    ready = true;
                }
                continue;
            }
        }
        uint32_t tbegin = node.first > gfirst ? node.first : gfirst;
        uint32_t tlimit = gfirst + gcount;
        uint32_t req_end = node.first + node.count;
        uint32_t tend = req_end < tlimit ? req_end : tlimit;
        bool truncated = node.first < gfirst || req_end < node.first || req_end > tlimit ||
                         node.count > kMaxLeafTriangles;
        uint32_t bend = tbegin + kMaxLeafTriangles;
        if (bend > tend) bend = tend;
        truncated = truncated || bend < tend;
        incomplete = incomplete || truncated;
        for (uint32_t i = tbegin; i < bend; ++i) {
            uint32_t ti = i < s.n_tri ? i : s.n_tri - 1u;
            const float4* tp = s.tri + 3ull * ti;
            float4 A = tp[0], B = tp[1], C = tp[2];
            st.tri();
            float3 a = make_float3(A.x, A.y, A.z), b = make_float3(B.x, B.y, B.z),
                   c = make_float3(C.x, C.y, C.z);
            if (culled(a, b, c, ld, flags)) continue;
            float3 r3 = intersect_tri(lo, ld, ray, a, b, c);
            if (r3.x > 0.000025f && r3.x < limit) {
                h.t = r3.x;
                h.tri = i;
                h.inst = ii;
                h.b1 = r3.y;
                h.b2 = r3.z;
                return true;
            }
        }
    }
    supported = !incomplete;
    return false;
}

// traceOccluded (traversal.slang:955-1046).  Result in h.err bits:
// kHitOccluded when blocked, kHitRoot when the root box was met,
// kErrUnsupported when the query could not be completed.
template <class Stack, class St>
__device__ __forceinline__ void trace_occluded(const DevScene& s, float3 o, float3 d, float limit,
                                               HitState& h, Stack& stk, St& st) {
    h.t = limit;
    h.tri = kEmpty;
    h.inst = kEmpty;
    h.err = 0;
    h.b1 = h.b2 = 0.0f;
    if (s.n_inst == 0) return;
    uint32_t count = 0;
    const SlabRay sr(s, o, d);
    Node node = load_node(s.tlas, s.n_tlas, 0);
    st.node(1);
    // the root's interval, when met, starts at or after 0
    bool root = sr.entry(node, limit) >= 0.0f;
    if (root) h.err |= kHitRoot;
    bool ready = root;
    if (!ready) return;
    uint32_t steps = 0, ninst = s.n_inst;
    bool incomplete = false;
    while (ready || count != 0) {
        if (++steps > kMaxSteps) {
            h.err |= kErrUnsupported;
            return;
        }
        if (!ready) {
            --count;
            node = load_node(s.tlas, s.n_tlas, stk.get(count).x);
            st.node(1);
        }
        ready = false;
        if (node.count == 0) {
            if (count + 2 > kStackCap) {
                node.first = 0;
                node.count = ninst;
                count = 0;
            } else {
                uint32_t li = node.first, ri = li + 1u;
                Node ln = load_node(s.tlas, s.n_tlas, li);
                Node rn = load_node(s.tlas, s.n_tlas, ri);
                st.node(2);
                float le = sr.entry(ln, limit);
                float rre = sr.entry(rn, limit);
                if (le >= 0.0f && rre >= 0.0f) {
                    if (le <= rre) {
                        stk.put(count, make_uint2(ri, 0u));
                        node = ln;
                    } else {
                        stk.put(count, make_uint2(li, 0u));
                        node = rn;
                    }
                    ++count;
                    ready = true;
                } else if (le >= 0.0f) {
                    node = ln;
                    ready = true;
                } else if (rre >= 0.0f) {
                    node = rn;
                    ready = true;
                }
                continue;
            }
        }
        uint32_t req_end = node.first + node.count;
        bool truncated = req_end < node.first || node.first > ninst || req_end > ninst ||
                         node.count > kMaxLeafInstances;
        uint32_t ib = node.first < ninst ? node.first : ninst;
        uint32_t ie = req_end < ninst ? req_end : ninst;
        uint32_t be = ib + kMaxLeafInstances;
        if (be > ie) be = ie;
        incomplete = incomplete || truncated || be < ie;
        for (uint32_t i = ib; i < be; ++i) {
            bool sup;
            bool blocked = occluded_instance(s, o, d, i, limit, sup, h, stk, count, st);
            incomplete = incomplete || !sup;
            if (blocked) {
                h.err |= kHitOccluded;
                return;
            }
        }
    }
    if (incomplete) h.err |= kErrUnsupported;
}


// ---------------------------------------------------------------- resumable walk
// The three traversals above restated as a per-lane state machine: one
// walk_step performs one iteration of the TLAS loop, starts one instance of
// a TLAS leaf, or performs one iteration of the BLAS loop (including its
// leaf's triangles).  Every ray executes exactly the operations, in exactly
// the order, of trace_closest / trace_instance (modes kWalkClosest and
// kWalkFirst) or trace_occluded / occluded_instance (kWalkOccluded), so
// results are bit-identical; only the interleaving across lanes changes,
// which lets a wave retire finished rays and refill their lanes.
constexpr uint32_t kWalkClosest = 0u, kWalkFirst = 1u, kWalkOccluded = 2u;
constexpr uint32_t kPhaseDone = 0u, kPhaseTlas = 1u, kPhaseLeaf = 2u, kPhaseBlas = 3u, kPhaseTri = 4u;

struct Walk {
    const DevScene* sp;      // the lane's scene (multi-scene launches)
    float3 o, d;
    float limit;             // occlusion limit (closest/first-hit use h.t)
    HitState h;
    uint32_t mode, phase;
    uint32_t tcount, tsteps, ti, tend;
    bool tready, incomplete;
    Node node;               // current TLAS node, then the current BLAS node
    uint32_t ii, flags, gfirst, gcount, bcount, bsteps;
    uint32_t ti_next, ti_end;  // BLAS leaf: next triangle, end (one triangle per step)
    bool bready, binc;
    float3 lo, ld;
    TriRay ray;
    // reciprocal slab divisions: scene bounds in range; world ray (TLAS) and
    // local ray (current BLAS) in range, and their direction reciprocals
    bool scene_fast, fast_t, fast_b;
    float3 yt, yb;
};

template <class St>
__device__ __forceinline__ void walk_begin(const DevScene& s, Walk& w, float3 o, float3 d, float tmax,
                                           uint32_t mode, St& st) {
    w.o = o;
    w.d = d;
    w.limit = tmax;
    w.mode = mode;
    w.h.t = tmax;
    w.h.tri = kEmpty;
    w.h.inst = kEmpty;
    w.h.err = 0;
    w.h.b1 = w.h.b2 = 0.0f;
    w.phase = kPhaseDone;
    w.incomplete = false;
    w.tcount = 0;
    w.tsteps = 0;
    w.fast_b = false;
    if (s.n_inst == 0) return;
    w.scene_fast = scene_in_range(s);
    w.fast_t = w.scene_fast && ray_in_range(o, d);
    if (w.fast_t) w.yt = reciprocal(d);
    w.node = load_node(s.tlas, s.n_tlas, 0);
    st.node(1);
    float re, rx;
    // box_entry_fast is box_interval's entry (or -1 on a miss) for in-range rays
    bool root;
    if (w.fast_t) {
        re = box_entry_fast(o, d, w.yt, w.node, tmax);
        root = re >= 0.0f;
    } else {
        root = box_interval(o, d, w.node.mn, w.node.mx, tmax, re, rx);
    }
    if (mode == kWalkOccluded && root) w.h.err |= kHitRoot;
    if (!root) return;
    w.tready = re >= 0.0f;
    if (mode == kWalkOccluded && !w.tready) return;
    w.phase = kPhaseTlas;
}

template <class Stack, class St>
__device__ __forceinline__ void walk_tlas(const DevScene& s, Walk& w, Stack& stk, St& st) {
    const bool occl = w.mode == kWalkOccluded;
    if (!(w.tready || w.tcount != 0)) {
        if (occl && w.incomplete) w.h.err |= kErrUnsupported;
        w.phase = kPhaseDone;
        return;
    }
    if (++w.tsteps > kMaxSteps) {
        w.h.err |= kErrUnsupported;
        w.phase = kPhaseDone;
        return;
    }
    if (!w.tready) {
        --w.tcount;
        uint2 ent = stk.get(w.tcount);
        if (!occl && __uint_as_float(ent.y) > w.h.t) return;
        w.node = load_node(s.tlas, s.n_tlas, ent.x);
        st.node(1);
    }
    w.tready = false;
    const uint32_t ninst = s.n_inst;
    if (w.node.count == 0) {
        if (w.tcount + 2 > kStackCap) {
            if (!occl) w.h.err |= kErrTlasOverflow;
            w.node.first = 0;
            w.node.count = ninst;
            w.tcount = 0;
        } else {
            uint32_t li = w.node.first, ri = li + 1u;
            Node ln = load_node(s.tlas, s.n_tlas, li);
            Node rn = load_node(s.tlas, s.n_tlas, ri);
            st.node(2);
            float lim = occl ? w.limit : w.h.t;
            float le, rre;
            if (w.fast_t) {
                le = box_entry_fast(w.o, w.d, w.yt, ln, lim);
                rre = box_entry_fast(w.o, w.d, w.yt, rn, lim);
            } else {
                le = box_entry(w.o, w.d, ln, lim);
                rre = box_entry(w.o, w.d, rn, lim);
            }
            if (le >= 0.0f && rre >= 0.0f) {
                if (le <= rre) {
                    stk.put(w.tcount, make_uint2(ri, occl ? 0u : __float_as_uint(rre)));
                    w.node = ln;
                } else {
                    stk.put(w.tcount, make_uint2(li, occl ? 0u : __float_as_uint(le)));
                    w.node = rn;
                }
                ++w.tcount;
                w.tready = true;
            } else if (le >= 0.0f) {
                w.node = ln;
                w.tready = true;
            } else if (rre >= 0.0f) {
                w.node = rn;
                w.tready = true;
            }
            return;
        }
    }
    uint32_t req_end = w.node.first + w.node.count;
    bool truncated = req_end < w.node.first || w.node.first > ninst || req_end > ninst ||
                     w.node.count > kMaxLeafInstances;
    uint32_t ib = w.node.first < ninst ? w.node.first : ninst;
    uint32_t ie = req_end < ninst ? req_end : ninst;
    uint32_t be = ib + kMaxLeafInstances;
    if (be > ie) be = ie;
    truncated = truncated || be < ie;
    if (occl) w.incomplete = w.incomplete || truncated;
    else if (truncated) w.h.err |= kErrUnsupported;
    w.ti = ib;
    w.tend = be;
    w.phase = kPhaseLeaf;
}

template <class St>
__device__ __forceinline__ void walk_instance(const DevScene& s, Walk& w, St& st) {
    if (w.ti >= w.tend) {
        w.phase = kPhaseTlas;
        return;
    }
    const uint32_t ii = w.ti;
    const float4* rec = s.inst + 8u * ii;
    float4 i0 = rec[0], i1 = rec[1], i2 = rec[2];
    uint4 g = *reinterpret_cast<const uint4*>(rec + 6);
    w.ii = ii;
    w.flags = g.x;
    w.gfirst = g.z;
    w.gcount = g.w;
    w.lo = xform_point(i0, i1, i2, w.o);
    w.ld = xform_dir(i0, i1, i2, w.d);
    w.fast_b = w.scene_fast && ray_in_range(w.lo, w.ld);
    if (w.fast_b) w.yb = reciprocal(w.ld);
    st.inst();
    w.bcount = 0;
    w.node = load_node(s.blas, s.n_blas, g.y);
    st.node(1);
    float e, x;
    float lim = w.mode == kWalkOccluded ? w.limit : w.h.t;
    if (w.fast_b) w.bready = box_entry_fast(w.lo, w.ld, w.yb, w.node, lim) >= 0.0f;
    else w.bready = box_interval(w.lo, w.ld, w.node.mn, w.node.mx, lim, e, x) && e >= 0.0f;
    w.ray.valid = false;
    w.ray.kx = w.ray.ky = w.ray.kz = 0;
    w.ray.sx = w.ray.sy = 0.0f;
    if (w.bready) w.ray = make_tri_ray(w.ld, w.fast_b ? &w.yb : nullptr);
    w.bsteps = 0;
    w.binc = false;
    w.phase = kPhaseBlas;
}

// End of the current instance's walk: the TLAS leaf loop resumes.
__device__ __forceinline__ void walk_instance_end(Walk& w, bool supported) {
    if (w.mode == kWalkOccluded) {
        w.incomplete = w.incomplete || !supported;
    } else if (w.mode == kWalkFirst && w.h.tri != kEmpty) {
        // an accepted hit proves an any-hit query (traversal.slang:786-792)
        w.h.err &= ~kErrUnsupported;
        w.tcount = 0;
        w.phase = kPhaseTlas;
        return;
    }
    ++w.ti;
    w.phase = kPhaseLeaf;
}

template <class Stack, class St>
__device__ __forceinline__ void walk_blas(const DevScene& s, Walk& w, Stack& stk, St& st) {
    const bool occl = w.mode == kWalkOccluded;
    if (!(w.bready || w.bcount != 0)) {
        walk_instance_end(w, !w.binc);
        return;
    }
    if (++w.bsteps > kMaxSteps) {
        if (!occl) w.h.err |= kErrUnsupported;
        walk_instance_end(w, false);
        return;
    }
    const uint32_t base = w.tcount;
    if (!w.bready) {
        --w.bcount;
        uint2 ent = stk.get(base + w.bcount);
        if (!occl && __uint_as_float(ent.y) > w.h.t) return;
        w.node = load_node(s.blas, s.n_blas, ent.x);
        st.node(1);
    }
    w.bready = false;
    if (w.node.count == 0) {
        if (w.bcount + 2 > kStackCap) {
            if (!occl) w.h.err |= kErrBlasOverflow;
            w.node.first = w.gfirst;
            w.node.count = w.gcount;
            w.bcount = 0;
        } else {
            uint32_t li = w.node.first, ri = li + 1u;
            Node ln = load_node(s.blas, s.n_blas, li);
            Node rn = load_node(s.blas, s.n_blas, ri);
            st.node(2);
            float lim = occl ? w.limit : w.h.t;
            float le, re;
            if (w.fast_b) {
                le = box_entry_fast(w.lo, w.ld, w.yb, ln, lim);
                re = box_entry_fast(w.lo, w.ld, w.yb, rn, lim);
            } else {
                le = box_entry(w.lo, w.ld, ln, lim);
                re = box_entry(w.lo, w.ld, rn, lim);
            }
            if (le >= 0.0f && re >= 0.0f) {
                if (le <= re) {
                    stk.put(base + w.bcount, make_uint2(ri, occl ? 0u : __float_as_uint(re)));
                    w.node = ln;
                } else {
                    stk.put(base + w.bcount, make_uint2(li, occl ? 0u : __float_as_uint(le)));
                    w.node = rn;
                }
                ++w.bcount;
                w.bready = true;
            } else if (le >= 0.0f) {
                w.node = ln;
                w.bready = true;
            } else if (re >= 0.0f) {
                w.node = rn;
                w.bready = true;
            }
            return;
        }
    }
    uint32_t tbegin = w.node.first > w.gfirst ? w.node.first : w.gfirst;
    uint32_t tlimit = w.gfirst + w.gcount;
    uint32_t req_end = w.node.first + w.node.count;
    uint32_t tend = req_end < tlimit ? req_end : tlimit;
    bool truncated = w.node.first < w.gfirst || req_end < w.node.first || req_end > tlimit ||
                     w.node.count > kMaxLeafTriangles;
    uint32_t bend = tbegin + kMaxLeafTriangles;
    if (bend > tend) bend = tend;
    truncated = truncated || bend < tend;
    if (occl) w.binc = w.binc || truncated;
    else if (truncated) w.h.err |= kErrUnsupported;
    // the leaf's triangles follow, one per step, before the BLAS loop resumes
    if (tbegin < bend) {
        w.ti_next = tbegin;
        w.ti_end = bend;
        w.phase = kPhaseTri;
    }
}

// One triangle of the current BLAS leaf (the leaf loop of trace_instance /
// occluded_instance, in order).
template <class St>
__device__ __forceinline__ void walk_tri(const DevScene& s, Walk& w, St& st) {
    const bool occl = w.mode == kWalkOccluded;
    const uint32_t i = w.ti_next++;
    if (w.ti_next >= w.ti_end) w.phase = kPhaseBlas;
    uint32_t ti = i < s.n_tri ? i : s.n_tri - 1u;
    const float4* tp = s.tri + 3ull * ti;
            float4 A = tp[0], B = tp[1], C = tp[2];
    st.tri();
    float3 a = make_float3(A.x, A.y, A.z), b = make_float3(B.x, B.y, B.z), c = make_float3(C.x, C.y, C.z);
    if (culled(a, b, c, w.ld, w.flags)) return;
    float3 r3 = intersect_tri(w.lo, w.ld, w.ray, a, b, c);
    float lim = occl ? w.limit : w.h.t;
    if (r3.x > 0.000025f && r3.x < lim) {
        w.h.t = r3.x;
        w.h.tri = i;
        w.h.inst = w.ii;
        w.h.b1 = r3.y;
        w.h.b2 = r3.z;
        if (occl) {
            w.h.err |= kHitOccluded;
            w.phase = kPhaseDone;
            return;
        }
    }
    if (w.mode == kWalkFirst && w.h.tri != kEmpty) walk_instance_end(w, true);
}

// One iteration; returns true once the ray is finished.  The phases run in
// walk order and each falls through to the next within the iteration: a ray
// that starts, or that ends its TLAS leaf's previous instance, enters its
// BLAS and takes its first node step at once, and a node step that reaches a
// leaf tests the leaf's first triangle at once.  A wave therefore runs the
// node-step and triangle bodies with more of its lanes in each, while every
// lane still performs its own steps in the same order.
template <class Stack, class St>
__device__ __forceinline__ bool walk_step(const DevScene& s, Walk& w, Stack& stk, St& st) {
    if (w.phase == kPhaseTlas) walk_tlas(s, w, stk, st);
    if (w.phase == kPhaseLeaf) walk_instance(s, w, st);
    if (w.phase == kPhaseBlas) walk_blas(s, w, stk, st);
    if (w.phase == kPhaseTri) walk_tri(s, w, st);
    return w.phase == kPhaseDone;
}

// Lane-refilling driver for resumable walks.  Src supplies the work items:
//   uint32_t count()                         number of items
//   bool begin(uint32_t i, Walk& w, St& st)   start item i; false when it
//                                            finished without a walk (its
//                                            result already written)
//   void finish(uint32_t i, const Walk& w)    store a finished walk
// All lanes run the loop together (ballots and shuffles see the whole
// wave); a wave refills its idle lanes with one atomic on *cursor once at
// least `regroup` lanes are idle, and leaves when the cursor is drained and
// none of its lanes holds an item.
template <class Src, class Stack, class St>
__device__ __forceinline__ void regroup_walks(const DevScene& s, Src& src, uint32_t* cursor,
                                              uint32_t regroup, uint32_t batch, Stack& stk, St& st) {
    const uint32_t lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t n = src.count();
    Walk w;
    uint32_t idx = 0;
    bool active = false, drained = false;
    while (true) {
        uint64_t idle = __ballot(!active);
        uint32_t nidle = (uint32_t)__popcll(idle);
        if (!drained && nidle >= regroup) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(cursor, nidle);
            base = __shfl(base, 0, 64);
            if (base >= n || n - base <= nidle) drained = true;
            if (!active) {
                idx = base + (uint32_t)__popcll(idle & below);
                if (idx < n) active = src.begin(idx, w, st);
            }
        }
        if (!__any(active)) {
            if (drained) break;
            continue;
        }
        for (uint32_t k = 0; k < batch; ++k) {
            if (active && walk_step(s, w, stk, st)) {
                src.finish(idx, w);
                active = false;
            }
        }
    }
}

// regroup_walks for items of several scenes: Src::begin also sets w.sp, and
// each lane steps through its own scene.
template <class Src, class Stack, class St>
__device__ __forceinline__ void regroup_walks_multi(Src& src, uint32_t* cursor, uint32_t regroup,
                                                    uint32_t batch, Stack& stk, St& st) {
    const uint32_t lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t n = src.count();
    Walk w;
    w.sp = nullptr;
    uint32_t idx = 0;
    bool active = false, drained = false;
    while (true) {
        uint64_t idle = __ballot(!active);
        uint32_t nidle = (uint32_t)__popcll(idle);
        if (!drained && nidle >= regroup) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(cursor, nidle);
            base = __shfl(base, 0, 64);
            if (base >= n || n - base <= nidle) drained = true;
            if (!active) {
                idx = base + (uint32_t)__popcll(idle & below);
                if (idx < n) active = src.begin(idx, w, st);
            }
        }
        if (!__any(active)) {
            if (drained) break;
            continue;
        }
        for (uint32_t k = 0; k < batch; ++k) {
            if (active && walk_step(*w.sp, w, stk, st)) {
                src.finish(idx, w);
                active = false;
            }
        }
    }
}

}  // namespace lux

// lux_screen.hip -- the screen-space passes after the resolve on gfx950:
// temporal (shaders/cascades.slang:310-622) and one denoise pass
// (:624-711), one thread per resolve texel in 8 x 8 tiles (one wave).
//
// Both are gather filters whose taps were dominated by decoding (IEEE
// divides and square roots of the packed normals and the world positions):
// for temporal a cache pass decodes each resolve texel once (the raster
// receiver words, normal, depth and world position of its representative
// pixel) and the 3 x 3 neighbourhood reads those (L2-resident); it then reads
// up to four previous texels with their metadata, receiver words and
// background.  Denoise (a 5 x 5 footprint without its four far corners)
// decodes the normal and world position of every texel of a 16 x 16 tile's
// footprint once into LDS (steps up to 8; larger steps decode in the taps)
// and its taps read the images' values from L2: a decode cache in global
// memory made the taps wait on its loads, staging the raw values as well
// cost occupancy, and staging without decoding left the decode as the cost.
//
// Restatement choices shared with the oracle: row-major matrices with
// h_i = ((m_i0 x + m_i1 y) + m_i2 z) + m_i3; dot products summed x, y, z;
// lerp(a, b, t) = a (1 - t) + b t; round = nearest-even; normalize =
// v / sqrt(v.v); pow(x, 32) by five squarings; expf.  The debug-only
// diagnostic counters of recordTemporalDiagnostics (:283-308) are not kept.
#include <cmath>
#include <cstring>
#include "lux_internal.h"

namespace {

constexpr uint32_t kTile = 8;

struct ScreenArgs {
    lux_screen_params p;
    uint32_t rw, rh, rs;
    int stride_x, stride_y;  // the neighbourhood stride (cellStride)
};

__device__ __forceinline__ bool fin(float f) { return (__float_as_uint(f) & 0x7f800000u) != 0x7f800000u; }
__device__ __forceinline__ float sat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float frc(float x) { return x - floorf(x); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 norm3(float3 a) {
    const float l = sqrtf(dot3(a, a));
    return make_float3(a.x / l, a.y / l, a.z / l);
}
__device__ __forceinline__ float lum(float r, float g, float b) { return (r * 0.2126f + g * 0.7152f) + b * 0.0722f; }

// encodeNormal / decodeNormal (shaders/common.slang:361-375)
__device__ float2 enc_normal(float3 n) {
    const float s = (fabsf(n.x) + fabsf(n.y)) + fabsf(n.z);
    n = make_float3(n.x / s, n.y / s, n.z / s);
    float px = n.x, py = n.y;
    if (n.z < 0.0f) {
        const float qx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        const float qy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = qx;
        py = qy;
    }
    return make_float2(px * 0.5f + 0.5f, py * 0.5f + 0.5f);
}

__device__ float3 dec_normal(float u, float v) {
    const float px = u * 2.0f - 1.0f, py = v * 2.0f - 1.0f;
    float3 n = make_float3(px, py, (1.0f - fabsf(px)) - fabsf(py));
    const float t = sat(-n.z);
    n.x += n.x >= 0.0f ? -t : t;
    n.y += n.y >= 0.0f ? -t : t;
    return norm3(n);
}

// shaders/receiver_metadata.slang
__device__ __forceinline__ float3 recv_normal(float4 m) {
    const uint32_t packed = (uint32_t)m.z;
    return dec_normal((float)(packed & 4095u) / 4095.0f, (float)(packed >> 12) / 4095.0f);
}
__device__ __forceinline__ float packed_normal(float3 n) {
    const float2 p = enc_normal(n);
    const uint32_t ox = (uint32_t)rintf(sat(p.x) * 4095.0f), oy = (uint32_t)rintf(sat(p.y) * 4095.0f);
    return (float)(ox | (oy << 12));
}
__device__ __forceinline__ uint32_t hist_len(float4 m) { return (uint32_t)m.w & 255u; }
__device__ __forceinline__ uint32_t support_mask(float4 m) { return ((uint32_t)m.w >> 16) & 1u; }
__device__ __forceinline__ float4 with_support(float4 m, uint32_t s) {
    uint32_t packed = (uint32_t)m.w;
    packed = (packed & ~(1u << 16)) | ((s & 1u) << 16);
    m.w = (float)packed;
    return m;
}
__device__ __forceinline__ bool recv_same(float4 a, float4 b) {
    return hist_len(a) > 0u && hist_len(b) > 0u && __float_as_uint(a.x) == __float_as_uint(b.x) &&
           __float_as_uint(a.y) == __float_as_uint(b.y);
}
// splitReceiverCompatible = recv_same && dot(normals) > 0.97; the callers
// decode each packed normal once and test the two parts in that order.

__device__ float3 geometric_normal(uint32_t packed, float3 fallback) {
    fallback = norm3(fallback);
    if ((packed & 0x80000000u) == 0u) return fallback;
    const float ox = (float)(packed & 0x7fffu) / 32767.0f * 2.0f - 1.0f;
    const float oy = (float)((packed >> 15) & 0x7fffu) / 32767.0f * 2.0f - 1.0f;
    float3 n = make_float3(ox, oy, (1.0f - fabsf(ox)) - fabsf(oy));
    if (n.z < 0.0f) {
        const float nx = (1.0f - fabsf(n.y)) * (n.x >= 0.0f ? 1.0f : -1.0f);
        const float ny = (1.0f - fabsf(n.x)) * (n.y >= 0.0f ? 1.0f : -1.0f);
        n.x = nx;
        n.y = ny;
    }
    n = norm3(n);
    return dot3(n, fallback) < 0.0f ? make_float3(-n.x, -n.y, -n.z) : n;
}

// splitRasterMetadata (shaders/cascades.slang:271-278)
__device__ float4 raster_meta(const uint4* gbuffer, const uint4* raster, uint64_t full) {
    const uint32_t gw = gbuffer[full].w;
    const uint4 r = raster[full];
    const float3 shading = dec_normal((float)(gw & 4095u) / 4095.0f, (float)((gw >> 12) & 4095u) / 4095.0f);
    return make_float4(__uint_as_float(r.x), __uint_as_float(r.y), packed_normal(geometric_normal(r.z, shading)),
                       1.0f);
}

__device__ __forceinline__ float4 mat_apply(const float* M, float x, float y, float z) {
    float h[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = ((M[4 * i] * x + M[4 * i + 1] * y) + M[4 * i + 2] * z) + M[4 * i + 3] * 1.0f;
    return make_float4(h[0], h[1], h[2], h[3]);
}

__device__ __forceinline__ float3 world_pos(const float* M, const lux_screen_params& p, float px, float py,
                                            float depth) {
    const float u = (px - (float)p.viewport[0]) / (float)p.viewport[2];
    const float v = (py - (float)p.viewport[1]) / (float)p.viewport[3];
    const float4 h = mat_apply(M, u * 2.0f - 1.0f, 1.0f - v * 2.0f, depth);
    return make_float3(h.x / h.w, h.y / h.w, h.z / h.w);
}

__device__ __forceinline__ float surface_tol(const lux_screen_params& p) {
    return fmaxf(p.normal_bias * 2.0f, p.probe_spacing * 0.02f);
}

__device__ __forceinline__ float plane_dist(float3 d, float3 tn, float3 sn) {
    const float3 a = make_float3(tn.x + sn.x, tn.y + sn.y, tn.z + sn.z);
    const float l2 = dot3(a, a);
    if (!(l2 > 1e-20f) || !fin(l2)) return fabsf(dot3(d, tn));
    const float r = 1.0f / sqrtf(l2);
    return fabsf(dot3(d, make_float3(a.x * r, a.y * r, a.z * r)));
}

__device__ __forceinline__ bool dirty(const lux_screen_params& p, float3 q) {
    return p.dirty_min[3] > 0.0f && q.x >= p.dirty_min[0] && q.y >= p.dirty_min[1] && q.z >= p.dirty_min[2] &&
           q.x <= p.dirty_max[0] && q.y <= p.dirty_max[1] && q.z <= p.dirty_max[2];
}

// This workgroup's 8 x 8 tile.  Workgroups go to the 8 XCDs in turn and each
// XCD has its own L2; numbering the tiles so that XCD k takes the k-th eighth
// of the image (row-major) keeps the texels that neighbouring tiles' taps
// share in one L2 instead of fetching them into all eight.
__device__ __forceinline__ void xcd_tile(uint32_t& bx, uint32_t& by) {
    const uint32_t gx = gridDim.x, n = gridDim.x * gridDim.y;
    const uint32_t id = blockIdx.y * gx + blockIdx.x;
    const uint32_t xcd = id % 8u, within = id / 8u, per = n / 8u, rem = n % 8u;
    const uint32_t t = xcd * per + (xcd < rem ? xcd : rem) + within;
    bx = t % gx;
    by = t / gx;
}

// The full-resolution representative of a resolve texel.
__device__ __forceinline__ uint64_t full_of(const ScreenArgs& a, int tx, int ty, uint32_t& fx, uint32_t& fy) {
    uint32_t lx = (uint32_t)tx * a.rs + a.rs / 2u, ly = (uint32_t)ty * a.rs + a.rs / 2u;
    lx = lx < a.p.viewport[2] - 1u ? lx : a.p.viewport[2] - 1u;
    ly = ly < a.p.viewport[3] - 1u ? ly : a.p.viewport[3] - 1u;
    fx = a.p.viewport[0] + lx;
    fy = a.p.viewport[1] + ly;
    return (uint64_t)fy * a.p.target[0] + fx;
}

__device__ __forceinline__ void store(float4* img, uint64_t i, float4 v) { img[i] = v; }

// Per resolve texel, once: the raster receiver words of its representative
// pixel, their decoded normal with the depth, and the world position (taps of
// the 3 x 3 neighbourhood read these instead of decoding again).
__global__ __launch_bounds__(64) void k_screen_cache_raster(ScreenArgs a, const float* __restrict__ depth,
                                                            const uint4* __restrict__ gbuffer,
                                                            const uint4* __restrict__ raster,
                                                            float4* __restrict__ cache) {
    uint32_t bx, by;
    xcd_tile(bx, by);
    const int tx = (int)(bx * kTile + threadIdx.x % kTile);
    const int ty = (int)(by * kTile + threadIdx.x / kTile);
    if (tx >= (int)a.rw || ty >= (int)a.rh) return;
    const uint64_t pix = (uint64_t)ty * a.rw + tx;
    uint32_t fx, fy;
    const uint64_t full = full_of(a, tx, ty, fx, fy);
    const float d = depth[full];
    const float4 m = raster_meta(gbuffer, raster, full);
    const float3 n = recv_normal(m);
    const float3 p = d > 0.0f ? world_pos(a.p.world_from_clip, a.p, (float)fx + 0.5f, (float)fy + 0.5f, d)
                              : make_float3(0.0f, 0.0f, 0.0f);
    cache[3 * pix] = m;
    cache[3 * pix + 1] = make_float4(n.x, n.y, n.z, d);
    cache[3 * pix + 2] = make_float4(p.x, p.y, p.z, 0.0f);
}


__global__ __launch_bounds__(64) void k_screen_temporal(
    ScreenArgs a, const float4* __restrict__ current, const float4* __restrict__ current_bg,
    const float* __restrict__ depth, const uint4* __restrict__ gbuffer, const uint4* __restrict__ raster,
    const float2* __restrict__ motion, const float4* __restrict__ prev, const float4* __restrict__ prev_meta,
    const float4* __restrict__ prev_recv, const float4* __restrict__ prev_bg, const float4* __restrict__ cache,
    float4* __restrict__ out_irr, float4* __restrict__ out_meta, float4* __restrict__ out_recv,
    float4* __restrict__ out_bg) {
    uint32_t bx, by;
    xcd_tile(bx, by);
    const int tx = (int)(bx * kTile + threadIdx.x % kTile);
    const int ty = (int)(by * kTile + threadIdx.x / kTile);
    if (tx >= (int)a.rw || ty >= (int)a.rh) return;
    const lux_screen_params& P = a.p;
    const uint32_t RW = a.rw, RH = a.rh, RS = a.rs;
    const uint64_t pix = (uint64_t)ty * RW + tx;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 center = current[pix];
    uint32_t fx, fy;
    const uint64_t full = full_of(a, tx, ty, fx, fy);
    const float4 receiver = cache[3 * pix];
    const float4 nd_c = cache[3 * pix + 1];
    const float d = nd_c.w;
    if (d <= 0.0f || hist_len(receiver) == 0u) {
        store(out_bg, pix, zero);
        store(out_irr, pix, zero);
        store(out_meta, pix, zero);
        store(out_recv, pix, zero);
        return;
    }
    const bool cur_finite = fin(center.w) && fin(center.x) && fin(center.y) && fin(center.z);
    const float4 bgs = current_bg ? current_bg[pix] : zero;
    const bool bg_sup = bgs.w > 0.0f && fin(bgs.x) && fin(bgs.y) && fin(bgs.z) && fin(bgs.w);
    const float bg_fresh = bg_sup ? sat(frc(bgs.w) * 4.0f) : 0.0f;
    uint32_t cur_mask = cur_finite && center.w > 0.0f ? ((uint32_t)floorf(center.w)) & 1u : 0u;
    bool cur_sup = cur_mask != 0u && cur_finite;
    float cur_fresh = cur_finite && cur_sup ? sat(frc(center.w) * 4.0f) : 0.0f;
    if (!cur_sup && cur_finite && center.w < 0.0f && bg_sup) {
        cur_sup = true;
        cur_mask = 1u;
        cur_fresh = bg_fresh;
    }
    if (!cur_sup) center.x = center.y = center.z = 0.0f;
    center.w = d;
    uint32_t out_mask = cur_mask;
    const float3 gn = make_float3(nd_c.x, nd_c.y, nd_c.z);
    const float3 normal = gn;
    const float4 pos_c = cache[3 * pix + 2];
    const float3 position = make_float3(pos_c.x, pos_c.y, pos_c.z);
    const float tol = surface_tol(P);
    float3 mean = make_float3(0.0f, 0.0f, 0.0f), square = mean;
    float count = 0.0f;
    for (int y = -1; y <= 1; ++y)
        for (int x = -1; x <= 1; ++x) {
            const int nx = tx + x * a.stride_x, ny = ty + y * a.stride_y;
            if (nx < 0 || ny < 0 || nx >= (int)RW || ny >= (int)RH) continue;
            const float4 value = current[(uint64_t)ny * RW + nx];
            if (value.w <= 0.0f) continue;
            const uint32_t nmask = ((uint32_t)floorf(value.w)) & 1u;
            if (cur_mask != 0u && nmask != cur_mask) continue;
            const uint64_t nt = (uint64_t)ny * RW + nx;
            const float4 ond = cache[3 * nt + 1];
            if (ond.w <= 0.0f) continue;
            const float4 orc = cache[3 * nt];
            if (!recv_same(receiver, orc)) continue;
            const float3 on = make_float3(ond.x, ond.y, ond.z);
            if (!(dot3(gn, on) > 0.97f)) continue;
            const float4 opc = cache[3 * nt + 2];
            const float3 op = make_float3(opc.x, opc.y, opc.z);
            if (dot3(gn, on) < 0.95f || plane_dist(sub3(op, position), gn, on) > tol) continue;
            mean.x += value.x; mean.y += value.y; mean.z += value.z;
            square.x += value.x * value.x; square.y += value.y * value.y; square.z += value.z * value.z;
            count += 1.0f;
        }
    const float cdiv = fmaxf(count, 1.0f);
    mean = make_float3(mean.x / cdiv, mean.y / cdiv, mean.z / cdiv);
    square = make_float3(square.x / cdiv, square.y / cdiv, square.z / cdiv);
    const float3 variance = make_float3(fmaxf(0.0f, square.x - mean.x * mean.x), fmaxf(0.0f, square.y - mean.y * mean.y),
                                        fmaxf(0.0f, square.z - mean.z * mean.z));
    const float3 deviation = make_float3(sqrtf(variance.x), sqrtf(variance.y), sqrtf(variance.z));
    float4 history = zero, metadata = zero, bg_hist = zero;
    float total = 0.0f, bg_weight = 0.0f;
    uint32_t hist_mask = 0u;
    float pfx, pfy;
    bool projected;
    const bool has_motion = P.stability[2] > 0.0f && (!P.persistent || P.stability[1] > 0.5f);
    if (has_motion) {
        const float2 mv = motion[full];
        pfx = (float)fx + 0.5f - mv.x * (float)P.viewport[2];
        pfy = (float)fy + 0.5f - mv.y * (float)P.viewport[3];
        projected = fin(mv.x) && fin(mv.y);
    } else {
        const float4 h = mat_apply(P.previous_clip_from_world, position.x, position.y, position.z);
        const float ndx = h.x / h.w, ndy = h.y / h.w;
        const float u = ndx * 0.5f + 0.5f, v = 0.5f - ndy * 0.5f;
        pfx = (float)P.viewport[0] + u * (float)P.viewport[2];
        pfy = (float)P.viewport[1] + v * (float)P.viewport[3];
        projected = h.w > 0.0f && fin(u) && fin(v);
    }
    const bool gradual = (P.source_flags & 1u) != 0u;
    const uint32_t hard = P.source_changes & ~(gradual ? 1u : 0u);
    const bool motion_expired = (P.source_flags & 2u) != 0u && P.frame - P.motion_epoch > 120u;
    if (projected && P.stability[0] > 0.0f && (hard & 1u) == 0u && !dirty(P, position) && P.history_blend > 0.0f &&
        P.history_frames > 0.0f && pfx >= (float)P.viewport[0] && pfy >= (float)P.viewport[1] &&
        pfx < (float)(P.viewport[0] + P.viewport[2]) && pfy < (float)(P.viewport[1] + P.viewport[3])) {
        const float lx = pfx - (float)P.viewport[0] - 0.5f, ly = pfy - (float)P.viewport[1] - 0.5f;
        const int bx = (int)floorf((lx - (float)(RS / 2u)) / (float)RS);
        const int by = (int)floorf((ly - (float)(RS / 2u)) / (float)RS);
        const float lastx = (float)(P.viewport[2] - 1u), lasty = (float)(P.viewport[3] - 1u);
        const float leftx = fminf((float)(bx * (int)RS + (int)(RS / 2u)), lastx);
        const float lefty = fminf((float)(by * (int)RS + (int)(RS / 2u)), lasty);
        const float rightx = fminf((float)((bx + 1) * (int)RS + (int)(RS / 2u)), lastx);
        const float righty = fminf((float)((by + 1) * (int)RS + (int)(RS / 2u)), lasty);
        const float frx = sat((lx - leftx) / fmaxf(rightx - leftx, 1.0f));
        const float fry = sat((ly - lefty) / fmaxf(righty - lefty, 1.0f));
        for (uint32_t corner = 0; corner < 4u; ++corner) {
            const int bitx = (int)(corner & 1u), bity = (int)(corner >> 1);
            const int ox = bx + bitx, oy = by + bity;
            if (ox < 0 || oy < 0 || ox >= (int)RW || oy >= (int)RH) continue;
            const uint64_t op = (uint64_t)oy * RW + ox;
            const float4 old = prev[op], meta = prev_meta[op], old_r = prev_recv[op];
            if (meta.z <= 0.0f || old.w <= 0.0f || !(fin(old.x) && fin(old.y) && fin(old.z)) ||
                dot3(normal, dec_normal(meta.x, meta.y)) < 0.95f || !recv_same(receiver, old_r))
                continue;
            const float3 old_n = recv_normal(old_r);
            if (!(dot3(gn, old_n) > 0.97f)) continue;
            uint32_t ofx, ofy;
            full_of(a, ox, oy, ofx, ofy);
            const float3 old_pos = world_pos(P.previous_world_from_clip, P, (float)ofx + 0.5f, (float)ofy + 0.5f, old.w);
            if (dirty(P, old_pos)) continue;
            if (has_motion) {
                const float4 h = mat_apply(P.previous_clip_from_world, position.x, position.y, position.z);
                const float expected = h.z / h.w;
                if (!fin(expected) || expected <= 0.0f) continue;
                float t = fmaxf(0.000002f, fabsf(expected) * 0.0001f);
                for (int axis = 0; axis < 2; ++axis) {
                    int nx = ox + (axis == 0 ? 1 : 0), ny = oy + (axis == 0 ? 0 : 1);
                    nx = nx < (int)RW - 1 ? nx : (int)RW - 1;
                    ny = ny < (int)RH - 1 ? ny : (int)RH - 1;
                    const uint64_t nb = (uint64_t)ny * RW + nx;
                    const float4 nr = prev_recv[nb];
                    const float nd = prev[nb].w;
                    if (recv_same(old_r, nr) && dot3(old_n, recv_normal(nr)) > 0.97f && nd > 0.0f)
                        t += 2.0f * fabsf(nd - old.w);
                }
                if (fabsf(expected - old.w) > t) continue;
            } else {
                if (plane_dist(sub3(old_pos, position), gn, old_n) > tol) continue;
            }
            const float wx = bitx ? frx : 1.0f - frx, wy = bity ? fry : 1.0f - fry;
            const float weight = wx * wy;
            if (weight <= 0.0f) continue;
            if (support_mask(old_r) == 0u) continue;
            const float4 obg = prev_bg ? prev_bg[op] : zero;
            if (P.dirty_min[3] <= 0.0f && obg.w > 0.0f && !(motion_expired && bg_fresh <= 0.00001f) && fin(obg.x) &&
                fin(obg.y) && fin(obg.z) && fin(obg.w)) {
                bg_hist.x += obg.x * weight; bg_hist.y += obg.y * weight;
                bg_hist.z += obg.z * weight; bg_hist.w += obg.w * weight;
                bg_weight += weight;
            }
            if (hard != 0u || (motion_expired && cur_fresh <= 0.00001f)) continue;
            hist_mask = 1u;
            history.x += old.x * weight; history.y += old.y * weight;
            history.z += old.z * weight; history.w += old.w * weight;
            metadata.x += meta.x * weight; metadata.y += meta.y * weight;
            metadata.z += meta.z * weight; metadata.w += meta.w * weight;
            total += weight;
        }
    }
    const float capacity = fmaxf(P.history_frames, 0.0f);
    float3 sbg = bg_sup ? make_float3(bgs.x, bgs.y, bgs.z) : make_float3(0.0f, 0.0f, 0.0f);
    float bg_len = bg_sup ? (P.persistent ? bg_fresh : 1.0f) : 0.0f;
    if (bg_weight > 0.00001f) {
        bg_hist.x /= bg_weight; bg_hist.y /= bg_weight; bg_hist.z /= bg_weight; bg_hist.w /= bg_weight;
        const float old_len = fminf(bg_hist.w, capacity);
        if (bg_sup && bg_fresh > 0.00001f) {
            float w = fminf(gradual ? fminf(P.history_blend, 0.9f) : P.history_blend, old_len / (old_len + 1.0f));
            if (P.persistent) w = w / fmaxf(w + bg_fresh * (1.0f - w), 1e-8f);
            sbg = make_float3(mixf(sbg.x, bg_hist.x, w), mixf(sbg.y, bg_hist.y, w), mixf(sbg.z, bg_hist.z, w));
            bg_len = fminf(capacity, old_len + bg_fresh);
        } else {
            bg_len = gradual ? fmaxf(0.0f, old_len - fmaxf(capacity, 1.0f) / 120.0f) : old_len;
            sbg = bg_len > 0.0f ? make_float3(bg_hist.x, bg_hist.y, bg_hist.z) : make_float3(0.0f, 0.0f, 0.0f);
        }
    }
    store(out_bg, pix, make_float4(sbg.x, sbg.y, sbg.z, bg_len));
    if (hard != 0u && bg_len > 0.0f) {
        center.x += sbg.x - (bg_sup ? bgs.x : 0.0f);
        center.y += sbg.y - (bg_sup ? bgs.y : 0.0f);
        center.z += sbg.z - (bg_sup ? bgs.z : 0.0f);
        if (!cur_sup && cur_finite) {
            cur_sup = true;
            cur_mask = 1u;
            out_mask = 1u;
            cur_fresh = fminf(bg_len, 1.0f);
        }
    }
    float retained = cur_sup ? (P.persistent ? cur_fresh : 1.0f) : 0.0f;
    float out_var = cur_sup ? lum(variance.x, variance.y, variance.z) : 0.0f;
    const float2 en = enc_normal(normal);
    if (total > 0.00001f) {
        history.x /= total; history.y /= total; history.z /= total; history.w /= total;
        metadata.x /= total; metadata.y /= total; metadata.z /= total; metadata.w /= total;
        out_mask = hist_mask;
        float old_len = fminf(metadata.z, capacity);
        if (gradual && cur_fresh <= 0.00001f) {
            old_len = fmaxf(0.0f, old_len - fmaxf(capacity, 1.0f) / 120.0f);
            if (old_len <= 0.0f) {
                store(out_irr, pix, zero);
                store(out_meta, pix, zero);
                store(out_recv, pix, with_support(receiver, 0u));
                return;
            }
        }
        if (cur_sup) {
            if (cur_fresh <= 0.00001f) {
                center.x = history.x; center.y = history.y; center.z = history.z;
                store(out_irr, pix, center);
                store(out_meta, pix, make_float4(en.x, en.y, old_len, metadata.w));
                store(out_recv, pix, with_support(receiver, out_mask));
                return;
            }
            const bool complete = !P.persistent || cur_fresh >= 1.0f;
            float3 clipped = make_float3(history.x, history.y, history.z);
            if (count >= 2.0f && complete) {
                const float dtc = lum(history.x, history.y, history.z) - lum(center.x, center.y, center.z);
                const float stv = sqrtf(fmaxf(metadata.w, dtc * dtc * 0.25f));
                const float rx = 2.5f * fmaxf(deviation.x, stv) + 0.0001f;
                const float ry = 2.5f * fmaxf(deviation.y, stv) + 0.0001f;
                const float rz = 2.5f * fmaxf(deviation.z, stv) + 0.0001f;
                clipped.x = fminf(fmaxf(clipped.x, fmaxf(0.0f, mean.x - rx)), mean.x + rx);
                clipped.y = fminf(fmaxf(clipped.y, fmaxf(0.0f, mean.y - ry)), mean.y + ry);
                clipped.z = fminf(fmaxf(clipped.z, fmaxf(0.0f, mean.z - rz)), mean.z + rz);
            }
            const float hl = lum(history.x, history.y, history.z), cl = lum(center.x, center.y, center.z);
            const float delta = hl - cl;
            const float clum = lum(clipped.x, clipped.y, clipped.z);
            const float change = fabsf(clum - cl) / fmaxf(fmaxf(clum, cl), 0.001f);
            const float reactive = !gradual && P.stability[1] > 0.0f && complete ? 1.0f - change * change : 1.0f;
            const float ref = fmaxf(fmaxf(hl, cl), 0.001f);
            const float rel_noise = sqrtf(fmaxf(metadata.w, 0.0f)) / ref;
            const float thr = fminf(fmaxf(3.0f * rel_noise, 0.05f), 0.25f);
            const bool step = cur_fresh > 0.0f && complete && !gradual && P.stability[1] > 0.5f && change > thr;
            if (step) old_len = 0.0f;
            float w = fminf(gradual ? fminf(P.history_blend, 0.9f) : P.history_blend, old_len / (old_len + 1.0f));
            if (P.persistent) w = w / fmaxf(w + cur_fresh * (1.0f - w), 1e-8f);
            w *= reactive;
            if (step) w = 0.0f;
            center.x = mixf(center.x, clipped.x, w);
            center.y = mixf(center.y, clipped.y, w);
            center.z = mixf(center.z, clipped.z, w);
            retained = fminf(capacity, old_len + cur_fresh);
            out_var = mixf(lum(variance.x, variance.y, variance.z), metadata.w, w) + w * (1.0f - w) * delta * delta;
        } else {
            center.x = history.x; center.y = history.y; center.z = history.z;
            retained = old_len;
            out_var = metadata.w;
        }
    }
    center.w = cur_sup || retained > 0.0f ? d : 0.0f;
    store(out_irr, pix, center);
    store(out_meta, pix, make_float4(en.x, en.y, retained, out_var));
    store(out_recv, pix, with_support(receiver, out_mask));
}

// the 5-tap binomial kernel {1, 4, 6, 4, 1} at offset -2..2
__device__ __forceinline__ float binomial(int o) { return o == 0 ? 6.0f : (o == 2 || o == -2 ? 1.0f : 4.0f); }

__device__ __forceinline__ float pow32(float x) {
    const float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
    return x16 * x16;
}

// Where a denoise tap reads its neighbour: the images themselves, or a tile of
// them staged in LDS (the images' values, so the taps compute the same).
struct ImageTaps {
    const ScreenArgs* a;
    const float4* irr;
    const float4* mom;
    const float4* recv;
    __device__ __forceinline__ float4 value(int ox, int oy) const { return irr[(uint64_t)oy * a->rw + ox]; }
    __device__ __forceinline__ float2 moment(int ox, int oy) const {
        const float4 m = mom[(uint64_t)oy * a->rw + ox];
        return make_float2(m.z, m.w);
    }
    __device__ __forceinline__ float4 receiver(int ox, int oy) const { return recv[(uint64_t)oy * a->rw + ox]; }
    __device__ __forceinline__ float3 normal(int, int, float4 orr) const { return recv_normal(orr); }
    __device__ __forceinline__ float3 position(int ox, int oy, float w) const {
        uint32_t ofx, ofy;
        full_of(*a, ox, oy, ofx, ofy);
        return world_pos(a->p.world_from_clip, a->p, (float)ofx + 0.5f, (float)ofy + 0.5f, w);
    }
};
// A tile's taps: the images' values, with each texel's receiver normal and
// world position decoded once when the tile is staged (the same expressions
// as ImageTaps evaluates).
struct TileTaps {
    ImageTaps image;
    const float4* geo;   // decoded normal x, y, z; world position x
    const float2* geo2;  // world position y, z
    int x0, y0, fw;      // image position of the tile's first slot, slots per row
    __device__ __forceinline__ int at(int ox, int oy) const { return (oy - y0) * fw + (ox - x0); }
    __device__ __forceinline__ float4 value(int ox, int oy) const { return image.value(ox, oy); }
    __device__ __forceinline__ float2 moment(int ox, int oy) const { return image.moment(ox, oy); }
    __device__ __forceinline__ float4 receiver(int ox, int oy) const { return image.receiver(ox, oy); }
    __device__ __forceinline__ float3 normal(int ox, int oy, float4) const {
        const float4 g = geo[at(ox, oy)];
        return make_float3(g.x, g.y, g.z);
    }
    __device__ __forceinline__ float3 position(int ox, int oy, float) const {
        const int i = at(ox, oy);
        return make_float3(geo[i].w, geo2[i].x, geo2[i].y);
    }
};

// One denoise pass at resolve texel (tx, ty) (shaders/cascades.slang:624-711).
template <class Taps>
__device__ __forceinline__ void denoise_texel(const ScreenArgs& a, int tx, int ty, int step_x, int step_y,
                                              const Taps& taps, const float4* __restrict__ in_irr,
                                              const float4* __restrict__ moments,
                                              const float4* __restrict__ in_recv, const float* __restrict__ depth,
                                              float4* __restrict__ out, float4* __restrict__ out_recv) {
    const lux_screen_params& P = a.p;
    const uint32_t RW = a.rw, RH = a.rh;
    const uint64_t pix = (uint64_t)ty * RW + tx;
    const float4 center = in_irr[pix];
    const bool has_light = center.w > 0.0f && fin(center.w) && fin(center.x) && fin(center.y) && fin(center.z);
    uint32_t fx, fy;
    const uint64_t full = full_of(a, tx, ty, fx, fy);
    const float d = has_light ? center.w : depth[full];
    if (d <= 0.0f || !fin(d)) {
        out[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        out_recv[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float4 receiver = in_recv[pix];
    const uint32_t cmask = support_mask(receiver);
    const float3 gn = recv_normal(receiver);
    const float3 position = world_pos(P.world_from_clip, P, (float)fx + 0.5f, (float)fy + 0.5f, d);
    const float4 mom = moments[pix];
    const float3 normal = gn;
    const bool csup = mom.z > 0.0f;
    const float variance = fmaxf(mom.w, 0.0f);
    const float tol = surface_tol(P);
    const float falloff = fmaxf(P.normal_bias * 2.0f, P.probe_spacing * 0.05f);
    const float center_lum = lum(center.x, center.y, center.z);
    float3 sum = make_float3(0.0f, 0.0f, 0.0f);
    float total = 0.0f;
    for (int y = -2; y <= 2; ++y) {
        const int xr = (y == 2 || y == -2) ? 1 : 2;
        for (int x = -xr; x <= xr; ++x) {
            const int ox = tx + x * step_x, oy = ty + y * step_y;
            if (ox < 0 || oy < 0 || ox >= (int)RW || oy >= (int)RH) continue;
            const float4 value = taps.value(ox, oy);
            if (value.w <= 0.0f) continue;
            const float2 om = taps.moment(ox, oy);
            if (om.x <= 0.0f) continue;
            const float4 orr = taps.receiver(ox, oy);
            if (!recv_same(receiver, orr)) continue;
            const float3 on = taps.normal(ox, oy, orr);
            if (!(dot3(gn, on) > 0.97f)) continue;
            const uint32_t omask = support_mask(orr);
            if (cmask != 0u && omask != cmask) continue;
            if (cmask == 0u && omask == 0u) continue;
            const float3 opos = taps.position(ox, oy, value.w);
            if (dot3(gn, on) < 0.95f) continue;
            const float pd = plane_dist(sub3(opos, position), gn, on);
            if (pd > tol) continue;
            float w = binomial(x) * binomial(y) * pow32(fmaxf(0.0f, dot3(normal, on))) * expf(-pd / falloff);
            if (csup) {
                const float scale = fmaxf(0.02f, 2.0f * sqrtf(fmaxf(variance + om.y, 0.0f)));
                w *= expf(-fabsf(lum(value.x, value.y, value.z) - center_lum) / scale);
            }
            sum.x += value.x * w; sum.y += value.y * w; sum.z += value.z * w;
            total += w;
        }
    }
    const bool fsup = total > 0.00001f;
    float4 o;
    o.x = fsup ? sum.x / total : center.x;
    o.y = fsup ? sum.y / total : center.y;
    o.z = fsup ? sum.z / total : center.z;
    o.w = has_light || csup || fsup ? d : 0.0f;
    out[pix] = o;
    out_recv[pix] = with_support(receiver, o.w > 0.0f ? 1u : 0u);
}

// Taps straight from the images: 8 x 8 tiles, one wave each (any step).
__global__ __launch_bounds__(64) void k_screen_denoise(ScreenArgs a, int step_x, int step_y,
                                                       const float4* __restrict__ in_irr,
                                                       const float4* __restrict__ moments,
                                                       const float4* __restrict__ in_recv,
                                                       const float* __restrict__ depth, float4* __restrict__ out,
                                                       float4* __restrict__ out_recv) {
    uint32_t bx, by;
    xcd_tile(bx, by);
    const int tx = (int)(bx * kTile + threadIdx.x % kTile);
    const int ty = (int)(by * kTile + threadIdx.x / kTile);
    if (tx >= (int)a.rw || ty >= (int)a.rh) return;
    const ImageTaps taps{&a, in_irr, moments, in_recv};
    denoise_texel(a, tx, ty, step_x, step_y, taps, in_irr, moments, in_recv, depth, out, out_recv);
}

// Taps with decoded geometry in LDS: a 16 x 16 tile of texels per workgroup
// decodes the receiver normal and the world position (at the texel's
// irradiance depth) of every texel of its footprint -- the tile and two steps
// on each side -- once, 24 bytes per texel, so a texel is decoded once per
// workgroup instead of by each of its up to 21 taps; the taps read the
// images' values (L2-resident).  Dynamic LDS of 24 bytes per slot.
constexpr int kDTile = 16;
__global__ __launch_bounds__(256) void k_screen_denoise_tile(ScreenArgs a, int step_x, int step_y,
                                                             const float4* __restrict__ in_irr,
                                                             const float4* __restrict__ moments,
                                                             const float4* __restrict__ in_recv,
                                                             const float* __restrict__ depth,
                                                             float4* __restrict__ out, float4* __restrict__ out_recv) {
    extern __shared__ float4 lds_taps[];
    uint32_t bx, by;
    xcd_tile(bx, by);
    const int fw = kDTile + 4 * step_x, n = fw * (kDTile + 4 * step_y);
    const int x0 = (int)bx * kDTile - 2 * step_x, y0 = (int)by * kDTile - 2 * step_y;
    float4* geo = lds_taps;
    float2* geo2 = reinterpret_cast<float2*>(geo + n);
    const ImageTaps image{&a, in_irr, moments, in_recv};
    for (int i = (int)threadIdx.x; i < n; i += 256) {
        const int ox = x0 + i % fw, oy = y0 + i / fw;
        if (ox >= 0 && oy >= 0 && ox < (int)a.rw && oy < (int)a.rh) {
            const float w = image.value(ox, oy).w;
            const float3 nn = image.normal(ox, oy, image.receiver(ox, oy));
            const float3 q = w > 0.0f ? image.position(ox, oy, w) : make_float3(0.0f, 0.0f, 0.0f);
            geo[i] = make_float4(nn.x, nn.y, nn.z, q.x);
            geo2[i] = make_float2(q.y, q.z);
        }
    }
    __syncthreads();
    const int tx = (int)bx * kDTile + (int)(threadIdx.x % kDTile);
    const int ty = (int)by * kDTile + (int)(threadIdx.x / kDTile);
    if (tx >= (int)a.rw || ty >= (int)a.rh) return;
    const TileTaps taps{image, geo, geo2, x0, y0, fw};
    denoise_texel(a, tx, ty, step_x, step_y, taps, in_irr, moments, in_recv, depth, out, out_recv);
}

int screen_args(lux_ctx* ctx, const lux_screen_params* p, const char* who, ScreenArgs* a) {
    if (!p) return lux_fail(ctx, LUX_ERR_INVALID, std::string(who) + ": null parameters");
    if (p->viewport[2] == 0 || p->viewport[3] == 0)
        return lux_fail(ctx, LUX_ERR_INVALID, std::string(who) + ": empty viewport");
    if ((uint64_t)p->viewport[0] + p->viewport[2] > p->target[0] ||
        (uint64_t)p->viewport[1] + p->viewport[3] > p->target[1])
        return lux_fail(ctx, LUX_ERR_INVALID, std::string(who) + ": viewport outside the target image");
    if (!p->persistent && (p->size[0] == 0 || p->size[1] == 0))
        return lux_fail(ctx, LUX_ERR_INVALID, std::string(who) + ": empty receiver grid");
    a->p = *p;
    a->rs = p->resolve_scale > 1u ? p->resolve_scale : 1u;
    a->rw = (p->viewport[2] + a->rs - 1u) / a->rs;
    a->rh = (p->viewport[3] + a->rs - 1u) / a->rs;
    int st[2] = {1, 1};
    if (!p->persistent)
        for (int q = 0; q < 2; ++q) {
            const float r = std::nearbyint((float)p->viewport[2 + q] / ((float)p->size[q] * (float)a->rs));
            const uint32_t u = (uint32_t)r;
            st[q] = (int)(u > 1u ? u : 1u);
        }
    a->stride_x = st[0];
    a->stride_y = st[1];
    return LUX_OK;
}

}  // namespace

extern "C" int lux_screen_extent(const lux_screen_params* p, uint32_t* width, uint32_t* height) {
    if (!p || !width || !height) return LUX_ERR_INVALID;
    const uint32_t s = p->resolve_scale > 1u ? p->resolve_scale : 1u;
    *width = (p->viewport[2] + s - 1u) / s;
    *height = (p->viewport[3] + s - 1u) / s;
    return LUX_OK;
}

extern "C" int lux_screen_temporal(lux_ctx* ctx, const lux_screen_params* p, const float* current,
                                   const float* current_background, const float* depth, const uint32_t* gbuffer,
                                   const uint32_t* raster, const float* motion, const float* previous,
                                   const float* previous_metadata, const float* previous_receiver,
                                   const float* previous_background, float* irradiance, float* metadata,
                                   float* receiver, float* background) {
    if (!ctx) return LUX_ERR_INVALID;
    ScreenArgs a;
    if (int rc = screen_args(ctx, p, "lux_screen_temporal", &a)) return rc;
    if (!current || !depth || !gbuffer || !raster || !previous || !previous_metadata || !previous_receiver ||
        !irradiance || !metadata || !receiver || !background)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_screen_temporal: null image");
    const bool has_motion = p->stability[2] > 0.0f && (!p->persistent || p->stability[1] > 0.5f);
    if (has_motion && !motion)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_screen_temporal: motion vectors needed (stability[2] > 0)");
    const float* ins[] = {current, current_background, previous, previous_metadata, previous_receiver,
                          previous_background};
    const float* outs[] = {irradiance, metadata, receiver, background};
    for (const float* o : outs)
        for (const float* i : ins)
            if (o == i) return lux_fail(ctx, LUX_ERR_INVALID, "lux_screen_temporal: outputs must not alias inputs");
    LUX_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = lux_ensure(ctx, ctx->screen_cache, 48ull * a.rw * a.rh)) return rc;
    float4* cache = static_cast<float4*>(ctx->screen_cache.ptr);
    const dim3 grid((a.rw + kTile - 1) / kTile, (a.rh + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_screen_cache_raster, grid, dim3(64), 0, ctx->stream, a, depth,
                       reinterpret_cast<const uint4*>(gbuffer), reinterpret_cast<const uint4*>(raster), cache);
    hipLaunchKernelGGL(k_screen_temporal, grid, dim3(64), 0, ctx->stream, a, reinterpret_cast<const float4*>(current),
                       reinterpret_cast<const float4*>(current_background), depth,
                       reinterpret_cast<const uint4*>(gbuffer), reinterpret_cast<const uint4*>(raster),
                       reinterpret_cast<const float2*>(motion), reinterpret_cast<const float4*>(previous),
                       reinterpret_cast<const float4*>(previous_metadata),
                       reinterpret_cast<const float4*>(previous_receiver),
                       reinterpret_cast<const float4*>(previous_background), cache,
                       reinterpret_cast<float4*>(irradiance),
                       reinterpret_cast<float4*>(metadata), reinterpret_cast<float4*>(receiver),
                       reinterpret_cast<float4*>(background));
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_screen_denoise(lux_ctx* ctx, const lux_screen_params* p, uint32_t pass, const float* irradiance,
                                  const float* moments, const float* receiver, const float* depth, float* out,
                                  float* out_receiver) {
    if (!ctx) return LUX_ERR_INVALID;
    ScreenArgs a;
    if (int rc = screen_args(ctx, p, "lux_screen_denoise", &a)) return rc;
    if (!irradiance || !moments || !receiver || !depth || !out || !out_receiver)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_screen_denoise: null image");
    if (pass > 8) return lux_fail(ctx, LUX_ERR_INVALID, "lux_screen_denoise: pass index above 8");
    const float* ins[] = {irradiance, moments, receiver};
    for (const float* i : ins)
        if (i == out || i == out_receiver)
            return lux_fail(ctx, LUX_ERR_INVALID, "lux_screen_denoise: outputs must not alias inputs");
    // step: 1 on pass 0, else max(round(cell stride / 4), 1) * 2^pass
    int step[2];
    const int cs[2] = {a.stride_x, a.stride_y};
    for (int q = 0; q < 2; ++q) {
        uint32_t b = (uint32_t)std::nearbyint((float)cs[q] * 0.25f);
        if (b < 1u) b = 1u;
        step[q] = pass == 0 ? 1 : (int)(b * (1u << pass));
    }
    LUX_HIP(ctx, hipSetDevice(ctx->device));
    // the staged tile while its footprint fits 64 KB of LDS (steps up to 8)
    const uint64_t lds = (kDTile + 4ull * (uint64_t)step[0]) * (kDTile + 4ull * (uint64_t)step[1]) * 24ull;
    if (lds <= 65536u) {
        const dim3 tgrid((a.rw + kDTile - 1) / kDTile, (a.rh + kDTile - 1) / kDTile);
        hipLaunchKernelGGL(k_screen_denoise_tile, tgrid, dim3(256), (size_t)lds, ctx->stream, a, step[0], step[1],
                           reinterpret_cast<const float4*>(irradiance), reinterpret_cast<const float4*>(moments),
                           reinterpret_cast<const float4*>(receiver), depth, reinterpret_cast<float4*>(out),
                           reinterpret_cast<float4*>(out_receiver));
        LUX_HIP(ctx, hipGetLastError());
        return LUX_OK;
    }
    const dim3 grid((a.rw + kTile - 1) / kTile, (a.rh + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_screen_denoise, grid, dim3(64), 0, ctx->stream, a, step[0], step[1],
                       reinterpret_cast<const float4*>(irradiance), reinterpret_cast<const float4*>(moments),
                       reinterpret_cast<const float4*>(receiver), depth, reinterpret_cast<float4*>(out),
                       reinterpret_cast<float4*>(out_receiver));
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

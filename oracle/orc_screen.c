/*
 * orc_screen.c -- CPU oracle of the screen-space temporal and denoise passes
 * (TEST INFRASTRUCTURE; the checker, never the product).
 *
 * Restates, in f32 with no contraction:
 *   temporal   shaders/cascades.slang:310-622 (without recordTemporalDiagnostics,
 *              :283-308, which only counts raw-output debug bins)
 *   denoise    shaders/cascades.slang:624-711
 *   helpers    splitRasterMetadata / splitReceiverCompatible (cascades.slang:271-282),
 *              shaders/receiver_metadata.slang (packed normal, history length,
 *              support bit, receiverSame, receiverGeometricNormal),
 *              shaders/common.slang: encodeNormal / decodeNormal (:361-375),
 *              luminance (:277), dirtyContains (:328), worldPosition /
 *              previousWorldPosition (:542-562), surfaceTolerance /
 *              surfacePlaneDistance (:583-591), giResolveScale / Extent (:95-100),
 *              deferredShadingNormal (:343-345)
 * Restatement choices (the same in the HIP kernels): matrices row-major with
 * h_i = ((m_i0 x + m_i1 y) + m_i2 z) + m_i3 w; dot products summed x, y, z in
 * order; lerp(a, b, t) = a (1 - t) + b t; round = nearest-even; normalize =
 * v / sqrt(v.v); rsqrt = 1 / sqrt; pow(x, 32) = five squarings; exp = the C
 * library's expf (the GPU's differs by ulps: tests compare with a tolerance).
 *
 * Pinned by: the temporal and denoise fixture of
 * src/transport_tests.rs:1235-1530, 854-1089 (tests/test_oracle_screen.py).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "lux_oracle.h"

typedef struct { float x, y, z, w; } f4;
typedef struct { float x, y, z; } f3;

static inline uint32_t fb(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline int fin(float f) { return (fb(f) & 0x7f800000u) != 0x7f800000u; }
static inline float sat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
static inline float frc(float x) { return x - floorf(x); }
static inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static inline f3 v3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline f3 sub3(f3 a, f3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 norm3(f3 a) {
    float l = sqrtf(dot3(a, a));
    return v3(a.x / l, a.y / l, a.z / l);
}
static inline float lum(float r, float g, float b) { return (r * 0.2126f + g * 0.7152f) + b * 0.0722f; }
static inline f4 ld4(const float* a, uint64_t i) {
    f4 r = {a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]};
    return r;
}
static inline void st4(float* a, uint64_t i, f4 v) {
    a[4 * i] = v.x; a[4 * i + 1] = v.y; a[4 * i + 2] = v.z; a[4 * i + 3] = v.w;
}
static inline f4 zero4(void) { f4 z = {0, 0, 0, 0}; return z; }

static void enc_normal(f3 n, float p[2]) {
    float s = (fabsf(n.x) + fabsf(n.y)) + fabsf(n.z);
    n = v3(n.x / s, n.y / s, n.z / s);
    float px = n.x, py = n.y;
    if (n.z < 0.0f) {
        float qx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        float qy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = qx;
        py = qy;
    }
    p[0] = px * 0.5f + 0.5f;
    p[1] = py * 0.5f + 0.5f;
}

static f3 dec_normal(float u, float v) {
    float px = u * 2.0f - 1.0f, py = v * 2.0f - 1.0f;
    f3 n = v3(px, py, (1.0f - fabsf(px)) - fabsf(py));
    float t = sat(-n.z);
    n.x += n.x >= 0.0f ? -t : t;
    n.y += n.y >= 0.0f ? -t : t;
    return norm3(n);
}

/* receiverNormal / receiverPackedNormal (receiver_metadata.slang:2-9) */
static f3 recv_normal(f4 m) {
    uint32_t packed = (uint32_t)m.z;
    return dec_normal((float)(packed & 4095u) / 4095.0f, (float)(packed >> 12) / 4095.0f);
}
static float packed_normal(f3 n) {
    float p[2];
    enc_normal(n, p);
    uint32_t ox = (uint32_t)rintf(sat(p[0]) * 4095.0f), oy = (uint32_t)rintf(sat(p[1]) * 4095.0f);
    return (float)(ox | (oy << 12));
}
static uint32_t hist_len(f4 m) { return (uint32_t)m.w & 255u; }
static uint32_t support_mask(f4 m) { return ((uint32_t)m.w >> 16) & 1u; }
static f4 with_support(f4 m, uint32_t s) {
    uint32_t packed = (uint32_t)m.w;
    packed = (packed & ~(1u << 16)) | ((s & 1u) << 16);
    m.w = (float)packed;
    return m;
}
static int recv_same(f4 a, f4 b) {
    return hist_len(a) > 0u && hist_len(b) > 0u && fb(a.x) == fb(b.x) && fb(a.y) == fb(b.y);
}
static int recv_compatible(f4 a, f4 b) { return recv_same(a, b) && dot3(recv_normal(a), recv_normal(b)) > 0.97f; }

/* receiverGeometricNormal (receiver_metadata.slang:10-23) */
static f3 geometric_normal(uint32_t packed, f3 fallback) {
    fallback = norm3(fallback);
    if ((packed & 0x80000000u) == 0u) return fallback;
    float ox = (float)(packed & 0x7fffu) / 32767.0f * 2.0f - 1.0f;
    float oy = (float)((packed >> 15) & 0x7fffu) / 32767.0f * 2.0f - 1.0f;
    f3 n = v3(ox, oy, (1.0f - fabsf(ox)) - fabsf(oy));
    if (n.z < 0.0f) {
        float nx = (1.0f - fabsf(n.y)) * (n.x >= 0.0f ? 1.0f : -1.0f);
        float ny = (1.0f - fabsf(n.x)) * (n.y >= 0.0f ? 1.0f : -1.0f);
        n.x = nx;
        n.y = ny;
    }
    n = norm3(n);
    return dot3(n, fallback) < 0.0f ? v3(-n.x, -n.y, -n.z) : n;
}

/* splitRasterMetadata (cascades.slang:271-278): raster identity, packed
 * geometric normal, history length 1. */
static f4 raster_meta(const orc_screen_params* P, const uint32_t* gbuffer, const uint32_t* raster, uint64_t full) {
    uint32_t gw = gbuffer[4 * full + 3];
    f3 shading = dec_normal((float)(gw & 4095u) / 4095.0f, (float)((gw >> 12) & 4095u) / 4095.0f);
    f3 g = geometric_normal(raster[4 * full + 2], shading);
    f4 r = {bf(raster[4 * full]), bf(raster[4 * full + 1]), packed_normal(g), 1.0f};
    (void)P;
    return r;
}

static f3 mat_point(const float* M, float x, float y, float z) {
    float h[4];
    for (int i = 0; i < 4; ++i) h[i] = ((M[4 * i] * x + M[4 * i + 1] * y) + M[4 * i + 2] * z) + M[4 * i + 3] * 1.0f;
    return v3(h[0] / h[3], h[1] / h[3], h[2] / h[3]);
}

static f3 world_pos(const float* M, const orc_screen_params* P, float px, float py, float depth) {
    float u = (px - (float)P->viewport[0]) / (float)P->viewport[2];
    float v = (py - (float)P->viewport[1]) / (float)P->viewport[3];
    return mat_point(M, u * 2.0f - 1.0f, 1.0f - v * 2.0f, depth);
}

static float surface_tol(const orc_screen_params* P) { return fmaxf(P->normal_bias * 2.0f, P->probe_spacing * 0.02f); }

static float plane_dist(f3 d, f3 tn, f3 sn) {
    f3 a = v3(tn.x + sn.x, tn.y + sn.y, tn.z + sn.z);
    float l2 = dot3(a, a);
    if (!(l2 > 1e-20f) || !fin(l2)) return fabsf(dot3(d, tn));
    float r = 1.0f / sqrtf(l2);
    return fabsf(dot3(d, v3(a.x * r, a.y * r, a.z * r)));
}

static int dirty(const orc_screen_params* P, f3 p) {
    return P->dirty_min[3] > 0.0f && p.x >= P->dirty_min[0] && p.y >= P->dirty_min[1] && p.z >= P->dirty_min[2] &&
           p.x <= P->dirty_max[0] && p.y <= P->dirty_max[1] && p.z <= P->dirty_max[2];
}

static uint32_t rscale(const orc_screen_params* P) { return P->resolve_scale > 1u ? P->resolve_scale : 1u; }

void orc_screen_extent(const orc_screen_params* P, uint32_t* w, uint32_t* h) {
    uint32_t s = rscale(P);
    *w = (P->viewport[2] + s - 1u) / s;
    *h = (P->viewport[3] + s - 1u) / s;
}

/* the full-resolution representative of a resolve texel */
static uint64_t full_of(const orc_screen_params* P, int tx, int ty, uint32_t* fx, uint32_t* fy) {
    uint32_t s = rscale(P);
    uint32_t lx = (uint32_t)tx * s + s / 2u, ly = (uint32_t)ty * s + s / 2u;
    if (lx > P->viewport[2] - 1u) lx = P->viewport[2] - 1u;
    if (ly > P->viewport[3] - 1u) ly = P->viewport[3] - 1u;
    *fx = P->viewport[0] + lx;
    *fy = P->viewport[1] + ly;
    return (uint64_t)*fy * P->target[0] + *fx;
}

static void cell_stride(const orc_screen_params* P, int st[2]) {
    if (P->persistent) {
        st[0] = st[1] = 1;
        return;
    }
    float s = (float)rscale(P);
    for (int a = 0; a < 2; ++a) {
        float r = rintf((float)P->viewport[2 + a] / ((float)P->size[a] * s));
        uint32_t u = (uint32_t)r;
        st[a] = (int)(u > 1u ? u : 1u);
    }
}

static void temporal_pixel(const orc_screen_params* P, int tx, int ty, uint32_t RW, uint32_t RH,
                           const float* current, const float* current_bg, const float* depth, const uint32_t* gbuffer,
                           const uint32_t* raster, const float* motion, const float* prev, const float* prev_meta,
                           const float* prev_recv, const float* prev_bg, float* out_irr, float* out_meta,
                           float* out_recv, float* out_bg) {
    const uint64_t pix = (uint64_t)ty * RW + tx;
    const uint32_t RS = rscale(P);
    f4 center = ld4(current, pix);
    uint32_t fx, fy;
    const uint64_t full = full_of(P, tx, ty, &fx, &fy);
    const f4 receiver = raster_meta(P, gbuffer, raster, full);
    const float d = depth[full];
    if (d <= 0.0f || hist_len(receiver) == 0u) {
        st4(out_bg, pix, zero4());
        st4(out_irr, pix, zero4());
        st4(out_meta, pix, zero4());
        st4(out_recv, pix, zero4());
        return;
    }
    const int cur_finite = fin(center.w) && fin(center.x) && fin(center.y) && fin(center.z);
    const f4 bgs = current_bg ? ld4(current_bg, pix) : zero4();
    const int bg_sup = bgs.w > 0.0f && fin(bgs.x) && fin(bgs.y) && fin(bgs.z) && fin(bgs.w);
    const float bg_fresh = bg_sup ? sat(frc(bgs.w) * 4.0f) : 0.0f;
    uint32_t cur_mask = cur_finite && center.w > 0.0f ? ((uint32_t)floorf(center.w)) & 1u : 0u;
    int cur_sup = cur_mask != 0u && cur_finite;
    float cur_fresh = cur_finite && cur_sup ? sat(frc(center.w) * 4.0f) : 0.0f;
    if (!cur_sup && cur_finite && center.w < 0.0f && bg_sup) {
        cur_sup = 1;
        cur_mask = 1u;
        cur_fresh = bg_fresh;
    }
    if (!cur_sup) center.x = center.y = center.z = 0.0f;
    center.w = d;
    uint32_t out_mask = cur_mask;
    const f3 gn = recv_normal(receiver);
    const f3 normal = gn;
    const f3 position = world_pos(P->world_from_clip, P, (float)fx + 0.5f, (float)fy + 0.5f, d);
    const float tol = surface_tol(P);
    float mean[3] = {0, 0, 0}, square[3] = {0, 0, 0}, count = 0.0f;
    int cs[2];
    cell_stride(P, cs);
    for (int y = -1; y <= 1; ++y)
        for (int x = -1; x <= 1; ++x) {
            int nx = tx + x * cs[0], ny = ty + y * cs[1];
            if (nx < 0 || ny < 0 || nx >= (int)RW || ny >= (int)RH) continue;
            f4 value = ld4(current, (uint64_t)ny * RW + nx);
            if (value.w <= 0.0f) continue;
            uint32_t nmask = ((uint32_t)floorf(value.w)) & 1u;
            if (cur_mask != 0u && nmask != cur_mask) continue;
            uint32_t ofx, ofy;
            uint64_t of = full_of(P, nx, ny, &ofx, &ofy);
            float od = depth[of];
            if (od <= 0.0f) continue;
            f4 orc = raster_meta(P, gbuffer, raster, of);
            if (!recv_compatible(receiver, orc)) continue;
            f3 on = recv_normal(orc);
            f3 op = world_pos(P->world_from_clip, P, (float)ofx + 0.5f, (float)ofy + 0.5f, od);
            if (dot3(gn, on) < 0.95f || plane_dist(sub3(op, position), gn, on) > tol) continue;
            mean[0] += value.x; mean[1] += value.y; mean[2] += value.z;
            square[0] += value.x * value.x; square[1] += value.y * value.y; square[2] += value.z * value.z;
            count += 1.0f;
        }
    const float cdiv = fmaxf(count, 1.0f);
    float variance[3], deviation[3];
    for (int c = 0; c < 3; ++c) {
        mean[c] /= cdiv;
        square[c] /= cdiv;
        variance[c] = fmaxf(0.0f, square[c] - mean[c] * mean[c]);
        deviation[c] = sqrtf(variance[c]);
    }
    f4 history = zero4(), metadata = zero4(), bg_hist = zero4();
    float total = 0.0f, bg_weight = 0.0f;
    uint32_t hist_mask = 0u;
    float pfx = 0.0f, pfy = 0.0f;
    int projected = 0;
    const int has_motion = P->stability[2] > 0.0f && (!P->persistent || P->stability[1] > 0.5f);
    if (has_motion) {
        float mx = motion[2 * full], my = motion[2 * full + 1];
        pfx = (float)fx + 0.5f - mx * (float)P->viewport[2];
        pfy = (float)fy + 0.5f - my * (float)P->viewport[3];
        projected = fin(mx) && fin(my);
    } else {
        const float* M = P->previous_clip_from_world;
        float h[4];
        for (int i = 0; i < 4; ++i)
            h[i] = ((M[4 * i] * position.x + M[4 * i + 1] * position.y) + M[4 * i + 2] * position.z) + M[4 * i + 3];
        float ndx = h[0] / h[3], ndy = h[1] / h[3];
        float u = ndx * 0.5f + 0.5f, v = 0.5f - ndy * 0.5f;
        pfx = (float)P->viewport[0] + u * (float)P->viewport[2];
        pfy = (float)P->viewport[1] + v * (float)P->viewport[3];
        projected = h[3] > 0.0f && fin(u) && fin(v);
    }
    const int gradual = (P->source_flags & 1u) != 0u;
    const uint32_t hard = P->source_changes & ~(gradual ? 1u : 0u);
    const int motion_expired = (P->source_flags & 2u) != 0u && P->frame - P->motion_epoch > 120u;
    if (projected && P->stability[0] > 0.0f && (hard & 1u) == 0u && !dirty(P, position) && P->history_blend > 0.0f &&
        P->history_frames > 0.0f && pfx >= (float)P->viewport[0] && pfy >= (float)P->viewport[1] &&
        pfx < (float)(P->viewport[0] + P->viewport[2]) && pfy < (float)(P->viewport[1] + P->viewport[3])) {
        const float lx = pfx - (float)P->viewport[0] - 0.5f, ly = pfy - (float)P->viewport[1] - 0.5f;
        const int bx = (int)floorf((lx - (float)(RS / 2u)) / (float)RS);
        const int by = (int)floorf((ly - (float)(RS / 2u)) / (float)RS);
        const float lastx = (float)(P->viewport[2] - 1u), lasty = (float)(P->viewport[3] - 1u);
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
            const f4 old = ld4(prev, op), meta = ld4(prev_meta, op), old_r = ld4(prev_recv, op);
            if (meta.z <= 0.0f || old.w <= 0.0f || !(fin(old.x) && fin(old.y) && fin(old.z)) ||
                dot3(normal, dec_normal(meta.x, meta.y)) < 0.95f || !recv_compatible(receiver, old_r))
                continue;
            uint32_t ofx, ofy;
            full_of(P, ox, oy, &ofx, &ofy);
            const f3 old_pos = world_pos(P->previous_world_from_clip, P, (float)ofx + 0.5f, (float)ofy + 0.5f, old.w);
            if (dirty(P, old_pos)) continue;
            if (has_motion) {
                const float* M = P->previous_clip_from_world;
                float h[4];
                for (int i = 0; i < 4; ++i)
                    h[i] = ((M[4 * i] * position.x + M[4 * i + 1] * position.y) + M[4 * i + 2] * position.z) +
                           M[4 * i + 3];
                const float expected = h[2] / h[3];
                if (!fin(expected) || expected <= 0.0f) continue;
                float t = fmaxf(0.000002f, fabsf(expected) * 0.0001f);
                for (int axis = 0; axis < 2; ++axis) {
                    int nx = ox + (axis == 0 ? 1 : 0), ny = oy + (axis == 0 ? 0 : 1);
                    if (nx > (int)RW - 1) nx = (int)RW - 1;
                    if (ny > (int)RH - 1) ny = (int)RH - 1;
                    const uint64_t nb = (uint64_t)ny * RW + nx;
                    const f4 nr = ld4(prev_recv, nb);
                    const float nd = prev[4 * nb + 3];
                    if (recv_compatible(old_r, nr) && nd > 0.0f) t += 2.0f * fabsf(nd - old.w);
                }
                if (fabsf(expected - old.w) > t) continue;
            } else {
                if (plane_dist(sub3(old_pos, position), gn, recv_normal(old_r)) > tol) continue;
            }
            const float wx = bitx ? frx : 1.0f - frx, wy = bity ? fry : 1.0f - fry;
            const float weight = wx * wy;
            if (weight <= 0.0f) continue;
            if (support_mask(old_r) == 0u) continue;
            const f4 obg = prev_bg ? ld4(prev_bg, op) : zero4();
            if (P->dirty_min[3] <= 0.0f && obg.w > 0.0f && !(motion_expired && bg_fresh <= 0.00001f) && fin(obg.x) &&
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
    const float capacity = fmaxf(P->history_frames, 0.0f);
    float sbg[3] = {bg_sup ? bgs.x : 0.0f, bg_sup ? bgs.y : 0.0f, bg_sup ? bgs.z : 0.0f};
    float bg_len = bg_sup ? (P->persistent ? bg_fresh : 1.0f) : 0.0f;
    if (bg_weight > 0.00001f) {
        bg_hist.x /= bg_weight; bg_hist.y /= bg_weight; bg_hist.z /= bg_weight; bg_hist.w /= bg_weight;
        const float old_len = fminf(bg_hist.w, capacity);
        if (bg_sup && bg_fresh > 0.00001f) {
            float w = fminf(gradual ? fminf(P->history_blend, 0.9f) : P->history_blend, old_len / (old_len + 1.0f));
            if (P->persistent) w = w / fmaxf(w + bg_fresh * (1.0f - w), 1e-8f);
            sbg[0] = mixf(sbg[0], bg_hist.x, w);
            sbg[1] = mixf(sbg[1], bg_hist.y, w);
            sbg[2] = mixf(sbg[2], bg_hist.z, w);
            bg_len = fminf(capacity, old_len + bg_fresh);
        } else {
            bg_len = gradual ? fmaxf(0.0f, old_len - fmaxf(capacity, 1.0f) / 120.0f) : old_len;
            sbg[0] = bg_len > 0.0f ? bg_hist.x : 0.0f;
            sbg[1] = bg_len > 0.0f ? bg_hist.y : 0.0f;
            sbg[2] = bg_len > 0.0f ? bg_hist.z : 0.0f;
        }
    }
    f4 ob = {sbg[0], sbg[1], sbg[2], bg_len};
    st4(out_bg, pix, ob);
    if (hard != 0u && bg_len > 0.0f) {
        center.x += sbg[0] - (bg_sup ? bgs.x : 0.0f);
        center.y += sbg[1] - (bg_sup ? bgs.y : 0.0f);
        center.z += sbg[2] - (bg_sup ? bgs.z : 0.0f);
        if (!cur_sup && cur_finite) {
            cur_sup = 1;
            cur_mask = 1u;
            out_mask = 1u;
            cur_fresh = fminf(bg_len, 1.0f);
        }
    }
    float retained = cur_sup ? (P->persistent ? cur_fresh : 1.0f) : 0.0f;
    float out_var = cur_sup ? lum(variance[0], variance[1], variance[2]) : 0.0f;
    float en[2];
    enc_normal(normal, en);
    if (total > 0.00001f) {
        history.x /= total; history.y /= total; history.z /= total; history.w /= total;
        metadata.x /= total; metadata.y /= total; metadata.z /= total; metadata.w /= total;
        out_mask = hist_mask;
        float old_len = fminf(metadata.z, capacity);
        if (gradual && cur_fresh <= 0.00001f) {
            old_len = fmaxf(0.0f, old_len - fmaxf(capacity, 1.0f) / 120.0f);
            if (old_len <= 0.0f) {
                st4(out_irr, pix, zero4());
                st4(out_meta, pix, zero4());
                st4(out_recv, pix, with_support(receiver, 0u));
                return;
            }
        }
        if (cur_sup) {
            if (cur_fresh <= 0.00001f) {
                center.x = history.x; center.y = history.y; center.z = history.z;
                f4 om = {en[0], en[1], old_len, metadata.w};
                st4(out_irr, pix, center);
                st4(out_meta, pix, om);
                st4(out_recv, pix, with_support(receiver, out_mask));
                return;
            }
            const int complete = !P->persistent || cur_fresh >= 1.0f;
            float clipped[3] = {history.x, history.y, history.z};
            if (count >= 2.0f && complete) {
                const float dtc = lum(history.x, history.y, history.z) - lum(center.x, center.y, center.z);
                const float tv = fmaxf(metadata.w, dtc * dtc * 0.25f);
                const float stv = sqrtf(tv);
                for (int c = 0; c < 3; ++c) {
                    const float radius = 2.5f * fmaxf(deviation[c], stv) + 0.0001f;
                    const float low = fmaxf(0.0f, mean[c] - radius), high = mean[c] + radius;
                    clipped[c] = fminf(fmaxf(clipped[c], low), high);
                }
            }
            const float hl = lum(history.x, history.y, history.z), cl = lum(center.x, center.y, center.z);
            const float delta = hl - cl;
            const float clum = lum(clipped[0], clipped[1], clipped[2]);
            const float change = fabsf(clum - cl) / fmaxf(fmaxf(clum, cl), 0.001f);
            const float reactive = !gradual && P->stability[1] > 0.0f && complete ? 1.0f - change * change : 1.0f;
            const float ref = fmaxf(fmaxf(hl, cl), 0.001f);
            const float rel_noise = sqrtf(fmaxf(metadata.w, 0.0f)) / ref;
            const float thr = fminf(fmaxf(3.0f * rel_noise, 0.05f), 0.25f);
            const int step = cur_fresh > 0.0f && complete && !gradual && P->stability[1] > 0.5f && change > thr;
            if (step) old_len = 0.0f;
            float w = fminf(gradual ? fminf(P->history_blend, 0.9f) : P->history_blend, old_len / (old_len + 1.0f));
            if (P->persistent) w = w / fmaxf(w + cur_fresh * (1.0f - w), 1e-8f);
            w *= reactive;
            if (step) w = 0.0f;
            center.x = mixf(center.x, clipped[0], w);
            center.y = mixf(center.y, clipped[1], w);
            center.z = mixf(center.z, clipped[2], w);
            retained = fminf(capacity, old_len + cur_fresh);
            out_var = mixf(lum(variance[0], variance[1], variance[2]), metadata.w, w) + w * (1.0f - w) * delta * delta;
        } else {
            center.x = history.x; center.y = history.y; center.z = history.z;
            retained = old_len;
            out_var = metadata.w;
        }
    }
    center.w = cur_sup || retained > 0.0f ? d : 0.0f;
    f4 om = {en[0], en[1], retained, out_var};
    st4(out_irr, pix, center);
    st4(out_meta, pix, om);
    st4(out_recv, pix, with_support(receiver, out_mask));
}

void orc_screen_temporal(const orc_screen_params* P, const float* current, const float* current_bg, const float* depth,
                         const uint32_t* gbuffer, const uint32_t* raster, const float* motion, const float* prev,
                         const float* prev_meta, const float* prev_recv, const float* prev_bg, float* out_irr,
                         float* out_meta, float* out_recv, float* out_bg) {
    uint32_t RW, RH;
    orc_screen_extent(P, &RW, &RH);
    for (uint32_t y = 0; y < RH; ++y)
        for (uint32_t x = 0; x < RW; ++x)
            temporal_pixel(P, (int)x, (int)y, RW, RH, current, current_bg, depth, gbuffer, raster, motion, prev,
                           prev_meta, prev_recv, prev_bg, out_irr, out_meta, out_recv, out_bg);
}

static float pow32(float x) {
    float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
    return x16 * x16;
}

void orc_screen_denoise(const orc_screen_params* P, uint32_t pass, const float* in_irr, const float* moments,
                        const float* in_recv, const float* depth, float* out, float* out_recv) {
    uint32_t RW, RH;
    orc_screen_extent(P, &RW, &RH);
    static const float kern[5] = {1.0f, 4.0f, 6.0f, 4.0f, 1.0f};
    int cs[2];
    cell_stride(P, cs);
    int stp[2];
    for (int a = 0; a < 2; ++a) {
        uint32_t b = (uint32_t)rintf((float)cs[a] * 0.25f);
        if (b < 1u) b = 1u;
        stp[a] = pass == 0u ? 1 : (int)(b * (1u << pass));
    }
    const float tol = surface_tol(P);
    const float falloff = fmaxf(P->normal_bias * 2.0f, P->probe_spacing * 0.05f);
    for (uint32_t ty = 0; ty < RH; ++ty)
        for (uint32_t tx = 0; tx < RW; ++tx) {
            const uint64_t pix = (uint64_t)ty * RW + tx;
            const f4 center = ld4(in_irr, pix);
            const int has_light = center.w > 0.0f && fin(center.w) && fin(center.x) && fin(center.y) && fin(center.z);
            uint32_t fx, fy;
            const uint64_t full = full_of(P, (int)tx, (int)ty, &fx, &fy);
            const float d = has_light ? center.w : depth[full];
            if (d <= 0.0f || !fin(d)) {
                st4(out, pix, zero4());
                st4(out_recv, pix, zero4());
                continue;
            }
            const f4 receiver = ld4(in_recv, pix);
            const uint32_t cmask = support_mask(receiver);
            const f3 gn = recv_normal(receiver);
            const f3 position = world_pos(P->world_from_clip, P, (float)fx + 0.5f, (float)fy + 0.5f, d);
            const f4 mom = ld4(moments, pix);
            const f3 normal = gn;
            const int csup = mom.z > 0.0f;
            const float variance = fmaxf(mom.w, 0.0f);
            float sum[3] = {0, 0, 0}, total = 0.0f;
            const float center_lum = lum(center.x, center.y, center.z);
            for (int y = -2; y <= 2; ++y) {
                const int xr = (y == 2 || y == -2) ? 1 : 2;
                for (int x = -xr; x <= xr; ++x) {
                    const int ox = (int)tx + x * stp[0], oy = (int)ty + y * stp[1];
                    if (ox < 0 || oy < 0 || ox >= (int)RW || oy >= (int)RH) continue;
                    const uint64_t op = (uint64_t)oy * RW + ox;
                    const f4 value = ld4(in_irr, op);
                    if (value.w <= 0.0f) continue;
                    const f4 om = ld4(moments, op);
                    if (om.z <= 0.0f) continue;
                    const f4 orr = ld4(in_recv, op);
                    if (!recv_compatible(receiver, orr)) continue;
                    const uint32_t omask = support_mask(orr);
                    if (cmask != 0u && omask != cmask) continue;
                    if (cmask == 0u && omask == 0u) continue;
                    const f3 on = recv_normal(orr);
                    uint32_t ofx, ofy;
                    full_of(P, ox, oy, &ofx, &ofy);
                    const f3 opos = world_pos(P->world_from_clip, P, (float)ofx + 0.5f, (float)ofy + 0.5f, value.w);
                    if (dot3(gn, on) < 0.95f) continue;
                    const float pd = plane_dist(sub3(opos, position), gn, on);
                    if (pd > tol) continue;
                    float w = kern[x + 2] * kern[y + 2] * pow32(fmaxf(0.0f, dot3(normal, on))) * expf(-pd / falloff);
                    if (csup) {
                        const float scale = fmaxf(0.02f, 2.0f * sqrtf(fmaxf(variance + om.w, 0.0f)));
                        w *= expf(-fabsf(lum(value.x, value.y, value.z) - center_lum) / scale);
                    }
                    sum[0] += value.x * w; sum[1] += value.y * w; sum[2] += value.z * w;
                    total += w;
                }
            }
            const int fsup = total > 0.00001f;
            f4 o;
            o.x = fsup ? sum[0] / total : center.x;
            o.y = fsup ? sum[1] / total : center.y;
            o.z = fsup ? sum[2] / total : center.z;
            o.w = has_light || csup || fsup ? d : 0.0f;
            st4(out, pix, o);
            st4(out_recv, pix, with_support(receiver, o.w > 0.0f ? 1u : 0u));
        }
}

float orc_screen_packed_normal(const float* n) { return packed_normal(v3(n[0], n[1], n[2])); }

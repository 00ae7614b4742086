/*
 * orc_transport.c -- oracle of the dense-grid interval trace, cascade merge,
 * probe packing and world resolve (TEST INFRASTRUCTURE, never the product).
 *
 * Restated from the reference (f32, no contraction, left-to-right sums):
 *   transportHash, r2, hashKey, cacheKey, rayOffset, decodeDirection,
 *   encodeNormal, decodeNormal, disneyDiffuseMultiplier
 *                           shaders/common.slang:282-429
 *   unitVector, transformDirection, triangleGeometricNormal
 *                           shaders/traversal.slang:170-254
 *   outgoingTransportRadianceTracked / emissiveResponseTracked
 *                           shaders/cascades.slang:60-134,196-268
 *   deposit (J, beta per interval, terminal tail)
 *                           shaders/split.slang:1056-1117
 *   reduce (sample mean)    shaders/split.slang:1386-1476
 *   mergeEntry without history: beginMergeGroup/addMergeChild/
 *   finishMergeGroup/mergeRadianceGroups/interpolateMergedChildSources
 *                           shaders/split.slang:389-479,1523-1622,1624-1725
 *   packProbes              shaders/split.slang:1788-1918
 *   irradianceLookup/diffuseProbePage/diffuseAtLod/resolveLodSource/
 *   resolveSplitDiffuse/rawGiSupportAlpha
 *                           shaders/world_resolve.slang:1-196, common.slang:324
 *   build_emissive_lights   src/scene.rs:2311-2372
 *
 * Dense-grid restatement (the reference keys sparse probes by hash; here a
 * probe is a dense grid cell, a missing entry is a cell without a supported
 * sample, beta = -1 marks "no sample").  The three radiance source groups
 * share one support in a history-free update, so they are merged as one RGB
 * group.  The BSDF is the diffuse albedo times the Disney multiplier / pi
 * (scatteringResponse's diffuse branch, shaders/environment.slang:2-6).
 * cos/sin of 2*pi*u use a fixed f64 Taylor evaluation (orc_cos_sin_2pi) so
 * that CPU and GPU directions are bit-identical.
 *
 * Pinned by (tests/test_oracle_transport.py): the four-cascade closed form of
 * src/split_merge_tests.rs:1403-1419 (level means and supports), the dense
 * direction round trip of src/transport_tests.rs:4040-4071, analytic cases of
 * the merge and pack formulas, the f64 twin (errors stated against it), and
 * the Cornell emitter irradiance against numerical quadrature.
 */
#include "lux_oracle.h"
#include <pthread.h>
#include <tgmath.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI_F 3.14159265358979323846f

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline int finitef_(float f) { return (fbits(f) & 0x7f800000u) != 0x7f800000u; }

/* ---------------- level geometry ---------------- */
uint32_t orc_gi_level_dims(const orc_gi_params* p, uint32_t level, uint32_t dims[3]) {
    uint32_t n = 1;
    for (int a = 0; a < 3; ++a) {
        uint32_t d = p->dims[a];
        for (uint32_t l = 0; l < level; ++l) d = (d + 1) / 2;
        dims[a] = d;
        n *= d;
    }
    return n;
}

uint32_t orc_gi_level_directions(const orc_gi_params* p, uint32_t level) {
    return (p->base_phi << level) * (p->base_y << level);
}

uint64_t orc_gi_level_offset(const orc_gi_params* p, uint32_t level) {
    uint64_t off = 0;
    for (uint32_t l = 0; l < level; ++l) {
        uint32_t d[3];
        off += (uint64_t)orc_gi_level_dims(p, l, d) * orc_gi_level_directions(p, l);
    }
    return off;
}

static float interval_end(const orc_gi_params* p, uint32_t level) {
    if (p->interval_mode == 1u)
        return p->interval_length * (float)((2u << level) - 1u);
    return p->interval_length * ldexpf(1.0f, (int)(2u * level));
}

void orc_gi_interval_bounds(const orc_gi_params* p, uint32_t level, float* start, float* end) {
    *end = interval_end(p, level);
    *start = level == 0 ? 0.0f : interval_end(p, level - 1);
}

/* ---------------- integer hashes and sequences (common.slang) ---------------- */
static inline uint32_t transport_hash(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

static inline void r2(uint32_t index, float out[2]) {
    uint32_t px = index * 3242174889u, py = index * 2447445414u;
    out[0] = (float)(px >> 8) / 16777216.0f;
    out[1] = (float)(py >> 8) / 16777216.0f;
}

static inline float fracf_(float x) { return x - floorf(x); }

static inline uint32_t hash_key(int32_t x, int32_t y, int32_t z, int32_t w, uint32_t buckets) {
    uint32_t h = (uint32_t)x * 0x8da6b343u ^ (uint32_t)y * 0xd8163841u ^
                 (uint32_t)z * 0xcb1ab31fu ^ (uint32_t)w * 0x165667b1u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    return h & (buckets - 1u);
}

/* ---------------- directions ---------------- */
/* cos(2 pi u), sin(2 pi u): quadrant split in exact f64, then Taylor series
 * to t^21 on [0, pi/2), rounded to f32. Only + and * (no contraction). */
void orc_cos_sin_2pi(float u, float* c, float* s) {
    double x = (double)u * 4.0;
    double q = floor(x);
    double r = x - q;
    int qi = ((int)q) & 3;
    double t = r * 1.5707963267948966;
    double t2 = t * t;
    double sn = 1.0 / 51090942171709440000.0;           /* 1/21! */
    sn = sn * t2 * -1.0 + 1.0 / 121645100408832000.0;   /* 1/19! */
    sn = sn * t2 * -1.0 + 1.0 / 355687428096000.0;      /* 1/17! */
    sn = sn * t2 * -1.0 + 1.0 / 1307674368000.0;        /* 1/15! */
    sn = sn * t2 * -1.0 + 1.0 / 6227020800.0;           /* 1/13! */
    sn = sn * t2 * -1.0 + 1.0 / 39916800.0;             /* 1/11! */
    sn = sn * t2 * -1.0 + 1.0 / 362880.0;               /* 1/9! */
    sn = sn * t2 * -1.0 + 1.0 / 5040.0;                 /* 1/7! */
    sn = sn * t2 * -1.0 + 1.0 / 120.0;                  /* 1/5! */
    sn = sn * t2 * -1.0 + 1.0 / 6.0;                    /* 1/3! */
    sn = sn * t2 * -1.0 + 1.0;
    sn = sn * t;
    double cs = 1.0 / 2432902008176640000.0;            /* 1/20! */
    cs = cs * t2 * -1.0 + 1.0 / 6402373705728000.0;     /* 1/18! */
    cs = cs * t2 * -1.0 + 1.0 / 20922789888000.0;       /* 1/16! */
    cs = cs * t2 * -1.0 + 1.0 / 87178291200.0;          /* 1/14! */
    cs = cs * t2 * -1.0 + 1.0 / 479001600.0;            /* 1/12! */
    cs = cs * t2 * -1.0 + 1.0 / 3628800.0;              /* 1/10! */
    cs = cs * t2 * -1.0 + 1.0 / 40320.0;                /* 1/8! */
    cs = cs * t2 * -1.0 + 1.0 / 720.0;                  /* 1/6! */
    cs = cs * t2 * -1.0 + 1.0 / 24.0;                   /* 1/4! */
    cs = cs * t2 * -1.0 + 1.0 / 2.0;                    /* 1/2! */
    cs = cs * t2 * -1.0 + 1.0;
    double oc, os;
    switch (qi) {
        case 0: oc = cs; os = sn; break;
        case 1: oc = -sn; os = cs; break;
        case 2: oc = -cs; os = -sn; break;
        default: oc = sn; os = -cs; break;
    }
    *c = (float)oc;
    *s = (float)os;
}

/* decodeDirection (common.slang:376-382) with the polynomial cos/sin. */
static void decode_direction(float ux, float uy, float out[3]) {
    float c, s;
    orc_cos_sin_2pi(ux, &c, &s);
    float y = uy * 2.0f - 1.0f;
    float q = 1.0f - y * y;
    float r = sqrtf(q > 0.0f ? q : 0.0f);
    out[0] = r * c;
    out[1] = y;
    out[2] = r * s;
}

/* Equal-area lat-long bin (binDirection, common.slang:395-400) generalised
 * to nphi x ny bins, with an in-bin offset (0.5, 0.5 = bin centre). */
void orc_bin_direction(uint32_t index, uint32_t nphi, uint32_t ny, float ox, float oy,
                       float out[3]) {
    float ux = ((float)(index % nphi) + ox) / (float)nphi;
    float uy = ((float)(index / nphi) + oy) / (float)ny;
    decode_direction(ux, uy, out);
}

/* Per-update in-bin offset of ray r: frac(0.5 + jitter + r2(r)) with
 * jitter = r2(transportHash(seed ^ 0x5a17c9e3)) (split.slang:1135-1143). */
static void ray_offset_uv(uint32_t seed, uint32_t r, float out[2]) {
    float j[2], q[2];
    r2(transport_hash(seed ^ 0x5a17c9e3u), j);
    r2(r, q);
    out[0] = fracf_(0.5f + j[0] + q[0]);
    out[1] = fracf_(0.5f + j[1] + q[1]);
}

/* rayOffset (common.slang:345-359). */
static void ray_offset(const float pos[3], const float n[3], float bias, float out[3]) {
    for (int a = 0; a < 3; ++a) {
        float origin = pos[a] + n[a] * bias;
        int32_t step = (int32_t)(n[a] * 256.0f);
        uint32_t bits = fbits(origin);
        uint32_t moved = bits + (uint32_t)(origin < 0.0f ? -step : step);
        float integer_offset = bitsf(moved);
        float shifted = fabsf(origin) < 0.03125f ? origin + n[a] / 65536.0f : integer_offset;
        out[a] = finitef_(shifted) ? shifted : pos[a];
    }
}

/* ---------------- emitter table (src/scene.rs:2311-2372) ---------------- */
uint32_t orc_build_lights(const float* const* mesh_pos, const uint32_t* const* mesh_mat,
                          const uint32_t* mesh_tri_count, uint32_t n_meshes, const float* inst_fwd,
                          const uint32_t* inst_mesh, const uint32_t* inst_flags, uint32_t n_inst,
                          const orc_material* mats, uint32_t n_mats, orc_light* out, uint32_t cap) {
    uint32_t n = 0;
    double* w = (double*)malloc(sizeof(double) * (cap ? cap : 1));
    for (uint32_t k = 0; k < n_inst; ++k) {
        uint32_t m = inst_mesh[k];
        if (m >= n_meshes) continue;
        const float* f = inst_fwd + 12ull * k;
        uint32_t cull = (inst_flags ? inst_flags[k] : 0u) & 3u;
        for (uint32_t t = 0; t < mesh_tri_count[m]; ++t) {
            uint32_t mi = mesh_mat && mesh_mat[m] ? mesh_mat[m][t] : 0u;
            if (n_mats == 0) continue;
            if (mi >= n_mats) mi = n_mats - 1;
            const orc_material* mat = &mats[mi];
            float em[3];
            for (int c = 0; c < 3; ++c) em[c] = mat->emission[c] > 0.0f ? mat->emission[c] : 0.0f;
            const float* p = mesh_pos[m] + 9ull * t;
            float v[3][3];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) {
                    const float* row = f + 4 * r;
                    v[c][r] = ((row[0] * p[3 * c] + row[1] * p[3 * c + 1]) + row[2] * p[3 * c + 2]) +
                              row[3];
                }
            float e1[3], e2[3];
            for (int r = 0; r < 3; ++r) { e1[r] = v[1][r] - v[0][r]; e2[r] = v[2][r] - v[0][r]; }
            float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                           e1[0] * e2[1] - e1[1] * e2[0]};
            float len = sqrtf((cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2]);
            float area = len * 0.5f;
            float lum = (em[0] * 0.2126f + em[1] * 0.7152f) + em[2] * 0.0722f;
            double power = (double)area * (double)lum * 1.0;
            if (!isfinite(power) || power <= 0.0) continue;
            if (n >= cap) continue;
            orc_light* L = &out[n];
            memset(L, 0, sizeof *L);
            L->position[0] = v[0][0]; L->position[1] = v[0][1]; L->position[2] = v[0][2];
            L->position[3] = cull == 0 ? -2.0f : -1.0f;
            L->color[0] = em[0]; L->color[1] = em[1]; L->color[2] = em[2]; L->color[3] = area;
            float inv = 1.0f / len;
            float sgn = cull == 2 ? -1.0f : 1.0f;
            for (int r = 0; r < 3; ++r) {
                L->direction[r] = (cr[r] * inv) * sgn;
                L->edge1[r] = e1[r];
                L->edge2[r] = e2[r];
            }
            L->source[0] = k;
            L->source[1] = t;
            w[n] = power;
            n++;
        }
    }
    if (n > 0) {
        const uint32_t steps = 1u << 24;
        double total = 0.0;
        for (uint32_t i = 0; i < n; ++i) total += w[i];
        double cum = 0.0;
        uint32_t prev = 0;
        for (uint32_t i = 0; i < n; ++i) {
            cum += w[i];
            uint32_t boundary;
            if (i + 1 == n) {
                boundary = steps;
            } else {
                double ideal_d = round(cum / total * (double)steps);
                uint32_t ideal = ideal_d < 0.0 ? 0u : (ideal_d > 4294967295.0 ? 0xffffffffu : (uint32_t)ideal_d);
                uint32_t lo = prev + 1, hi = steps - (n - i - 1);
                boundary = ideal < lo ? lo : (ideal > hi ? hi : ideal);
            }
            out[i].direction[3] = (float)boundary / (float)steps;
            out[i].edge1[3] = (float)(boundary - prev) / (float)steps;
            prev = boundary;
        }
    }
    free(w);
    return n;
}

/* ---------------- shared f32 helpers ---------------- */
static inline float dot3f(const float a[3], const float b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline float satf(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

static void unit_vector_f(const float v[3], const float fb[3], float out[3]) {
    float sc = fmaxf(fabsf(v[0]), fmaxf(fabsf(v[1]), fabsf(v[2])));
    if (sc == 0.0f || !finitef_(sc)) { memcpy(out, fb, 12); return; }
    float div = fmaxf(sc * 0.25f, 1.17549435e-38f);
    float w[3] = {v[0] / div, v[1] / div, v[2] / div};
    float len = sqrtf(dot3f(w, w));
    out[0] = w[0] / len; out[1] = w[1] / len; out[2] = w[2] / len;
}

/* triangleGeometricNormal (traversal.slang:248-251) through the inverse rows. */
static void geometric_normal(const orc_instance* in, const float* tri, float out[3]) {
    const float* a = tri;
    const float* b = tri + 4;
    const float* c = tri + 8;
    float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    /* transformDirection(base, cr, transpose = true) */
    float r[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i][j] = in->inv[4 * i + j];
    float mag[3];
    for (int j = 0; j < 3; ++j) mag[j] = fmaxf(fabsf(r[0][j]), fmaxf(fabsf(r[1][j]), fabsf(r[2][j])));
    float sc = fmaxf(mag[0], fmaxf(mag[1], mag[2]));
    float tv[3] = {0.0f, 0.0f, 0.0f};
    if (sc != 0.0f) {
        sc = fmaxf(sc * 0.25f, 1.17549435e-38f);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r[i][j] = r[i][j] / sc;
        float zero[3] = {0.0f, 0.0f, 0.0f}, u[3];
        unit_vector_f(cr, zero, u);
        for (int j = 0; j < 3; ++j) tv[j] = (r[0][j] * u[0] + r[1][j] * u[1]) + r[2][j] * u[2];
    }
    float up[3] = {0.0f, 1.0f, 0.0f}, n[3];
    unit_vector_f(tv, up, n);
    float sg = (in->flags & 4u) != 0u ? -1.0f : 1.0f;
    out[0] = sg * n[0]; out[1] = sg * n[1]; out[2] = sg * n[2];
}

/* ---------------- the REAL-generic parts (float product, double twin) ---------- */
/* splitLodBlendForParameters (shaders/common.slang:431-461).  floor(log2(x))
 * of x >= 1 is read from the exponent field (exact, where a log2
 * approximation may round across a power of two); exp2 of an integer is a
 * power of two.  max(lodScale, 1) maps NaN to 1 as IEEE maxNum does. */
void orc_lod_select(const orc_lod_view* v, const float* position, uint32_t* low, uint32_t* high,
                    float* blend) {
    const float* pp = v->clip_row3;
    int perspective = finitef_(pp[0]) && finitef_(pp[1]) && finitef_(pp[2]) &&
                      ((pp[0] * pp[0] + pp[1] * pp[1]) + pp[2] * pp[2]) > 1e-12f;
    const float* vp = v->clip_row1;
    float focal = 0.5f * (float)v->viewport_height * sqrtf((vp[0] * vp[0] + vp[1] * vp[1]) + vp[2] * vp[2]);
    uint32_t target = v->target_pixels > 1u ? v->target_pixels : 1u;
    float ts = v->probe_spacing;
    float scale;
    if (perspective) {
        float ox = fabsf(position[0] - v->camera_position[0]);
        float oy = fabsf(position[1] - v->camera_position[1]);
        float oz = fabsf(position[2] - v->camera_position[2]);
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
    uint32_t e = ((fbits(m) >> 23) & 0xffu) - 127u; /* m >= 1: floor(log2 m); inf: 128 */
    uint32_t lo = e < 24u ? e : 24u;
    *low = lo;
    *high = lo;
    *blend = 0.0f;
    if (lo < 24u) {
        float next = ldexpf(1.0f, (int)lo + 1);
        float start = 0.9f * next;
        if (scale >= start && scale < next) {
            *high = lo + 1u;
            float b = (scale - start) / (next - start);
            *blend = b < 0.0f ? 0.0f : (b > 1.0f ? 1.0f : b);
        }
    }
}

/* ---------------- dependencies (f32) ----------------
 * DependencyContext (shaders/common.slang:138-242): the union of world
 * segments that established a value, known or unknown, measured-empty or
 * bounded.  A ray's segment is its part inside the scene's root box, from
 * max(root entry, 0) to the hit (closest-hit rays that hit) or to the root
 * exit clipped at the ray's limit (traceRayTracked / traceOccludedTracked,
 * shaders/traversal.slang:696-810,1097-1150); endpoints are widened by
 * max(scale * 1e-6, 2^-23), scale = max(1, |coords|) (dependencyIncludeSegment). */
typedef struct {
    float lo[3], hi[3];
    int known, has;
} orc_dep;

static inline float dmin_(float a, float b) { return b < a ? b : a; }
static inline float dmax_(float a, float b) { return b > a ? b : a; }

static orc_dep dep_empty(void) {
    orc_dep d;
    memset(&d, 0, sizeof d);
    d.known = 1;
    return d;
}
static orc_dep dep_unknown(void) {
    orc_dep d = dep_empty();
    d.known = 0;
    return d;
}
static int finite3_(const float v[3]) { return finitef_(v[0]) && finitef_(v[1]) && finitef_(v[2]); }
static void dep_include_bounds(orc_dep* d, const float mn[3], const float mx[3]) {
    if (!d->known) return;
    if (!finite3_(mn) || !finite3_(mx) || mn[0] > mx[0] || mn[1] > mx[1] || mn[2] > mx[2]) {
        *d = dep_unknown();
        return;
    }
    if (!d->has) {
        memcpy(d->lo, mn, 12);
        memcpy(d->hi, mx, 12);
        d->has = 1;
    } else {
        for (int a = 0; a < 3; ++a) {
            d->lo[a] = dmin_(d->lo[a], mn[a]);
            d->hi[a] = dmax_(d->hi[a], mx[a]);
        }
    }
}
static void dep_include_segment(orc_dep* d, const float st[3], const float en[3]) {
    if (!finite3_(st) || !finite3_(en)) {
        *d = dep_unknown();
        return;
    }
    float scale = dmax_(1.0f, dmax_(dmax_(fabsf(st[0]), fabsf(st[1])), fabsf(st[2])));
    scale = dmax_(scale, dmax_(dmax_(fabsf(en[0]), fabsf(en[1])), fabsf(en[2])));
    const float r = dmax_(scale * 1e-6f, 1.1920929e-7f);
    float mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = dmin_(st[a], en[a]) - r;
        mx[a] = dmax_(st[a], en[a]) + r;
    }
    dep_include_bounds(d, mn, mx);
}
static void dep_union(orc_dep* d, const orc_dep* o) {
    if (!d->known || !o->known) {
        *d = dep_unknown();
        return;
    }
    if (o->has) dep_include_bounds(d, o->lo, o->hi);
}
static void dep_ray(orc_dep* d, const orc_scene* s, const float o[3], const float dir[3], float limit, int hit,
                    float t) {
    if (s->n_inst == 0) return;
    float e, x;
    if (!orc_box_interval(o, dir, s->tlas[0].mn, s->tlas[0].mx, limit, &e, &x)) return;
    const float s0 = dmax_(e, 0.0f), s1 = hit ? dmin_(t, x) : x;
    float a[3], b[3];
    for (int q = 0; q < 3; ++q) {
        a[q] = o[q] + dir[q] * s0;
        b[q] = o[q] + dir[q] * s1;
    }
    dep_include_segment(d, a, b);
}
static void dep_store(float* out, const orc_dep* d) {
    out[0] = d->lo[0]; out[1] = d->lo[1]; out[2] = d->lo[2]; out[3] = d->known ? 1.0f : 0.0f;
    out[4] = d->hi[0]; out[5] = d->hi[1]; out[6] = d->hi[2]; out[7] = d->has ? 1.0f : 0.0f;
}
static orc_dep dep_load(const float* in) {
    orc_dep d;
    memcpy(d.lo, in, 12);
    memcpy(d.hi, in + 4, 12);
    d.known = in[3] > 0.5f;
    d.has = in[7] > 0.5f;
    if (d.known && d.has && (!finite3_(d.lo) || !finite3_(d.hi) || d.lo[0] > d.hi[0] || d.lo[1] > d.hi[1] ||
                             d.lo[2] > d.hi[2]))
        d.known = 0;
    return d;
}

#define REAL float
#define SFX(name) name##_f32
#include "orc_transport_impl.h"
#undef REAL
#undef SFX
#define REAL double
#define SFX(name) name##_f64
#include "orc_transport_impl.h"
#undef REAL
#undef SFX

void orc_gi_trace(const orc_scene* s, const orc_light* lights, uint32_t n_lights,
                  const orc_material* mats, uint32_t n_mats, const orc_gi_params* p,
                  float* intervals, int threads) {
    gi_trace_f32(s, lights, n_lights, mats, n_mats, p, intervals, NULL, threads, 1, NULL);
}
/* The interval trace with, per interval, its supported rays and the
 * dependency of this trace (8 floats: lo.xyz, known, hi.xyz, has bounds):
 * each supported ray's segment, then its NEE occlusion segments, in order. */
void orc_gi_trace_dep(const orc_scene* s, const orc_light* lights, uint32_t n_lights,
                      const orc_material* mats, uint32_t n_mats, const orc_gi_params* p,
                      float* intervals, uint32_t* samples, float* dep, int threads) {
    gi_trace_f32(s, lights, n_lights, mats, n_mats, p, intervals, samples, threads, 1, dep);
}
/* Every stride-th interval (g = 0, stride, 2 stride, ...) only; the others
 * are left as they are (CPU timing of a bounded fraction of an update). */
void orc_gi_trace_stride(const orc_scene* s, const orc_light* lights, uint32_t n_lights,
                         const orc_material* mats, uint32_t n_mats, const orc_gi_params* p,
                         float* intervals, int threads, uint64_t stride) {
    gi_trace_f32(s, lights, n_lights, mats, n_mats, p, intervals, NULL, threads, stride, NULL);
}
void orc_gi_trace_f64(const orc_scene* s, const orc_light* lights, uint32_t n_lights,
                      const orc_material* mats, uint32_t n_mats, const orc_gi_params* p,
                      double* intervals, int threads) {
    gi_trace_f64(s, lights, n_lights, mats, n_mats, p, intervals, NULL, threads, 1, NULL);
}
void orc_gi_trace_samples(const orc_scene* s, const orc_light* lights, uint32_t n_lights,
                          const orc_material* mats, uint32_t n_mats, const orc_gi_params* p,
                          float* intervals, uint32_t* samples, int threads) {
    gi_trace_f32(s, lights, n_lights, mats, n_mats, p, intervals, samples, threads, 1, NULL);
}
void orc_gi_merge_history(const orc_gi_params* p, const float* intervals, const float* history, float* merged) {
    gi_merge_f32(p, intervals, history, merged);
}

static inline float hsat(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }
static inline float hmin(float a, float b) { return b < a ? b : a; }
static inline float hmax(float a, float b) { return b > a ? b : a; }
static inline float hmass(float m) { return (finitef_(m) && m >= 0.0f) ? m : 0.0f; }

/* Temporal history of intervals [0, count) (mergeEntry's history part,
 * split.slang:1624-1705; mergeRawRadianceGroup :1478-1521): reuse when the
 * history is valid, alive (persistent, or sampled within history_frames
 * frames, common.slang:85-89) and the dependency box (probe centre +-
 * interval end joined with the emitter bounds; unbounded when terminal)
 * meets no dirty box; per channel only when that channel's blend is
 * positive.  Masses: retention = blend^n, new mass (1 - retention) /
 * (1 - blend), old mass min(old, limit) * retention, sample-mass weighted
 * means; an update with no sample keeps the old value and mass.  pow(blend,
 * n) for an integer n is a product of n factors.  history: 8 floats per
 * interval (J.rgb, radiance mass, beta, transmittance mass, last sampled
 * frame bits, 0).  emit: lo.xyz, hi.xyz; boxes: lo.xyz, hi.xyz each. */
void orc_gi_history(const orc_gi_params* p, const orc_history_params* h, const float* boxes,
                    uint32_t n_boxes, uint64_t count, float* intervals, const uint32_t* samples, const float* dep,
                    float* history) {
    uint64_t off[17];
    for (uint32_t l = 0; l <= p->levels; ++l) off[l] = orc_gi_level_offset(p, l);
    const float rb = hmin(hsat(h->radiance_blend), 0.999f), tb = hmin(hsat(h->transmittance_blend), 0.999f);
    const float rlim = 1.0f / (1.0f - rb), tlim = 1.0f / (1.0f - tb);
    for (uint64_t g = 0; g < count; ++g) {
        uint32_t L = 0;
        while (L + 1 < p->levels && g >= off[L + 1]) ++L;
        float* iv = intervals + 4 * g;
        float* hs = history + 16 * g;
        uint32_t n = samples[g];
        uint32_t last;
        memcpy(&last, hs + 6, 4);
        int reuse = h->valid && h->history_frames > 0;
        if (reuse) reuse = h->persistent || p->frame - last <= h->history_frames;
        /* dependencyIntersectsDirty (split.slang:13-25) per dirty box: an
         * unknown dependency meets every region, a measured-empty one none */
        const orc_dep old = dep_load(hs + 8);
        if (reuse && h->root_grew && old.known && !old.has) reuse = 0;
        if (reuse && n_boxes) {
            if (!old.known) reuse = 0;
            for (uint32_t b = 0; b < n_boxes && reuse && old.has; ++b) {
                const float* bx = boxes + 6 * b;
                if (old.lo[0] <= bx[3] && old.hi[0] >= bx[0] && old.lo[1] <= bx[4] && old.hi[1] >= bx[1] &&
                    old.lo[2] <= bx[5] && old.hi[2] >= bx[2])
                    reuse = 0;
            }
        }
        int reuse_r = reuse && h->radiance_blend > 0.0f, reuse_t = reuse && h->transmittance_blend > 0.0f;
        float om = hmass(hs[3]), otm = hmass(hs[5]);
        float J[3] = {iv[0], iv[1], iv[2]}, beta = iv[3], mass, ts;
        if (n > 0) {
            float ret = 1.0f;
            for (uint32_t i = 0; i < n; ++i) ret = ret * rb;
            mass = (1.0f - ret) * rlim;
            float o = hmin(om, rlim) * ret;
            if (reuse_r && o > 0.0f) {
                float t = mass + o;
                for (int q = 0; q < 3; ++q) J[q] = (J[q] * mass + hs[q] * o) / t;
                mass = mass + o;
            }
            float tret = 1.0f;
            for (uint32_t i = 0; i < n; ++i) tret = tret * tb;
            ts = (1.0f - tret) * tlim;
            float ot = hmin(otm, tlim) * tret;
            if (reuse_t && ot > 0.0f) {
                beta = (beta * ts + hs[4] * ot) / (ts + ot);
                ts = ts + ot;
            }
            last = p->frame;
        } else {
            if (reuse_r && om > 0.0f) {
                J[0] = hs[0]; J[1] = hs[1]; J[2] = hs[2];
                mass = om;
            } else {
                J[0] = J[1] = J[2] = 0.0f;
                mass = 0.0f;
            }
            if (reuse_t && otm > 0.0f) {
                beta = hs[4];
                ts = otm;
            } else {
                beta = 1.0f;
                ts = 0.0f;
            }
        }
        iv[0] = J[0]; iv[1] = J[1]; iv[2] = J[2]; iv[3] = beta;
        hs[0] = J[0]; hs[1] = J[1]; hs[2] = J[2]; hs[3] = mass;
        hs[4] = beta; hs[5] = ts;
        memcpy(hs + 6, &last, 4);
        hs[7] = 0.0f;
        /* the retained value depends on this trace's segments and, when the
         * history was reused, on everything the history depended on */
        orc_dep cur = dep_load(dep + 8 * g);
        if (reuse) {
            orc_dep u = old;
            dep_union(&u, &cur);
            cur = u;
        }
        dep_store(hs + 8, &cur);
    }
}

void orc_gi_merge(const orc_gi_params* p, const float* intervals, float* merged) {
    gi_merge_f32(p, intervals, NULL, merged);
}
void orc_gi_merge_f64(const orc_gi_params* p, const double* intervals, double* merged) {
    gi_merge_f64(p, intervals, NULL, merged);
}
void orc_gi_pack(const orc_gi_params* p, const float* merged, float* probes) {
    gi_pack_f32(p, merged, probes);
}
void orc_gi_pack_f64(const orc_gi_params* p, const double* merged, double* probes) {
    gi_pack_f64(p, merged, probes);
}
void orc_world_resolve(const orc_gi_params* p, const float* probes, const float* receivers,
                       uint32_t n, float* out) {
    world_resolve_f32(p, probes, receivers, n, out);
}
void orc_world_resolve_f64(const orc_gi_params* p, const double* probes, const float* receivers,
                           uint32_t n, double* out) {
    world_resolve_f64(p, probes, receivers, n, out);
}
void orc_world_resolve_lods(const orc_gi_params* const* grids, const float* const* probes, uint32_t count,
                            const orc_lod_view* view, const float* receivers, uint32_t n, float* out) {
    world_resolve_lods_f32(grids, probes, count, view, receivers, n, out);
}
void orc_world_resolve_lods_f64(const orc_gi_params* const* grids, const double* const* probes,
                                uint32_t count, const orc_lod_view* view, const float* receivers, uint32_t n,
                                double* out) {
    world_resolve_lods_f64(grids, probes, count, view, receivers, n, out);
}

/* Surrogate query of every (probe, bin) of a level, in merged-array order:
 * probe centre, the update's first jittered ray direction of the bin, and
 * that direction again as the normal channel (9 floats each). */
void orc_gi_level_queries(const orc_gi_params* p, uint32_t level, float* out) {
    uint32_t dims[3];
    uint64_t probes = orc_gi_level_dims(p, level, dims);
    uint32_t D = orc_gi_level_directions(p, level);
    uint32_t nphi = p->base_phi << level, ny = p->base_y << level;
    float sl = p->spacing * ldexpf(1.0f, (int)level);
    float ouv[2];
    ray_offset_uv(p->seed, 0u, ouv);
    for (uint64_t probe = 0; probe < probes; ++probe) {
        uint32_t ix = (uint32_t)(probe % dims[0]);
        uint32_t iy = (uint32_t)((probe / dims[0]) % dims[1]);
        uint32_t iz = (uint32_t)(probe / ((uint64_t)dims[0] * dims[1]));
        float c[3] = {p->origin[0] + ((float)ix + 0.5f) * sl, p->origin[1] + ((float)iy + 0.5f) * sl,
                      p->origin[2] + ((float)iz + 0.5f) * sl};
        for (uint32_t k = 0; k < D; ++k) {
            float d[3];
            orc_bin_direction(k, nphi, ny, ouv[0], ouv[1], d);
            float* o = out + 9 * (probe * D + k);
            o[0] = c[0]; o[1] = c[1]; o[2] = c[2];
            o[3] = d[0]; o[4] = d[1]; o[5] = d[2];
            o[6] = d[0]; o[7] = d[1]; o[8] = d[2];
        }
    }
}

/* Index within its level of training sample i: a multiply-shift of
 * transportHash(i + transportHash(seed ^ 0x3c6ef372)) onto [0, count). */
uint64_t orc_gi_sample_index(uint32_t seed, uint32_t i, uint64_t count) {
    uint32_t h = transport_hash(i + transport_hash(seed ^ 0x3c6ef372u));
    return ((uint64_t)h * count) >> 32;
}

/* ---------------- receiver trace (traceTransport / deposit / reduce) ---------------- */
/* atan of |t| <= tan(pi/8): Taylor series to t^47 in f64 (+ and * only). */
static double atan_small(double t) {
    double t2 = t * t;
    double s = 1.0 / 47.0;
    for (int k = 22; k >= 0; --k) s = s * t2 + ((k & 1) ? -1.0 : 1.0) / (double)(2 * k + 1);
    return s * t;
}

/* atan2 from a fixed f64 series (the same arithmetic on CPU and GPU, so the
 * azimuth bin is deterministic); IEEE atan2's signed-zero conventions. */
static double atan2_fixed(double y, double x) {
    double ax = fabs(x), ay = fabs(y);
    double mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    double a = mx > 0.0 ? mn / mx : 0.0;
    double r = a > 0.41421356237309503 ? 0.7853981633974483 + atan_small((a - 1.0) / (a + 1.0)) : atan_small(a);
    if (ay > ax) r = 1.5707963267948966 - r;
    if (signbit(x)) r = 3.141592653589793 - r;
    if (signbit(y)) r = -r;
    return r;
}

/* directionIndex (common.slang:384-394): encodeDirection's azimuth
 * frac(atan2(z, x) / 2pi + 1) and saturate(y / 2 + 1/2), binned over
 * nphi x ny (2 theta x theta in the reference). */
uint32_t orc_direction_index(const float d[3], uint32_t nphi, uint32_t ny) {
    float u = (float)(atan2_fixed((double)d[2], (double)d[0]) * 0.15915494309189535);
    u = fracf_(u + 1.0f);
    float v = d[1] * 0.5f + 0.5f;
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    uint32_t x = (uint32_t)(u * (float)nphi), y = (uint32_t)(v * (float)ny);
    if (x > nphi - 1u) x = nphi - 1u;
    if (y > ny - 1u) y = ny - 1u;
    return x + y * nphi;
}

/* traceTransport (split.slang:1123-1163) for each ticket's selected
 * variants, finishTransportVariant (:1189-1218) with this path's shading,
 * deposit (:1056-1118: J / beta per level up to the terminating one) and
 * reduce (:1386-1476: mean over the update's samples of a key), the keys
 * (cell, direction bin, level, LOD) addressed in the dense grid of their
 * LOD.  The deposits of a key are in ray order (tickets, then variants,
 * then levels); their sum is fixed as chunks of 64 summed by a pairwise tree,
 * the chunk sums added in order (the reference's order is that of its
 * hash-bucket lists, left open by its atomics). */
void orc_rx_trace(const orc_scene* s, const orc_light* lights, uint32_t n_lights, const orc_material* mats,
                  uint32_t n_mats, const orc_gi_params* const* grids, uint32_t lod_count, const float* receivers,
                  uint32_t count, const uint32_t* lods, const uint32_t* sequence, const uint32_t* tickets,
                  uint32_t n_tickets, uint32_t frame, float* const* intervals, uint32_t* const* samples,
                  uint32_t stats[4]) {
    orc_gi_params p0c = *grids[0];
    p0c.frame = frame; /* params.frame.x drives the NEE sequence too */
    const orc_gi_params* p0 = &p0c;
    const uint32_t levels = p0->levels;
    stats[0] = stats[1] = stats[2] = stats[3] = 0;
    /* deposits in ray order: (grid, slot) and J.rgb / beta */
    size_t cap = 1024, nd = 0;
    uint32_t* dg = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    uint64_t* ds = (uint64_t*)malloc(sizeof(uint64_t) * cap);
    float* dv = (float*)malloc(sizeof(float) * 4 * cap);
    float jit[2];
    r2(transport_hash(frame ^ 0x5a17c9e3u), jit);
    for (uint32_t t = 0; t < n_tickets; ++t) {
        const uint32_t smp = tickets[t];
        float pos[3], n[3];
        if (!orc_rx_surface(receivers + 6ull * smp, pos, n)) continue;
        for (uint32_t v = 0; v < 2; ++v) {
            const uint32_t seq = sequence[(uint64_t)v * count + smp];
            if (seq == ORC_EMPTY) continue;
            const uint32_t lod = lods[(uint64_t)v * count + smp];
            stats[0]++;
            float q[2], d[3], o[3];
            r2(seq + 1u, q);
            decode_direction(fracf_((0.5f + q[0]) + jit[0]), fracf_((0.5f + q[1]) + jit[1]), d);
            if (dot3f(d, n) < 0.0f) { d[0] = -d[0]; d[1] = -d[1]; d[2] = -d[2]; }
            ray_offset(pos, n, p0->normal_bias, o);
            orc_ray ray;
            ray.ox = o[0]; ray.oy = o[1]; ray.oz = o[2];
            ray.tmax = 1e20f;
            ray.dx = d[0]; ray.dy = d[1]; ray.dz = d[2];
            ray.flags = 0;
            orc_hit h;
            uint32_t ts = 0, is = 0;
            orc_trace_ray(s, &ray, &h, NULL, &ts, &is);
            if (h.flags & ORC_ERR_UNSUPPORTED) { stats[1]++; continue; }
            float rad[3] = {0.0f, 0.0f, 0.0f};
            const int hit = h.triangle != ORC_EMPTY;
            if (hit && !shade_f32(s, lights, n_lights, mats, n_mats, p0, o, d, h.t, ts, is, rad, NULL)) {
                stats[1]++;
                continue;
            }
            const orc_gi_params* g = lod < lod_count ? grids[lod] : NULL;
            const float lod_scale = ldexpf(1.0f, (int)(lod < 24u ? lod : 24u));
            for (uint32_t L = 0; L < levels; ++L) {
                float start = L == 0 ? 0.0f : interval_end(p0, L - 1u) * lod_scale;
                float end = interval_end(p0, L) * lod_scale;
                int terminal = p0->terminal_tail && L + 1u == levels;
                int terminates = terminal || (hit && h.t > start && h.t <= end);
                int32_t key[4];
                orc_rx_cell(pos, p0->spacing, L, lod, key);
                int kept = 0;
                if (g) {
                    uint32_t dims[3];
                    orc_gi_level_dims(g, L, dims);
                    float sl = g->spacing * ldexpf(1.0f, (int)L);
                    int64_t idx[3];
                    int inside = 1;
                    for (int a = 0; a < 3; ++a) {
                        idx[a] = (int64_t)key[a] - (int64_t)floorf(g->origin[a] / sl);
                        inside = inside && idx[a] >= 0 && idx[a] < (int64_t)dims[a];
                    }
                    if (inside) {
                        uint32_t D = orc_gi_level_directions(g, L);
                        uint32_t dir = orc_direction_index(d, g->base_phi << L, g->base_y << L);
                        uint64_t probe = (uint64_t)idx[0] + (uint64_t)dims[0] * ((uint64_t)idx[1] + (uint64_t)dims[1] * (uint64_t)idx[2]);
                        if (nd == cap) {
                            cap *= 2;
                            dg = (uint32_t*)realloc(dg, sizeof(uint32_t) * cap);
                            ds = (uint64_t*)realloc(ds, sizeof(uint64_t) * cap);
                            dv = (float*)realloc(dv, sizeof(float) * 4 * cap);
                        }
                        dg[nd] = lod;
                        ds[nd] = orc_gi_level_offset(g, L) + probe * D + dir;
                        for (int c = 0; c < 3; ++c) dv[4 * nd + c] = terminates ? rad[c] : 0.0f;
                        dv[4 * nd + 3] = terminates ? 0.0f : 1.0f;
                        nd++;
                        kept = 1;
                    }
                }
                stats[kept ? 2 : 3]++;
                if (terminates) break;
            }
        }
    }
    /* reduce: per slot its deposits in ray order (a stable counting sort),
     * summed in chunks of 64 by a pairwise tree (x[i] += x[i + s] for i a
     * multiple of 2s, s = 1, 2, ..., 32), chunk sums added in order */
    for (uint32_t k = 0; k < lod_count; ++k) {
        uint64_t n = orc_gi_level_offset(grids[k], grids[k]->levels);
        uint32_t* start = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
        for (size_t e = 0; e < nd; ++e)
            if (dg[e] == k) start[ds[e] + 1]++;
        for (uint64_t g = 0; g < n; ++g) start[g + 1] += start[g];
        uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
        memcpy(fill, start, sizeof(uint32_t) * ((size_t)n + 1));
        uint32_t total = start[n];
        uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (total ? total : 1));
        for (size_t e = 0; e < nd; ++e)
            if (dg[e] == k) order[fill[ds[e]]++] = (uint32_t)e;
        for (uint64_t g = 0; g < n; ++g) {
            float* iv = intervals[k] + 4 * g;
            uint32_t c = start[g + 1] - start[g];
            samples[k][g] = c;
            if (c == 0) {
                iv[0] = iv[1] = iv[2] = 0.0f;
                iv[3] = -1.0f;
                continue;
            }
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t b = 0; b < c; b += 64u) {
                uint32_t m = c - b < 64u ? c - b : 64u;
                float x[64][4];
                for (uint32_t t = 0; t < m; ++t)
                    for (int q = 0; q < 4; ++q) x[t][q] = dv[4 * (size_t)order[start[g] + b + t] + q];
                for (uint32_t st = 1; st < 64u; st <<= 1)
                    for (uint32_t t = 0; t + st < m; t += 2u * st)
                        for (int q = 0; q < 4; ++q) x[t][q] = x[t][q] + x[t + st][q];
                for (int q = 0; q < 4; ++q) acc[q] = b == 0 ? x[0][q] : acc[q] + x[0][q];
            }
            float fc = (float)c;
            for (int q = 0; q < 4; ++q) iv[q] = acc[q] / fc;
        }
        free(start);
        free(fill);
        free(order);
    }
    free(dg);
    free(ds);
    free(dv);
}

/*
 * orc_bvh.c -- oracle BVH builders (TEST INFRASTRUCTURE, never the product).
 *
 * 1. Morton-30 LBVH, restated exactly as the HIP builder computes it:
 *    primitive centre c = lo*0.5 + hi*0.5 (the centre formula of
 *    src/scene.rs:2461-2466), quantised to 10 bits per axis inside the centre
 *    bounds, bit-interleaved x>y>z, stable-sorted by (key, index) -- the
 *    gpuSortLess contract of shaders/split.slang:3-10 -- then a Karras (2012)
 *    radix tree whose subtrees of <= leaf_max primitives become leaves.
 *    Output layout: the reference's 32-byte BvhNode (src/scene.rs:76-83);
 *    Karras internal node k stores its two children in slots 1+2k and 2+2k.
 *    Pinned by: std-sort comparison of the order (tests), tree invariants
 *    (every primitive covered once, children contained, ranges contiguous).
 * 2. Binned SAH build_tree (src/scene.rs:2398-2554): 12 bins, f64 bin index,
 *    in-place swap partition, leaves <= 4, depth <= 48.  Box and centroid
 *    folds and bin joins are glam 0.33.12's Vec3::min/max (the reference's
 *    `min.min(a)`), restated as the per-component select a < b ? a : b
 *    (a > b ? a : b): of equal values the later one is kept, which shows only
 *    on signed zeros (positions are finite; non-finite ones are rejected by
 *    the builders).  This tie rule is parity unpinned: no reference test or
 *    fixture holds a signed-zero tie, and glam is not vendored.  The median fallback
 *    for coincident centroids uses Rust's select_nth_unstable_by (by
 *    f32::total_cmp), which for a slice of at most 10 elements is a stable
 *    insertion sort (insertion_sort_shift_left); here it is that insertion
 *    sort for every size.  Pinned on 5-10 coincident triangles whose order a
 *    parent's partition has permuted (tests/test_oracle_bvh.py); larger tied
 *    groups (quickselect in the reference) are parity unpinned.
 */
#include "lux_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
/* Box unions use compare-selects (the GPU builder uses the same form), so the
 * stored bounds never depend on how a min/max instruction orders -0 and +0. */
static inline float sel_min(float a, float b) { return b < a ? b : a; }
static inline float sel_max(float a, float b) { return b > a ? b : a; }

static inline uint32_t spread10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

static inline uint32_t quantize10(float c, float lo, float ext) {
    if (!(ext > 0.0f)) return 0u;
    float q = (c - lo) / ext * 1024.0f;
    if (!(q > 0.0f)) return 0u; /* also maps NaN to 0 */
    if (q >= 1023.0f) return 1023u;
    return (uint32_t)q;
}

void orc_morton_codes(const float* lo, const float* hi, uint32_t n, uint32_t* codes) {
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float c = lo[4 * i + a] * 0.5f + hi[4 * i + a] * 0.5f;
            cmin[a] = fminf(cmin[a], c);
            cmax[a] = fmaxf(cmax[a], c);
        }
    float ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = cmax[a] - cmin[a];
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            float c = lo[4 * i + a] * 0.5f + hi[4 * i + a] * 0.5f;
            q[a] = quantize10(c, cmin[a], ext[a]);
        }
        codes[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    }
}

static const uint32_t* g_sort_keys;
static int pair_cmp(const void* a, const void* b) {
    uint32_t i = *(const uint32_t*)a, j = *(const uint32_t*)b;
    uint32_t ki = g_sort_keys[i], kj = g_sort_keys[j];
    if (ki != kj) return ki < kj ? -1 : 1;
    return i < j ? -1 : (i > j ? 1 : 0);
}

void orc_stable_sort_pairs(const uint32_t* keys, uint32_t n, uint32_t* perm) {
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    g_sort_keys = keys;
    qsort(perm, n, sizeof(uint32_t), pair_cmp);
}

static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

static inline int kdelta(const uint32_t* k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

static void box_union(const float* alo, const float* ahi, const float* blo, const float* bhi,
                      float* olo, float* ohi) {
    for (int a = 0; a < 3; ++a) {
        olo[a] = sel_min(alo[a], blo[a]);
        ohi[a] = sel_max(ahi[a], bhi[a]);
    }
}

uint32_t orc_lbvh_build(const float* lo, const float* hi, uint32_t n, uint32_t leaf_max,
                        uint32_t* codes, uint32_t* perm, orc_node* nodes) {
    if (n == 0) return 0;
    if (leaf_max < 1) leaf_max = 1;
    orc_morton_codes(lo, hi, n, codes);
    orc_stable_sort_pairs(codes, n, perm);
    uint32_t* sk = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; ++i) sk[i] = codes[perm[i]];
    /* sorted primitive boxes */
    float* plo = (float*)malloc(sizeof(float) * 3 * n);
    float* phi = (float*)malloc(sizeof(float) * 3 * n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            plo[3 * i + a] = lo[4 * perm[i] + a];
            phi[3 * i + a] = hi[4 * perm[i] + a];
        }
    uint32_t node_count = n > 1 ? 2 * n - 1 : 1;
    if (n <= leaf_max) {
        float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = 0; i < n; ++i) box_union(blo, bhi, plo + 3 * i, phi + 3 * i, blo, bhi);
        memset(nodes, 0, sizeof(orc_node) * node_count);
        memcpy(nodes[0].mn, blo, 12);
        memcpy(nodes[0].mx, bhi, 12);
        nodes[0].first = 0;
        nodes[0].count = n;
        free(sk); free(plo); free(phi);
        return node_count;
    }
    uint32_t m = n - 1; /* internal nodes */
    uint32_t* left = (uint32_t*)malloc(sizeof(uint32_t) * m);
    uint32_t* rangelo = (uint32_t*)malloc(sizeof(uint32_t) * m);
    uint32_t* rangehi = (uint32_t*)malloc(sizeof(uint32_t) * m);
    for (uint32_t ii = 0; ii < m; ++ii) {
        int64_t i = ii, N = n;
        int d = kdelta(sk, N, i, i + 1) - kdelta(sk, N, i, i - 1) > 0 ? 1 : -1;
        int dmin = kdelta(sk, N, i, i - d);
        int64_t lmax = 2;
        while (kdelta(sk, N, i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (kdelta(sk, N, i, i + (l + t) * d) > dmin) l += t;
        int64_t j = i + l * d;
        int dnode = kdelta(sk, N, i, j);
        int64_t s = 0, div = 2, t;
        do {
            t = (l + div - 1) / div;
            if (kdelta(sk, N, i, i + (s + t) * d) > dnode) s += t;
            div *= 2;
        } while (t > 1);
        int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
        left[ii] = (uint32_t)gamma;
        rangelo[ii] = (uint32_t)(i < j ? i : j);
        rangehi[ii] = (uint32_t)(i < j ? j : i);
    }
    /* boxes of internal nodes, computed from their contiguous ranges
     * (min/max is exact, so any evaluation order gives the same bits) */
    float* ilo = (float*)malloc(sizeof(float) * 3 * m);
    float* ihi = (float*)malloc(sizeof(float) * 3 * m);
    /* process internal nodes by increasing range size so children are done */
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * m);
    {
        /* counting sort by size */
        uint32_t* cnt = (uint32_t*)calloc(n + 1, sizeof(uint32_t));
        for (uint32_t k = 0; k < m; ++k) cnt[rangehi[k] - rangelo[k] + 1]++;
        uint32_t acc = 0;
        for (uint32_t s2 = 0; s2 <= n; ++s2) { uint32_t c = cnt[s2]; cnt[s2] = acc; acc += c; }
        for (uint32_t k = 0; k < m; ++k) order[cnt[rangehi[k] - rangelo[k] + 1]++] = k;
        free(cnt);
    }
    for (uint32_t oi = 0; oi < m; ++oi) {
        uint32_t k = order[oi];
        uint32_t g = left[k];
        const float *alo, *ahi, *blo, *bhi;
        if (g == rangelo[k]) { alo = plo + 3 * g; ahi = phi + 3 * g; }
        else { alo = ilo + 3 * g; ahi = ihi + 3 * g; }
        if (g + 1 == rangehi[k]) { blo = plo + 3 * (g + 1); bhi = phi + 3 * (g + 1); }
        else { blo = ilo + 3 * (g + 1); bhi = ihi + 3 * (g + 1); }
        box_union(alo, ahi, blo, bhi, ilo + 3 * k, ihi + 3 * k);
    }
    memset(nodes, 0, sizeof(orc_node) * node_count);
    /* root */
    memcpy(nodes[0].mn, ilo, 12);
    memcpy(nodes[0].mx, ihi, 12);
    nodes[0].first = 1;
    nodes[0].count = 0;
    for (uint32_t k = 0; k < m; ++k) {
        uint32_t g = left[k];
        /* children: left covers [rangelo, g], right covers [g+1, rangehi] */
        uint32_t cl[2] = {rangelo[k], g + 1}, ch[2] = {g, rangehi[k]};
        uint32_t cidx[2] = {g, g + 1};
        for (int c = 0; c < 2; ++c) {
            orc_node* out = &nodes[1 + 2 * k + c];
            uint32_t size = ch[c] - cl[c] + 1;
            if (size == 1) {
                memcpy(out->mn, plo + 3 * cl[c], 12);
                memcpy(out->mx, phi + 3 * cl[c], 12);
            } else {
                memcpy(out->mn, ilo + 3 * cidx[c], 12);
                memcpy(out->mx, ihi + 3 * cidx[c], 12);
            }
            if (size <= leaf_max) {
                out->first = cl[c];
                out->count = size;
            } else {
                out->first = 1 + 2 * cidx[c];
                out->count = 0;
            }
        }
    }
    free(sk); free(plo); free(phi); free(left); free(rangelo); free(rangehi);
    free(ilo); free(ihi); free(order);
    return node_count;
}

/* ---------------- binned SAH (src/scene.rs:2398-2554) ---------------- */
typedef struct { float mn[3], mx[3]; size_t count; } sah_bin;

/* glam Vec3::min / max per component */
static inline float vmin(float a, float b) { return a < b ? a : b; }
static inline float vmax(float a, float b) { return a > b ? a : b; }

static double bounds_area(const float* mn, const float* mx) {
    double sx = (double)mx[0] - (double)mn[0];
    double sy = (double)mx[1] - (double)mn[1];
    double sz = (double)mx[2] - (double)mn[2];
    return 2.0 * (sx * sy + sy * sz + sz * sx);
}
static void bin_empty(sah_bin* b) {
    for (int a = 0; a < 3; ++a) { b->mn[a] = INFINITY; b->mx[a] = -INFINITY; }
    b->count = 0;
}
static void bin_join(sah_bin* o, const sah_bin* x) {
    for (int a = 0; a < 3; ++a) { o->mn[a] = vmin(o->mn[a], x->mn[a]); o->mx[a] = vmax(o->mx[a], x->mx[a]); }
    o->count += x->count;
}
static double bin_cost(const sah_bin* b) { return b->count == 0 ? 0.0 : bounds_area(b->mn, b->mx) * (double)b->count; }

typedef struct {
    const float* lo;
    const float* hi;
    uint32_t* perm;
    orc_node* nodes;
    uint32_t n_nodes;
} sah_ctx;

static float centre(const sah_ctx* c, uint32_t e, int axis) {
    return c->lo[4 * e + axis] * 0.5f + c->hi[4 * e + axis] * 0.5f;
}

/* f32::total_cmp as an unsigned key: -NaN < -inf < ... < -0 < +0 < ... < NaN */
static uint32_t total_key(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* Stable insertion sort of el[0..count) by the total-order centre on `axis`:
 * each element shifts left past the elements of strictly greater key. */
static void median_sort(const sah_ctx* c, uint32_t* el, uint32_t count, int axis) {
    for (uint32_t i = 1; i < count; ++i) {
        uint32_t e = el[i];
        uint32_t k = total_key(centre(c, e, axis));
        uint32_t j = i;
        while (j > 0 && k < total_key(centre(c, el[j - 1], axis))) {
            el[j] = el[j - 1];
            --j;
        }
        el[j] = e;
    }
}

static void sah_rec(sah_ctx* c, uint32_t index, uint32_t first, uint32_t count, uint32_t depth) {
    uint32_t* el = c->perm + first;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < count; ++i)
        for (int a = 0; a < 3; ++a) {
            mn[a] = vmin(mn[a], c->lo[4 * el[i] + a]);
            mx[a] = vmax(mx[a], c->hi[4 * el[i] + a]);
        }
    orc_node* nd = &c->nodes[index];
    memcpy(nd->mn, mn, 12);
    memcpy(nd->mx, mx, 12);
    nd->first = first;
    nd->count = count;
    if (count <= 4 || depth == 48) return;
    enum { BINS = 12 };
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < count; ++i)
        for (int a = 0; a < 3; ++a) {
            float ce = centre(c, el[i], a);
            cmin[a] = vmin(cmin[a], ce);
            cmax[a] = vmax(cmax[a], ce);
        }
    double best = INFINITY;
    int split_axis = -1, split_bin = 0;
    double split_span = 0.0;
    for (int axis = 0; axis < 3; ++axis) {
        double span = (double)cmax[axis] - (double)cmin[axis];
        if (span <= 0.0) continue;
        sah_bin bins[BINS], right[BINS], sum, leftb;
        for (int b = 0; b < BINS; ++b) bin_empty(&bins[b]);
        for (uint32_t i = 0; i < count; ++i) {
            double f = ((double)centre(c, el[i], axis) - (double)cmin[axis]) / span * (double)BINS;
            size_t bi = (size_t)f;
            if (bi > BINS - 1) bi = BINS - 1;
            sah_bin one;
            for (int a = 0; a < 3; ++a) { one.mn[a] = c->lo[4 * el[i] + a]; one.mx[a] = c->hi[4 * el[i] + a]; }
            one.count = 1;
            bin_join(&bins[bi], &one);
        }
        bin_empty(&sum);
        for (int i = BINS - 1; i >= 0; --i) { bin_join(&sum, &bins[i]); right[i] = sum; }
        bin_empty(&leftb);
        for (int i = 0; i < BINS - 1; ++i) {
            bin_join(&leftb, &bins[i]);
            double cost = bin_cost(&leftb) + bin_cost(&right[i + 1]);
            if (leftb.count > 0 && right[i + 1].count > 0 && cost < best) {
                best = cost;
                split_axis = axis;
                split_bin = i;
                split_span = span;
            }
        }
    }
    uint32_t mid;
    if (split_axis >= 0) {
        uint32_t l = 0, r = count;
        while (l < r) {
            double f = ((double)centre(c, el[l], split_axis) - (double)cmin[split_axis]) / split_span *
                       (double)BINS;
            size_t bi = (size_t)f;
            if (bi > BINS - 1) bi = BINS - 1;
            if ((int)bi <= split_bin) {
                l++;
            } else {
                r--;
                uint32_t t = el[l]; el[l] = el[r]; el[r] = t;
            }
        }
        mid = l;
    } else {
        float sx = cmax[0] - cmin[0], sy = cmax[1] - cmin[1], sz = cmax[2] - cmin[2];
        int axis = (sx >= sy && sx >= sz) ? 0 : (sy >= sz ? 1 : 2);
        median_sort(c, el, count, axis);
        mid = count / 2;
    }
    uint32_t children = c->n_nodes;
    c->n_nodes += 2;
    nd = &c->nodes[index];
    nd->first = children;
    nd->count = 0;
    sah_rec(c, children, first, mid, depth + 1);
    sah_rec(c, children + 1, first + mid, count - mid, depth + 1);
}

uint32_t orc_sah_build(const float* lo, const float* hi, uint32_t n, uint32_t* perm,
                       orc_node* nodes) {
    if (n == 0) return 0;
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    sah_ctx c = {lo, hi, perm, nodes, 1};
    sah_rec(&c, 0, 0, n, 0);
    return c.n_nodes;
}

/* tree_cost (src/scene.rs:2383-2396) over the nodes reachable from the root
 * (the LBVH layout keeps unreachable slots inside collapsed subtrees). */
double orc_tree_cost(const orc_node* nodes, uint32_t n_nodes) {
    if (n_nodes == 0) return 0.0;
    double area = bounds_area(nodes[0].mn, nodes[0].mx);
    if (area <= 0.0) return 0.0;
    double sum = 0.0;
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (n_nodes + 2));
    uint32_t top = 0;
    stack[top++] = 0;
    while (top) {
        uint32_t i = stack[--top];
        if (i >= n_nodes) continue;
        uint32_t c = nodes[i].count > 1 ? nodes[i].count : 1;
        sum += bounds_area(nodes[i].mn, nodes[i].mx) * (double)c;
        if (nodes[i].count == 0 && top + 2 <= n_nodes) {
            stack[top++] = nodes[i].first;
            stack[top++] = nodes[i].first + 1;
        }
    }
    free(stack);
    return sum / area;
}

static uint32_t depth_rec(const orc_node* nodes, uint32_t n, uint32_t i, uint32_t d) {
    if (i >= n || d > 256) return d;
    if (nodes[i].count > 0) return d;
    uint32_t a = depth_rec(nodes, n, nodes[i].first, d + 1);
    uint32_t b = depth_rec(nodes, n, nodes[i].first + 1, d + 1);
    return a > b ? a : b;
}
uint32_t orc_tree_depth(const orc_node* nodes, uint32_t n_nodes) {
    if (n_nodes == 0) return 0;
    return depth_rec(nodes, n_nodes, 0, 0);
}

/* ---------------- whole scene (mirrors lux_bvh_build) ---------------- */
void orc_scene_capacity(const uint32_t* mesh_tri_count, uint32_t n_meshes, uint32_t n_inst,
                        uint32_t* tri_cap, uint32_t* blas_cap, uint32_t* tlas_cap) {
    uint32_t t = 0, b = 0;
    for (uint32_t m = 0; m < n_meshes; ++m) {
        t += mesh_tri_count[m];
        b += mesh_tri_count[m] > 1 ? 2 * mesh_tri_count[m] - 1 : 1;
    }
    *tri_cap = t;
    *blas_cap = b;
    *tlas_cap = n_inst > 1 ? 2 * n_inst - 1 : 1;
}

static void tri_bounds(const float* p, float* lo, float* hi) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = sel_min(sel_min(p[a], p[3 + a]), p[6 + a]);
        hi[a] = sel_max(sel_max(p[a], p[3 + a]), p[6 + a]);
    }
    lo[3] = 0.0f;
    hi[3] = 0.0f;
}

/* World bounds of an instance: the 8 corners of its BLAS root box through
 * the forward rows, padded by 4 ulp-scale units so that f32 rounding of the
 * corner transform can never make the box tighter than the geometry. */
static void instance_bounds(const float* fwd, const orc_node* root, float* lo, float* hi) {
    for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
    for (int k = 0; k < 8; ++k) {
        float p[3] = {(k & 1) ? root->mx[0] : root->mn[0], (k & 2) ? root->mx[1] : root->mn[1],
                      (k & 4) ? root->mx[2] : root->mn[2]};
        for (int r = 0; r < 3; ++r) {
            const float* w = fwd + 4 * r;
            float v = ((w[0] * p[0] + w[1] * p[1]) + w[2] * p[2]) + w[3] * 1.0f;
            lo[r] = sel_min(lo[r], v);
            hi[r] = sel_max(hi[r], v);
        }
    }
    for (int a = 0; a < 3; ++a) {
        float m = fmaxf(fabsf(lo[a]), fabsf(hi[a]));
        float pad = m * 4.76837158e-7f + 1e-30f;
        lo[a] = lo[a] - pad;
        hi[a] = hi[a] + pad;
    }
    lo[3] = hi[3] = 0.0f;
}

void orc_instance_bounds(const float* fwd, const orc_node* root, float* lo, float* hi) {
    instance_bounds(fwd, root, lo, hi);
}

/* Post-order refit of the subtree at `slot` (see orc_tlas_refit). */
static void refit_node(orc_node* nodes, uint32_t n_nodes, uint32_t slot, const float* lo, const float* hi,
                       const uint32_t* order, uint32_t n_inst) {
    if (slot >= n_nodes) return;
    orc_node* nd = &nodes[slot];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (nd->count > 0) {
        for (uint32_t j = nd->first; j < nd->first + nd->count && j < n_inst; ++j) {
            const float* l = lo + 4ull * order[j];
            const float* h = hi + 4ull * order[j];
            for (int a = 0; a < 3; ++a) {
                mn[a] = sel_min(mn[a], l[a]);
                mx[a] = sel_max(mx[a], h[a]);
            }
        }
    } else {
        if (nd->first + 1u >= n_nodes) return;
        refit_node(nodes, n_nodes, nd->first, lo, hi, order, n_inst);
        refit_node(nodes, n_nodes, nd->first + 1u, lo, hi, order, n_inst);
        const orc_node* a = &nodes[nd->first];
        const orc_node* b = &nodes[nd->first + 1u];
        for (int k = 0; k < 3; ++k) {
            mn[k] = sel_min(a->mn[k], b->mn[k]);
            mx[k] = sel_max(a->mx[k], b->mx[k]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        nd->mn[a] = mn[a];
        nd->mx[a] = mx[a];
    }
}

/* refit (src/scene.rs:1040-1058): every node reachable from the root gets
 * the union of its instances' boxes (leaf; TLAS slot j holds instance
 * order[j]) or of its children's boxes, children first.  lo/hi: 4 floats
 * per instance.  Topology and unreachable slots are unchanged. */
void orc_tlas_refit(orc_node* nodes, uint32_t n_nodes, const float* lo, const float* hi, const uint32_t* order,
                    uint32_t n_inst) {
    if (n_nodes > 0) refit_node(nodes, n_nodes, 0u, lo, hi, order, n_inst);
}

int orc_scene_build(const float* const* mesh_pos, const uint32_t* const* mesh_mat,
                    const uint32_t* mesh_tri_count, uint32_t n_meshes, const float* inst_inv,
                    const float* inst_fwd, const uint32_t* inst_mesh, const uint32_t* inst_flags,
                    uint32_t n_inst, uint32_t leaf_max, int sah, float* tri, orc_node* blas,
                    orc_node* tlas, orc_instance* inst, orc_scene_sizes* sizes) {
    uint32_t tri_base = 0, blas_base = 0;
    uint32_t* mesh_tri_base = (uint32_t*)malloc(sizeof(uint32_t) * (n_meshes + 1));
    uint32_t* mesh_blas_base = (uint32_t*)malloc(sizeof(uint32_t) * (n_meshes + 1));
    for (uint32_t m = 0; m < n_meshes; ++m) {
        uint32_t n = mesh_tri_count[m];
        mesh_tri_base[m] = tri_base;
        mesh_blas_base[m] = blas_base;
        if (n == 0) { continue; }
        float* lo = (float*)malloc(sizeof(float) * 4 * n);
        float* hi = (float*)malloc(sizeof(float) * 4 * n);
        for (uint32_t i = 0; i < n; ++i) tri_bounds(mesh_pos[m] + 9ull * i, lo + 4 * i, hi + 4 * i);
        uint32_t* codes = (uint32_t*)malloc(sizeof(uint32_t) * n);
        uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * n);
        uint32_t cnt = sah ? orc_sah_build(lo, hi, n, perm, blas + blas_base)
                           : orc_lbvh_build(lo, hi, n, leaf_max, codes, perm, blas + blas_base);
        for (uint32_t k = 0; k < cnt; ++k) {
            orc_node* nd = &blas[blas_base + k];
            /* unused slots stay all-zero (no internal node has first == 0) */
            if (nd->count == 0 && nd->first == 0) continue;
            if (nd->count == 0) nd->first += blas_base;
            else nd->first += tri_base;
        }
        for (uint32_t i = 0; i < n; ++i) {
            const float* p = mesh_pos[m] + 9ull * perm[i];
            float* t = tri + 12ull * (tri_base + i);
            t[0] = p[0]; t[1] = p[1]; t[2] = p[2]; t[3] = bitsf(perm[i]);
            t[4] = p[3]; t[5] = p[4]; t[6] = p[5];
            t[7] = bitsf(mesh_mat && mesh_mat[m] ? mesh_mat[m][perm[i]] : 0u);
            t[8] = p[6]; t[9] = p[7]; t[10] = p[8]; t[11] = 0.0f;
        }
        free(lo); free(hi); free(codes); free(perm);
        tri_base += n;
        /* sah 1: BLASes packed; sah 2: each in its LBVH-sized region (2n - 1) */
        blas_base += sah == 1 ? cnt : (n > 1 ? 2 * n - 1 : 1);
    }
    mesh_tri_base[n_meshes] = tri_base;
    mesh_blas_base[n_meshes] = blas_base;
    /* instances */
    uint32_t n_tlas = 0;
    if (n_inst > 0) {
        float* lo = (float*)malloc(sizeof(float) * 4 * n_inst);
        float* hi = (float*)malloc(sizeof(float) * 4 * n_inst);
        for (uint32_t i = 0; i < n_inst; ++i) {
            uint32_t m = inst_mesh[i];
            if (m >= n_meshes || mesh_tri_count[m] == 0) {
                lo[4 * i] = lo[4 * i + 1] = lo[4 * i + 2] = INFINITY;
                hi[4 * i] = hi[4 * i + 1] = hi[4 * i + 2] = -INFINITY;
                lo[4 * i + 3] = hi[4 * i + 3] = 0.0f;
                continue;
            }
            instance_bounds(inst_fwd + 12ull * i, &blas[mesh_blas_base[m]], lo + 4 * i, hi + 4 * i);
        }
        uint32_t* codes = (uint32_t*)malloc(sizeof(uint32_t) * n_inst);
        uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * n_inst);
        n_tlas = sah == 1 ? orc_sah_build(lo, hi, n_inst, perm, tlas)
                     : orc_lbvh_build(lo, hi, n_inst, leaf_max, codes, perm, tlas);
        if (sah != 1) n_tlas = n_inst > 1 ? 2 * n_inst - 1 : 1;
        for (uint32_t k = 0; k < n_inst; ++k) {
            uint32_t src = perm[k];
            orc_instance* in = &inst[k];
            memset(in, 0, sizeof *in);
            memcpy(in->inv, inst_inv + 12ull * src, 48);
            memcpy(in->fwd, inst_fwd + 12ull * src, 48);
            uint32_t m = inst_mesh[src];
            in->flags = inst_flags ? inst_flags[src] : 0u;
            in->mesh = m;
            in->orig = src;
            if (m < n_meshes) {
                in->root = mesh_blas_base[m];
                in->first = mesh_tri_base[m];
                in->count = mesh_tri_count[m];
            }
        }
        free(lo); free(hi); free(codes); free(perm);
    }
    sizes->n_tri = tri_base;
    sizes->n_blas = blas_base;
    sizes->n_tlas = n_tlas;
    sizes->n_inst = n_inst;
    free(mesh_tri_base);
    free(mesh_blas_base);
    return 0;
}

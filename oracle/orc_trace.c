/*
 * orc_trace.c -- oracle traversal (TEST INFRASTRUCTURE, never the product).
 *
 * Restates, operation for operation in IEEE f32 with no contraction:
 *   boxRayInterval        shaders/traversal.slang:349-369
 *   orientedEdge          shaders/traversal.slang:379-392
 *   triangleRay           shaders/traversal.slang:399-413
 *   intersectTriangle     shaders/traversal.slang:415-443
 *   traceTriangle         shaders/traversal.slang:445-473 (opaque materials)
 *   traceInstance         shaders/traversal.slang:474-566
 *   traceRayTracked       shaders/traversal.slang:696-811 (no dependency/cone),
 *                         closest hit or, with ORC_RAY_FIRST_HIT, its anyHit mode
 *   occludedTriangle      shaders/traversal.slang:826-843
 *   occludedInstance      shaders/traversal.slang:845-932
 *   traceOccluded         shaders/traversal.slang:955-1046
 *   brute_force           src/traversal_tests.rs:978-1038
 *   gpuSortLinkedList     shaders/linked_merge_sort.slang:3-37
 * Dot products are summed left to right; "dot(row, (p,1))" keeps the + w*1
 * term and "dot(row, (v,0))" keeps the + w*0 term, as the shader writes them.
 *
 * Pinned by (tests/test_oracle_*.py): the shared-edge fixture of
 * src/traversal_tests.rs:2909-2998, the f64 brute force over the reference's
 * ray families (src/traversal_tests.rs:1040-1125), the forced deep-chain
 * fixtures with their expected error bits (:1521-1593, :2270-2283), and the
 * linked-sort fixture (:2824-2905).
 */
#include "lux_oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <pthread.h>

#define MAX_TRAVERSAL_STEPS 4096u
#define MAX_LEAF_TRIANGLES 256u
#define MAX_LEAF_INSTANCES 4096u
#define STACK_CAP 49u

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline int finitef_(float f) { return (fbits(f) & 0x7f800000u) != 0x7f800000u; }
static inline float fmin_(float a, float b) { return fminf(a, b); }
static inline float fmax_(float a, float b) { return fmaxf(a, b); }

int orc_box_interval(const float o[3], const float d[3], const float mn[3], const float mx[3],
                     float limit, float* entry, float* exit_) {
    float near_ = 0.0f, far_ = limit;
    for (int axis = 0; axis < 3; ++axis) {
        if (d[axis] == 0.0f) {
            if (o[axis] < mn[axis] || o[axis] > mx[axis]) {
                *entry = 0.0f;
                *exit_ = 0.0f;
                return 0;
            }
        } else {
            float a = (mn[axis] - o[axis]) / d[axis];
            float b = (mx[axis] - o[axis]) / d[axis];
            float axis_exit = fmax_(a, b);
            near_ = fmax_(near_, fmin_(a, b));
            float pad = fabsf(axis_exit) * 3.5762791e-7f;
            far_ = fmin_(far_, axis_exit + pad);
        }
    }
    *entry = near_;
    *exit_ = far_;
    return near_ <= far_;
}

static inline float box_entry(const float o[3], const float d[3], const orc_node* n, float limit) {
    float e, x;
    return orc_box_interval(o, d, n->mn, n->mx, limit, &e, &x) ? e : -1.0f;
}

static inline float oriented_edge(float ax, float ay, float bx, float by) {
    int reverse = ax > bx || (ax == bx && ay > by);
    float fx = reverse ? bx : ax, fy = reverse ? by : ay;
    float sx = reverse ? ax : bx, sy = reverse ? ay : by;
    float pa = fx * sy;
    float pb = fy * sx;
    float edge = pa - pb;
    float err = (fabsf(pa) + fabsf(pb)) * 2.3841858e-7f;
    if (fabsf(edge) <= err) edge = 0.0f;
    return reverse ? -edge : edge;
}

typedef struct {
    int kx, ky, kz;
    float sx, sy;
    int valid;
} tri_ray;

static tri_ray make_tri_ray(const float d[3]) {
    tri_ray r;
    memset(&r, 0, sizeof r);
    float m0 = fabsf(d[0]), m1 = fabsf(d[1]), m2 = fabsf(d[2]);
    r.kz = m0 > m1 ? (m0 > m2 ? 0 : 2) : (m1 > m2 ? 1 : 2);
    r.valid = d[r.kz] != 0.0f;
    if (!r.valid) return r;
    r.kx = (r.kz + 1) % 3;
    r.ky = (r.kx + 1) % 3;
    if (d[r.kz] < 0.0f) { int s = r.kx; r.kx = r.ky; r.ky = s; }
    r.sx = d[r.kx] / d[r.kz];
    r.sy = d[r.ky] / d[r.kz];
    return r;
}

static void intersect_tri(const float o[3], const float d[3], const tri_ray* r, const float* ta,
                          const float* tb, const float* tc, float out[3]) {
    out[0] = -1.0f; out[1] = 0.0f; out[2] = 0.0f;
    if (!r->valid) return;
    float A[3], B[3], C[3];
    for (int i = 0; i < 3; ++i) { A[i] = ta[i] - o[i]; B[i] = tb[i] - o[i]; C[i] = tc[i] - o[i]; }
    int kx = r->kx, ky = r->ky, kz = r->kz;
    float sx = r->sx, sy = r->sy;
    float pax = A[kx] - sx * A[kz], pay = A[ky] - sy * A[kz];
    float pbx = B[kx] - sx * B[kz], pby = B[ky] - sy * B[kz];
    float pcx = C[kx] - sx * C[kz], pcy = C[ky] - sy * C[kz];
    float largest = fmax_(fmax_(fmax_(fabsf(pax), fabsf(pay)), fmax_(fabsf(pbx), fabsf(pby))),
                          fmax_(fabsf(pcx), fabsf(pcy)));
    if (largest == 0.0f || !finitef_(largest)) return;
    uint32_t exponent = (fbits(largest) >> 23) & 255u;
    uint32_t ce = exponent < 1u ? 1u : (exponent > 253u ? 253u : exponent);
    float scale = bitsf((254u - ce) << 23);
    pax *= scale; pay *= scale; pbx *= scale; pby *= scale; pcx *= scale; pcy *= scale;
    float e0 = oriented_edge(pbx, pby, pcx, pcy);
    float e1 = oriented_edge(pcx, pcy, pax, pay);
    float e2 = oriented_edge(pax, pay, pbx, pby);
    int anyneg = e0 < 0.0f || e1 < 0.0f || e2 < 0.0f;
    int anypos = e0 > 0.0f || e1 > 0.0f || e2 > 0.0f;
    if (anyneg && anypos) return;
    float det = (e0 + e1) + e2;
    if (det == 0.0f || !finitef_(det)) return;
    float w0 = e0 / det, w1 = e1 / det, w2 = e2 / det;
    float dist = ((w0 * A[kz] + w1 * B[kz]) + w2 * C[kz]) / d[kz];
    if (!finitef_(dist)) return;
    out[0] = dist; out[1] = w1; out[2] = w2;
}

void orc_intersect_triangle(const float o[3], const float d[3], const float a[3], const float b[3],
                            const float c[3], float out[3]) {
    tri_ray r = make_tri_ray(d);
    intersect_tri(o, d, &r, a, b, c, out);
}

static inline void load_node(const orc_node* base, uint32_t n, uint32_t idx, orc_node* out) {
    /* The reference clamps scene word reads (traversal.slang:114-118); an
     * out-of-range node reads as the last valid one. */
    if (n == 0) { memset(out, 0, sizeof *out); return; }
    *out = base[idx < n ? idx : n - 1];
}

static inline void xform_point(const float* rows, const float p[3], float out[3]) {
    for (int r = 0; r < 3; ++r) {
        const float* w = rows + 4 * r;
        out[r] = ((w[0] * p[0] + w[1] * p[1]) + w[2] * p[2]) + w[3] * 1.0f;
    }
}
static inline void xform_dir(const float* rows, const float v[3], float out[3]) {
    for (int r = 0; r < 3; ++r) {
        const float* w = rows + 4 * r;
        out[r] = ((w[0] * v[0] + w[1] * v[1]) + w[2] * v[2]) + w[3] * 0.0f;
    }
}

static inline int culled(const float* ta, const float* tb, const float* tc, const float d[3],
                         uint32_t flags) {
    if ((flags & 3u) == 0u) return 0;
    float e1x = tb[0] - ta[0], e1y = tb[1] - ta[1], e1z = tb[2] - ta[2];
    float e2x = tc[0] - ta[0], e2y = tc[1] - ta[1], e2z = tc[2] - ta[2];
    float cx = e1y * e2z - e1z * e2y;
    float cy = e1z * e2x - e1x * e2z;
    float cz = e1x * e2y - e1y * e2x;
    float facing = ((cx * d[0] + cy * d[1]) + cz * d[2]) * ((flags & 4u) != 0u ? -1.0f : 1.0f);
    return ((flags & 1u) != 0u && facing >= 0.0f) || ((flags & 2u) != 0u && facing <= 0.0f);
}

typedef struct {
    float t;
    uint32_t tri, inst, err;
    float b1, b2;
} hit_state;

/* traceInstance, closest hit (shaders/traversal.slang:474-566). */
static void trace_instance(const orc_scene* s, const float o[3], const float d[3], uint32_t ii,
                           hit_state* h, orc_stats* st, int any_hit) {
    const orc_instance* in = &s->inst[ii];
    if (st) st->inst_visits++;
    float lo[3], ld[3];
    xform_point(in->inv, o, lo);
    xform_dir(in->inv, d, ld);
    uint32_t flags = in->flags;
    uint32_t stk_node[STACK_CAP];
    float stk_t[STACK_CAP];
    uint32_t count = 0;
    orc_node node;
    load_node(s->blas, s->n_blas, in->root, &node);
    if (st) st->node_fetches++;
    float e, x;
    int ready = orc_box_interval(lo, ld, node.mn, node.mx, h->t, &e, &x) && e >= 0.0f;
    tri_ray ray;
    memset(&ray, 0, sizeof ray);
    if (ready) ray = make_tri_ray(ld);
    uint32_t steps = 0;
    while (ready || count != 0) {
        if (++steps > MAX_TRAVERSAL_STEPS) { h->err |= ORC_ERR_UNSUPPORTED; break; }
        if (!ready) {
            --count;
            if (stk_t[count] > h->t) continue;
            load_node(s->blas, s->n_blas, stk_node[count], &node);
            if (st) st->node_fetches++;
        }
        ready = 0;
        if (node.count == 0) {
            if (count + 2 > STACK_CAP) {
                h->err |= ORC_ERR_BLAS_OVERFLOW;
                node.first = in->first;
                node.count = in->count;
                count = 0;
            } else {
                uint32_t li = node.first, ri = li + 1;
                orc_node ln, rn;
                load_node(s->blas, s->n_blas, li, &ln);
                load_node(s->blas, s->n_blas, ri, &rn);
                if (st) st->node_fetches += 2;
                float le = box_entry(lo, ld, &ln, h->t);
                float re = box_entry(lo, ld, &rn, h->t);
                if (le >= 0.0f && re >= 0.0f) {
                    if (le <= re) {
                        stk_node[count] = ri; stk_t[count] = re; count++;
                        node = ln;
                    } else {
                        stk_node[count] = li; stk_t[count] = le; count++;
                        node = rn;
                    }
                    ready = 1;
                } else if (le >= 0.0f) {
                    node = ln; ready = 1;
                } else if (re >= 0.0f) {
                    node = rn; ready = 1;
                }
                continue;
            }
        }
        uint32_t gfirst = in->first, gcount = in->count;
        uint32_t tbegin = node.first > gfirst ? node.first : gfirst;
        uint32_t tlimit = gfirst + gcount;
        uint32_t req_end = node.first + node.count;
        uint32_t tend = req_end < tlimit ? req_end : tlimit;
        int truncated = node.first < gfirst || req_end < node.first || req_end > tlimit ||
                        node.count > MAX_LEAF_TRIANGLES;
        uint32_t bend = tbegin + MAX_LEAF_TRIANGLES;
        if (bend > tend) bend = tend;
        truncated = truncated || bend < tend;
        if (truncated) h->err |= ORC_ERR_UNSUPPORTED;
        for (uint32_t i = tbegin; i < bend; ++i) {
            const float* t = s->tri + 12ull * (i < s->n_tri ? i : s->n_tri - 1);
            if (st) st->tri_tests++;
            if (culled(t, t + 4, t + 8, ld, flags)) continue;
            float r3[3];
            intersect_tri(lo, ld, &ray, t, t + 4, t + 8, r3);
            if (r3[0] > 0.000025f && r3[0] < h->t) {
                h->t = r3[0];
                h->tri = i;
                h->inst = ii;
                h->b1 = r3[1];
                h->b2 = r3[2];
            }
            if (any_hit && h->tri != ORC_EMPTY) return;
        }
    }
}

static void trace_closest(const orc_scene* s, const orc_ray* r, hit_state* h, orc_stats* st,
                          int any_hit) {
    h->t = r->tmax; h->tri = ORC_EMPTY; h->inst = ORC_EMPTY; h->err = 0; h->b1 = h->b2 = 0.0f;
    if (s->n_inst == 0) return;
    const float o[3] = {r->ox, r->oy, r->oz};
    const float d[3] = {r->dx, r->dy, r->dz};
    uint32_t stk_node[STACK_CAP];
    float stk_t[STACK_CAP];
    uint32_t count = 0;
    orc_node node;
    load_node(s->tlas, s->n_tlas, 0, &node);
    if (st) st->node_fetches++;
    float re, rx;
    if (!orc_box_interval(o, d, node.mn, node.mx, h->t, &re, &rx)) return;
    int ready = re >= 0.0f;
    uint32_t steps = 0, ninst = s->n_inst;
    while (ready || count != 0) {
        if (++steps > MAX_TRAVERSAL_STEPS) { h->err |= ORC_ERR_UNSUPPORTED; break; }
        if (!ready) {
            --count;
            if (stk_t[count] > h->t) continue;
            load_node(s->tlas, s->n_tlas, stk_node[count], &node);
            if (st) st->node_fetches++;
        }
        ready = 0;
        if (node.count == 0) {
            if (count + 2 > STACK_CAP) {
                h->err |= ORC_ERR_TLAS_OVERFLOW;
                node.first = 0; node.count = ninst; count = 0;
            } else {
                uint32_t li = node.first, ri = li + 1;
                orc_node ln, rn;
                load_node(s->tlas, s->n_tlas, li, &ln);
                load_node(s->tlas, s->n_tlas, ri, &rn);
                if (st) st->node_fetches += 2;
                float le = box_entry(o, d, &ln, h->t);
                float rre = box_entry(o, d, &rn, h->t);
                if (le >= 0.0f && rre >= 0.0f) {
                    if (le <= rre) {
                        stk_node[count] = ri; stk_t[count] = rre; count++;
                        node = ln;
                    } else {
                        stk_node[count] = li; stk_t[count] = le; count++;
                        node = rn;
                    }
                    ready = 1;
                } else if (le >= 0.0f) {
                    node = ln; ready = 1;
                } else if (rre >= 0.0f) {
                    node = rn; ready = 1;
                }
                continue;
            }
        }
        uint32_t req_end = node.first + node.count;
        int truncated = req_end < node.first || node.first > ninst || req_end > ninst ||
                        node.count > MAX_LEAF_INSTANCES;
        uint32_t ib = node.first < ninst ? node.first : ninst;
        uint32_t ie = req_end < ninst ? req_end : ninst;
        uint32_t be = ib + MAX_LEAF_INSTANCES;
        if (be > ie) be = ie;
        truncated = truncated || be < ie;
        if (truncated) h->err |= ORC_ERR_UNSUPPORTED;
        for (uint32_t i = ib; i < be; ++i) {
            trace_instance(s, o, d, i, h, st, any_hit);
            if (any_hit && h->tri != ORC_EMPTY) {
                /* an accepted hit proves an any-hit query (traversal.slang:786-792) */
                h->err &= ~ORC_ERR_UNSUPPORTED;
                count = 0;
                break;
            }
        }
    }
}

/* occludedInstance (shaders/traversal.slang:845-932). Returns blocked. */
static int occluded_instance(const orc_scene* s, const float o[3], const float d[3], uint32_t ii,
                             float limit, int* supported, hit_state* h, orc_stats* st) {
    *supported = 1;
    const orc_instance* in = &s->inst[ii];
    if (st) st->inst_visits++;
    float lo[3], ld[3];
    xform_point(in->inv, o, lo);
    xform_dir(in->inv, d, ld);
    uint32_t flags = in->flags;
    uint32_t stk[STACK_CAP];
    uint32_t count = 0;
    orc_node node;
    load_node(s->blas, s->n_blas, in->root, &node);
    if (st) st->node_fetches++;
    float e, x;
    int ready = orc_box_interval(lo, ld, node.mn, node.mx, limit, &e, &x) && e >= 0.0f;
    tri_ray ray;
    memset(&ray, 0, sizeof ray);
    if (ready) ray = make_tri_ray(ld);
    uint32_t steps = 0;
    int incomplete = 0;
    while (ready || count != 0) {
        if (++steps > MAX_TRAVERSAL_STEPS) { *supported = 0; return 0; }
        if (!ready) {
            load_node(s->blas, s->n_blas, stk[--count], &node);
            if (st) st->node_fetches++;
        }
        ready = 0;
        if (node.count == 0) {
            if (count + 2 > STACK_CAP) {
                node.first = in->first; node.count = in->count; count = 0;
            } else {
                uint32_t li = node.first, ri = li + 1;
                orc_node ln, rn;
                load_node(s->blas, s->n_blas, li, &ln);
                load_node(s->blas, s->n_blas, ri, &rn);
                if (st) st->node_fetches += 2;
                float le = box_entry(lo, ld, &ln, limit);
                float re = box_entry(lo, ld, &rn, limit);
                if (le >= 0.0f && re >= 0.0f) {
                    if (le <= re) { stk[count++] = ri; node = ln; }
                    else { stk[count++] = li; node = rn; }
                    ready = 1;
                } else if (le >= 0.0f) {
                    node = ln; ready = 1;
                } else if (re >= 0.0f) {
                    node = rn; ready = 1;
                }
                continue;
            }
        }
        uint32_t gfirst = in->first, gcount = in->count;
        uint32_t tbegin = node.first > gfirst ? node.first : gfirst;
        uint32_t tlimit = gfirst + gcount;
        uint32_t req_end = node.first + node.count;
        uint32_t tend = req_end < tlimit ? req_end : tlimit;
        int truncated = node.first < gfirst || req_end < node.first || req_end > tlimit ||
                        node.count > MAX_LEAF_TRIANGLES;
        uint32_t bend = tbegin + MAX_LEAF_TRIANGLES;
        if (bend > tend) bend = tend;
        truncated = truncated || bend < tend;
        incomplete = incomplete || truncated;
        for (uint32_t i = tbegin; i < bend; ++i) {
            const float* t = s->tri + 12ull * (i < s->n_tri ? i : s->n_tri - 1);
            if (st) st->tri_tests++;
            if (culled(t, t + 4, t + 8, ld, flags)) continue;
            float r3[3];
            intersect_tri(lo, ld, &ray, t, t + 4, t + 8, r3);
            if (r3[0] > 0.000025f && r3[0] < limit) {
                h->t = r3[0]; h->tri = i; h->inst = ii; h->b1 = r3[1]; h->b2 = r3[2];
                return 1;
            }
        }
    }
    *supported = !incomplete;
    return 0;
}

/* traceOccluded (shaders/traversal.slang:955-1046). */
static void trace_occluded(const orc_scene* s, const orc_ray* r, hit_state* h, orc_stats* st) {
    h->t = r->tmax; h->tri = ORC_EMPTY; h->inst = ORC_EMPTY; h->err = 0; h->b1 = h->b2 = 0.0f;
    if (s->n_inst == 0) return;
    const float o[3] = {r->ox, r->oy, r->oz};
    const float d[3] = {r->dx, r->dy, r->dz};
    float limit = r->tmax;
    uint32_t stk[STACK_CAP];
    uint32_t count = 0;
    orc_node node;
    load_node(s->tlas, s->n_tlas, 0, &node);
    if (st) st->node_fetches++;
    float re, rx;
    int root = orc_box_interval(o, d, node.mn, node.mx, limit, &re, &rx);
    if (root) h->err |= ORC_HIT_ROOT;
    int ready = root && re >= 0.0f;
    if (!ready) return;
    uint32_t steps = 0, ninst = s->n_inst;
    int incomplete = 0;
    while (ready || count != 0) {
        if (++steps > MAX_TRAVERSAL_STEPS) { h->err |= ORC_ERR_UNSUPPORTED; return; }
        if (!ready) {
            load_node(s->tlas, s->n_tlas, stk[--count], &node);
            if (st) st->node_fetches++;
        }
        ready = 0;
        if (node.count == 0) {
            if (count + 2 > STACK_CAP) {
                node.first = 0; node.count = ninst; count = 0;
            } else {
                uint32_t li = node.first, ri = li + 1;
                orc_node ln, rn;
                load_node(s->tlas, s->n_tlas, li, &ln);
                load_node(s->tlas, s->n_tlas, ri, &rn);
                if (st) st->node_fetches += 2;
                float le = box_entry(o, d, &ln, limit);
                float rre = box_entry(o, d, &rn, limit);
                if (le >= 0.0f && rre >= 0.0f) {
                    if (le <= rre) { stk[count++] = ri; node = ln; }
                    else { stk[count++] = li; node = rn; }
                    ready = 1;
                } else if (le >= 0.0f) {
                    node = ln; ready = 1;
                } else if (rre >= 0.0f) {
                    node = rn; ready = 1;
                }
                continue;
            }
        }
        uint32_t req_end = node.first + node.count;
        int truncated = req_end < node.first || node.first > ninst || req_end > ninst ||
                        node.count > MAX_LEAF_INSTANCES;
        uint32_t ib = node.first < ninst ? node.first : ninst;
        uint32_t ie = req_end < ninst ? req_end : ninst;
        uint32_t be = ib + MAX_LEAF_INSTANCES;
        if (be > ie) be = ie;
        incomplete = incomplete || truncated || be < ie;
        for (uint32_t i = ib; i < be; ++i) {
            int sup;
            int blocked = occluded_instance(s, o, d, i, limit, &sup, h, st);
            incomplete = incomplete || !sup;
            if (blocked) {
                h->err |= ORC_HIT_OCCLUDED;
                return;
            }
        }
    }
    if (incomplete) h->err |= ORC_ERR_UNSUPPORTED;
}

typedef struct {
    const orc_scene* s;
    const orc_ray* rays;
    orc_hit* hits;
    float* bary;
    orc_stats stats;
    int want_stats;
    uint32_t begin, end;
} trace_job;

static void* trace_worker(void* arg) {
    trace_job* j = (trace_job*)arg;
    orc_stats* st = j->want_stats ? &j->stats : NULL;
    for (uint32_t i = j->begin; i < j->end; ++i) {
        hit_state h;
        const orc_ray* r = &j->rays[i];
        if (r->flags & ORC_RAY_ANY_HIT) trace_occluded(j->s, r, &h, st);
        else trace_closest(j->s, r, &h, st, (r->flags & ORC_RAY_FIRST_HIT) != 0);
        orc_hit* o = &j->hits[i];
        o->t = h.t;
        /* Report caller-facing identities: the triangle's index within its
         * mesh (stored in A.w) and the caller's instance index. */
        o->triangle = h.tri == ORC_EMPTY ? ORC_EMPTY : fbits(j->s->tri[12ull * h.tri + 3]);
        o->instance = h.inst == ORC_EMPTY ? ORC_EMPTY : j->s->inst[h.inst].orig;
        o->flags = h.err;
        if (j->bary) { j->bary[2 * i] = h.b1; j->bary[2 * i + 1] = h.b2; }
    }
    return NULL;
}

void orc_trace(const orc_scene* s, const orc_ray* rays, uint32_t n, orc_hit* hits, float* bary,
               orc_stats* stats, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((uint32_t)threads > n) threads = n ? (int)n : 1;
    trace_job jobs[256];
    pthread_t tids[256];
    uint32_t per = (n + (uint32_t)threads - 1) / (uint32_t)threads;
    for (int t = 0; t < threads; ++t) {
        jobs[t].s = s; jobs[t].rays = rays; jobs[t].hits = hits; jobs[t].bary = bary;
        memset(&jobs[t].stats, 0, sizeof(orc_stats));
        jobs[t].want_stats = stats != NULL;
        jobs[t].begin = per * (uint32_t)t < n ? per * (uint32_t)t : n;
        jobs[t].end = jobs[t].begin + per < n ? jobs[t].begin + per : n;
    }
    if (threads == 1) {
        trace_worker(&jobs[0]);
    } else {
        for (int t = 0; t < threads; ++t) pthread_create(&tids[t], NULL, trace_worker, &jobs[t]);
        for (int t = 0; t < threads; ++t) pthread_join(tids[t], NULL);
    }
    if (stats) {
        memset(stats, 0, sizeof *stats);
        for (int t = 0; t < threads; ++t) {
            stats->node_fetches += jobs[t].stats.node_fetches;
            stats->tri_tests += jobs[t].stats.tri_tests;
            stats->inst_visits += jobs[t].stats.inst_visits;
        }
    }
}

/* ---------------- f64 brute force (src/traversal_tests.rs:978-1038) ------------- */
typedef struct { double t; uint32_t tri, inst; } bf_hit;
static int bf_cmp(const void* a, const void* b) {
    double x = ((const bf_hit*)a)->t, y = ((const bf_hit*)b)->t;
    return x < y ? -1 : (x > y ? 1 : 0);
}

uint32_t orc_brute_force(const orc_scene* s, const orc_ray* ray, double* ts, uint32_t* tris,
                         uint32_t* insts, uint32_t cap) {
    uint32_t found = 0, stored = 0;
    bf_hit* buf = (bf_hit*)malloc(sizeof(bf_hit) * (cap ? cap : 1) * 4 + 64);
    uint32_t bcap = (cap ? cap : 1) * 4 + 4;
    double o[3] = {ray->ox, ray->oy, ray->oz}, d[3] = {ray->dx, ray->dy, ray->dz};
    for (uint32_t ii = 0; ii < s->n_inst; ++ii) {
        const orc_instance* in = &s->inst[ii];
        for (uint32_t ti = in->first; ti < in->first + in->count; ++ti) {
            const float* t = s->tri + 12ull * ti;
            double v[3][3];
            for (int k = 0; k < 3; ++k) {
                const float* p = t + 4 * k;
                for (int r = 0; r < 3; ++r) {
                    const float* w = in->fwd + 4 * r;
                    v[k][r] = (double)w[0] * p[0] + (double)w[1] * p[1] + (double)w[2] * p[2] +
                              (double)w[3];
                }
            }
            double e1[3], e2[3];
            for (int r = 0; r < 3; ++r) { e1[r] = v[1][r] - v[0][r]; e2[r] = v[2][r] - v[0][r]; }
            double nrm[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                             e1[0] * e2[1] - e1[1] * e2[0]};
            double facing = nrm[0] * d[0] + nrm[1] * d[1] + nrm[2] * d[2];
            uint32_t cull = in->flags & 3u;
            if ((cull == 1 && facing >= 0.0) || (cull == 2 && facing <= 0.0)) continue;
            double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2],
                           d[0] * e2[1] - d[1] * e2[0]};
            double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
            if (det == 0.0) continue;
            double rel[3] = {o[0] - v[0][0], o[1] - v[0][1], o[2] - v[0][2]};
            double u = (rel[0] * p[0] + rel[1] * p[1] + rel[2] * p[2]) / det;
            double q[3] = {rel[1] * e1[2] - rel[2] * e1[1], rel[2] * e1[0] - rel[0] * e1[2],
                           rel[0] * e1[1] - rel[1] * e1[0]};
            double vv = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) / det;
            double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / det;
            double nl2 = nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2];
            double gu[3] = {(e2[1] * nrm[2] - e2[2] * nrm[1]) / nl2,
                            (e2[2] * nrm[0] - e2[0] * nrm[2]) / nl2,
                            (e2[0] * nrm[1] - e2[1] * nrm[0]) / nl2};
            double gv[3] = {(nrm[1] * e1[2] - nrm[2] * e1[1]) / nl2,
                            (nrm[2] * e1[0] - nrm[0] * e1[2]) / nl2,
                            (nrm[0] * e1[1] - nrm[1] * e1[0]) / nl2};
            double err[3];
            for (int r = 0; r < 3; ++r)
                err[r] = (fabs(o[r]) + fabs(d[r]) * fabs(tt) + fabs(v[0][r]) + fabs(v[1][r]) +
                          fabs(v[2][r])) * (8.0 * 1.1920928955078125e-07);
            double tu = fabs(gu[0]) * err[0] + fabs(gu[1]) * err[1] + fabs(gu[2]) * err[2];
            double tv = fabs(gv[0]) * err[0] + fabs(gv[1]) * err[1] + fabs(gv[2]) * err[2];
            double tuv = fabs(gu[0] + gv[0]) * err[0] + fabs(gu[1] + gv[1]) * err[1] +
                         fabs(gu[2] + gv[2]) * err[2];
            if (u < -tu || vv < -tv || u + vv > 1.0 + tuv) continue;
            if (tt > 0.000025 && tt < (double)ray->tmax) {
                if (found < bcap) {
                    buf[found].t = tt; buf[found].tri = ti; buf[found].inst = ii;
                }
                found++;
            }
        }
    }
    uint32_t m = found < bcap ? found : bcap;
    qsort(buf, m, sizeof(bf_hit), bf_cmp);
    for (uint32_t i = 0; i < m && stored < cap; ++i, ++stored) {
        ts[stored] = buf[i].t; tris[stored] = buf[i].tri; insts[stored] = buf[i].inst;
    }
    free(buf);
    return stored;
}

/* ---------------- linked merge sort (shaders/linked_merge_sort.slang:3-37) -------- */
static inline uint32_t ls_next(const uint32_t* nodes, uint32_t i) { return nodes[2u * i + 1u]; }
static inline void ls_set(uint32_t* nodes, uint32_t i, uint32_t n) { nodes[2u * i + 1u] = n; }
static inline int ls_less(const uint32_t* nodes, uint32_t l, uint32_t r) {
    int32_t a = (int32_t)nodes[2u * l], b = (int32_t)nodes[2u * r];
    return a < b || (a == b && l < r);
}

uint32_t orc_linked_sort(uint32_t* nodes, uint32_t head) {
    const uint32_t end = 0xffffffffu;
    if (head == end) return head;
    uint32_t width = 1;
    for (uint32_t pass = 0; pass < 32; ++pass) {
        uint32_t input = head, output = end, tail = end, runs = 0;
        while (input != end) {
            ++runs;
            uint32_t left = input, right = input, left_count = 0;
            for (uint32_t i = 0; i < width && right != end; ++i) {
                ++left_count;
                right = ls_next(nodes, right);
            }
            uint32_t right_count = width;
            while (left_count > 0 || (right_count > 0 && right != end)) {
                uint32_t sel;
                if (left_count == 0) {
                    sel = right; right = ls_next(nodes, right); --right_count;
                } else if (right_count == 0 || right == end || !ls_less(nodes, right, left)) {
                    sel = left; left = ls_next(nodes, left); --left_count;
                } else {
                    sel = right; right = ls_next(nodes, right); --right_count;
                }
                if (tail == end) output = sel;
                else ls_set(nodes, tail, sel);
                tail = sel;
            }
            input = right;
        }
        ls_set(nodes, tail, end);
        head = output;
        if (runs <= 1) return head;
        width *= 2;
    }
    return head;
}

/* Single-ray entry for the transport oracle: returns the caller-facing hit and
 * the BVH slots of the hit triangle and instance (ORC_EMPTY on a miss). */
void orc_trace_ray(const orc_scene* s, const orc_ray* r, orc_hit* out, float* bary2,
                   uint32_t* tri_slot, uint32_t* inst_slot) {
    hit_state h;
    if (r->flags & ORC_RAY_ANY_HIT) trace_occluded(s, r, &h, NULL);
    else trace_closest(s, r, &h, NULL, (r->flags & ORC_RAY_FIRST_HIT) != 0);
    out->t = h.t;
    out->triangle = h.tri == ORC_EMPTY ? ORC_EMPTY : fbits(s->tri[12ull * h.tri + 3]);
    out->instance = h.inst == ORC_EMPTY ? ORC_EMPTY : s->inst[h.inst].orig;
    out->flags = h.err;
    if (bary2) { bary2[0] = h.b1; bary2[1] = h.b2; }
    if (tri_slot) *tri_slot = h.tri;
    if (inst_slot) *inst_slot = h.inst;
}

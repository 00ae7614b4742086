/*
 * lux_oracle.h -- CPU oracle for the transport hot path (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it.  It restates, in plain C with IEEE f32 arithmetic and no floating-point
 * contraction (-ffp-contract=off), the algorithms of the reference's Split
 * transport path so that the HIP kernels can be compared bit for bit:
 *
 *   traversal      shaders/traversal.slang:349-443 (box/triangle tests),
 *                  :474-566 (traceInstance), :696-811 (traceRayTracked),
 *                  :826-1046 (occludedTriangle/occludedInstance/traceOccluded)
 *   SAH build      src/scene.rs:2374-2554 (build_bvh/build_tree/tree_cost)
 *   Morton LBVH    new (the reference has no Morton build); order contract is
 *                  gpuSortLess, shaders/split.slang:3-10
 *   linked sort    shaders/linked_merge_sort.slang:3-37
 *   brute force    src/traversal_tests.rs:978-1038 (f64 Moller-Trumbore)
 *   transport      shaders/split.slang, shaders/cascades.slang:60-268,
 *                  shaders/world_resolve.slang (see orc_transport.c)
 *
 * Parity pins (what each function is checked against) are listed per file.
 */
#ifndef LUX_ORACLE_H
#define LUX_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_EMPTY 0xffffffffu
#define ORC_ERR_BLAS_OVERFLOW 1u
#define ORC_ERR_TLAS_OVERFLOW 2u
#define ORC_ERR_UNSUPPORTED 4u
#define ORC_HIT_OCCLUDED 0x100u
#define ORC_HIT_ROOT 0x200u
#define ORC_RAY_ANY_HIT 1u   /* traceOccluded */
#define ORC_RAY_FIRST_HIT 2u /* traceRay(anyHit = true) */

/* 32-byte BVH node, reference layout (src/scene.rs:76-83): count > 0 is a
 * leaf over [first, first+count); count == 0 has children first, first+1. */
typedef struct {
    float mn[3];
    uint32_t first;
    float mx[3];
    uint32_t count;
} orc_node;

/* 128-byte instance record in TLAS order.  inv/fwd are row-major 3x4
 * (local_from_world rows, world_from_local rows) as in the reference's
 * 9-word instance record (src/scene.rs:1071-1122). */
typedef struct {
    float inv[12];
    float fwd[12];
    uint32_t flags, root, first, count;
    uint32_t orig, mesh;
    uint32_t material; /* 0: the triangles' materials; m + 1: material m for the whole instance */
    uint32_t pad1;
} orc_instance;

typedef struct {
    const orc_node* tlas;
    uint32_t n_tlas;
    const orc_instance* inst;
    uint32_t n_inst;
    const orc_node* blas;
    uint32_t n_blas;
    const float* tri; /* 12 floats per triangle: A.xyzw B.xyzw C.xyzw */
    uint32_t n_tri;
} orc_scene;

typedef struct {
    float ox, oy, oz, tmax;
    float dx, dy, dz;
    uint32_t flags;
} orc_ray;

typedef struct {
    float t;
    uint32_t triangle;
    uint32_t instance;
    uint32_t flags;
} orc_hit;

typedef struct {
    uint64_t node_fetches;
    uint64_t tri_tests;
    uint64_t inst_visits;
} orc_stats;

/* ---- geometric primitives (bit-exact restatements) ---- */
int orc_box_interval(const float o[3], const float d[3], const float mn[3], const float mx[3],
                     float limit, float* entry, float* exit_);
/* out = (t, w1, w2); t = -1 on miss */
void orc_intersect_triangle(const float o[3], const float d[3], const float a[3], const float b[3],
                            const float c[3], float out[3]);

/* ---- traversal ---- */
void orc_trace(const orc_scene* s, const orc_ray* rays, uint32_t n, orc_hit* hits,
               float* bary /* nullable, 2 per ray */, orc_stats* stats /* nullable */,
               int threads);

void orc_trace_ray(const orc_scene* s, const orc_ray* r, orc_hit* out, float* bary2,
                   uint32_t* tri_slot, uint32_t* inst_slot);

/* ---- LBVH (Morton-30) build, restated exactly as the GPU builds it ---- */
/* boxes: lo[4*n], hi[4*n] (xyz used).  Outputs: codes[n] (unsorted Morton),
 * perm[n] (sorted order), nodes[max(1,2n-1)].  Returns node count. */
uint32_t orc_lbvh_build(const float* lo, const float* hi, uint32_t n, uint32_t leaf_max,
                        uint32_t* codes, uint32_t* perm, orc_node* nodes);
void orc_morton_codes(const float* lo, const float* hi, uint32_t n, uint32_t* codes);
void orc_stable_sort_pairs(const uint32_t* keys, uint32_t n, uint32_t* perm);

/* Whole-scene build mirroring lux_bvh_build: meshes of non-indexed
 * triangles (9 floats each), instances with inv/fwd transforms. */
typedef struct {
    uint32_t n_tri;
    uint32_t n_blas;
    uint32_t n_tlas;
    uint32_t n_inst;
} orc_scene_sizes;
/* World bounds of an instance from its forward rows and BLAS root (the
 * boxes the TLAS is built from); TLAS refit to new instance boxes. */
void orc_instance_bounds(const float* fwd, const orc_node* root, float* lo, float* hi);
void orc_tlas_refit(orc_node* nodes, uint32_t n_nodes, const float* lo, const float* hi, const uint32_t* order,
                    uint32_t n_inst);
int orc_scene_build(const float* const* mesh_pos, const uint32_t* const* mesh_mat,
                    const uint32_t* mesh_tri_count, uint32_t n_meshes, const float* inst_inv,
                    const float* inst_fwd, const uint32_t* inst_mesh, const uint32_t* inst_flags,
                    uint32_t n_inst, uint32_t leaf_max, int sah,
                    /* outputs, caller-allocated with capacities from orc_scene_capacity */
                    float* tri, orc_node* blas, orc_node* tlas, orc_instance* inst,
                    orc_scene_sizes* sizes);
void orc_scene_capacity(const uint32_t* mesh_tri_count, uint32_t n_meshes, uint32_t n_inst,
                        uint32_t* tri_cap, uint32_t* blas_cap, uint32_t* tlas_cap);

/* ---- binned SAH build_tree restatement (src/scene.rs:2398-2554) ---- */
/* Reorders perm in place; returns node count (<= 2n-1). */
uint32_t orc_sah_build(const float* lo, const float* hi, uint32_t n, uint32_t* perm,
                       orc_node* nodes);
double orc_tree_cost(const orc_node* nodes, uint32_t n_nodes);
uint32_t orc_tree_depth(const orc_node* nodes, uint32_t n_nodes);

/* ---- f64 brute force, world space (src/traversal_tests.rs:978-1038) ---- */
/* Writes up to cap candidate hits sorted by t: (t, global triangle slot,
 * TLAS-order instance).  Returns the number of candidates. */
uint32_t orc_brute_force(const orc_scene* s, const orc_ray* ray, double* ts, uint32_t* tris,
                         uint32_t* insts, uint32_t cap);

/* ---- linked merge sort (shaders/linked_merge_sort.slang:3-37) ---- */
/* nodes: 2 words per node (key as int32 bits, next). Sorts by (key, index). */
uint32_t orc_linked_sort(uint32_t* nodes, uint32_t head);

/* ---- transport: dense-grid interval trace, merge, pack, resolve ---- */
/* Same layout as lux_gi_params in include/lux.h. */
typedef struct {
    float origin[3];
    float spacing;
    uint32_t dims[3];
    uint32_t levels;
    uint32_t base_phi, base_y;
    uint32_t interval_mode; /* 0: end = len*4^L; 1: end = len*(2^(L+1)-1) */
    float interval_length;
    float normal_bias;
    uint32_t emissive_samples;
    uint32_t frame;
    uint32_t seed;
    uint32_t rays_per_interval;
    uint32_t terminal_tail;
    uint32_t flags;
    uint32_t query_limit; /* unused by the oracle (the GPU's admission budget) */
} orc_gi_params;

/* Emissive triangle record (GiLight, src/scene.rs:90-108), 96 B. */
typedef struct {
    float position[4];  /* a.xyz, -2 double-sided / -1 one-sided */
    float color[4];     /* emission.rgb, area */
    float direction[4]; /* unit normal.xyz, CDF boundary */
    float edge1[4];     /* b - a, discrete probability */
    float edge2[4];     /* c - a, 0 */
    uint32_t source[4]; /* TLAS instance slot, triangle slot, 0, 0 */
} orc_light;

typedef struct {
    float albedo[3];
    float roughness;
    float emission[3];
    float pad;
} orc_material;

uint32_t orc_gi_level_dims(const orc_gi_params* p, uint32_t level, uint32_t dims[3]);
uint64_t orc_gi_level_offset(const orc_gi_params* p, uint32_t level);
uint32_t orc_gi_level_directions(const orc_gi_params* p, uint32_t level);
void orc_gi_interval_bounds(const orc_gi_params* p, uint32_t level, float* start, float* end);
/* Emitter table of a built scene (build_emissive_lights, src/scene.rs:2311-2372).
 * Returns the light count (<= cap). */
uint32_t orc_build_lights(const float* const* mesh_pos, const uint32_t* const* mesh_mat,
                          const uint32_t* mesh_tri_count, uint32_t n_meshes, const float* inst_fwd,
                          const uint32_t* inst_mesh, const uint32_t* inst_flags, uint32_t n_inst,
                          const orc_material* mats, uint32_t n_mats, orc_light* out, uint32_t cap);
/* Surrogate queries of a level (float[9 * intervals of the level]) and the
 * level-local index of training sample i. */
void orc_gi_level_queries(const orc_gi_params* p, uint32_t level, float* out);
uint64_t orc_gi_sample_index(uint32_t seed, uint32_t i, uint64_t count);
void orc_bin_direction(uint32_t index, uint32_t nphi, uint32_t ny, float ox, float oy,
                       float out[3]);
void orc_cos_sin_2pi(float u, float* c, float* s);
/* Intervals (J.rgb, beta; beta = -1 where no supported sample) of every level. */
void orc_gi_trace(const orc_scene* s, const orc_light* lights, uint32_t n_lights,
                  const orc_material* mats, uint32_t n_mats, const orc_gi_params* p,
                  float* intervals, int threads);
void orc_gi_trace_stride(const orc_scene* s, const orc_light* lights, uint32_t n_lights,
                         const orc_material* mats, uint32_t n_mats, const orc_gi_params* p,
                         float* intervals, int threads, uint64_t stride);
void orc_gi_merge(const orc_gi_params* p, const float* intervals, float* merged);
/* Temporal history (mergeEntry's history part): settings and state. */
typedef struct {
    float radiance_blend, transmittance_blend;
    uint32_t history_frames, persistent;
    uint32_t valid;  /* a previous update wrote the history */
    uint32_t root_grew; /* the scene's root box grew: measured-empty dependencies are dropped */
} orc_history_params;
void orc_gi_trace_samples(const orc_scene* s, const orc_light* lights, uint32_t n_lights,
                          const orc_material* mats, uint32_t n_mats, const orc_gi_params* p,
                          float* intervals, uint32_t* samples, int threads);
void orc_gi_history(const orc_gi_params* p, const orc_history_params* h, const float* boxes,
                    uint32_t n_boxes, uint64_t count, float* intervals, const uint32_t* samples, const float* dep,
                    float* history);
void orc_gi_trace_dep(const orc_scene* s, const orc_light* lights, uint32_t n_lights,
                      const orc_material* mats, uint32_t n_mats, const orc_gi_params* p,
                      float* intervals, uint32_t* samples, float* dep, int threads);
void orc_gi_merge_history(const orc_gi_params* p, const float* intervals, const float* history, float* merged);
void orc_gi_merge_f64(const orc_gi_params* p, const double* intervals, double* merged);
/* 36 octahedral texels (irradiance/pi rgb, coverage) per level-0 probe. */
void orc_gi_pack(const orc_gi_params* p, const float* merged, float* probes);
void orc_gi_pack_f64(const orc_gi_params* p, const double* merged, double* probes);
/* receivers: 6 floats (position, normal); out: rgba (alpha = support code). */
void orc_world_resolve(const orc_gi_params* p, const float* probes, const float* receivers,
                       uint32_t n, float* out);
/* Camera terms of the split LOD choice (splitLodBlendForParameters,
 * shaders/common.slang:431-461); the layout of lux_lod_view. */
typedef struct {
    float camera_position[3];
    float probe_spacing;       /* tuning.x */
    float clip_row1[3];        /* clip_from_world row 1 (clip y), xyz */
    uint32_t viewport_height;  /* viewport.w */
    float clip_row3[3];        /* clip_from_world row 3 (clip w), xyz; zero: orthographic */
    uint32_t target_pixels;    /* ownerSequence.x (split_lod_target_pixels) */
} orc_lod_view;
void orc_lod_select(const orc_lod_view* v, const float* position, uint32_t* low, uint32_t* high,
                    float* blend);
/* resolveSplitDiffuse over a pyramid of dense grids (grid k = LOD k). */
void orc_world_resolve_lods(const orc_gi_params* const* grids, const float* const* probes, uint32_t count,
                            const orc_lod_view* view, const float* receivers, uint32_t n, float* out);
void orc_world_resolve_lods_f64(const orc_gi_params* const* grids, const double* const* probes,
                                uint32_t count, const orc_lod_view* view, const float* receivers, uint32_t n,
                                double* out);
void orc_world_resolve_f64(const orc_gi_params* p, const double* probes, const float* receivers,
                           uint32_t n, double* out);
/* ---- receiver-driven probe selection (orc_receivers.c) and the receiver
 * trace (orc_transport.c) ---- */
#define ORC_RX_DISABLED 1u /* ownerSequence.w bit 30: updates disabled */
#define ORC_RX_VALID 1u    /* descriptor flags: seeded */
#define ORC_RX_RAY 2u      /* descriptor flags: padding.x ray marker */
typedef struct {
    uint32_t width, height; /* receiver grid (params.size.xy); count = width * height */
    uint32_t budget;        /* sampling.x: 0 = every receiver */
    uint32_t fair_cursor;   /* ownerSequence.w low 26 bits: fair window start (rank) */
    uint32_t fair_quota;    /* ownerSequence.z */
    uint32_t frame;         /* params.frame.x */
    uint32_t flags;         /* ORC_RX_DISABLED */
    uint32_t reserved;
} orc_receiver_params;
uint32_t orc_candidate_rank(uint32_t sample, uint32_t width, uint32_t height, uint32_t count);
uint32_t orc_candidate_from_rank(uint32_t rank, uint32_t width, uint32_t height, uint32_t count);
int orc_candidate_selected(uint32_t sample, const orc_receiver_params* rp);
int orc_rx_surface(const float* receiver, float pos[3], float n[3]);
void orc_rx_cell(const float pos[3], float base_spacing, uint32_t level, uint32_t lod, int32_t key[4]);
uint32_t orc_rx_directory_capacity(uint32_t count);
/* Descriptors (variant, receiver, level) -> keys int32[4], flags, and the
 * LOD of each (variant, receiver). view may be NULL (LOD 0 everywhere). */
void orc_rx_seed(const float* receivers, const orc_receiver_params* rp, const orc_lod_view* view,
                 float base_spacing, uint32_t levels, int cold, int32_t* keys, uint32_t* flags, uint32_t* lods);
uint32_t orc_cold_candidate_rank(uint32_t sample, uint32_t width, uint32_t height, uint32_t count, uint32_t frame);
uint32_t orc_cold_candidate_from_rank(uint32_t rank, uint32_t width, uint32_t height, uint32_t count,
                                      uint32_t frame);
/* rank (with its cold-priority phases when cold)/parents/prefix/
 * generateDirections over 2 * count * levels descriptors.
 * Per descriptor: canonical index (EMPTY if not seeded), its probe's ray
 * count and offset; per (variant, receiver): sequence (EMPTY: no ray); the
 * ticket list.  Returns the root total. */
uint32_t orc_rx_select(uint32_t count, uint32_t levels, const int32_t* keys, const uint32_t* flags,
                       const orc_receiver_params* rp, int cold, uint32_t* canonical, uint32_t* probe_count,
                       uint32_t* probe_offset, uint32_t* sequence, uint32_t* tickets, uint32_t* n_tickets,
                       uint32_t* flags_out /* nullable: the flags after rank */);
/* directionIndex (common.slang:384-394) over nphi x ny bins. */
uint32_t orc_direction_index(const float d[3], uint32_t nphi, uint32_t ny);
/* Receiver trace + deposit + reduce into the dense grids of a LOD pyramid
 * (grid k = LOD k).  stats: [0] rays traced, [1] unsupported, [2] deposits
 * kept, [3] deposits outside the grids. */
void orc_rx_trace(const orc_scene* s, const orc_light* lights, uint32_t n_lights, const orc_material* mats,
                  uint32_t n_mats, const orc_gi_params* const* grids, uint32_t lod_count, const float* receivers,
                  uint32_t count, const uint32_t* lods, const uint32_t* sequence, const uint32_t* tickets,
                  uint32_t n_tickets, uint32_t frame, float* const* intervals, uint32_t* const* samples,
                  uint32_t stats[4]);

/* Shading in f64 (FP64 twin): same hits and visibility, f64 radiance. */
void orc_gi_trace_f64(const orc_scene* s, const orc_light* lights, uint32_t n_lights,
                      const orc_material* mats, uint32_t n_mats, const orc_gi_params* p,
                      double* intervals, int threads);

/* ---- screen-space temporal and denoise passes (orc_screen.c) ---- */
typedef struct {
    float world_from_clip[16];          /* row-major */
    float previous_clip_from_world[16];
    float previous_world_from_clip[16];
    uint32_t viewport[4];               /* x, y, width, height */
    uint32_t target[2];                 /* full-resolution image width, height */
    uint32_t size[2];                   /* receiver grid */
    uint32_t resolve_scale;
    uint32_t persistent;
    float probe_spacing, normal_bias;
    float history_blend, history_frames;
    float stability[4];
    uint32_t source_changes, source_flags, motion_epoch, frame;
    float dirty_min[4], dirty_max[4];
} orc_screen_params;
void orc_screen_extent(const orc_screen_params* P, uint32_t* w, uint32_t* h);
void orc_screen_temporal(const orc_screen_params* P, const float* current, const float* current_bg, const float* depth,
                         const uint32_t* gbuffer, const uint32_t* raster, const float* motion, const float* prev,
                         const float* prev_meta, const float* prev_recv, const float* prev_bg, float* out_irr,
                         float* out_meta, float* out_recv, float* out_bg);
void orc_screen_denoise(const orc_screen_params* P, uint32_t pass, const float* in_irr, const float* moments,
                        const float* in_recv, const float* depth, float* out, float* out_recv);
float orc_screen_packed_normal(const float* n);

#ifdef __cplusplus
}
#endif
#endif

/*
 * lux.h -- C ABI of the MI355X transport hot path (liblux_hip.so).
 *
 * This is the boundary a Rust FFI crate (lux-hip-sys) binds with
 * `extern "C"`.  Every entry point takes plain pointers and sizes; no
 * PyTorch or HIP types cross it (streams are opaque void*).
 *
 * It replaces, inside the reference's render() Split branch
 * (src/render.rs:9553-9801), the sequence of wgpu compute dispatches that
 * trace, reduce, merge, pack and resolve radiance intervals, and the CPU BVH
 * build of src/scene.rs:2374-2554 / :2014-2095.  The Bevy-facing types
 * (RadianceCascadesPlugin, RadianceCascades, GiMethod,
 * RadianceCascadesGeometry; src/lib.rs:305-307,431-518,631-669,795-900) are
 * unchanged and sit above this ABI.
 *
 * Status codes: 0 = OK, negative = error (lux_last_error has the text).
 * A context is thread-compatible: one thread at a time; several contexts may
 * coexist, each with its own stream.
 */
#ifndef LUX_H
#define LUX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUX_OK 0
#define LUX_ERR_INVALID -1
#define LUX_ERR_OOM -2
#define LUX_ERR_HIP -3
#define LUX_ERR_STATE -4

/* Per-ray result flags.  Bits 1, 2 and 4 are the reference's Hit.errors bits
 * (shaders/traversal.slang:12-16,500,515,743). */
#define LUX_HIT_BLAS_OVERFLOW 1u
#define LUX_HIT_TLAS_OVERFLOW 2u
#define LUX_HIT_UNSUPPORTED 4u
#define LUX_HIT_OCCLUDED 0x100u /* any-hit: an accepted blocker was found */
#define LUX_HIT_ROOT 0x200u     /* any-hit: the ray met the scene root box */

/* ray.flags: 0 = closest hit (traceRay), LUX_RAY_ANY_HIT = occlusion query
 * (traceOccluded), LUX_RAY_FIRST_HIT = first accepted hit (traceRay with
 * anyHit = true, shaders/traversal.slang:696-811). */
#define LUX_RAY_ANY_HIT 1u
#define LUX_RAY_FIRST_HIT 2u

#define LUX_EMPTY 0xffffffffu

/* 32 B ray: origin + tmax, direction + flags. */
typedef struct {
    float ox, oy, oz, tmax;
    float dx, dy, dz;
    uint32_t flags;
} lux_ray;

/* 16 B hit: closest (or first accepted, any-hit) distance, the triangle's
 * index within its mesh, the caller's instance index, flags. */
typedef struct {
    float t;
    uint32_t triangle;
    uint32_t instance;
    uint32_t flags;
} lux_hit;

/* Reference BvhNode layout (src/scene.rs:76-83), 32 B. */
typedef struct {
    float mn[3];
    uint32_t first;
    float mx[3];
    uint32_t count;
} lux_bvh_node;

/* One mesh of non-indexed triangles: positions[9*i .. 9*i+8] = a, b, c. */
typedef struct {
    const float* positions;
    const uint32_t* materials; /* per triangle material index, or NULL = 0 */
    uint32_t triangle_count;
} lux_mesh_desc;

/* One rigid instance.  Both transforms are row-major 3x4 and supplied by
 * the caller, as the reference's instance record stores both
 * (src/scene.rs:1071-1122).  flags: 1 cull back, 2 cull front, 4 flip
 * winding, 16 unlit (reference instance flags). */
typedef struct {
    float local_from_world[12];
    float world_from_local[12];
    uint32_t mesh;
    uint32_t flags;
} lux_instance_desc;

/* Constant material (configs use untextured, opaque materials). */
typedef struct {
    float albedo[3];
    float roughness;
    float emission[3];
    float pad;
} lux_material;

typedef struct lux_ctx lux_ctx;

/* Full traversal launch configuration.  regroup_threshold 0 runs each wave's
 * 64 rays to completion; 1..64 runs resumable walks and refills idle lanes
 * once that many are idle, checking every steps_per_refill walk steps
 * (0 = 8).  lds_stack_depth: traversal-stack slots per lane kept in LDS
 * (4, 8, 12 or 16; 0 = 8), deeper slots spill to private memory.
 * waves_per_cu: persistent grid size (0 = 28; 1000 = one ray per thread,
 * no persistence).  A new context starts at {28, 12, 6, 24}.  Results do not
 * depend on the configuration. */
typedef struct {
    uint32_t waves_per_cu;
    uint32_t lds_stack_depth;
    uint32_t regroup_threshold;
    uint32_t steps_per_refill;
    uint32_t reserved[4];
} lux_trace_config;

/* ---- context (replaces render::install's device/pipeline set-up,
 *      src/render.rs:1641-1760) ---- */
int lux_ctx_create(int device, lux_ctx** out);
void lux_ctx_destroy(lux_ctx* ctx);
const char* lux_last_error(const lux_ctx* ctx);
/* The context's HIP stream as an opaque handle (hipStream_t). */
void* lux_ctx_stream(lux_ctx* ctx);
int lux_ctx_synchronize(lux_ctx* ctx);

/* ---- device buffers owned by the caller (rays, hits, receivers) ----
 * Plain HIP allocations on the context's device; copies are synchronous
 * with respect to the context stream. */
int lux_device_alloc(lux_ctx* ctx, size_t bytes, void** out);
int lux_device_free(lux_ctx* ctx, void* ptr);
int lux_copy_to_device(lux_ctx* ctx, void* dst, const void* src, size_t bytes);
int lux_copy_to_host(lux_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy on the context stream (asynchronous). */
int lux_device_copy(lux_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Zero device memory with a fill kernel on the context stream. */
int lux_device_zero(lux_ctx* ctx, void* ptr, size_t bytes);

/* ---- scene (replaces scene.rs triangle/instance packing and uploads,
 *      src/scene.rs:809-823,1025-1122,2096-2117) ---- */
int lux_scene_upload_soa(lux_ctx* ctx, const lux_mesh_desc* meshes, uint32_t mesh_count,
                         const lux_instance_desc* instances, uint32_t instance_count,
                         const lux_material* materials, uint32_t material_count);
/* Replace instance transforms (same count/meshes); the next lux_bvh_build
 * rebuilds only the TLAS when the BLASes are already built. */
int lux_scene_set_transforms(lux_ctx* ctx, const float* local_from_world,
                             const float* world_from_local, uint32_t instance_count);

/* ---- BVH build (replaces build_bvh/build_tree, src/scene.rs:2374-2554) ---- */
#define LUX_BVH_MORTON30 0u
/* or-ed into the mode: rebuild every BLAS too, not only what changed */
#define LUX_BVH_FORCE_BLAS 0x100u
/* or-ed into the mode: when only instance transforms changed since the last
 * TLAS build, keep its topology and instance order and refit its boxes to
 * the new instance bounds (the reference's per-frame TLAS refit,
 * src/scene.rs:1040-1058,2033-2047); otherwise build as usual. */
#define LUX_BVH_REFIT 0x200u
/* or-ed into the mode: build each BLAS with the reference's binned SAH
 * (build_tree, src/scene.rs:2398-2554) on the GPU -- the reference's own
 * tree node for node (fewer node visits than Morton-30), every mesh in one
 * level-by-level pass -- instead of Morton-30; the TLAS stays Morton-30.
 * Triangle positions must be finite (LUX_ERR_INVALID otherwise).  A change
 * of builder rebuilds the BLASes. */
#define LUX_BVH_SAH 0x1u
/* or-ed into the mode instead of LUX_BVH_SAH: the same binned-SAH tree built
 * on the host from the uploaded positions (meshes on up to 16 threads, a
 * large mesh's top subtrees in parallel), for comparison with the GPU build. */
#define LUX_BVH_SAH_HOST 0x2u
/* Builds what is stale: the BLASes after lux_scene_upload_soa, the TLAS after
 * lux_scene_set_transforms (Morton-30 over instance boxes; no host wait). */
int lux_bvh_build(lux_ctx* ctx, uint32_t mode);
/* 128 B instance record of a built scene, in TLAS order (the reference's
 * 9-word record, src/scene.rs:1071-1122, minus layer/material words). */
typedef struct {
    float local_from_world[12];
    float world_from_local[12];
    uint32_t flags, blas_root, tri_first, tri_count;
    uint32_t orig_index, mesh;
    uint32_t material; /* 0: the triangles' materials; m + 1: material m for every triangle */
    uint32_t reserved1;
} lux_instance_record;

/* Install an already-built scene (packed-scene interchange): BLAS/TLAS nodes
 * in the reference layout, triangles as 12 floats each (A.xyz, index within
 * mesh as uint bits; B.xyz, material as uint bits; C.xyz, 0), instance
 * records in TLAS order.  Node and triangle indices are absolute. */
int lux_scene_upload_built(lux_ctx* ctx, const lux_bvh_node* blas, uint32_t blas_count,
                           const lux_bvh_node* tlas, uint32_t tlas_count, const float* triangles,
                           uint32_t triangle_count, const lux_instance_record* instances,
                           uint32_t instance_count);
/* Install a scene from the reference's packed inline scene and geometry
 * words (16-byte words; src/scene.rs:809-823 geometry_words, :1025-1038
 * node_words, :1071-1122 pack_instances, :2102-2116 the prefix header
 * (materials end, instance base, instance count, 9)).  The node array
 * (words 1 .. instance base) is installed as it is -- TLAS root 0, BLAS
 * roots and child indices absolute; each instance's transforms, flags
 * (word 6.w), material (word 7.w, an address into the material block that
 * follows the instances: material index = (address - block) / 8, taken from
 * `materials`), BLAS root and triangle range (word 8); triangle t's vertex
 * indices at geometry word 1 + t and vertex positions at vertex_base +
 * stride * index.  The geometry words are repacked on the device into the
 * triangle records; a hit reports t as its triangle.  Layer masks, textures
 * and paged words are not read.  The scene counts as built: transforms
 * cannot change (upload new words); the GI emitter table follows the
 * instance materials. */
int lux_scene_upload_words(lux_ctx* ctx, const uint32_t* scene_words, uint32_t n_scene_words,
                           const uint32_t* geometry_words, uint32_t n_geometry_words,
                           const lux_material* materials, uint32_t material_count);

/* Node/instance counts of the built scene. */
int lux_bvh_info(lux_ctx* ctx, uint32_t* blas_nodes, uint32_t* tlas_nodes,
                 uint32_t* triangles, uint32_t* instances);
/* Copy the built BLAS/TLAS nodes, the sorted-triangle permutation and the
 * TLAS instance order to host memory (for parity checks and interchange). */
int lux_bvh_download(lux_ctx* ctx, lux_bvh_node* blas, lux_bvh_node* tlas, uint32_t* tri_order,
                     uint32_t* inst_order);

/* Stable (key, index) sort of 32-bit keys on the GPU: the gpuSortLess order
 * contract (shaders/split.slang:3-10).  Device pointers. */
int lux_sort_pairs_device(lux_ctx* ctx, const uint32_t* keys_in, uint32_t n, uint32_t* keys_out,
                          uint32_t* index_out);
/* Morton-30 codes of primitive boxes (device pointers, lo/hi float4). */
int lux_morton_codes_device(lux_ctx* ctx, const float* lo4, const float* hi4, uint32_t n,
                            uint32_t* codes);

/* ---- traversal (replaces traceRayTracked/traceOccluded dispatches,
 *      shaders/traversal.slang:696-1046; split.slang trace/traceVisibility) ---- */
/* Host buffers: copies in, traces, copies out, synchronizes. */
int lux_trace_rays(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits);
/* Device buffers, asynchronous on the context stream. bary may be NULL
 * (else 2 floats per ray: w1, w2). */
int lux_trace_rays_device(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits,
                          float* bary);
/* lux_trace_rays_device that allocates nothing, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  The ray cursor lives in a 256-byte
 * library-owned device array per GPU (held until the process ends) instead
 * of context memory; no timing events are recorded (lux_last_kernel_ms keeps
 * the last timed trace).  n <= 2^31 (LUX_ERR_INVALID beyond).  Calls that
 * overlap on one GPU share the cursor and race: serialise them. */
int lux_trace_rays_device_noalloc(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits,
                                  float* bary);

/* Self-check of the traversal's slab division (no reference counterpart):
 * the resumable walk divides by a per-ray reciprocal with one correction
 * (Markstein) when the scene's node bounds and the ray lie in the range where
 * that quotient is the correctly rounded one; this counts, over `pairs`
 * seeded operand pairs of that range (random and all-ones / all-zeros /
 * near-boundary significands), those whose quotient differs from the IEEE
 * division in any bit.  Synchronous. */
int lux_check_reciprocal_division(lux_ctx* ctx, uint64_t pairs, uint32_t seed, uint64_t* mismatches);

/* Instrumented trace (same results): counters[0] += node fetches,
 * counters[1] += triangle tests, counters[2] += instance visits, summed over
 * all rays.  counters is a device pointer to 3 uint64 (zeroed by caller). */
int lux_trace_stats_device(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits,
                           unsigned long long* counters);

/* ---- transport: interval trace, cascade merge, probe packing, world resolve
 *      (replaces the Split dispatches trace/reduce/merge x L/packProbes/
 *      worldResolve, src/render.rs:9640-9801; shaders/split.slang:1056-1918,
 *      shaders/cascades.slang:60-268, shaders/world_resolve.slang) ---- */
#define LUX_INTERVAL_POW4 0u     /* interval end = length * 4^L (split.slang:224-231) */
#define LUX_INTERVAL_DOUBLING 1u /* interval end = length * (2^(L+1) - 1) */

/* Dense probe grid.  Level L has ceil(dims / 2^L) probes at spacing * 2^L
 * and (base_phi * 2^L) x (base_y * 2^L) equal-area direction bins
 * (reference preset: 8 x 4 = 32 at L0).  One interval = one (probe,
 * direction, level); its ray starts at the probe centre + start * dir. */
typedef struct {
    float origin[3];           /* world position of the grid corner */
    float spacing;             /* level-0 probe spacing (RadianceCascades probe spacing) */
    uint32_t dims[3];          /* level-0 probe counts */
    uint32_t levels;           /* cascades, 1..8 */
    uint32_t base_phi, base_y; /* level-0 direction bins */
    uint32_t interval_mode;    /* LUX_INTERVAL_* */
    float interval_length;     /* level-0 interval end */
    float normal_bias;         /* surface offset of visibility rays */
    uint32_t emissive_samples; /* SamplingSettings emissive samples; NEE uses max(1, n/2) */
    uint32_t frame;            /* frame index (NEE sequence) */
    uint32_t seed;             /* direction jitter seed */
    uint32_t rays_per_interval;
    uint32_t terminal_tail;    /* 1: the last level owns the unbounded tail */
    uint32_t flags;            /* reserved, 0 */
    uint32_t query_limit;      /* SamplingSettings::max_traversal_queries: per-update
                                  ticketed budget; interval and NEE traversals are
                                  admitted up to max(limit / 2, 1), a denied query
                                  makes its sample unsupported; 0 = unbounded
                                  (admitTransportQuery, shaders/traversal.slang:20-49;
                                  traversal_query_caps, src/render.rs:290-300) */
} lux_gi_params;

/* Allocate the per-level interval/merged arrays (float4 per interval:
 * J.rgb + beta, beta = -1 where no supported sample; I.rgb + support), the
 * packed probes (36 float4 per level-0 probe) and build the emitter table
 * from the uploaded scene (build_emissive_lights, src/scene.rs:2311-2372). */
int lux_gi_configure(lux_ctx* ctx, const lux_gi_params* params);
int lux_gi_set_frame(lux_ctx* ctx, uint32_t frame, uint32_t seed);
int lux_gi_trace(lux_ctx* ctx);
int lux_gi_merge(lux_ctx* ctx);
int lux_gi_pack(lux_ctx* ctx);
/* trace + merge + pack; *ms (nullable) = GPU time of the three phases. */
int lux_gi_update(lux_ctx* ctx, float* ms);
/* Opt-in stage timing of the wavefront interval trace (HIP events between
 * the stages; each chunk synchronizes): lux_gi_stage_times gives the last
 * traced update's milliseconds of {primary walks, NEE setup + occlusion-ray
 * packing, occlusion walks, accumulation}.  No reference counterpart. */
int lux_gi_set_stage_timing(lux_ctx* ctx, int on);
int lux_gi_stage_times(lux_ctx* ctx, float* ms);
/* One instrumented interval trace of the configured grid (wavefront, the
 * same results as lux_gi_trace; synchronous), counting the work that
 * defines its algorithmic bytes: counters[0] primary walks, [1..3] their
 * node fetches, triangle tests and instance visits, [4] occlusion rays,
 * [5..7] their node fetches, triangle tests and instance visits. */
int lux_gi_trace_stats(lux_ctx* ctx, uint64_t* counters);
/* Total intervals of all levels and level-0 probe count. */
int lux_gi_sizes(lux_ctx* ctx, uint64_t* intervals, uint32_t* probes, uint32_t* lights);
/* Host copies (each nullable): intervals/merged float[4*intervals],
 * probes float[4*36*probes], lights float[24*lights]. */
int lux_gi_download(lux_ctx* ctx, float* intervals, float* merged, float* probes, float* lights);
/* Replace the intervals (host float[4*intervals]) before lux_gi_merge. */
int lux_gi_upload_intervals(lux_ctx* ctx, const float* intervals);
/* Device addresses of the context-owned arrays (for zero-copy interop). */
int lux_gi_device_arrays(lux_ctx* ctx, float** intervals, float** merged, float** probes);
/* Resolve receivers (device float[6*n]: position, normal) against the packed
 * probes into device float[4*n]: diffuse irradiance/pi rgb and the support
 * code (0 not eligible, -1 unsupported, 1 + 0.25 coverage). */
int lux_world_resolve_device(lux_ctx* ctx, const float* receivers, uint32_t n, float* out);

/* Camera terms of the split LOD choice (splitLodBlendForParameters,
 * shaders/common.slang:431-461).  The clip rows are rows 1 (clip y) and 3
 * (clip w) of clip_from_world in column-vector convention, xyz only; row 3 is
 * zero for an orthographic camera.  48 bytes. */
typedef struct {
    float camera_position[3];
    float probe_spacing;        /* RadianceCascades::probe_spacing (tuning.x) */
    float clip_row1[3];
    uint32_t viewport_height;   /* pixels (viewport.w) */
    float clip_row3[3];
    uint32_t target_pixels;     /* split_lod_target_pixels (src/render.rs:181-183), >= 1 */
} lux_lod_view;

/* resolveSplitDiffuse (shaders/world_resolve.slang:184-219, worldResolve
 * :281-313) over a pyramid of dense probe grids: lods[k] is a configured
 * context whose grid is LOD k (spacing and interval length scaled by 2^k,
 * shaders/common.slang:414-416, shaders/split.slang:224-227).  Per receiver:
 * the LOD pair and blend from `view`, each LOD's diffuse lookup (LODs at or
 * beyond lod_count have no probes), resolveLodSource, the support code of
 * lux_world_resolve_device.  Runs on lods[0]'s stream after waiting for the
 * other contexts' streams; lod_count <= 8. */
int lux_world_resolve_lods(lux_ctx* const* lods, uint32_t lod_count, const lux_lod_view* view,
                           const float* receivers, uint32_t n, float* out);

/* Temporal history of the raw intervals (mergeEntry's history part,
 * shaders/split.slang:1624-1705, mergeRawRadianceGroup :1478-1521): each
 * update blends every traced interval's (J, beta) with its history by
 * sample masses, mass limits 1 / (1 - blend), and merges with the blended
 * values.  radiance_blend = RadianceCascades::temporal_blend (history.x),
 * transmittance_blend = split_transmittance_blend (intervalHistory.x),
 * history_frames (history.w; 0: nothing is reused), persistent (world
 * intervals never expire; intervalHistory.y, set by the reference's renderer;
 * otherwise an interval last sampled more than history_frames frames ago is
 * not reused).  16 bytes. */
typedef struct {
    float radiance_blend;
    float transmittance_blend;
    uint32_t history_frames;
    uint32_t persistent;
} lux_gi_history_params;
/* NULL or both blends 0: history off (the default).  Turning it on, and
 * every lux_gi_configure, starts from an empty history. */
int lux_gi_set_history(lux_ctx* ctx, const lux_gi_history_params* params);
int lux_gi_reset_history(lux_ctx* ctx);
/* Dirty region (world AABB): the next update reuses no history of an
 * interval whose retained dependency meets it (dependencyIntersectsDirty,
 * shaders/split.slang:13-25).  An interval's dependency is the union of the
 * world segments its supported rays traversed (from the probe-side start to
 * the hit, clipped to the scene's root box) and, for rays that hit, their NEE
 * occlusion segments, joined over the updates its history kept
 * (DependencyContext, shaders/common.slang:138-242; deposit,
 * shaders/split.slang:1085-1105); receiver-driven updates record an unknown
 * dependency, which any region meets.  More than 1024 pending regions drop the
 * whole history.  lux_scene_set_transforms marks the old and new boxes of every
 * moved instance; a moved instance with emissive triangles drops the whole
 * history (the light every NEE sample may reach moved), and a moved box
 * leaving the previous scene root makes the next pass drop measured-empty
 * dependencies. */
int lux_gi_invalidate_history(lux_ctx* ctx, const float* lo, const float* hi);
/* Device addresses of the history (16 floats per interval: J.rgb, radiance
 * mass, beta, transmittance mass, last sampled frame bits, 0, then the
 * retained dependency lo.xyz, known, hi.xyz, has bounds) and of the
 * supported-ray counts of the last update; NULL when history is off. */
int lux_gi_history_arrays(lux_ctx* ctx, float** history, uint32_t** samples);
/* Intervals whose history the last history pass reused (valid, alive, with
 * a dependency meeting no dirty region, and holding mass); synchronous. */
int lux_gi_history_reused(lux_ctx* ctx, uint32_t* reused);

/* ---- receiver-driven updates (the sparse Split path, shaders/split.slang:
 *      567-1476, src/render.rs:9640-9794 for a frame's dispatches) ----
 * Receivers (device float[6 * width * height], row-major: position, normal;
 * a zero normal marks no surface) choose the probes and rays of an update:
 * seed / rank / parents / prefix / generateDirections, then one jittered ray
 * per selected receiver, deposits per level up to the terminating one, the
 * mean per interval, and each LOD grid's history blend, merge and pack.
 * lods[k] is the dense grid of LOD k (spacing and interval length x 2^k,
 * origins on the coarsest level's cells, same levels and directions); a
 * deposit outside the grids is dropped.  Orders the reference leaves to its
 * atomics are the sequential ones (lowest descriptor index canonical,
 * ordinals and tickets in receiver order, roots by canonical index).
 * Cold priority (coldPriorityEnabled, split.slang:685-705, rank :836-888):
 * when lods[0] has a persistent history with history_frames > 0 that is not
 * valid (never blended since lux_gi_set_history / lux_gi_configure /
 * lux_gi_reset_history or a whole-history invalidation), a budgeted update
 * first admits one nominee per probe (its receiver of lowest frame-seeded
 * rank), probes of larger receiver populations first, then fills the budget
 * from the fair window; admission and fill go in ascending fair rank, a
 * nominee's ordinal is 0.  The caller then keeps its fair cursor
 * (stats->cold). */
typedef struct {
    uint32_t width, height; /* receiver grid (params.size.xy) */
    uint32_t budget;        /* sampling.x: receivers traced per update, 0 = all */
    uint32_t fair_cursor;   /* ownerSequence.w: fair window start (lux_candidate_schedule_phase) */
    uint32_t fair_quota;    /* ownerSequence.z: window length */
    uint32_t frame;         /* params.frame.x: LOD phase, direction jitter, NEE sequence */
    uint32_t flags;         /* 1: updates disabled (no rays) */
    uint32_t reserved;
} lux_receiver_params;

typedef struct {
    uint32_t receivers;   /* eligible receivers */
    uint32_t tickets;     /* receivers tracing this update */
    uint32_t rays;        /* rays traced */
    uint32_t unsupported; /* rays without a sample (traversal or visibility unsupported) */
    uint32_t deposits;    /* deposits kept in a grid */
    uint32_t outside;     /* deposits outside every grid */
    uint32_t root_rays;   /* ray total of the prefix roots (= rays selected) */
    uint32_t overflow;    /* descriptors the probe directory could not hold */
    float ms;             /* device time of the update */
    uint32_t cold;        /* 1: cold-priority selection (do not advance the fair cursor) */
    uint32_t nominees;    /* cold nominees admitted */
    uint32_t reserved;
} lux_receiver_stats;

int lux_gi_update_receivers(lux_ctx* const* lods, uint32_t lod_count, const lux_lod_view* view,
                            const float* receivers, const lux_receiver_params* params,
                            lux_receiver_stats* stats /* nullable */);
/* The last receiver update's selection (on lods[0]'s context; any pointer may
 * be NULL): per descriptor (2 * count * levels, (variant * count + receiver)
 * * levels + level) its key (int32[4]), flags (1 seeded, 2 ray, 4 admitted
 * cold nominee), canonical
 * descriptor, probe ray count and offset; per (variant, receiver) its LOD and
 * sequence (0xffffffff: no ray); the tickets and their number. */
int lux_receivers_download(lux_ctx* ctx, int32_t* keys, uint32_t* flags, uint32_t* canonical,
                           uint32_t* probe_count, uint32_t* probe_offset, uint32_t* lods, uint32_t* sequence,
                           uint32_t* tickets, uint32_t* n_tickets);
/* Samples per interval of the last update that counted them (receiver
 * updates always; plain updates with history on). */
int lux_gi_samples_download(lux_ctx* ctx, uint32_t* samples);

/* Traversal query counters of the current update: attempts and admitted
 * (admitted <= the budget; both count every interval and NEE query).
 * lux_gi_update and lux_gi_update_surrogate reset them at their start, as
 * the reference's cache-clear pass does once per frame. */
int lux_gi_query_counters(lux_ctx* ctx, uint32_t* attempts, uint32_t* admitted);
int lux_gi_reset_query_counters(lux_ctx* ctx);

/* Trace only levels first..last (inclusive) of the configured grid into the
 * interval array; lux_gi_trace is first = 0, last = levels - 1. */
int lux_gi_trace_levels(lux_ctx* ctx, uint32_t first, uint32_t last);
/* Interval trace organisation (same results either way): 0 (default)
 * wavefront -- primary rays, shading setup, compacted NEE occlusion rays and
 * an ordered accumulation, the traversals as regrouped walks under the
 * context's trace configuration; 1 fused -- one thread per interval. */
#define LUX_GI_TRACE_WAVEFRONT 0u
#define LUX_GI_TRACE_FUSED 1u
int lux_gi_set_trace_mode(lux_ctx* ctx, uint32_t mode);
/* Launch configuration of the wavefront stages' walks (same fields and
 * limits as lux_trace_config; waves_per_cu 1..32; a new context starts at
 * {24, 12, 12, 24}) and its autotuner, which times full interval traces of the
 * configured grid and installs the fastest (overwrites the interval array).
 * No counterpart in the reference. */
int lux_gi_set_trace_config(lux_ctx* ctx, const lux_trace_config* cfg);
int lux_gi_get_trace_config(lux_ctx* ctx, lux_trace_config* cfg);
int lux_autotune_gi(lux_ctx* ctx, uint32_t repeats, lux_trace_config* best, float* best_ms,
                    uint32_t* evaluations);

/* ---- learned far-cascade surrogate.  No counterpart in the reference: its
 *      README defers online neural radiance caching (README.md:384-385).
 *      A multiresolution hash-grid encoder (16 levels x 2 FP16 features,
 *      2^19 entries per level, base resolution 16, growth 1.5), 8-band real
 *      SH of the direction and the normal feed an MLP 128-64-64-16 (ReLU,
 *      BF16 weights, MFMA with FP32 accumulation).  Outputs 0-2 radiance,
 *      output 3 log of the predicted squared error. ---- */
typedef struct lux_surrogate lux_surrogate;
typedef struct {
    uint32_t levels;          /* hash levels: 16 (the only supported value) */
    uint32_t features;        /* features per level: 2 */
    uint32_t log2_table;      /* log2 entries per level: 19 */
    float base_resolution;    /* level-0 grid resolution, 16 */
    float growth;             /* per-level resolution factor, 1.5 */
    float box_min[3];         /* encoder domain (world units); positions are clamped to it */
    float box_max[3];
    float learning_rate;      /* Adam step, 1e-2 (beta 0.9 / 0.99, eps 1e-15) */
    float error_weight;       /* loss weight of the error head, 0.1 */
    uint32_t seed;            /* initialisation seed */
    uint32_t freeze_updates;  /* 1: accumulate gradients only, no Adam step */
    float ema_decay;          /* inference reads an exponential moving average of the
                                 parameters with this decay (0.95); 0 = the parameters */
    float table_learning_rate; /* Adam step of the hash table; 0 = learning_rate */
    uint32_t reserved[3];
} lux_surrogate_params;

void lux_surrogate_default_params(lux_surrogate_params* params);
int lux_surrogate_create(lux_ctx* ctx, const lux_surrogate_params* params, lux_surrogate** out);
void lux_surrogate_destroy(lux_surrogate* s);
/* Re-initialise parameters and optimiser state from the seed. */
int lux_surrogate_reset(lux_surrogate* s);
/* Device buffers.  queries: 9 floats each (position, unit direction, normal).
 * encode writes the BF16 bit patterns of the 128 network inputs per query;
 * infer writes 16 floats per query.  Both read the inference parameters
 * (the moving average when ema_decay > 0). */
int lux_surrogate_encode(lux_surrogate* s, const float* queries, uint32_t n, uint16_t* inputs);
int lux_surrogate_infer(lux_surrogate* s, const float* queries, uint32_t n, float* outputs);
/* One optimisation step over device samples of 12 floats (query + target
 * radiance; a non-finite target is skipped).  *loss (host, nullable) receives
 * the step's loss and makes the call synchronous. */
int lux_surrogate_train_step(lux_surrogate* s, const float* samples, uint32_t n, float* loss);
int lux_surrogate_zero_grad(lux_surrogate* s);
int lux_surrogate_sizes(lux_surrogate* s, uint64_t* mlp_params, uint64_t* table_params,
                        uint32_t* steps);
/* Host copies (each nullable): MLP weights W1 (128x64) | W2 (64x64) |
 * W3 (64x16), [in][out] row-major, their gradients, the FP32 table
 * [16][2^19][2] and its gradient. */
int lux_surrogate_download(lux_surrogate* s, float* mlp, float* mlp_grad, float* table,
                           float* table_grad);
int lux_surrogate_upload(lux_surrogate* s, const float* mlp, const float* table);

/* Training samples from level `level` of the configured grid (device float
 * [12 * n]): n (probe, direction) pairs drawn by hashing (sample_seed, i);
 * query = probe centre, the frame's jittered bin direction, normal = that
 * direction; target = the radiance beyond the level's interval start, either
 * traced now as the level's unbounded terminal interval (source
 * LUX_SAMPLES_TRACE) or read from the merged array of the last plain update
 * (LUX_SAMPLES_MERGED).  Unsupported targets are NaN. */
#define LUX_SAMPLES_TRACE 0u
#define LUX_SAMPLES_MERGED 1u
int lux_gi_surrogate_samples(lux_ctx* ctx, uint32_t level, uint32_t n, uint32_t sample_seed,
                             uint32_t source, float* samples);
/* The queries of every (probe, direction) of `level` in merged-array order
 * (device float[9 * count]); *count (nullable) = intervals of the level. */
int lux_gi_surrogate_queries(lux_ctx* ctx, uint32_t level, float* queries, uint64_t* count);

typedef struct {
    uint32_t level;            /* first replaced cascade */
    uint32_t probes;           /* probes of that level */
    uint32_t fallback_probes;  /* probes traced because of the error bound */
    uint32_t queries;          /* network queries */
    float ms_total;            /* GPU time of the whole update */
    float ms_trace;            /* levels below `level` */
    float ms_infer;            /* queries, network, predicted error */
    float ms_fallback;         /* traced fallback probes */
    float ms_merge_pack;       /* merge of the traced levels + packing */
    float reserved[3];
} lux_gi_surrogate_stats;

/* Surrogate-assisted GI update: trace levels 0..level-1; predict merged
 * radiance of every (probe, direction) of `level` with support 1; a probe
 * whose predicted radiance RMSE over its directions exceeds error_bound
 * (radiance units; negative: every probe, +inf: none) is traced instead as
 * the unbounded terminal interval of `level`; merge levels level-1..0 and
 * pack.  Levels above `level` are neither traced nor merged.  With every
 * probe traced the result equals the plain update of the same grid with
 * levels = level + 1 and terminal_tail = 1.  stats nullable; when given the
 * call synchronizes. */
int lux_gi_update_surrogate(lux_ctx* ctx, lux_surrogate* s, uint32_t level, float error_bound,
                            lux_gi_surrogate_stats* stats);
/* Per-probe predicted RMSE of the last surrogate update (host float[probes]). */
int lux_gi_surrogate_probe_error(lux_ctx* ctx, float* out, uint32_t count);

/* Receivers of a pinhole camera for lux_world_resolve_device (device
 * float[6 * width * height], row-major, y down): per pixel one closest-hit
 * ray; hit position + geometric normal facing the camera, or a zero normal
 * (not eligible) on a miss.  camera: origin[3], forward[3] (unit), right[3]
 * and up[3] scaled to the image plane's half extents at unit distance; the
 * ray of pixel (x, y) is normalize(forward + s right + t up) with
 * s = (x + 0.5) / width * 2 - 1, t = 1 - (y + 0.5) / height * 2.  Stands in
 * for the depth/normal prepass whose receivers the reference resolves. */
int lux_camera_receivers(lux_ctx* ctx, const float* camera, uint32_t width, uint32_t height,
                         float* receivers);

/* Kernel timing of the last device trace (HIP events on the context stream). */
int lux_last_kernel_ms(lux_ctx* ctx, float* ms);

/* Launch tuning: persistent-grid waves per CU (0 = 28) and refill threshold
 * (idle lanes before a wave fetches new rays; 0 = run each wave's 64 rays to
 * completion).  Keeps the LDS stack depth and steps per refill. */
int lux_set_trace_tuning(lux_ctx* ctx, uint32_t waves_per_cu, uint32_t regroup_threshold);

int lux_set_trace_config(lux_ctx* ctx, const lux_trace_config* cfg);
int lux_get_trace_config(lux_ctx* ctx, lux_trace_config* cfg);
/* Autotuner: times candidate configurations on the caller's device rays
 * (coordinate descent over waves per CU {12..32}, LDS stack depth {4,8,16},
 * refill threshold {0..32}, steps per refill {2..16}; median of `repeats`
 * launches each) and installs the fastest.  best, best_ms and evaluations
 * are nullable.  No counterpart in the reference (its traversal is a fixed
 * wgpu dispatch); every configuration returns identical hits. */
int lux_autotune_trace(lux_ctx* ctx, const lux_ray* rays, uint32_t n, lux_hit* hits, uint32_t repeats,
                       lux_trace_config* best, float* best_ms, uint32_t* evaluations);

/* ------------------------------------------------------------------ batch
 * One GI update of many configured scenes of uneven size at once.  The
 * reference runs the Split dispatch sequence once per view and scene
 * (src/render.rs:9553-9801); a batch runs every scene's traced levels in ONE
 * persistent closest-hit launch and ONE persistent occlusion launch whose
 * lanes are refilled across scenes, then every scene's merge and pack, and
 * can be captured once as a HIP graph and replayed per frame.  Each scene's
 * intervals, merged radiance, probes and query counters are exactly what its
 * own lux_gi_update (or lux_gi_update_surrogate with error_bound = +inf)
 * writes.  The batch works on the scenes' own device arrays through a table
 * built at create / refresh: after lux_gi_configure, lux_scene_upload_* or a
 * BVH build of a member scene, call lux_batch_refresh (which drops a captured
 * graph).  Batch work runs on the batch's own stream: call
 * lux_batch_synchronize before reading a scene's results with per-scene calls,
 * and do not update a member scene concurrently.  lux_batch_last_error gives
 * the message of a failed batch call.  No counterpart in the reference. */
typedef struct lux_batch lux_batch;
int lux_batch_create(lux_ctx* const* scenes, uint32_t n, lux_batch** out);
void lux_batch_destroy(lux_batch* b);
const char* lux_batch_last_error(const lux_batch* b);
/* Predict `level` of scene `scene` with surrogate s (of that scene's context;
 * NULL: trace every level), as lux_gi_update_surrogate with no fallback. */
int lux_batch_set_surrogate(lux_batch* b, uint32_t scene, lux_surrogate* s, uint32_t level);
/* The traced per-probe fallback of that scene's surrogate level, as
 * lux_gi_update_surrogate(..., error_bound): probes whose predicted RMSE is
 * above the bound (or NaN) are traced inside the batch launch (and a
 * captured graph).  +inf (the default, and after lux_batch_set_surrogate):
 * none.  Needs a surrogate on the scene.  Both calls drop a captured graph
 * (lux_batch_capture again). */
int lux_batch_set_surrogate_bound(lux_batch* b, uint32_t scene, float error_bound);
/* Rebuild the table after member scenes changed; the fallback words of a
 * scene with a finite bound follow its current probe count.  Fails with
 * LUX_ERR_STATE when a scene's surrogate level is no longer in
 * 1..levels-1 (set or clear that scene's surrogate, then refresh). */
int lux_batch_refresh(lux_batch* b);
/* Frame index and seed of every scene (kept in device memory: a captured
 * graph picks them up); the member contexts' frame and seed follow. */
int lux_batch_set_frame(lux_batch* b, uint32_t frame, uint32_t seed);
/* Enqueue one update of every scene; ms nullable (when given, synchronizes
 * and returns the update's device time). */
int lux_batch_update(lux_batch* b, float* ms);
/* Capture the update's launch sequence as a HIP graph / replay it. */
int lux_batch_capture(lux_batch* b);
int lux_batch_launch(lux_batch* b, float* ms);
int lux_batch_synchronize(lux_batch* b);
/* Totals over the batch: closest-hit items (interval x ray), traced
 * intervals, level-0 probes, scenes.  Each pointer nullable. */
int lux_batch_sizes(lux_batch* b, uint64_t* items, uint64_t* intervals, uint64_t* probes, uint32_t* scenes);
/* Launch configuration of the two persistent walks (lux_trace_config
 * meaning; a new batch starts at {24, 12, 48, 16}); setting it drops a
 * captured graph. */
int lux_batch_set_config(lux_batch* b, const lux_trace_config* cfg);
int lux_batch_get_config(lux_batch* b, lux_trace_config* cfg);
/* Autotuner over whole batch updates (coordinate descent over waves per CU,
 * LDS stack depth, refill threshold, steps per refill; median of `repeats`);
 * installs the fastest and drops a captured graph.  Every configuration
 * writes identical results.  The timed updates skip the temporal history
 * pass: scenes' histories, their validity and pending dirty regions are left
 * as they were (interval, merged and probe arrays hold a tuning update's). */
int lux_batch_autotune(lux_batch* b, uint32_t repeats, lux_trace_config* best, float* best_ms,
                       uint32_t* evaluations);

/* ---- screen-space temporal and denoise passes (f4's screen filters) ----
 * temporal (shaders/cascades.slang:310-622) and denoise (:624-711) over the
 * resolve extent ((viewport w + s - 1) / s) x ((viewport h + s - 1) / s), s =
 * max(resolve_scale, 1) (giResolveScale / giResolveExtent, common.slang:95-100).
 * Resolve-extent images are float4 row-major; full-resolution images (the
 * render target, target[0] x target[1], viewport inside it) are depth
 * (float, reverse-Z, > 0 on surfaces), the G-buffer (uint4; .w holds the
 * packed shading normal), the raster receiver (uint4: identity .xy, packed
 * geometric normal .z) and motion (float2, NDC units per viewport; read only
 * when stability[2] > 0 and, for persistent fields, stability[1] > 0.5).
 * All pointers are device pointers; outputs must not alias inputs.  The
 * passes run on ctx's stream.  Matrices are row-major (h = M (x, y, z, 1)).
 * Field mapping to the reference's Parameters: history_blend = history.x,
 * history_frames = history.w, persistent = intervalHistory.y > 0.5,
 * stability = stability.xyz(w), source_changes / source_flags /
 * motion_epoch = sourceHistory.yzw, frame = frame.x, probe_spacing /
 * normal_bias = tuning.x / tuning.z, size = size.xy, resolve_scale =
 * sampling.w, dirty_min / dirty_max = dirtyMin / dirtyMax (w > 0: on).
 * The debug-only temporal diagnostic counters are not produced.  312 bytes. */
typedef struct {
    float world_from_clip[16];
    float previous_clip_from_world[16];
    float previous_world_from_clip[16];
    uint32_t viewport[4];
    uint32_t target[2];
    uint32_t size[2];
    uint32_t resolve_scale;
    uint32_t persistent;
    float probe_spacing, normal_bias;
    float history_blend, history_frames;
    float stability[4];
    uint32_t source_changes, source_flags, motion_epoch, frame;
    float dirty_min[4], dirty_max[4];
} lux_screen_params;
int lux_screen_extent(const lux_screen_params* params, uint32_t* width, uint32_t* height);
/* temporal: current raw resolve (alpha: 0 not eligible, -1 unsupported, 1 +
 * 0.25 coverage), its background (nullable: none), the previous frame's
 * temporal irradiance / metadata / receiver words / background (the last one
 * nullable) -> this frame's four. */
int lux_screen_temporal(lux_ctx* ctx, const lux_screen_params* params, const float* current,
                        const float* current_background, const float* depth, const uint32_t* gbuffer,
                        const uint32_t* raster, const float* motion, const float* previous,
                        const float* previous_metadata, const float* previous_receiver,
                        const float* previous_background, float* irradiance, float* metadata, float* receiver,
                        float* background);
/* One denoise pass (pass index = the reference's frame.y: step 1, then
 * max(round(cell stride / 4), 1) * 2^pass): the previous pass's (or
 * temporal's) irradiance and receiver words, temporal's metadata as moments
 * -> filtered irradiance and receiver words.  Ping-pong between passes. */
int lux_screen_denoise(lux_ctx* ctx, const lux_screen_params* params, uint32_t pass, const float* irradiance,
                       const float* moments, const float* receiver, const float* depth, float* out,
                       float* out_receiver);

#ifdef __cplusplus
}
#endif
#endif

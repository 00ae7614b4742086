// lux_internal.h -- host-side context shared by the liblux_hip translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/lux.h"
#include "lux_device.h"

struct lux_buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct lux_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string error;
    int cu_count = 256;

    // scene input (device copies of caller arrays)
    std::vector<uint32_t> mesh_tri_count, mesh_tri_base, mesh_blas_base;
    uint32_t n_meshes = 0, n_inst = 0, n_tri = 0, n_materials = 0;
    std::vector<lux_instance_desc> instances;
    lux_buffer raw_pos;   // float[9*n_tri], input order, meshes concatenated
    lux_buffer raw_mat;   // uint32[n_tri]
    lux_buffer inst_desc; // float[24*n_inst] (inv, fwd) + uint32 mesh/flags
    lux_buffer materials; // lux_material[n_materials]

    // built scene
    bool blas_built = false, tlas_built = false;
    bool tlas_topology = false;  // a TLAS built by lux_bvh_build for the current instance set
    bool prebuilt = false;       // scene installed with its BVH (upload_built / upload_words)
    std::vector<uint32_t> inst_material;  // per-instance material (upload_words), else empty
    uint32_t n_blas = 0, n_tlas = 0;
    lux_buffer tris;              // float4[3 * n_tri]: A, B, C of each sorted triangle
    lux_buffer tri_order;         // uint32[n_tri] sorted -> input triangle (global)
    lux_buffer blas;              // float4[2*n_blas]
    lux_buffer tlas;              // float4[2*n_tlas]
    lux_buffer inst;              // float4[8*n_inst]
    lux_buffer inst_order;        // uint32[n_inst]
    lux_buffer range_ok;          // uint32[2]: BLAS / TLAS node bounds in the walk's division range

    // scratch
    lux_buffer scratch;
    lux_buffer ray_buf, hit_buf;

    // host copies kept for the emitter table (CPU-side scene packing, as in
    // the reference's build_emissive_lights)
    std::vector<std::vector<float>> host_pos;
    std::vector<std::vector<uint32_t>> host_mat;
    std::vector<lux_material> host_materials;
    std::vector<float> mesh_bounds;  // per mesh: local box (lo, hi) of host_pos, computed on demand
    std::vector<std::vector<uint32_t>> mesh_emissive;  // per mesh: its emissive triangles, computed on demand

    // transport state
    bool gi_ready = false;
    lux_gi_params gi{};
    uint64_t gi_total = 0;
    uint64_t gi_offset[17] = {};
    uint32_t gi_probes0 = 0, gi_lights = 0;
    lux_buffer gi_intervals, gi_merged, gi_probes, gi_light_buf;
    // surrogate-assisted update: level queries, per-probe predicted error,
    // fallback probe list + count
    lux_buffer gi_sur_queries, gi_sur_err, gi_sur_list;
    uint32_t gi_sur_probes = 0;
    lux_buffer gi_counters;  // traversal query attempts, admitted (per update)
    bool gi_lights_dirty = false;  // instance transforms changed since the emitter table
    lux_buffer tlas_tmp;     // TLAS build temporaries (instance boxes, permutation)
    // wavefront interval trace: mode 0 wavefront, 1 fused; per-item records,
    // NEE slots, slot visibility, compacted occlusion rays, counters
    uint32_t gi_trace_mode = 0;
    // launch configuration of the wavefront stages' walks (own tuning: their
    // item setup is heavier than a ray-list load)
    uint32_t gi_wpc = 24, gi_depth = 12, gi_regroup = 12, gi_batch = 24;
    lux_buffer gi_w_prim, gi_w_shade, gi_w_slots, gi_w_vis, gi_w_shadow, gi_w_counters;
    lux_buffer gi_w_dep;     // per item: NEE segments' dependency (history on)
    lux_buffer mesh_info;    // uint32 per mesh: BLAS base | triangle base | count
    hipEvent_t ev_phase[5] = {};
    // opt-in stage timing of the wavefront trace (lux_gi_set_stage_timing):
    // primary walks, setup + occlusion-ray packing, occlusion walks,
    // accumulation, summed over the chunks of the last traced update
    bool gi_stage_timing = false;
    hipEvent_t ev_stage[5] = {};
    float gi_stage_ms[4] = {};
    // instrumented pass (lux_gi_trace_stats): device counters, occlusion rays
    lux_buffer gi_stats_buf;
    uint64_t gi_stats_rays = 0;
    // temporal history of raw intervals: on when a blend is positive; sample
    // counts, two float4 per interval, state words (valid, dirty boxes)
    uint32_t blas_builder = 0;  // builder of the built BLASes: 0 Morton-30, 1 SAH on the GPU, 2 SAH on the host
    bool gi_hist_on = false;
    bool gi_hist_valid = false;  // a history pass ran since the last reset
    lux_gi_history_params gi_hist{};
    lux_buffer gi_samples, gi_history, gi_hist_state;
    lux_buffer gi_dep;  // per interval: this update's dependency (2 float4)
    uint32_t gi_dirty_n = 0;                 // dirty boxes pending on the device
    // receiver-driven updates (lux_gi_update_receivers, state on the first
    // LOD's context): descriptor keys / flags / directory slots, per
    // (variant, receiver) LOD, ordinal and sequence, the probe directory
    // (owner, canonical, count, offset per level), receiver tickets and ray
    // list, per-ray deposit slots, compacted deposit values and long-run list,
    // sort and scan scratch, counters
    lux_buffer rx_keys, rx_flags, rx_slot, rx_lod, rx_ord, rx_seq, rx_dir, rx_tickets;
    lux_buffer rx_dep_val, rx_pack, rx_traced, rx_sort, rx_scan, rx_counters, rx_cold;
    // screen passes: per resolve texel its decoded receiver words, normal and
    // world position (temporal: 3 float4; denoise: 2 float4)
    lux_buffer screen_cache;
    uint32_t rx_count = 0, rx_levels = 0, rx_cap = 0, rx_tickets_n = 0;

    // trace tuning
    // traversal launch configuration (defaults: the autotuner's choice on the
    // 262,144-triangle colonnade microbenchmark)
    uint32_t waves_per_cu = 28, regroup = 6, trace_depth = 12, trace_batch = 24;
    float last_ms = 0.0f;

    lux::DevScene dev_scene() const;
};

// Autotuners keep the incumbent launch configuration unless a candidate's
// median time is more than 1.5 % lower (run-to-run noise of a timed launch).
constexpr float kTuneMargin = 0.985f;

int lux_fail(lux_ctx* ctx, int code, const std::string& msg);
// Looks up this device's copy of the allocation-free trace's scratch (once
// per device; ctx->device must be current).
int lux_trace_scratch_init(lux_ctx* ctx);
int lux_hip_check(lux_ctx* ctx, hipError_t e, const char* what);
int lux_ensure(lux_ctx* ctx, lux_buffer& b, size_t bytes);
void lux_release(lux_buffer& b);

#define LUX_HIP(ctx, call)                                         \
    do {                                                           \
        hipError_t _e = (call);                                    \
        if (_e != hipSuccess) return lux_hip_check((ctx), _e, #call); \
    } while (0)

// LBVH over primitive boxes (device lo/hi float4[n]).  Writes nodes
// (float4[2*(2n-1)] or 2 for n<=leaf_max... always 2n-1 slots) and perm.
// node_first_offset is added to internal 'first', leaf_first_offset to
// leaf 'first'.
int lux_lbvh_build_device(lux_ctx* ctx, const float4* lo, const float4* hi, uint32_t n,
                          uint32_t leaf_max, float4* nodes, uint32_t* perm,
                          uint32_t node_first_offset, uint32_t leaf_first_offset);
// Stable LSD radix sort of n keys on their low `bits` bits (values = input
// positions); hist holds lux_sort_hist_words(n) words.
int lux_sort_pairs_impl(lux_ctx* ctx, const uint32_t* keys_in, uint32_t n, uint32_t* keys_out,
                        uint32_t* vals_out, uint32_t* tmp_keys, uint32_t* tmp_vals,
                        uint32_t* hist, uint32_t bits);
// The same sort carrying vals_in (nullptr: input positions) as values.
int lux_sort_pairs_vals_impl(lux_ctx* ctx, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t n,
                             uint32_t* keys_out, uint32_t* vals_out, uint32_t* tmp_keys, uint32_t* tmp_vals,
                             uint32_t* hist, uint32_t bits);
size_t lux_sort_hist_words(uint32_t n);
// Every mesh's BLAS in one pass over all n triangles (device positions
// float[9n], materials u32[n] or null).  seg_info (device) holds n_seg BLAS
// node bases, then n_seg triangle bases, then n_seg triangle counts (each
// count > 0).  Each mesh's nodes, and its triangles in leaf order (tris
// float4[3n], order u32[n]), are those a single LBVH build of that mesh
// writes with node_first_offset = its BLAS base and leaf_first_offset = its
// triangle base.  tiles (device, n_tiles + 1 words) is the bottom-up tile
// table lux_segment_tiles makes from the same counts.  Stream-ordered; no
// host wait.
int lux_blas_build_segmented(lux_ctx* ctx, const float* pos, const uint32_t* mat, uint32_t n, uint32_t n_seg,
                             const uint32_t* seg_info, const uint32_t* tiles, uint32_t n_tiles, uint32_t leaf_max,
                             float4* nodes, float4* tris, uint32_t* order);
// Triangle boxes (compare-select a.min(b).min(c)), segment ids and
// per-segment centre bounds of n triangles (device positions float[9n]).
int lux_seg_prepare(lux_ctx* ctx, const float* pos, uint32_t n, const uint32_t* tri_base, uint32_t n_seg,
                    float4* lo, float4* hi, uint32_t* seg_of, uint32_t* bounds);
// Triangles in leaf order from a permutation of global triangle indices:
// tris float4[3n] (w: index in the mesh, material, 0), order u32[n].
int lux_seg_gather(lux_ctx* ctx, const float* pos, const uint32_t* mat, const uint32_t* perm, const uint32_t* seg_of,
                   const uint32_t* tri_base, uint32_t n, float4* tris, uint32_t* order);
// Every mesh's binned-SAH BLAS on the GPU (the reference's build_tree;
// nodes, triangle order and tris as the host builder writes them);
// seg_info as lux_blas_build_segmented.  Positions must be finite.
int lux_blas_build_sah(lux_ctx* ctx, const float* pos, const uint32_t* mat, uint32_t n, uint32_t n_seg,
                       const uint32_t* seg_info, uint32_t n_blas, float4* nodes, float4* tris, uint32_t* order);
// Bottom-up tile starts over segments of the given counts, then the total.
void lux_segment_tiles(const uint32_t* count, uint32_t n_seg, std::vector<uint32_t>& tiles);
// Exclusive scan of n words in place on the context stream; sums holds
// lux_scan_sums_words(n) words, *total (device) receives the sum.
int lux_scan_exclusive_impl(lux_ctx* ctx, uint32_t* data, uint32_t n, uint32_t* sums, uint32_t* total);
// The same on stream st (no host synchronisation; capturable in a graph).
int lux_scan_exclusive_on(lux_ctx* ctx, hipStream_t st, uint32_t* data, uint32_t n, uint32_t* sums,
                          uint32_t* total);
size_t lux_scan_sums_words(uint32_t n);

// Surrogate forward pass over device queries (9 floats each); *Y receives a
// device pointer to n x 16 outputs owned by the surrogate, valid until its
// next call.
int lux_surrogate_forward_internal(lux_surrogate* s, const float* queries, uint32_t n, const float** Y);
lux_ctx* lux_surrogate_context(lux_surrogate* s);
// Bytes of caller-owned scratch for an inference pass over n queries, and
// that pass on a caller's stream into it (*Y points into the scratch).
// Allocates nothing and keeps no pointer, so the launches can be captured
// into a graph that outlives later training steps.
size_t lux_surrogate_scratch_bytes(uint32_t n);
int lux_surrogate_forward_into(lux_surrogate* s, hipStream_t st, const float* queries, uint32_t n, void* scratch,
                               const float** Y);
// GI state check of a context (configured, BVH built, emitter table current).
int lux_gi_check_internal(lux_ctx* ctx);
// Dirty boxes (6 floats each: lo, hi) for the next history pass.
int lux_gi_invalidate_boxes(lux_ctx* ctx, const float* boxes, uint32_t count);
// The scene's root box grew: the next history pass drops measured-empty dependencies.
int lux_gi_root_grew(lux_ctx* ctx);
// This update's dependency of every interval set to unknown (updates whose
// deposits record no segments: receiver-driven updates).
int lux_gi_dep_unknown(lux_ctx* ctx, hipStream_t st);
// History blend (when on), merge and pack of a context's intervals on stream st.
int lux_gi_finish_internal(lux_ctx* ctx, hipStream_t st);
uint64_t lux_sur_fallback_words(const lux_ctx* ctx, uint32_t level);
int lux_sur_fallback_on(lux_ctx* ctx, hipStream_t st, uint32_t level, const float* Y, float bound,
                        uint32_t* words);
// Receiver rays of a ray list through the wavefront stages on the context
// stream: per ray (rad.rgb, hit t) and status (1 hit, 2 unsupported).
int lux_rx_wave_trace(lux_ctx* ctx, const float* recv, uint32_t count, const uint32_t* rays, uint32_t n_rays,
                      const uint32_t* seq, uint32_t frame, float4* out, uint32_t* status);

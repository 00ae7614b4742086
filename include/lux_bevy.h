/*
 * lux_bevy.h -- C view of the Bevy-facing RadianceCascades component
 * (src/lib.rs:431-518, defaults :671-698, low_end :700-722, sanitized
 * :731-759) exported by liblux_hip.so, so a binding can reuse the reference's
 * defaults and clamps verbatim and map a camera onto lux_gi_params.
 */
#ifndef LUX_BEVY_H
#define LUX_BEVY_H
#include <stdint.h>
#include "lux.h"

#ifdef __cplusplus
extern "C" {
#endif

/* GiMethod (src/lib.rs:631-669) */
#define LUX_GI_SPLIT_RADIANCE_CASCADES 0u
#define LUX_GI_RAY_TRACED 1u
#define LUX_GI_PXTGI 2u

typedef struct lux_radiance_cascades {
    uint32_t method;
    uint32_t history_epoch;
    float probe_spacing;
    float interval_length;
    uint32_t resolution_divisor;
    uint32_t max_rays;
    float normal_bias;
    float intensity;
    float environment[3];
    float temporal_blend;
    float split_transmittance_blend;
    uint32_t history_frames;
    uint32_t spatial_filter_steps;
    uint32_t emissive_samples;
    /* SamplingSettings: grid_divisor, max_receiver_cells, max_updates_per_frame,
     * max_cache_entries, max_traversal_queries, resolve_scale */
    uint32_t sampling[6];
    /* 1 multi_bounce, 2 cached_hit_sky, 4 hardware_ray_queries,
     * 8 ray_traced_sun_shadows, 16 artistic_occlusion */
    uint32_t flags;
} lux_radiance_cascades;

/* RadianceCascades::default() (src/lib.rs:671-698) */
int lux_radiance_cascades_default(lux_radiance_cascades* out);
/* RadianceCascades::low_end() (src/lib.rs:700-722) */
int lux_radiance_cascades_low_end(lux_radiance_cascades* out);
/* RadianceCascades::sanitized() (src/lib.rs:731-759) */
int lux_radiance_cascades_sanitized(const lux_radiance_cascades* in, lux_radiance_cascades* out);
/* Map a (sanitized) camera component and a dense probe volume onto the
 * transport parameters: probe_spacing -> spacing, interval_length,
 * normal_bias, emissive_samples, history_epoch -> direction jitter seed. */
int lux_radiance_cascades_gi_params(const lux_radiance_cascades* in, const float origin[3],
                                    const uint32_t dims[3], uint32_t cascades, uint32_t base_phi,
                                    uint32_t base_y, uint32_t interval_mode, uint32_t frame,
                                    lux_gi_params* out);
/* The interval history of a (sanitized) camera component:
 * temporal_blend -> radiance_blend, split_transmittance_blend ->
 * transmittance_blend, history_frames, persistent world intervals (the
 * renderer's intervalHistory.y = 1, src/render.rs:8455-8462). */
int lux_radiance_cascades_history_params(const lux_radiance_cascades* in, lux_gi_history_params* out);

/* resolution() (src/render.rs:5350-5380): the receiver grid (candidate
 * count = width * height) of a viewport. */
int lux_radiance_cascades_receiver_grid(const lux_radiance_cascades* in, uint32_t viewport_width,
                                        uint32_t viewport_height, uint32_t* width, uint32_t* height);
/* runtime_update_budget (src/render.rs:243-255): receivers traced per
 * update; has_cap = 0 is the reference's None. */
int lux_radiance_cascades_update_budget(const lux_radiance_cascades* in, uint32_t has_cap, uint32_t runtime_cap,
                                        uint32_t* budget);

/* CandidateSchedule (src/render.rs:5172-5211): zero-initialise, then per
 * update phase() -> lux_receiver_params.fair_cursor / fair_quota and, after
 * a submitted update, advance(!lux_receiver_stats.cold). */
typedef struct lux_candidate_schedule {
    uint32_t width, height, count, cursor, pending;
} lux_candidate_schedule;
int lux_candidate_schedule_phase(lux_candidate_schedule* s, uint32_t width, uint32_t height, uint32_t count,
                                 uint32_t budget, uint32_t* cursor, uint32_t* quota);
int lux_candidate_schedule_advance(lux_candidate_schedule* s, uint32_t consume_fair_cursor);
/* cold_priority_enabled (src/render.rs:128-131): history frames > 0 and the
 * history not valid (the renderer's world intervals are persistent). */
int lux_cold_priority_enabled(uint32_t history_frames, uint32_t history_was_valid);

#ifdef __cplusplus
}
#endif
#endif

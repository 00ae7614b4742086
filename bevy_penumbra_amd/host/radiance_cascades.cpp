// radiance_cascades.cpp -- defaults, clamps and the backend glue of the
// Bevy-facing surface (src/lib.rs), plus a C export of the component
// defaults/sanitization so that tests can check them through the ABI.
#include "radiance_cascades.hpp"
#include "../../include/lux_bevy.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bevy_lux {

namespace {
template <class T>
T clampv(T v, T lo, T hi) { return std::min(std::max(v, lo), hi); }

// `finite` helper of RadianceCascades::sanitized (src/lib.rs:732-738).
float finite_or(float value, float fallback, float lo, float hi) {
    return std::isfinite(value) ? clampv(value, lo, hi) : fallback;
}
}  // namespace

SamplingSettings SamplingSettings::unlimited() { return SamplingSettings{}; }

SamplingSettings SamplingSettings::low_end() {
    SamplingSettings s;
    s.grid_divisor = 8;
    s.max_receiver_cells = 16384;
    s.max_updates_per_frame = 1024;
    s.max_cache_entries = 16384;
    s.max_traversal_queries = 16384;
    s.resolve_scale = 2;
    return s;
}

SamplingSettings SamplingSettings::sanitized() const {
    SamplingSettings s = *this;
    if (s.grid_divisor != 0) s.grid_divisor = clampv<uint32_t>(s.grid_divisor, 1, 256);
    if (s.max_receiver_cells != 0) s.max_receiver_cells = clampv<uint32_t>(s.max_receiver_cells, 1, 131072);
    if (s.max_updates_per_frame != 0)
        s.max_updates_per_frame = clampv<uint32_t>(s.max_updates_per_frame, 1, 131072);
    if (s.max_cache_entries != 0) s.max_cache_entries = clampv<uint32_t>(s.max_cache_entries, 1, 262144);
    if (s.max_traversal_queries != 0)
        s.max_traversal_queries = clampv<uint32_t>(s.max_traversal_queries, 1, 131072);
    s.resolve_scale = clampv<uint32_t>(s.resolve_scale, 1, 8);
    return s;
}

RadianceCascades RadianceCascades::low_end() {
    RadianceCascades r;
    r.sampling = SamplingSettings::low_end();
    r.sampling.max_cache_entries = 65536;
    r.probe_spacing = 1.0f;
    r.interval_length = 1.6f;
    r.resolution_divisor = 8;
    r.max_rays = 1024;
    r.emissive_samples = 4;
    r.history_frames = 24;
    r.temporal_blend = 0.95f;
    r.spatial_filter_steps = 3;
    return r;
}

void RadianceCascades::reset_history() { history_epoch = history_epoch + 1u; }

RadianceCascades RadianceCascades::sanitized() const {
    RadianceCascades r = *this;
    r.probe_spacing = finite_or(r.probe_spacing, 0.25f, 0.001f, 10000.0f);
    r.interval_length = finite_or(r.interval_length, 0.4f, 0.001f, 100000.0f);
    r.normal_bias = finite_or(r.normal_bias, 0.005f, 0.00001f, r.probe_spacing);
    r.intensity = finite_or(r.intensity, 1.0f, 0.0f, 100.0f);
    r.temporal_blend = finite_or(r.temporal_blend, 0.88f, 0.0f, 0.99f);
    r.split_transmittance_blend = finite_or(r.split_transmittance_blend, 0.98f, 0.0f, 0.999f);
    r.history_frames = std::min<uint32_t>(r.history_frames, 64);
    r.spatial_filter_steps = std::min<uint32_t>(r.spatial_filter_steps, 4);
    r.emissive_samples = clampv<uint32_t>(r.emissive_samples, 1, 32);
    r.sampling = r.sampling.sanitized();
    r.resolution_divisor = clampv<uint32_t>(r.resolution_divisor, 1, 64);
    r.max_rays = clampv<uint32_t>(r.max_rays, 1, 131072);
    r.environment.red = finite_or(r.environment.red, 0.0f, 0.0f, 100000.0f);
    r.environment.green = finite_or(r.environment.green, 0.0f, 0.0f, 100000.0f);
    r.environment.blue = finite_or(r.environment.blue, 0.0f, 0.0f, 100000.0f);
    r.environment.alpha = 1.0f;
    return r;
}

bool RadianceCascades::radiance_transport_active() const {
    return intensity > 0.0f && method != GiMethod::Pxtgi;
}

void RadianceCascadesStatus::record_failure(const std::string& message) {
    failures.push_back(message);
    revision += 1;
}

RadianceCascadesPlugin RadianceCascadesPlugin::low_end() {
    RadianceCascadesPlugin p;
    p.scene_gpu_budget_bytes = 256ull * 1024 * 1024;
    p.texture_atlas_budget_bytes = 256ull * 1024 * 1024;
    p.view_gpu_budget_bytes = 512ull * 1024 * 1024;
    p.max_concurrent_pipeline_compiles = 1;
    return p;
}

lux_gi_params gi_params_for(const RadianceCascades& camera, const ProbeVolume& volume,
                            uint32_t frame) {
    RadianceCascades c = camera.sanitized();
    lux_gi_params p;
    std::memset(&p, 0, sizeof p);
    for (int a = 0; a < 3; ++a) {
        p.origin[a] = volume.origin[a];
        p.dims[a] = volume.dims[a];
    }
    p.spacing = c.probe_spacing;
    p.levels = volume.cascades;
    p.base_phi = volume.base_phi;
    p.base_y = volume.base_y;
    p.interval_mode = volume.interval_mode;
    p.interval_length = c.interval_length;
    p.normal_bias = c.normal_bias;
    p.emissive_samples = c.emissive_samples;
    p.frame = frame;
    p.seed = c.history_epoch;
    p.rays_per_interval = 1;
    p.terminal_tail = 1;
    p.query_limit = c.sampling.max_traversal_queries;
    return p;
}

Renderer::Renderer(const RadianceCascadesPlugin& plugin, int device) : plugin_(plugin) {
    int rc = lux_ctx_create(device, &ctx_);
    if (rc != LUX_OK) {
        ctx_ = nullptr;
        status_.record_failure("MI355X backend unavailable: lux_ctx_create failed");
    }
}

Renderer::~Renderer() {
    if (ctx_) lux_ctx_destroy(ctx_);
}

bool Renderer::check(int rc, const char* what) {
    if (rc == LUX_OK) return true;
    std::string msg = std::string(what) + ": " + (ctx_ ? lux_last_error(ctx_) : "no context");
    status_.record_failure(msg);
    return false;
}

bool Renderer::update_scene(const std::vector<lux_mesh_desc>& meshes,
                            const std::vector<lux_instance_desc>& instances,
                            const std::vector<lux_material>& materials) {
    if (!ctx_) return false;
    uint64_t bytes = 0;
    for (const lux_mesh_desc& m : meshes) bytes += 48ull * m.triangle_count + 64ull * m.triangle_count;
    bytes += 128ull * instances.size();
    if (bytes > plugin_.scene_gpu_budget_bytes) {
        status_.record_failure("scene exceeds scene_gpu_budget_bytes");
        return false;
    }
    if (!check(lux_scene_upload_soa(ctx_, meshes.data(), (uint32_t)meshes.size(), instances.data(),
                                    (uint32_t)instances.size(), materials.data(),
                                    (uint32_t)materials.size()),
               "lux_scene_upload_soa"))
        return false;
    return check(lux_bvh_build(ctx_, LUX_BVH_MORTON30), "lux_bvh_build");
}

bool Renderer::render_view(const RadianceCascades& camera, const ProbeVolume& volume,
                           uint32_t frame, float* ms) {
    if (!ctx_) return false;
    RadianceCascades c = camera.sanitized();
    if (!c.radiance_transport_active()) return true;
    if (c.method != GiMethod::SplitRadianceCascades) {
        status_.record_failure("only GiMethod::SplitRadianceCascades runs on the MI355X backend");
        return false;
    }
    lux_gi_params p = gi_params_for(c, volume, frame);
    if (!check(lux_gi_configure(ctx_, &p), "lux_gi_configure")) return false;
    return check(lux_gi_update(ctx_, ms), "lux_gi_update");
}

}  // namespace bevy_lux

// ---- C export of the component surface (for bindings and tests) ----
extern "C" {



static void to_c(const bevy_lux::RadianceCascades& r, lux_radiance_cascades* o) {
    o->method = (uint32_t)r.method;
    o->history_epoch = r.history_epoch;
    o->probe_spacing = r.probe_spacing;
    o->interval_length = r.interval_length;
    o->resolution_divisor = r.resolution_divisor;
    o->max_rays = r.max_rays;
    o->normal_bias = r.normal_bias;
    o->intensity = r.intensity;
    o->environment[0] = r.environment.red;
    o->environment[1] = r.environment.green;
    o->environment[2] = r.environment.blue;
    o->temporal_blend = r.temporal_blend;
    o->split_transmittance_blend = r.split_transmittance_blend;
    o->history_frames = r.history_frames;
    o->spatial_filter_steps = r.spatial_filter_steps;
    o->emissive_samples = r.emissive_samples;
    const bevy_lux::SamplingSettings& s = r.sampling;
    uint32_t sv[6] = {s.grid_divisor, s.max_receiver_cells, s.max_updates_per_frame,
                      s.max_cache_entries, s.max_traversal_queries, s.resolve_scale};
    std::memcpy(o->sampling, sv, sizeof sv);
    o->flags = (r.multi_bounce ? 1u : 0u) | (r.cached_hit_sky ? 2u : 0u) |
               (r.hardware_ray_queries ? 4u : 0u) | (r.ray_traced_sun_shadows ? 8u : 0u) |
               (r.artistic_occlusion ? 16u : 0u);
}

static bevy_lux::RadianceCascades from_c(const lux_radiance_cascades* o) {
    bevy_lux::RadianceCascades r;
    r.method = (bevy_lux::GiMethod)o->method;
    r.history_epoch = o->history_epoch;
    r.probe_spacing = o->probe_spacing;
    r.interval_length = o->interval_length;
    r.resolution_divisor = o->resolution_divisor;
    r.max_rays = o->max_rays;
    r.normal_bias = o->normal_bias;
    r.intensity = o->intensity;
    r.environment.red = o->environment[0];
    r.environment.green = o->environment[1];
    r.environment.blue = o->environment[2];
    r.temporal_blend = o->temporal_blend;
    r.split_transmittance_blend = o->split_transmittance_blend;
    r.history_frames = o->history_frames;
    r.spatial_filter_steps = o->spatial_filter_steps;
    r.emissive_samples = o->emissive_samples;
    r.sampling.grid_divisor = o->sampling[0];
    r.sampling.max_receiver_cells = o->sampling[1];
    r.sampling.max_updates_per_frame = o->sampling[2];
    r.sampling.max_cache_entries = o->sampling[3];
    r.sampling.max_traversal_queries = o->sampling[4];
    r.sampling.resolve_scale = o->sampling[5];
    r.multi_bounce = (o->flags & 1u) != 0;
    r.cached_hit_sky = (o->flags & 2u) != 0;
    r.hardware_ray_queries = (o->flags & 4u) != 0;
    r.ray_traced_sun_shadows = (o->flags & 8u) != 0;
    r.artistic_occlusion = (o->flags & 16u) != 0;
    return r;
}

int lux_radiance_cascades_default(lux_radiance_cascades* out) {
    if (!out) return LUX_ERR_INVALID;
    to_c(bevy_lux::RadianceCascades{}, out);
    return LUX_OK;
}

int lux_radiance_cascades_low_end(lux_radiance_cascades* out) {
    if (!out) return LUX_ERR_INVALID;
    to_c(bevy_lux::RadianceCascades::low_end(), out);
    return LUX_OK;
}

int lux_radiance_cascades_sanitized(const lux_radiance_cascades* in, lux_radiance_cascades* out) {
    if (!in || !out) return LUX_ERR_INVALID;
    to_c(from_c(in).sanitized(), out);
    return LUX_OK;
}

int lux_radiance_cascades_gi_params(const lux_radiance_cascades* in, const float origin[3],
                                    const uint32_t dims[3], uint32_t cascades, uint32_t base_phi,
                                    uint32_t base_y, uint32_t interval_mode, uint32_t frame,
                                    lux_gi_params* out) {
    if (!in || !origin || !dims || !out) return LUX_ERR_INVALID;
    bevy_lux::ProbeVolume v;
    for (int a = 0; a < 3; ++a) {
        v.origin[a] = origin[a];
        v.dims[a] = dims[a];
    }
    v.cascades = cascades;
    v.base_phi = base_phi;
    v.base_y = base_y;
    v.interval_mode = interval_mode;
    *out = bevy_lux::gi_params_for(from_c(in), v, frame);
    return LUX_OK;
}

int lux_radiance_cascades_history_params(const lux_radiance_cascades* in, lux_gi_history_params* out) {
    if (!in || !out) return LUX_ERR_INVALID;
    bevy_lux::RadianceCascades r = from_c(in).sanitized();
    out->radiance_blend = r.temporal_blend;
    out->transmittance_blend = r.split_transmittance_blend;
    out->history_frames = r.history_frames;
    out->persistent = 1u;
    return LUX_OK;
}

// resolution() (src/render.rs:5350-5380): the receiver (candidate) grid of
// a viewport -- the smallest divisor >= the configured one whose grid holds at
// most max_receiver_cells receivers (131072 when 0).
int lux_radiance_cascades_receiver_grid(const lux_radiance_cascades* in, uint32_t viewport_width,
                                        uint32_t viewport_height, uint32_t* width, uint32_t* height) {
    if (!in || !width || !height) return LUX_ERR_INVALID;
    auto size_at = [&](uint32_t d, uint32_t& w, uint32_t& h) {
        w = std::max<uint32_t>(1u, viewport_width / d + (viewport_width % d != 0));
        h = std::max<uint32_t>(1u, viewport_height / d + (viewport_height % d != 0));
    };
    const uint64_t max_receivers = in->sampling[1] == 0 ? 131072u : in->sampling[1];
    const uint32_t divisor = in->sampling[0] != 0 ? in->sampling[0] : in->resolution_divisor;
    uint32_t low = std::max<uint32_t>(divisor, 1u);
    uint32_t high = std::max(std::max(viewport_width, viewport_height), low);
    while (low < high) {
        uint32_t mid = low + (high - low) / 2u, w, h;
        size_at(mid, w, h);
        if ((uint64_t)w * h > max_receivers) low = mid + 1u;
        else high = mid;
    }
    size_at(low, *width, *height);
    return LUX_OK;
}

// selected_update_budget / runtime_update_budget (src/render.rs:243-255):
// max_updates_per_frame (max_rays when 0) capped by max_rays, at least 1,
// then by the runtime cap when has_cap (a cap of 0 counts as 1).
int lux_radiance_cascades_update_budget(const lux_radiance_cascades* in, uint32_t has_cap, uint32_t runtime_cap,
                                        uint32_t* budget) {
    if (!in || !budget) return LUX_ERR_INVALID;
    const uint32_t ceiling = std::max<uint32_t>(in->max_rays, 1u);
    const uint32_t requested = in->sampling[2] == 0 ? ceiling : in->sampling[2];
    uint32_t b = std::max<uint32_t>(std::min(requested, ceiling), 1u);
    if (has_cap) b = std::min(b, std::max<uint32_t>(runtime_cap, 1u));
    *budget = b;
    return LUX_OK;
}

// CandidateSchedule (src/render.rs:5172-5211): the fair window's cursor over
// the stable rank permutation, advanced only after a submitted update.
int lux_candidate_schedule_phase(lux_candidate_schedule* s, uint32_t width, uint32_t height, uint32_t count,
                                 uint32_t budget, uint32_t* cursor, uint32_t* quota) {
    if (!s || !cursor || !quota) return LUX_ERR_INVALID;
    if (s->width != width || s->height != height || s->count != count) {
        s->width = width;
        s->height = height;
        s->count = count;
        s->cursor = 0;
    }
    s->cursor = count > 0 ? s->cursor % count : 0u;
    s->pending = std::min(budget, count);
    *cursor = s->cursor;
    *quota = s->pending;
    return LUX_OK;
}

int lux_candidate_schedule_advance(lux_candidate_schedule* s, uint32_t consume_fair_cursor) {
    if (!s) return LUX_ERR_INVALID;
    if (consume_fair_cursor && s->count > 0)
        s->cursor = (uint32_t)(((uint64_t)s->cursor + s->pending) % s->count);
    return LUX_OK;
}

int lux_cold_priority_enabled(uint32_t history_frames, uint32_t history_was_valid) {
    return history_frames > 0 && !history_was_valid ? 1 : 0;
}

}  // extern "C"

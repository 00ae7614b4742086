// radiance_cascades.hpp -- the Bevy-facing surface of bevy_lux, kept as-is
// above the lux C ABI.
//
// Mirrors src/lib.rs of the reference: RadianceCascadesGeometry (:305-307),
// SamplingSettings (:344-407), RadianceCascades (:431-518, defaults :671-698,
// low_end :700-722, reset_history/sanitized :724-759), GiMethod (:631-669),
// RadianceCascadesStatus::record_failure (:225-230) and
// RadianceCascadesPlugin (:795-900).  Field names, defaults and clamps are
// the reference's; Plugin::build installs the MI355X backend (a lux context)
// instead of wgpu pipelines, and a view's GI update maps the camera component
// onto lux_gi_params.  Rust has no toolchain in this build image, so the
// surface is restated in C++; the Rust binding a maintainer would add is the
// extern "C" block over include/lux.h.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lux.h"

namespace bevy_lux {

// src/lib.rs:305-307 -- marker: meshes with this component enter the GI scene.
struct RadianceCascadesGeometry {};

// src/lib.rs:631-669
enum class GiMethod : uint32_t { SplitRadianceCascades = 0, RayTraced = 1, Pxtgi = 2 };

// src/lib.rs:344-407
struct SamplingSettings {
    uint32_t grid_divisor = 0;
    uint32_t max_receiver_cells = 0;
    uint32_t max_updates_per_frame = 0;
    uint32_t max_cache_entries = 0;
    uint32_t max_traversal_queries = 0;
    uint32_t resolve_scale = 1;

    static SamplingSettings unlimited();
    static SamplingSettings low_end();
    SamplingSettings sanitized() const;
};

struct LinearRgba {
    float red = 0.0f, green = 0.0f, blue = 0.0f, alpha = 1.0f;
};

// src/lib.rs:431-518 (RayTraced/Pxtgi settings are not part of this path and
// are omitted; `method` other than Split is rejected by the backend).
struct RadianceCascades {
    GiMethod method = GiMethod::SplitRadianceCascades;
    SamplingSettings sampling = SamplingSettings::unlimited();
    uint32_t history_epoch = 0;
    float probe_spacing = 0.25f;
    float interval_length = 0.4f;
    uint32_t resolution_divisor = 2;
    uint32_t max_rays = 131072;
    float normal_bias = 0.005f;
    float intensity = 1.0f;
    bool multi_bounce = false;
    bool cached_hit_sky = false;
    bool hardware_ray_queries = false;
    bool ray_traced_sun_shadows = false;
    bool artistic_occlusion = true;
    LinearRgba environment{0.0f, 0.0f, 0.0f, 1.0f};
    float temporal_blend = 0.88f;
    float split_transmittance_blend = 0.98f;
    uint32_t history_frames = 24;
    uint32_t spatial_filter_steps = 3;
    uint32_t emissive_samples = 16;

    static RadianceCascades low_end();
    void reset_history();
    RadianceCascades sanitized() const;
    bool radiance_transport_active() const;
};

// src/lib.rs:38-274 (failure recording subset).
struct RadianceCascadesStatus {
    std::vector<std::string> failures;
    uint64_t revision = 0;
    void record_failure(const std::string& message);
};

// Dense probe volume a view updates (the configs' world-space probe grid).
struct ProbeVolume {
    float origin[3] = {0.0f, 0.0f, 0.0f};
    uint32_t dims[3] = {16, 16, 16};
    uint32_t cascades = 4;
    uint32_t base_phi = 8, base_y = 4;  // reference preset: 32 level-0 directions
    uint32_t interval_mode = LUX_INTERVAL_POW4;
};

// src/lib.rs:795-900
struct RadianceCascadesPlugin {
    uint64_t texture_atlas_budget_bytes = 256ull * 1024 * 1024;
    uint64_t scene_gpu_budget_bytes = 256ull * 1024 * 1024;
    uint64_t view_gpu_budget_bytes = 2ull * 1024 * 1024 * 1024;
    uint64_t geometry_upload_bytes_per_frame = 2ull * 1024 * 1024;
    uint32_t max_concurrent_pipeline_compiles = 2;

    static RadianceCascadesPlugin low_end();
};

// The render-world side: one lux context per device, driven per frame.
class Renderer {
public:
    explicit Renderer(const RadianceCascadesPlugin& plugin, int device = 0);
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    bool ok() const { return ctx_ != nullptr; }
    RadianceCascadesStatus& status() { return status_; }
    lux_ctx* context() { return ctx_; }

    // update_scene: meshes tagged RadianceCascadesGeometry and their instances.
    bool update_scene(const std::vector<lux_mesh_desc>& meshes,
                      const std::vector<lux_instance_desc>& instances,
                      const std::vector<lux_material>& materials);
    // One Split GI update for a view; returns false (and records a failure)
    // when the backend rejects it.  ms receives the GPU time when non-null.
    bool render_view(const RadianceCascades& camera, const ProbeVolume& volume, uint32_t frame,
                     float* ms = nullptr);

private:
    bool check(int rc, const char* what);
    RadianceCascadesPlugin plugin_;
    RadianceCascadesStatus status_;
    lux_ctx* ctx_ = nullptr;
};

// Map a sanitized camera component and probe volume onto lux_gi_params.
lux_gi_params gi_params_for(const RadianceCascades& camera, const ProbeVolume& volume,
                            uint32_t frame);

}  // namespace bevy_lux

// lux_api.hip -- context, scene upload and BVH orchestration behind the C ABI.
//
// The context owns every device allocation (the reference's ViewBuffers /
// SceneData ownership, src/render.rs:5077-5160); caller host arrays are
// copied on call.  Errors never abort: they return a negative status and the
// text is kept for lux_last_error (forwarded to
// RadianceCascadesStatus::record_failure, src/lib.rs:225-230).
#include <algorithm>
#include <atomic>
#include <thread>
#include <cstring>
#include <cstdio>
#include "lux_internal.h"

int lux_fail(lux_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->error = msg;
    return code;
}

int lux_hip_check(lux_ctx* ctx, hipError_t e, const char* what) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s failed: %s", what, hipGetErrorString(e));
    return lux_fail(ctx, e == hipErrorOutOfMemory ? LUX_ERR_OOM : LUX_ERR_HIP, buf);
}

void lux_release(lux_buffer& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

int lux_ensure(lux_ctx* ctx, lux_buffer& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return LUX_OK;
    if (b.ptr) {
        // the buffer may still be read by queued work
        LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        lux_release(b);
    }
    size_t cap = bytes + bytes / 8;
    LUX_HIP(ctx, hipMalloc(&b.ptr, cap));
    b.bytes = cap;
    return LUX_OK;
}

namespace {

__device__ __forceinline__ float sel_min(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float sel_max(float a, float b) { return b > a ? b : a; }

// Gather one mesh's triangles into BLAS leaf order (SAH builds).
__global__ void k_tri_gather(const float* pos, const uint32_t* mat, const uint32_t* perm, uint32_t n,
                             float4* tri, uint32_t* order, uint32_t tri_base) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t s = perm[j];
    const float* p = pos + 9ull * s;
    float4* t = tri + 3ull * j;
    t[0] = make_float4(p[0], p[1], p[2], __uint_as_float(s));
    t[1] = make_float4(p[3], p[4], p[5], __uint_as_float(mat ? mat[s] : 0u));
    t[2] = make_float4(p[6], p[7], p[8], 0.0f);
    order[j] = tri_base + s;
}

struct InstDesc {
    float inv[12];
    float fwd[12];
    uint32_t mesh, flags;
};

__global__ void k_inst_boxes(const InstDesc* d, uint32_t n, const float4* blas,
                             const uint32_t* mesh_blas_base, float4* lo, float4* hi) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const InstDesc& in = d[i];
    uint32_t root = mesh_blas_base[in.mesh];
    float4 a = blas[2u * root], b = blas[2u * root + 1u];
    float mn[3] = {a.x, a.y, a.z}, mx[3] = {b.x, b.y, b.z};
    float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 8; ++k) {
        float p[3] = {(k & 1) ? mx[0] : mn[0], (k & 2) ? mx[1] : mn[1], (k & 4) ? mx[2] : mn[2]};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float* w = in.fwd + 4 * r;
            float v = ((w[0] * p[0] + w[1] * p[1]) + w[2] * p[2]) + w[3] * 1.0f;
            l[r] = sel_min(l[r], v);
            h[r] = sel_max(h[r], v);
        }
    }
    float4 lo4, hi4;
    float* lp = &lo4.x;
    float* hp = &hi4.x;
#pragma unroll
    for (int a2 = 0; a2 < 3; ++a2) {
        float m = fmaxf(fabsf(l[a2]), fabsf(h[a2]));
        float pad = m * 4.76837158e-7f + 1e-30f;
        lp[a2] = l[a2] - pad;
        hp[a2] = h[a2] + pad;
    }
    lo4.w = hi4.w = 0.0f;
    lo[i] = lo4;
    hi[i] = hi4;
}

__global__ void k_inst_records(const InstDesc* d, const uint32_t* perm, uint32_t n,
                               const uint32_t* mesh_blas_base, const uint32_t* mesh_tri_base,
                               const uint32_t* mesh_tri_count, float4* rec, uint32_t* order) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t s = perm[k];
    const InstDesc& in = d[s];
    float4* r = rec + 8u * k;
    for (int i = 0; i < 3; ++i) {
        r[i] = make_float4(in.inv[4 * i], in.inv[4 * i + 1], in.inv[4 * i + 2], in.inv[4 * i + 3]);
        r[3 + i] = make_float4(in.fwd[4 * i], in.fwd[4 * i + 1], in.fwd[4 * i + 2], in.fwd[4 * i + 3]);
    }
    uint32_t m = in.mesh;
    r[6] = make_float4(__uint_as_float(in.flags), __uint_as_float(mesh_blas_base[m]),
                       __uint_as_float(mesh_tri_base[m]), __uint_as_float(mesh_tri_count[m]));
    r[7] = make_float4(__uint_as_float(s), __uint_as_float(m), 0.0f, 0.0f);
    order[k] = s;
}

// Refit of a built TLAS to new instance boxes (the reference's refit,
// src/scene.rs:1040-1058): topology and instance order stay, every node
// reachable from the root gets the union of its children's boxes, a leaf the
// union of its instances' (TLAS slot j holds caller instance order[j]).
// One thread, post-order with an explicit stack (children may sit at lower
// slots than their parent in the LBVH layout).
// The same refit with nodes and instance boxes staged in LDS (dynamic
// shared memory: 32 bytes per node, then 32 per instance): the block loads
// them, one thread refits, the block writes the nodes back.
__global__ __launch_bounds__(256) void k_tlas_refit_lds(float4* nodes, uint32_t n_nodes, const float4* lo,
                                                        const float4* hi, const uint32_t* order, uint32_t n_inst) {
    extern __shared__ float4 sm[];
    float4* sn = sm;                  // 2 per node
    float4* sb = sm + 2u * n_nodes;   // lo, hi per TLAS slot (instance order[j])
    for (uint32_t i = threadIdx.x; i < 2u * n_nodes; i += blockDim.x) sn[i] = nodes[i];
    for (uint32_t j = threadIdx.x; j < n_inst; j += blockDim.x) {
        uint32_t c = order[j];
        sb[2u * j] = lo[c];
        sb[2u * j + 1u] = hi[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t stack[128];
        uint32_t sp = 0;
        stack[sp++] = 0u;
        while (sp > 0) {
            uint32_t top = stack[sp - 1];
            uint32_t slot = top & 0x7fffffffu;
            if (slot >= n_nodes) {
                --sp;
                continue;
            }
            float4 a = sn[2u * slot], b = sn[2u * slot + 1u];
            uint32_t first = __float_as_uint(a.w), count = __float_as_uint(b.w);
            float4 mn = make_float4(INFINITY, INFINITY, INFINITY, a.w);
            float4 mx = make_float4(-INFINITY, -INFINITY, -INFINITY, b.w);
            if (count > 0u) {
                for (uint32_t j = first; j < first + count && j < n_inst; ++j) {
                    float4 l = sb[2u * j], h = sb[2u * j + 1u];
                    mn = make_float4(sel_min(mn.x, l.x), sel_min(mn.y, l.y), sel_min(mn.z, l.z), a.w);
                    mx = make_float4(sel_max(mx.x, h.x), sel_max(mx.y, h.y), sel_max(mx.z, h.z), b.w);
                }
            } else if (!(top & 0x80000000u) && sp + 2 <= 128u && first + 1u < n_nodes) {
                stack[sp - 1] = top | 0x80000000u;
                stack[sp++] = first;
                stack[sp++] = first + 1u;
                continue;
            } else if (first + 1u < n_nodes) {
                float4 la = sn[2u * first], lb = sn[2u * first + 1u];
                float4 ra = sn[2u * first + 2u], rb = sn[2u * first + 3u];
                mn = make_float4(sel_min(la.x, ra.x), sel_min(la.y, ra.y), sel_min(la.z, ra.z), a.w);
                mx = make_float4(sel_max(lb.x, rb.x), sel_max(lb.y, rb.y), sel_max(lb.z, rb.z), b.w);
            } else {
                --sp;
                continue;
            }
            sn[2u * slot] = mn;
            sn[2u * slot + 1u] = mx;
            --sp;
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 2u * n_nodes; i += blockDim.x) nodes[i] = sn[i];
}

__global__ void k_tlas_refit(float4* nodes, uint32_t n_nodes, const float4* lo, const float4* hi,
                             const uint32_t* order, uint32_t n_inst) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t stack[128];
    uint32_t sp = 0;
    stack[sp++] = 0u;  // bit 31: children done
    while (sp > 0) {
        uint32_t top = stack[sp - 1];
        uint32_t slot = top & 0x7fffffffu;
        if (slot >= n_nodes) {
            --sp;
            continue;
        }
        float4 a = nodes[2u * slot], b = nodes[2u * slot + 1u];
        uint32_t first = __float_as_uint(a.w), count = __float_as_uint(b.w);
        float4 mn = make_float4(INFINITY, INFINITY, INFINITY, a.w), mx = make_float4(-INFINITY, -INFINITY, -INFINITY, b.w);
        if (count > 0u) {
            for (uint32_t j = first; j < first + count && j < n_inst; ++j) {
                uint32_t s = order[j];
                float4 l = lo[s], h = hi[s];
                mn = make_float4(sel_min(mn.x, l.x), sel_min(mn.y, l.y), sel_min(mn.z, l.z), a.w);
                mx = make_float4(sel_max(mx.x, h.x), sel_max(mx.y, h.y), sel_max(mx.z, h.z), b.w);
            }
        } else if (!(top & 0x80000000u) && sp + 2 <= 128u && first + 1u < n_nodes) {
            stack[sp - 1] = top | 0x80000000u;
            stack[sp++] = first;
            stack[sp++] = first + 1u;
            continue;
        } else if (first + 1u < n_nodes) {
            float4 la = nodes[2u * first], lb = nodes[2u * first + 1u];
            float4 ra = nodes[2u * first + 2u], rb = nodes[2u * first + 3u];
            mn = make_float4(sel_min(la.x, ra.x), sel_min(la.y, ra.y), sel_min(la.z, ra.z), a.w);
            mx = make_float4(sel_max(lb.x, rb.x), sel_max(lb.y, rb.y), sel_max(lb.z, rb.z), b.w);
        } else {
            --sp;
            continue;
        }
        nodes[2u * slot] = mn;
        nodes[2u * slot + 1u] = mx;
        --sp;
    }
}

inline uint32_t nblk(uint32_t n, uint32_t t) { return (n + t - 1) / t; }

}  // namespace

extern "C" int lux_ctx_create(int device, lux_ctx** out) {
    if (!out) return LUX_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return LUX_ERR_HIP;
    if (device < 0 || device >= count) return LUX_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return LUX_ERR_HIP;
    lux_ctx* ctx = new lux_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        lux_ctx_destroy(ctx);
        return LUX_ERR_HIP;
    }
    for (hipEvent_t& e : ctx->ev_phase)
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            lux_ctx_destroy(ctx);
            return LUX_ERR_HIP;
        }
    for (hipEvent_t& e : ctx->ev_stage)
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            lux_ctx_destroy(ctx);
            return LUX_ERR_HIP;
        }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->cu_count = prop.multiProcessorCount;
    if (int rc = lux_trace_scratch_init(ctx)) {
        lux_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return LUX_OK;
}

extern "C" void lux_ctx_destroy(lux_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    lux_buffer* bufs[] = {&ctx->raw_pos, &ctx->raw_mat, &ctx->inst_desc, &ctx->materials,
                          &ctx->tris, &ctx->tri_order, &ctx->blas,
                          &ctx->tlas, &ctx->inst, &ctx->inst_order, &ctx->scratch,
                          &ctx->ray_buf, &ctx->hit_buf, &ctx->gi_intervals, &ctx->gi_merged,
                          &ctx->gi_probes, &ctx->gi_light_buf, &ctx->gi_sur_queries,
                          &ctx->gi_sur_err, &ctx->gi_sur_list, &ctx->gi_counters,
                          &ctx->tlas_tmp, &ctx->mesh_info, &ctx->gi_w_prim, &ctx->gi_w_shade,
                          &ctx->gi_w_slots, &ctx->gi_w_vis, &ctx->gi_w_shadow, &ctx->gi_w_counters,
                          &ctx->gi_samples, &ctx->gi_history, &ctx->gi_hist_state,
                          &ctx->rx_keys, &ctx->rx_flags, &ctx->rx_slot, &ctx->rx_lod, &ctx->rx_ord,
                          &ctx->rx_seq, &ctx->rx_dir, &ctx->rx_tickets, &ctx->rx_dep_val, &ctx->rx_pack, &ctx->rx_traced,
                          &ctx->rx_sort, &ctx->rx_scan, &ctx->rx_counters, &ctx->rx_cold,
                          &ctx->screen_cache, &ctx->range_ok, &ctx->gi_stats_buf, &ctx->gi_dep, &ctx->gi_w_dep};
    for (lux_buffer* b : bufs) lux_release(*b);
    for (hipEvent_t e : ctx->ev_phase)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->ev_stage)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char* lux_last_error(const lux_ctx* ctx) {
    return ctx ? ctx->error.c_str() : "null context";
}

extern "C" void* lux_ctx_stream(lux_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int lux_ctx_synchronize(lux_ctx* ctx) {
    if (!ctx) return LUX_ERR_INVALID;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LUX_OK;
}

extern "C" int lux_scene_upload_soa(lux_ctx* ctx, const lux_mesh_desc* meshes, uint32_t mesh_count,
                                    const lux_instance_desc* instances, uint32_t instance_count,
                                    const lux_material* materials, uint32_t material_count) {
    if (!ctx) return LUX_ERR_INVALID;
    if ((mesh_count && !meshes) || (instance_count && !instances) || (material_count && !materials))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_soa: null array");
    uint64_t total = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) {
        if (meshes[m].triangle_count == 0)
            return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_soa: mesh without triangles");
        if (!meshes[m].positions)
            return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_soa: mesh without positions");
        total += meshes[m].triangle_count;
    }
    if (total >= (1ull << 31))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_soa: too many triangles");
    for (uint32_t i = 0; i < instance_count; ++i)
        if (instances[i].mesh >= mesh_count)
            return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_soa: instance mesh out of range");
    LUX_HIP(ctx, hipSetDevice(ctx->device));
    ctx->n_meshes = mesh_count;
    ctx->n_tri = (uint32_t)total;
    ctx->n_inst = instance_count;
    ctx->n_materials = material_count;
    ctx->mesh_tri_count.assign(mesh_count, 0);
    ctx->mesh_tri_base.assign(mesh_count, 0);
    ctx->mesh_blas_base.assign(mesh_count, 0);
    uint32_t tb = 0, bb = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) {
        uint32_t n = meshes[m].triangle_count;
        ctx->mesh_tri_count[m] = n;
        ctx->mesh_tri_base[m] = tb;
        ctx->mesh_blas_base[m] = bb;
        tb += n;
        bb += n > 1 ? 2 * n - 1 : 1;
    }
    ctx->n_blas = bb;
    if (int rc = lux_ensure(ctx, ctx->raw_pos, 36ull * total)) return rc;
    if (int rc = lux_ensure(ctx, ctx->raw_mat, 4ull * total)) return rc;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (uint32_t m = 0; m < mesh_count; ++m) {
        uint32_t n = meshes[m].triangle_count;
        float* dp = static_cast<float*>(ctx->raw_pos.ptr) + 9ull * ctx->mesh_tri_base[m];
        uint32_t* dm = static_cast<uint32_t*>(ctx->raw_mat.ptr) + ctx->mesh_tri_base[m];
        LUX_HIP(ctx, hipMemcpy(dp, meshes[m].positions, 36ull * n, hipMemcpyHostToDevice));
        if (meshes[m].materials) {
            LUX_HIP(ctx, hipMemcpy(dm, meshes[m].materials, 4ull * n, hipMemcpyHostToDevice));
        } else {
            LUX_HIP(ctx, hipMemset(dm, 0, 4ull * n));
        }
    }
    ctx->host_pos.assign(mesh_count, {});
    ctx->mesh_bounds.clear();
    ctx->mesh_emissive.clear();
    ctx->host_mat.assign(mesh_count, {});
    for (uint32_t m = 0; m < mesh_count; ++m) {
        uint32_t n = meshes[m].triangle_count;
        ctx->host_pos[m].assign(meshes[m].positions, meshes[m].positions + 9ull * n);
        if (meshes[m].materials) ctx->host_mat[m].assign(meshes[m].materials, meshes[m].materials + n);
        else ctx->host_mat[m].assign(n, 0u);
    }
    ctx->host_materials.assign(materials, materials + material_count);
    ctx->gi_ready = false;
    ctx->prebuilt = false;
    ctx->inst_material.clear();
    ctx->instances.assign(instances, instances + instance_count);
    std::vector<InstDesc> idesc(instance_count);
    for (uint32_t i = 0; i < instance_count; ++i) {
        memcpy(idesc[i].inv, instances[i].local_from_world, 48);
        memcpy(idesc[i].fwd, instances[i].world_from_local, 48);
        idesc[i].mesh = instances[i].mesh;
        idesc[i].flags = instances[i].flags;
    }
    if (int rc = lux_ensure(ctx, ctx->inst_desc, sizeof(InstDesc) * instance_count + 16)) return rc;
    if (instance_count)
        LUX_HIP(ctx, hipMemcpy(ctx->inst_desc.ptr, idesc.data(), sizeof(InstDesc) * instance_count,
                               hipMemcpyHostToDevice));
    if (int rc = lux_ensure(ctx, ctx->materials, sizeof(lux_material) * material_count + 16)) return rc;
    if (material_count)
        LUX_HIP(ctx, hipMemcpy(ctx->materials.ptr, materials, sizeof(lux_material) * material_count,
                               hipMemcpyHostToDevice));
    ctx->blas_built = false;
    ctx->tlas_built = false;
    return LUX_OK;
}

extern "C" int lux_scene_set_transforms(lux_ctx* ctx, const float* local_from_world,
                                        const float* world_from_local, uint32_t instance_count) {
    if (!ctx) return LUX_ERR_INVALID;
    if (ctx->prebuilt)
        return lux_fail(ctx, LUX_ERR_STATE, "lux_scene_set_transforms: scene installed built (upload new words)");
    if (instance_count != ctx->n_inst)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_set_transforms: instance count changed");
    if (instance_count && (!local_from_world || !world_from_local))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_set_transforms: null array");
    // dirty regions for the GI history: the world boxes of every moved
    // instance before and after the move.  A measured-empty dependency (a ray
    // that missed the scene's root box) stays valid only while the root does
    // not grow: a moved box leaving the previous root marks the next history
    // pass to drop measured-empty dependencies (the reference's root-revision
    // precondition, shaders/split.slang:19-21).  A moved instance that emits
    // changes the light every NEE sample may reach, sampled or not: the
    // whole history goes.
    std::vector<float> dirty;
    bool root_grew = false, reset_history = false;
    if (ctx->gi_hist_on) {
        if (ctx->mesh_bounds.size() != 6ull * ctx->host_pos.size()) {
            ctx->mesh_bounds.assign(6ull * ctx->host_pos.size(), 0.0f);
            for (size_t m = 0; m < ctx->host_pos.size(); ++m) {
                const std::vector<float>& pos = ctx->host_pos[m];
                float* mb = &ctx->mesh_bounds[6 * m];
                for (int q = 0; q < 3; ++q) {
                    mb[q] = pos.empty() ? 0.0f : pos[q];
                    mb[3 + q] = mb[q];
                }
                for (size_t v = 0; v < pos.size(); v += 3)
                    for (int q = 0; q < 3; ++q) {
                        mb[q] = std::min(mb[q], pos[v + q]);
                        mb[3 + q] = std::max(mb[3 + q], pos[v + q]);
                    }
            }
        }
        auto world_box = [&](uint32_t m, const float* f, float* out) {
            const float* mb = &ctx->mesh_bounds[6ull * m];
            for (int r = 0; r < 3; ++r) {
                out[r] = INFINITY;
                out[3 + r] = -INFINITY;
            }
            for (int c = 0; c < 8; ++c) {
                const float pl[3] = {(c & 1) ? mb[3] : mb[0], (c & 2) ? mb[4] : mb[1], (c & 4) ? mb[5] : mb[2]};
                for (int r = 0; r < 3; ++r) {
                    const float w = ((f[4 * r] * pl[0] + f[4 * r + 1] * pl[1]) + f[4 * r + 2] * pl[2]) + f[4 * r + 3];
                    out[r] = std::min(out[r], w);
                    out[3 + r] = std::max(out[3 + r], w);
                }
            }
        };
        float root[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = 0; i < instance_count; ++i) {
            const uint32_t m = ctx->instances[i].mesh;
            if (m >= ctx->host_pos.size() || ctx->host_pos[m].empty()) continue;
            float wb[6];
            world_box(m, ctx->instances[i].world_from_local, wb);
            for (int q = 0; q < 3; ++q) {
                root[q] = std::min(root[q], wb[q]);
                root[3 + q] = std::max(root[3 + q], wb[3 + q]);
            }
        }
        for (uint32_t i = 0; i < instance_count; ++i) {
            const float* nf = world_from_local + 12ull * i;
            if (memcmp(ctx->instances[i].world_from_local, nf, 48) == 0) continue;
            const uint32_t m = ctx->instances[i].mesh;
            if (m >= ctx->host_pos.size() || ctx->host_pos[m].empty()) continue;
            if (m < ctx->host_mat.size())
                for (uint32_t mat : ctx->host_mat[m]) {
                    if (mat >= ctx->host_materials.size()) continue;
                    const float* e = ctx->host_materials[mat].emission;
                    if (e[0] > 0.0f || e[1] > 0.0f || e[2] > 0.0f) {
                        reset_history = true;
                        break;
                    }
                }
            float ob[6], nb[6];
            world_box(m, ctx->instances[i].world_from_local, ob);
            world_box(m, nf, nb);
            for (int q = 0; q < 3; ++q) {
                // outside the previous root beyond a relative margin (the
                // device rounds its own root box)
                const float pad = 1e-5f * std::max(1.0f, std::max(std::fabs(root[q]), std::fabs(root[3 + q])));
                if (!(nb[q] >= root[q] - pad && nb[3 + q] <= root[3 + q] + pad)) root_grew = true;
            }
            dirty.insert(dirty.end(), ob, ob + 6);
            dirty.insert(dirty.end(), nb, nb + 6);
        }
    }
    std::vector<InstDesc> idesc(instance_count);
    for (uint32_t i = 0; i < instance_count; ++i) {
        memcpy(idesc[i].inv, local_from_world + 12ull * i, 48);
        memcpy(idesc[i].fwd, world_from_local + 12ull * i, 48);
        idesc[i].mesh = ctx->instances[i].mesh;
        idesc[i].flags = ctx->instances[i].flags;
        memcpy(ctx->instances[i].local_from_world, local_from_world + 12ull * i, 48);
        memcpy(ctx->instances[i].world_from_local, world_from_local + 12ull * i, 48);
    }
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (instance_count)
        LUX_HIP(ctx, hipMemcpy(ctx->inst_desc.ptr, idesc.data(), sizeof(InstDesc) * instance_count,
                               hipMemcpyHostToDevice));
    ctx->tlas_built = false;
    ctx->gi_lights_dirty = true;
    if (reset_history) return lux_gi_reset_history(ctx);
    if (root_grew)
        if (int rc = lux_gi_root_grew(ctx)) return rc;
    return lux_gi_invalidate_boxes(ctx, dirty.data(), (uint32_t)(dirty.size() / 6));
}

namespace {
// Binned SAH over one mesh's triangle boxes on the host: build_tree
// (src/scene.rs:2398-2554) -- box and centroid folds with glam's Vec3::min /
// max per component (a < b ? a : b: the later of equal values wins), 12 bins
// on the centroid span per axis with an
// f64 bin index, bin costs area(f64) x count, the cheapest split with both
// sides non-empty (axis order x, y, z, first strictly cheaper wins), Hoare
// partition by bin, leaves of at most 4 and depth 48, and for coincident
// centroids a median split on the widest centroid axis.  The reference takes
// that median with select_nth_unstable_by (by f32::total_cmp), which sorts a
// slice of up to 10 elements by stable insertion; here the elements are
// stably sorted by the total-order centre alone, the same order for those
// slices (elements of equal key keep their current order).  Nodes are
// written depth-first as the reference appends them (children pairs after
// their parent).  A failed allocation in a worker thread is recorded in
// `failed` and reported by the caller.
struct SahBuilder {
    const float* lo;  // 4 floats per element
    const float* hi;
    uint32_t* perm;
    std::vector<lux_bvh_node> nodes;
    std::atomic<bool>* failed = nullptr;

    float centre(uint32_t e, int a) const { return lo[4 * e + a] * 0.5f + hi[4 * e + a] * 0.5f; }
    static float vmin(float a, float b) { return a < b ? a : b; }
    static float vmax(float a, float b) { return a > b ? a : b; }

    static double area(const float* mn, const float* mx) {
        const double sx = (double)mx[0] - (double)mn[0], sy = (double)mx[1] - (double)mn[1];
        const double sz = (double)mx[2] - (double)mn[2];
        return 2.0 * (sx * sy + sy * sz + sz * sx);
    }

    struct Bin {
        float mn[3], mx[3];
        size_t count;
        void clear() {
            for (int a = 0; a < 3; ++a) {
                mn[a] = INFINITY;
                mx[a] = -INFINITY;
            }
            count = 0;
        }
        void add(const Bin& o) {
            for (int a = 0; a < 3; ++a) {
                mn[a] = vmin(mn[a], o.mn[a]);
                mx[a] = vmax(mx[a], o.mx[a]);
            }
            count += o.count;
        }
        double cost() const { return count == 0 ? 0.0 : area(mn, mx) * (double)count; }
    };

    static size_t bin_of(float c, float cmin, double span) {
        const size_t b = (size_t)(((double)c - (double)cmin) / span * 12.0);
        return b < 11 ? b : 11;
    }

    void build(uint32_t index, uint32_t first, uint32_t count, uint32_t depth) {
        uint32_t* el = perm + first;
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = 0; i < count; ++i)
            for (int a = 0; a < 3; ++a) {
                mn[a] = vmin(mn[a], lo[4 * el[i] + a]);
                mx[a] = vmax(mx[a], hi[4 * el[i] + a]);
            }
        lux_bvh_node& nd = nodes[index];
        std::memcpy(nd.mn, mn, 12);
        std::memcpy(nd.mx, mx, 12);
        nd.first = first;
        nd.count = count;
        if (count <= 4 || depth == 48) return;
        float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = 0; i < count; ++i)
            for (int a = 0; a < 3; ++a) {
                const float c = centre(el[i], a);
                cmin[a] = vmin(cmin[a], c);
                cmax[a] = vmax(cmax[a], c);
            }
        double best = INFINITY, best_span = 0.0;
        int best_axis = -1;
        size_t best_bin = 0;
        for (int axis = 0; axis < 3; ++axis) {
            const double span = (double)cmax[axis] - (double)cmin[axis];
            if (span <= 0.0) continue;
            Bin bins[12], right[12], sum, left;
            for (Bin& b : bins) b.clear();
            for (uint32_t i = 0; i < count; ++i) {
                Bin one;
                for (int a = 0; a < 3; ++a) {
                    one.mn[a] = lo[4 * el[i] + a];
                    one.mx[a] = hi[4 * el[i] + a];
                }
                one.count = 1;
                bins[bin_of(centre(el[i], axis), cmin[axis], span)].add(one);
            }
            sum.clear();
            for (int b = 11; b >= 0; --b) {
                sum.add(bins[b]);
                right[b] = sum;
            }
            left.clear();
            for (size_t b = 0; b < 11; ++b) {
                left.add(bins[b]);
                const double cost = left.cost() + right[b + 1].cost();
                if (left.count > 0 && right[b + 1].count > 0 && cost < best) {
                    best = cost;
                    best_axis = axis;
                    best_bin = b;
                    best_span = span;
                }
            }
        }
        uint32_t mid;
        if (best_axis >= 0) {
            uint32_t l = 0, r = count;
            while (l < r) {
                if (bin_of(centre(el[l], best_axis), cmin[best_axis], best_span) <= best_bin) {
                    ++l;
                } else {
                    --r;
                    std::swap(el[l], el[r]);
                }
            }
            mid = l;
        } else {
            const float sx = cmax[0] - cmin[0], sy = cmax[1] - cmin[1], sz = cmax[2] - cmin[2];
            const int axis = (sx >= sy && sx >= sz) ? 0 : (sy >= sz ? 1 : 2);
            // select_nth_unstable_by's f32::total_cmp, as an unsigned key
            auto key = [&](uint32_t e) {
                uint32_t u;
                const float c = centre(e, axis);
                std::memcpy(&u, &c, 4);
                return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
            };
            std::stable_sort(el, el + count, [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
            mid = count / 2;
        }
        const uint32_t children = (uint32_t)nodes.size();
        nodes.resize(nodes.size() + 2);
        nodes[index].first = children;
        nodes[index].count = 0;
        if (depth < 3 && count >= 65536) {
            // large top-level subtrees in parallel, each into its own array,
            // then spliced in the sequential pre-order: the left subtree's
            // nodes after the pair, then the right's (local k -> base + k - 1)
            SahBuilder left{lo, hi, perm, {}, failed}, right{lo, hi, perm, {}, failed};
            left.nodes.resize(1);
            right.nodes.resize(1);
            auto guarded = [this](SahBuilder& sub, uint32_t f, uint32_t c, uint32_t d) {
                try {
                    sub.build(0, f, c, d);
                } catch (...) {
                    failed->store(true);
                }
            };
            std::thread t;
            try {
                t = std::thread([&] { guarded(left, first, mid, depth + 1); });
            } catch (...) {
                guarded(left, first, mid, depth + 1);  // no thread: build it here
            }
            guarded(right, first + mid, count - mid, depth + 1);
            if (t.joinable()) t.join();
            if (failed->load()) return;
            const uint32_t base_l = (uint32_t)nodes.size();
            const uint32_t base_r = base_l + (uint32_t)left.nodes.size() - 1u;
            auto moved = [](lux_bvh_node nd, uint32_t base) {
                if (nd.count == 0) nd.first = base + nd.first - 1u;
                return nd;
            };
            nodes[children] = moved(left.nodes[0], base_l);
            nodes[children + 1] = moved(right.nodes[0], base_r);
            nodes.reserve(nodes.size() + left.nodes.size() + right.nodes.size());
            for (size_t k = 1; k < left.nodes.size(); ++k) nodes.push_back(moved(left.nodes[k], base_l));
            for (size_t k = 1; k < right.nodes.size(); ++k) nodes.push_back(moved(right.nodes[k], base_r));
            return;
        }
        build(children, first, mid, depth + 1);
        build(children + 1, first + mid, count - mid, depth + 1);
    }
};

// One mesh's SAH BLAS: nodes with absolute child (node_base) and triangle
// (tri_base) indices, and the triangle order.
void sah_mesh(const float* pos, uint32_t n, uint32_t node_base, uint32_t tri_base, std::vector<lux_bvh_node>& nodes,
              std::vector<uint32_t>& perm, std::atomic<bool>& failed) {
    std::vector<float> lo(4ull * n), hi(4ull * n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = pos + 9ull * i;
        for (int a = 0; a < 3; ++a) {
            // Triangle::bounds: a.min(b).min(c) as compare-selects
            const float m01 = p[3 + a] < p[a] ? p[3 + a] : p[a];
            lo[4 * i + a] = p[6 + a] < m01 ? p[6 + a] : m01;
            const float x01 = p[3 + a] > p[a] ? p[3 + a] : p[a];
            hi[4 * i + a] = p[6 + a] > x01 ? p[6 + a] : x01;
        }
        lo[4 * i + 3] = hi[4 * i + 3] = 0.0f;
    }
    perm.resize(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    SahBuilder b{lo.data(), hi.data(), perm.data(), {}, &failed};
    b.nodes.resize(1);
    b.build(0, 0, n, 0);
    if (failed.load()) return;
    for (lux_bvh_node& nd : b.nodes) nd.first += nd.count == 0 ? node_base : tri_base;
    nodes.swap(b.nodes);
}
}  // namespace

namespace {
__global__ void k_range_flag(uint32_t* flag, uint32_t v) { *flag = v; }

// Clears *flag when a node bound lies outside the range of the walk's
// reciprocal slab divisions (lux::coord_in_range).
__global__ void k_range_check(const float4* nodes, uint32_t n, uint32_t* flag) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float4 a = nodes[2ull * i], b = nodes[2ull * i + 1];
    if (!(lux::coord_in_range(a.x) && lux::coord_in_range(a.y) && lux::coord_in_range(a.z) &&
          lux::coord_in_range(b.x) && lux::coord_in_range(b.y) && lux::coord_in_range(b.z)))
        *flag = 0u;
}
}  // namespace

// Range flags of the built nodes (DevScene::range_ok): word 0 for the BLAS
// nodes, word 1 for the TLAS nodes; stream-ordered, no host wait.
static int range_check(lux_ctx* ctx, bool blas, bool tlas) {
    if (!ctx->range_ok.ptr) {
        if (int rc = lux_ensure(ctx, ctx->range_ok, 16)) return rc;
        hipLaunchKernelGGL(k_range_flag, dim3(1), dim3(1), 0, ctx->stream, static_cast<uint32_t*>(ctx->range_ok.ptr), 0u);
        hipLaunchKernelGGL(k_range_flag, dim3(1), dim3(1), 0, ctx->stream,
                           static_cast<uint32_t*>(ctx->range_ok.ptr) + 1, 0u);
    }
    uint32_t* flags = static_cast<uint32_t*>(ctx->range_ok.ptr);
    const bool which[2] = {blas, tlas};
    const float4* nodes[2] = {static_cast<const float4*>(ctx->blas.ptr), static_cast<const float4*>(ctx->tlas.ptr)};
    const uint32_t count[2] = {ctx->n_blas, ctx->n_tlas};
    for (int k = 0; k < 2; ++k) {
        if (!which[k]) continue;
        hipLaunchKernelGGL(k_range_flag, dim3(1), dim3(1), 0, ctx->stream, flags + k, 1u);
        if (count[k] && nodes[k])
            hipLaunchKernelGGL(k_range_check, dim3(nblk(count[k], 256)), dim3(256), 0, ctx->stream, nodes[k], count[k],
                               flags + k);
    }
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

static int build_blas(lux_ctx* ctx, uint32_t builder) {
    hipStream_t st = ctx->stream;
    uint32_t nt = ctx->n_tri;
    if (int rc = lux_ensure(ctx, ctx->tris, 48ull * nt)) return rc;
    if (int rc = lux_ensure(ctx, ctx->tri_order, 4ull * nt)) return rc;
    if (int rc = lux_ensure(ctx, ctx->blas, 32ull * ctx->n_blas)) return rc;
    // per-mesh info for the TLAS build (BLAS base | triangle base | count)
    std::vector<uint32_t> info;
    info.insert(info.end(), ctx->mesh_blas_base.begin(), ctx->mesh_blas_base.end());
    info.insert(info.end(), ctx->mesh_tri_base.begin(), ctx->mesh_tri_base.end());
    info.insert(info.end(), ctx->mesh_tri_count.begin(), ctx->mesh_tri_count.end());
    // then the segmented build's bottom-up tiles
    std::vector<uint32_t> tiles;
    lux_segment_tiles(ctx->mesh_tri_count.data(), ctx->n_meshes, tiles);
    const size_t tiles_at = info.size();
    if (builder == 0) info.insert(info.end(), tiles.begin(), tiles.end());
    if (int rc = lux_ensure(ctx, ctx->mesh_info, 4ull * info.size() + 16)) return rc;
    LUX_HIP(ctx, hipStreamSynchronize(st));
    LUX_HIP(ctx, hipMemcpy(ctx->mesh_info.ptr, info.data(), 4ull * info.size(), hipMemcpyHostToDevice));
    int rc = LUX_OK;
    if (builder == 1) {
        // binned SAH on the GPU: every mesh's tree level by level
        rc = lux_blas_build_sah(ctx, static_cast<const float*>(ctx->raw_pos.ptr),
                                static_cast<const uint32_t*>(ctx->raw_mat.ptr), nt, ctx->n_meshes,
                                static_cast<const uint32_t*>(ctx->mesh_info.ptr), ctx->n_blas,
                                static_cast<float4*>(ctx->blas.ptr), static_cast<float4*>(ctx->tris.ptr),
                                static_cast<uint32_t*>(ctx->tri_order.ptr));
        if (rc == LUX_OK) {
            const hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) rc = lux_hip_check(ctx, e, "build_blas");
        }
        if (rc == LUX_OK) rc = range_check(ctx, true, false);
        if (rc == LUX_OK) ctx->blas_built = true;
        return rc;
    }
    if (builder == 0) {
        // Morton-30: every mesh's tree in one segmented pass
        rc = lux_blas_build_segmented(ctx, static_cast<const float*>(ctx->raw_pos.ptr),
                                      static_cast<const uint32_t*>(ctx->raw_mat.ptr), nt, ctx->n_meshes,
                                      static_cast<const uint32_t*>(ctx->mesh_info.ptr),
                                      static_cast<const uint32_t*>(ctx->mesh_info.ptr) + tiles_at,
                                      (uint32_t)tiles.size() - 1u, 4u,
                                      static_cast<float4*>(ctx->blas.ptr), static_cast<float4*>(ctx->tris.ptr),
                                      static_cast<uint32_t*>(ctx->tri_order.ptr));
        if (rc == LUX_OK) {
            const hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) rc = lux_hip_check(ctx, e, "build_blas");
        }
        if (rc == LUX_OK) rc = range_check(ctx, true, false);
        if (rc == LUX_OK) ctx->blas_built = true;
        return rc;
    }
    // SAH on the host: every mesh's tree first, meshes spread over threads,
    // then uploaded mesh by mesh with the largest mesh's permutation buffer
    for (const std::vector<float>& hp : ctx->host_pos)
        for (float v : hp)
            if (!(std::fabs(v) <= 3.402823466e38f))
                return lux_fail(ctx, LUX_ERR_INVALID, "lux_bvh_build: SAH needs finite triangle positions");
    uint32_t maxn = 1;
    for (uint32_t n : ctx->mesh_tri_count) maxn = n > maxn ? n : maxn;
    lux_buffer& tmp = ctx->tlas_tmp;
    if (int rc = lux_ensure(ctx, tmp, 4ull * maxn + 64)) return rc;
    uint32_t* perm = static_cast<uint32_t*>(tmp.ptr);
    std::vector<std::vector<lux_bvh_node>> sah_nodes;
    std::vector<std::vector<uint32_t>> sah_perm;
    {
        sah_nodes.resize(ctx->n_meshes);
        sah_perm.resize(ctx->n_meshes);
        const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const uint32_t nth = std::min<uint32_t>(hw, ctx->n_meshes);
        std::atomic<uint32_t> next{0};
        std::atomic<bool> failed{false};
        // an exception (a failed allocation) stops every worker and is
        // reported as LUX_ERR_OOM; none escapes a thread or this function
        auto work = [&]() {
            try {
                for (uint32_t m = next++; m < ctx->n_meshes && !failed.load(); m = next++)
                    sah_mesh(ctx->host_pos[m].data(), ctx->mesh_tri_count[m], ctx->mesh_blas_base[m],
                             ctx->mesh_tri_base[m], sah_nodes[m], sah_perm[m], failed);
            } catch (...) {
                failed.store(true);
            }
        };
        std::vector<std::thread> pool;
        try {
            pool.reserve(nth);
            for (uint32_t t = 1; t < nth; ++t) pool.emplace_back(work);
        } catch (...) {
            // no more threads: the ones started and this one finish the work
        }
        work();
        for (std::thread& t : pool) t.join();
        if (failed.load()) return lux_fail(ctx, LUX_ERR_OOM, "lux_bvh_build: host SAH build ran out of memory");
    }
    for (uint32_t m = 0; m < ctx->n_meshes && rc == LUX_OK; ++m) {
        uint32_t n = ctx->mesh_tri_count[m], tb = ctx->mesh_tri_base[m], bb = ctx->mesh_blas_base[m];
        const float* pos = static_cast<const float*>(ctx->raw_pos.ptr) + 9ull * tb;
        const uint32_t* mat = static_cast<const uint32_t*>(ctx->raw_mat.ptr) + tb;
        float4* nodes = static_cast<float4*>(ctx->blas.ptr) + 2ull * bb;
        // host build from the uploaded positions; unused slots of the
        // mesh's 2n - 1 region stay zero
        const std::vector<lux_bvh_node>& hn = sah_nodes[m];
        const std::vector<uint32_t>& hp = sah_perm[m];
        const uint32_t region = n > 1 ? 2 * n - 1 : 1;
        LUX_HIP(ctx, hipStreamSynchronize(st));
        LUX_HIP(ctx, hipMemset(nodes, 0, 32ull * region));
        LUX_HIP(ctx, hipMemcpy(nodes, hn.data(), 32ull * hn.size(), hipMemcpyHostToDevice));
        LUX_HIP(ctx, hipMemcpy(perm, hp.data(), 4ull * n, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_tri_gather, dim3(nblk(n, 256)), dim3(256), 0, st, pos, mat, perm, n,
                           static_cast<float4*>(ctx->tris.ptr) + 3ull * tb,
                           static_cast<uint32_t*>(ctx->tri_order.ptr) + tb, tb);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = lux_hip_check(ctx, e, "build_blas launch");
    }
    (void)hipStreamSynchronize(st);
    if (rc == LUX_OK) rc = range_check(ctx, true, false);
    if (rc == LUX_OK) ctx->blas_built = true;
    return rc;
}

static int build_tlas(lux_ctx* ctx) {
    hipStream_t st = ctx->stream;
    uint32_t ni = ctx->n_inst;
    ctx->n_tlas = ni > 1 ? 2 * ni - 1 : 1;
    if (int rc = lux_ensure(ctx, ctx->tlas, 32ull * ctx->n_tlas)) return rc;
    if (int rc = lux_ensure(ctx, ctx->inst, 128ull * (ni ? ni : 1))) return rc;
    if (int rc = lux_ensure(ctx, ctx->inst_order, 4ull * (ni ? ni : 1))) return rc;
    if (ni == 0) {
        ctx->tlas_built = true;
        return range_check(ctx, false, true);
    }
    // persistent temporaries (a rigid-motion frame rebuilds the TLAS every
    // frame: no allocation and no host wait here)
    if (int rc = lux_ensure(ctx, ctx->tlas_tmp, 36ull * ni + 64)) return rc;
    float4* lo = static_cast<float4*>(ctx->tlas_tmp.ptr);
    float4* hi = lo + ni;
    uint32_t* perm = reinterpret_cast<uint32_t*>(hi + ni);
    const uint32_t* mb = static_cast<const uint32_t*>(ctx->mesh_info.ptr);
    const uint32_t* mt = mb + ctx->n_meshes;
    const uint32_t* mc = mt + ctx->n_meshes;
    int rc = LUX_OK;
    do {
        hipError_t e;
        const InstDesc* d = static_cast<const InstDesc*>(ctx->inst_desc.ptr);
        hipLaunchKernelGGL(k_inst_boxes, dim3(nblk(ni, 128)), dim3(128), 0, st, d, ni,
                           static_cast<const float4*>(ctx->blas.ptr), mb, lo, hi);
        rc = lux_lbvh_build_device(ctx, lo, hi, ni, 4u, static_cast<float4*>(ctx->tlas.ptr), perm,
                                   0u, 0u);
        if (rc) break;
        hipLaunchKernelGGL(k_inst_records, dim3(nblk(ni, 128)), dim3(128), 0, st, d, perm, ni, mb, mt,
                           mc, static_cast<float4*>(ctx->inst.ptr),
                           static_cast<uint32_t*>(ctx->inst_order.ptr));
        if ((e = hipGetLastError()) != hipSuccess) rc = lux_hip_check(ctx, e, "build_tlas launch");
    } while (0);
    if (rc == LUX_OK) rc = range_check(ctx, false, true);
    if (rc == LUX_OK) ctx->tlas_built = true;
    return rc;
}

// Refit path of lux_bvh_build: new instance boxes and records in the
// existing TLAS order, boxes refitted (no allocation, no host wait).
static int refit_tlas(lux_ctx* ctx) {
    hipStream_t st = ctx->stream;
    const uint32_t ni = ctx->n_inst;
    float4* lo = static_cast<float4*>(ctx->tlas_tmp.ptr);
    float4* hi = lo + ni;
    const uint32_t* mb = static_cast<const uint32_t*>(ctx->mesh_info.ptr);
    const uint32_t* mt = mb + ctx->n_meshes;
    const uint32_t* mc = mt + ctx->n_meshes;
    const InstDesc* d = static_cast<const InstDesc*>(ctx->inst_desc.ptr);
    const uint32_t* order = static_cast<const uint32_t*>(ctx->inst_order.ptr);
    hipLaunchKernelGGL(k_inst_boxes, dim3(nblk(ni, 128)), dim3(128), 0, st, d, ni,
                       static_cast<const float4*>(ctx->blas.ptr), mb, lo, hi);
    const size_t lds = 32ull * ctx->n_tlas + 32ull * ni;
    if (lds <= 60ull * 1024)
        hipLaunchKernelGGL(k_tlas_refit_lds, dim3(1), dim3(256), lds, st, static_cast<float4*>(ctx->tlas.ptr),
                           ctx->n_tlas, static_cast<const float4*>(lo), static_cast<const float4*>(hi), order, ni);
    else
        hipLaunchKernelGGL(k_tlas_refit, dim3(1), dim3(64), 0, st, static_cast<float4*>(ctx->tlas.ptr), ctx->n_tlas,
                           static_cast<const float4*>(lo), static_cast<const float4*>(hi), order, ni);
    // the records take the new transforms in the same order (order[k] stays)
    hipLaunchKernelGGL(k_inst_records, dim3(nblk(ni, 128)), dim3(128), 0, st, d, order, ni, mb, mt, mc,
                       static_cast<float4*>(ctx->inst.ptr), static_cast<uint32_t*>(ctx->tlas_tmp.ptr) + 8ull * ni);
    LUX_HIP(ctx, hipGetLastError());
    if (int rc = range_check(ctx, false, true)) return rc;
    ctx->tlas_built = true;
    return LUX_OK;
}

extern "C" int lux_bvh_build(lux_ctx* ctx, uint32_t mode) {
    if (!ctx) return LUX_ERR_INVALID;
    if ((mode & ~(LUX_BVH_FORCE_BLAS | LUX_BVH_REFIT | LUX_BVH_SAH | LUX_BVH_SAH_HOST)) != LUX_BVH_MORTON30 ||
        ((mode & LUX_BVH_SAH) && (mode & LUX_BVH_SAH_HOST)))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_bvh_build: unknown mode");
    const uint32_t builder = (mode & LUX_BVH_SAH) ? 1u : ((mode & LUX_BVH_SAH_HOST) ? 2u : 0u);
    // a change of BLAS builder rebuilds the BLASes
    if (ctx->blas_built && ctx->blas_builder != builder) ctx->blas_built = false;
    if (mode & LUX_BVH_FORCE_BLAS) ctx->blas_built = false;
    if (ctx->n_meshes == 0 && ctx->n_inst == 0)
        return lux_fail(ctx, LUX_ERR_STATE, "lux_bvh_build: no scene uploaded");
    if (ctx->prebuilt) return LUX_OK;  // installed built: nothing to build
    LUX_HIP(ctx, hipSetDevice(ctx->device));
    if (!ctx->blas_built) {
        if (builder == 2 && ctx->host_pos.size() != ctx->n_meshes)
            return lux_fail(ctx, LUX_ERR_STATE, "lux_bvh_build: SAH needs the scene's host positions");
        int rc;
        try {
            rc = build_blas(ctx, builder);
        } catch (...) {  // host containers of the build; nothing may leave a C entry point
            rc = lux_fail(ctx, LUX_ERR_OOM, "lux_bvh_build: host allocation failed");
        }
        if (rc) return rc;
        ctx->blas_builder = builder;
        ctx->tlas_built = false;
        ctx->tlas_topology = false;
    }
    if (!ctx->tlas_built) {
        // refit when only transforms changed since a TLAS build of this scene
        if ((mode & LUX_BVH_REFIT) && ctx->tlas_topology && ctx->n_inst > 1) return refit_tlas(ctx);
        if (int rc = build_tlas(ctx)) return rc;
        ctx->tlas_topology = true;
    }
    return LUX_OK;
}

extern "C" int lux_bvh_info(lux_ctx* ctx, uint32_t* blas_nodes, uint32_t* tlas_nodes,
                            uint32_t* triangles, uint32_t* instances) {
    if (!ctx) return LUX_ERR_INVALID;
    if (blas_nodes) *blas_nodes = ctx->n_blas;
    if (tlas_nodes) *tlas_nodes = ctx->n_tlas;
    if (triangles) *triangles = ctx->n_tri;
    if (instances) *instances = ctx->n_inst;
    return LUX_OK;
}

extern "C" int lux_bvh_download(lux_ctx* ctx, lux_bvh_node* blas, lux_bvh_node* tlas,
                                uint32_t* tri_order, uint32_t* inst_order) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->tlas_built) return lux_fail(ctx, LUX_ERR_STATE, "lux_bvh_download: BVH not built");
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (blas) LUX_HIP(ctx, hipMemcpy(blas, ctx->blas.ptr, 32ull * ctx->n_blas, hipMemcpyDeviceToHost));
    if (tlas && ctx->n_inst)
        LUX_HIP(ctx, hipMemcpy(tlas, ctx->tlas.ptr, 32ull * ctx->n_tlas, hipMemcpyDeviceToHost));
    if (tri_order)
        LUX_HIP(ctx, hipMemcpy(tri_order, ctx->tri_order.ptr, 4ull * ctx->n_tri, hipMemcpyDeviceToHost));
    if (inst_order && ctx->n_inst)
        LUX_HIP(ctx, hipMemcpy(inst_order, ctx->inst_order.ptr, 4ull * ctx->n_inst, hipMemcpyDeviceToHost));
    return LUX_OK;
}

extern "C" int lux_sort_pairs_device(lux_ctx* ctx, const uint32_t* keys_in, uint32_t n,
                                     uint32_t* keys_out, uint32_t* index_out) {
    if (!ctx) return LUX_ERR_INVALID;
    if (n == 0) return LUX_OK;
    if (!keys_in || !keys_out || !index_out)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_sort_pairs_device: null buffer");
    size_t hw = lux_sort_hist_words(n);
    lux_buffer tmp;
    if (hipMalloc(&tmp.ptr, 8ull * n + 4ull * hw + 64) != hipSuccess)
        return lux_fail(ctx, LUX_ERR_OOM, "lux_sort_pairs_device: allocation failed");
    uint32_t* tk = static_cast<uint32_t*>(tmp.ptr);
    uint32_t* tv = tk + n;
    uint32_t* hist = tv + n;
    int rc = lux_sort_pairs_impl(ctx, keys_in, n, keys_out, index_out, tk, tv, hist, 32);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp.ptr);
    return rc;
}

extern "C" int lux_scene_upload_built(lux_ctx* ctx, const lux_bvh_node* blas, uint32_t blas_count,
                                      const lux_bvh_node* tlas, uint32_t tlas_count,
                                      const float* triangles, uint32_t triangle_count,
                                      const lux_instance_record* instances,
                                      uint32_t instance_count) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!blas || !tlas || !triangles || !instances || !blas_count || !tlas_count ||
        !triangle_count || !instance_count)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_built: empty or null array");
    static_assert(sizeof(lux_instance_record) == 128, "instance record layout");
    static_assert(sizeof(lux_bvh_node) == 32, "node layout");
    // shading reads the hit triangle unclamped: every instance's range must lie inside the table
    for (uint32_t k = 0; k < instance_count; ++k) {
        uint32_t g[4];
        std::memcpy(g, reinterpret_cast<const uint8_t*>(instances + k) + 96, 16);
        if ((uint64_t)g[2] + g[3] > triangle_count || g[1] >= blas_count)
            return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_built: instance triangles or root out of range");
    }
    LUX_HIP(ctx, hipSetDevice(ctx->device));
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (int rc = lux_ensure(ctx, ctx->blas, 32ull * blas_count)) return rc;
    if (int rc = lux_ensure(ctx, ctx->tlas, 32ull * tlas_count)) return rc;
    if (int rc = lux_ensure(ctx, ctx->inst, 128ull * instance_count)) return rc;
    if (int rc = lux_ensure(ctx, ctx->tris, 48ull * triangle_count)) return rc;
    LUX_HIP(ctx, hipMemcpy(ctx->blas.ptr, blas, 32ull * blas_count, hipMemcpyHostToDevice));
    LUX_HIP(ctx, hipMemcpy(ctx->tlas.ptr, tlas, 32ull * tlas_count, hipMemcpyHostToDevice));
    LUX_HIP(ctx, hipMemcpy(ctx->inst.ptr, instances, 128ull * instance_count, hipMemcpyHostToDevice));
    // 12 floats per triangle: the device layout as given
    LUX_HIP(ctx, hipMemcpy(ctx->tris.ptr, triangles, 48ull * triangle_count, hipMemcpyHostToDevice));
    ctx->n_blas = blas_count;
    ctx->n_tlas = tlas_count;
    ctx->n_inst = instance_count;
    ctx->n_tri = triangle_count;
    ctx->n_meshes = 0;
    ctx->mesh_tri_count.clear();
    ctx->mesh_tri_base.clear();
    ctx->mesh_blas_base.clear();
    ctx->instances.clear();
    ctx->inst_material.clear();
    if (int rc = range_check(ctx, true, true)) return rc;
    ctx->blas_built = true;
    ctx->tlas_built = true;
    ctx->tlas_topology = false;
    ctx->prebuilt = true;
    return LUX_OK;
}

namespace {
// Triangle records from the reference's geometry words: triangle t's vertex
// indices at word 1 + t, vertex v's position at word vertex_base + stride v
// (clamped reads, as traversal.slang:51-62 clamps every word address).
__global__ void k_words_triangles(const uint4* words, uint32_t n_words, uint32_t n_tri, uint32_t vertex_base,
                                  uint32_t stride, float4* tri) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tri) return;
    const uint32_t last = n_words - 1u;
    uint32_t iw = 1u + t;
    uint4 idx = words[iw < last ? iw : last];
    uint32_t v[3] = {idx.x, idx.y, idx.z};
    float4* out = tri + 3ull * t;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        uint64_t w = (uint64_t)vertex_base + (uint64_t)stride * v[c];
        uint4 p = words[w < last ? w : last];
        out[c] = make_float4(__uint_as_float(p.x), __uint_as_float(p.y), __uint_as_float(p.z),
                             c == 0 ? __uint_as_float(t) : 0.0f);
    }
}

__global__ void k_zero_bytes(uint8_t* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0;
}
}  // namespace

// Packed-scene interchange: the reference's inline scene and geometry words
// (src/scene.rs:809-823 geometry_words, :1025-1038 node_words, :1071-1122
// pack_instances, :2102-2116 the prefix header) installed as a built scene.
extern "C" int lux_scene_upload_words(lux_ctx* ctx, const uint32_t* scene_words, uint32_t n_scene_words,
                                      const uint32_t* geometry_words, uint32_t n_geometry_words,
                                      const lux_material* materials, uint32_t material_count) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!scene_words || !geometry_words || n_scene_words < 4 || n_geometry_words < 2)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_words: empty or null words");
    if (material_count && !materials) return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_words: null materials");
    auto word = [&](const uint32_t* w, uint64_t i, int c) { return w[4ull * i + c]; };
    const uint32_t inst_base = word(scene_words, 0, 1), n_inst = word(scene_words, 0, 2),
                   stride_i = word(scene_words, 0, 3);
    if (stride_i != 9u) return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_words: instance stride is not 9 words");
    if (n_inst == 0 || inst_base < 3 || (uint64_t)inst_base + 9ull * n_inst > n_scene_words)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_words: instance block outside the scene words");
    const uint32_t n_nodes = (inst_base - 1u) / 2u;
    const uint32_t tri_base = word(geometry_words, 0, 0), vertex_base = word(geometry_words, 0, 1),
                   vstride = word(geometry_words, 0, 2), n_vert = word(geometry_words, 0, 3);
    if (tri_base != 1u || vstride == 0 || vertex_base < 1 ||
        (uint64_t)vertex_base + (uint64_t)vstride * n_vert > n_geometry_words)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_words: malformed geometry header");
    const uint64_t material_base = (uint64_t)inst_base + 9ull * n_inst;
    // instances: rows, flags, material, BVH root and triangle range
    std::vector<float> rec(32ull * n_inst);
    std::vector<uint32_t> inst_mat(n_inst);
    std::vector<std::pair<uint32_t, uint32_t>> ranges;
    std::vector<uint32_t> inst_mesh(n_inst);
    uint32_t n_tri = 0;
    for (uint32_t k = 0; k < n_inst; ++k) {
        const uint64_t b = (uint64_t)inst_base + 9ull * k;
        float* r = rec.data() + 32ull * k;
        for (int row = 0; row < 6; ++row)
            for (int c = 0; c < 4; ++c) {
                uint32_t u = word(scene_words, b + row, c);
                std::memcpy(&r[4 * row + c], &u, 4);
            }
        uint32_t flags = word(scene_words, b + 6, 3), mat_addr = word(scene_words, b + 7, 3);
        uint32_t root = word(scene_words, b + 8, 0), first = word(scene_words, b + 8, 1),
                 count = word(scene_words, b + 8, 2);
        if (root >= n_nodes) return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_words: instance root out of range");
        if ((uint64_t)first + count > (uint64_t)vertex_base - 1u)
            return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_words: instance triangles outside the geometry");
        n_tri = std::max(n_tri, first + count);
        uint64_t mi = mat_addr >= material_base ? (mat_addr - material_base) / 8u : 0u;
        inst_mat[k] = material_count ? (uint32_t)std::min<uint64_t>(mi, material_count - 1u) : 0u;
        uint32_t m = 0;
        while (m < ranges.size() && ranges[m] != std::make_pair(first, count)) ++m;
        if (m == ranges.size()) ranges.emplace_back(first, count);
        inst_mesh[k] = m;
        uint32_t w6[4] = {flags, root, first, count}, w7[4] = {k, m, inst_mat[k] + 1u, 0u};
        std::memcpy(&r[24], w6, 16);
        std::memcpy(&r[28], w7, 16);
    }
    if (n_tri == 0) return lux_fail(ctx, LUX_ERR_INVALID, "lux_scene_upload_words: no triangles");
    LUX_HIP(ctx, hipSetDevice(ctx->device));
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // nodes: one array serves as TLAS (root 0) and BLAS (absolute roots)
    if (int rc = lux_ensure(ctx, ctx->blas, 32ull * n_nodes)) return rc;
    if (int rc = lux_ensure(ctx, ctx->tlas, 32ull * n_nodes)) return rc;
    if (int rc = lux_ensure(ctx, ctx->inst, 128ull * n_inst)) return rc;
    if (int rc = lux_ensure(ctx, ctx->tris, 48ull * n_tri)) return rc;
    if (int rc = lux_ensure(ctx, ctx->materials, sizeof(lux_material) * (material_count ? material_count : 1))) return rc;
    LUX_HIP(ctx, hipMemcpy(ctx->blas.ptr, scene_words + 4, 32ull * n_nodes, hipMemcpyHostToDevice));
    LUX_HIP(ctx, hipMemcpy(ctx->tlas.ptr, scene_words + 4, 32ull * n_nodes, hipMemcpyHostToDevice));
    LUX_HIP(ctx, hipMemcpy(ctx->inst.ptr, rec.data(), 128ull * n_inst, hipMemcpyHostToDevice));
    if (material_count)
        LUX_HIP(ctx, hipMemcpy(ctx->materials.ptr, materials, sizeof(lux_material) * material_count,
                               hipMemcpyHostToDevice));
    // triangles: the geometry words go up once and are repacked on the device
    if (int rc = lux_ensure(ctx, ctx->scratch, 16ull * n_geometry_words)) return rc;
    LUX_HIP(ctx, hipMemcpy(ctx->scratch.ptr, geometry_words, 16ull * n_geometry_words, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_words_triangles, dim3((n_tri + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const uint4*>(ctx->scratch.ptr), n_geometry_words, n_tri, vertex_base, vstride,
                       static_cast<float4*>(ctx->tris.ptr));
    LUX_HIP(ctx, hipGetLastError());
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // host copies for the emitter table: one mesh per distinct triangle range
    ctx->host_pos.assign(ranges.size(), {});
    ctx->mesh_bounds.clear();
    ctx->mesh_emissive.clear();
    ctx->host_mat.assign(ranges.size(), {});
    for (size_t m = 0; m < ranges.size(); ++m) {
        std::vector<float>& hp = ctx->host_pos[m];
        hp.resize(9ull * ranges[m].second);
        for (uint32_t t = 0; t < ranges[m].second; ++t) {
            uint32_t idx[3] = {word(geometry_words, 1ull + ranges[m].first + t, 0),
                               word(geometry_words, 1ull + ranges[m].first + t, 1),
                               word(geometry_words, 1ull + ranges[m].first + t, 2)};
            for (int c = 0; c < 3; ++c) {
                uint64_t w = std::min<uint64_t>((uint64_t)vertex_base + (uint64_t)vstride * idx[c], n_geometry_words - 1u);
                for (int a = 0; a < 3; ++a) {
                    uint32_t u = word(geometry_words, w, a);
                    std::memcpy(&hp[9ull * t + 3 * c + a], &u, 4);
                }
            }
        }
        ctx->host_mat[m].assign(ranges[m].second, 0u);
    }
    ctx->host_materials.assign(materials, materials + material_count);
    ctx->instances.assign(n_inst, lux_instance_desc{});
    for (uint32_t k = 0; k < n_inst; ++k) {
        lux_instance_desc& d = ctx->instances[k];
        std::memcpy(d.local_from_world, rec.data() + 32ull * k, 48);
        std::memcpy(d.world_from_local, rec.data() + 32ull * k + 12, 48);
        d.mesh = inst_mesh[k];
        d.flags = word(scene_words, (uint64_t)inst_base + 9ull * k + 6, 3);
    }
    ctx->inst_material = inst_mat;
    ctx->n_blas = n_nodes;
    ctx->n_tlas = n_nodes;
    ctx->n_inst = n_inst;
    ctx->n_tri = n_tri;
    ctx->n_materials = material_count;
    ctx->n_meshes = 0;
    ctx->mesh_tri_count.clear();
    ctx->mesh_tri_base.clear();
    ctx->mesh_blas_base.clear();
    if (int rc = range_check(ctx, true, true)) return rc;
    ctx->blas_built = true;
    ctx->tlas_built = true;
    ctx->tlas_topology = false;
    ctx->prebuilt = true;
    ctx->gi_ready = false;
    ctx->gi_lights_dirty = true;
    return LUX_OK;
}

extern "C" int lux_device_alloc(lux_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return LUX_ERR_INVALID;
    *out = nullptr;
    LUX_HIP(ctx, hipSetDevice(ctx->device));
    LUX_HIP(ctx, hipMalloc(out, bytes ? bytes : 16));
    return LUX_OK;
}

extern "C" int lux_device_free(lux_ctx* ctx, void* ptr) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ptr) return LUX_OK;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipFree(ptr));
    return LUX_OK;
}

extern "C" int lux_copy_to_device(lux_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return LUX_ERR_INVALID;
    if (!bytes) return LUX_OK;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return LUX_OK;
}

extern "C" int lux_copy_to_host(lux_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return LUX_ERR_INVALID;
    if (!bytes) return LUX_OK;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return LUX_OK;
}

extern "C" int lux_device_zero(lux_ctx* ctx, void* ptr, size_t bytes) {
    if (!ctx || (bytes && !ptr)) return LUX_ERR_INVALID;
    if (!bytes) return LUX_OK;
    size_t blocks = (bytes + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero_bytes, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream,
                       static_cast<uint8_t*>(ptr), bytes);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_device_copy(lux_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return LUX_ERR_INVALID;
    if (!bytes) return LUX_OK;
    LUX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return LUX_OK;
}

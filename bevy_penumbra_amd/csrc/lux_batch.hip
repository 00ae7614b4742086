// lux_batch.hip -- the GI update of many scenes of uneven size at once.
//
// A batch references configured scene contexts (BVH, emitter table,
// interval / merged / probe arrays) through a device table of per-scene
// kernel arguments and runs one update of every scene as
//   k_b_zero      admission counters of every scene, batch counters
//   k_b_primary   ONE persistent launch: regrouped closest-hit walks of every
//                 (interval, ray) item of every scene -- each lane walks its
//                 own item's scene, lanes are refilled across scenes
//   k_b_setup     shading setup of every item (NEE slots, occlusion rays)
//   k_b_shadow    ONE persistent launch: occlusion walks of every scene
//   k_b_accum     accumulation of every traced interval
//   surrogate     per scene with a network: queries, forward, merged level S
//   k_b_merge     one launch per level over every scene's intervals of it
//   k_b_pack      every scene's level-0 probes
// Per-frame values (frame, seed) live in the device table, so the sequence
// is captured once as a hipGraph and replayed (lux_batch_capture/launch).
// The per-scene arithmetic is the single-scene code (lux_gi_device.h), so a
// batch update gives every scene exactly its own lux_gi_update result.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>
#include <cmath>
#include "lux_gi_device.h"

namespace {

constexpr uint32_t kMaxBatchLevels = 8;
// Scenes whose unbounded admission counts a walk workgroup tallies in LDS.
constexpr uint32_t kTallyScenes = 256;
typedef __attribute__((address_space(3))) uint32_t LdsU32;


struct BatchScene {
    GiArgs a;
    uint32_t R;             // rays per interval
    uint32_t trace_levels;  // levels traced (all, or those below the surrogate level)
};

struct BatchWave {
    const BatchScene* sc;
    uint32_t n_scenes;
    const uint64_t* item_base;      // [n_scenes + 1] (interval, ray) items
    const uint64_t* interval_base;  // [n_scenes + 1] traced intervals
    const uint64_t* probe_base;     // [n_scenes + 1] level-0 probes
    const uint64_t* pack_block_base;  // [n_scenes + 1] k_b_pack blocks (kPackBlockProbes probes each)
    const uint64_t* merge_base;     // [kMaxBatchLevels][n_scenes + 1] intervals merged per level
    uint32_t nee_max;
    uint32_t stride;  // items per NEE sample row of slots / vis (sample-major: smp * stride + item)
    uint4* prim;
    float4* shade;
    float4* slots;
    uint32_t* vis;
    float4* shadow;
    uint32_t* counters;  // [0] primary cursor, [1] occlusion rays, [2] occlusion cursor
    float4* slab;        // per setup wave: its occlusion rays (64 * nee_max)
    uint32_t* wave_cnt;  // per setup wave: its ray count, then scanned into offsets
    uint32_t regroup, batch;
    float4* dep_item;    // per item of a history scene: its NEE segments' dependency (null: no history)
};

// Largest s with base[s] <= x (base[0] = 0, base[n] = total; empty scenes
// share their successor's base and are never returned for an index).
__device__ __forceinline__ uint32_t find_scene(const uint64_t* base, uint32_t n, uint64_t x) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) / 2;
        if (base[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// admit_query for lanes of possibly different scenes: the active lanes are
// served in groups of one scene each (one atomic per group on that scene's
// counters), with admitTransportQuery's semantics per scene.  A scene with
// an unbounded budget admits every query and only counts it: its groups add
// to the workgroup's LDS tally instead (scenes below kTallyScenes), which
// tally_flush adds to the scene's counters once, with the same totals.  The
// global atomics would make every refill of a wave wait for their return.
__device__ bool admit_query_scene(const BatchWave& w, uint32_t s, LdsU32* tally) {
    const GiArgs& a = w.sc[s].a;
    const uint32_t lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    bool done = false, ok = true;
    while (true) {
        uint64_t pending = __ballot(!done);
        if (pending == 0) break;
        uint32_t leader = (uint32_t)(__ffsll((unsigned long long)pending) - 1);
        uint32_t s0 = __shfl(s, (int)leader, 64);
        bool mine = !done && s == s0;
        if (mine) {
            uint64_t group = __ballot(1);
            uint32_t count = (uint32_t)__popcll(group);
            uint32_t rank = (uint32_t)__popcll(group & below);
            if (a.admit_limit == 0u && s < kTallyScenes) {
                if (lane == leader)
                    __hip_atomic_fetch_add(&tally[s], count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                done = true;
                continue;
            }
            uint32_t base = 0;
            if (lane == leader) {
                atomicAdd(&a.counters[0], count);
                base = atomicAdd(&a.counters[1], count);
            }
            if (a.admit_limit != 0u) {
                base = __shfl(base, (int)leader, 64);
                ok = (uint64_t)base + rank < (uint64_t)a.admit_limit;
                uint64_t denied = __ballot(!ok);
                if (lane == leader && denied) atomicSub(&a.counters[1], (uint32_t)__popcll(denied));
            }
            done = true;
        }
    }
    return ok;
}

__device__ __forceinline__ void tally_clear(uint32_t* tally) {
    for (uint32_t k = threadIdx.x; k < kTallyScenes; k += kWg) tally[k] = 0u;
    __syncthreads();
}

// Called by every thread of the workgroup once its walks are done.
__device__ __forceinline__ void tally_flush(const BatchWave& w, const uint32_t* tally) {
    __syncthreads();
    const uint32_t n = w.n_scenes < kTallyScenes ? w.n_scenes : kTallyScenes;
    for (uint32_t k = threadIdx.x; k < n; k += kWg) {
        const uint32_t v = tally[k];
        if (v) {
            atomicAdd(&w.sc[k].a.counters[0], v);
            atomicAdd(&w.sc[k].a.counters[1], v);
        }
    }
}

struct BatchPrimary {
    const BatchWave* w;
    uint32_t n;
    LdsU32* tally;
    __device__ uint32_t count() const { return n; }
    template <class St>
    __device__ bool begin(uint32_t i, lux::Walk& wk, St& st) {
        uint32_t s = find_scene(w->item_base, w->n_scenes, i);
        const BatchScene& S = w->sc[s];
        uint64_t local = i - w->item_base[s];
        float3 o, d;
        float lim;
        bool term;
        item_ray(S.a, local / S.R, (uint32_t)(local % S.R), false, o, d, lim, term);
        if (!admit_query_scene(*w, s, tally)) {
            w->prim[i] = make_uint4(0u, lux::kEmpty, lux::kEmpty, lux::kErrUnsupported);
            return false;
        }
        wk.sp = &S.a.scene;
        lux::walk_begin(S.a.scene, wk, o, d, lim, lux::kWalkClosest, st);
        if (wk.phase != lux::kPhaseDone) return true;
        finish(i, wk);
        return false;
    }
    __device__ void finish(uint32_t i, const lux::Walk& wk) {
        w->prim[i] = make_uint4(__float_as_uint(wk.h.t), wk.h.tri, wk.h.inst, wk.h.err);
    }
};

template <int Depth>
__global__ __launch_bounds__(kWg, lux::walk_min_waves(Depth, kWg, 4 * kTallyScenes)) void k_b_primary(BatchWave w) {
    __shared__ uint2 lds[Depth][kWg];
    lux::LaneStack<Depth, kWg> stk;
    stk.bind(&lds[0][threadIdx.x]);
    lux::NoStats st;
    __shared__ uint32_t tally[kTallyScenes];
    tally_clear(tally);
    BatchPrimary src{&w, (uint32_t)w.item_base[w.n_scenes], (LdsU32*)tally};
    lux::regroup_walks_multi(src, &w.counters[0], w.regroup, w.batch, stk, st);
    tally_flush(w, tally);
}

__global__ __launch_bounds__(kWg) void k_b_setup(BatchWave w) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    const uint32_t gw = blockIdx.x * (kWg / 64u) + (threadIdx.x >> 6);
    const bool live = i < w.item_base[w.n_scenes];
    uint4 pr = live ? w.prim[i] : make_uint4(0u, lux::kEmpty, lux::kEmpty, lux::kErrUnsupported);
    float4 sh = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t running = 0;  // rays of this wave so far (the same in every lane of a ballot)
    if ((pr.w & lux::kErrUnsupported) == 0u && pr.y != lux::kEmpty) {
        uint32_t s = find_scene(w.item_base, w.n_scenes, i);
        const BatchScene& S = w.sc[s];
        uint64_t local = i - w.item_base[s];
        float3 o, d;
        float lim;
        bool term;
        item_ray(S.a, local / S.R, (uint32_t)(local % S.R), false, o, d, lim, term);
        lux::HitState h;
        h.t = __uint_as_float(pr.x);
        h.tri = pr.y;
        h.inst = pr.z;
        h.err = pr.w;
        h.b1 = h.b2 = 0.0f;
        float3 rad;
        NeeCtx c;
        bool nee = shade_begin(S.a, o, d, h, rad, c);
        sh = make_float4(rad.x, rad.y, rad.z, nee ? 1.0f : 0.0f);
        if (nee) {
            const uint32_t lane = __lane_id();
            const bool track = w.dep_item != nullptr && S.a.dep != nullptr;
            Dep nd = dep_empty();
            for (uint32_t smp = 0; smp < c.count; ++smp) {
                float3 origin = make_float3(0.0f, 0.0f, 0.0f), dir = origin, t3 = origin;
                float sl = 0.0f;
                bool q = nee_sample(S.a, c, smp, origin, dir, sl, t3);
                if (q && track) dep_ray(nd, S.a.scene, origin, dir, sl, false, 0.0f);
                uint32_t slot = smp * w.stride + i;
                w.slots[slot] = make_float4(t3.x, t3.y, t3.z, q ? 1.0f : 0.0f);
                uint64_t mask = __ballot(q);
                if (q) {
                    const uint64_t at = (uint64_t)gw * 64u * w.nee_max + running +
                                        (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
                    w.slab[2ull * at] = make_float4(origin.x, origin.y, origin.z, sl);
                    w.slab[2ull * at + 1] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(slot));
                }
                running += (uint32_t)__popcll(mask);
            }
            if (track) dep_store(w.dep_item + 2ull * i, nd);
        }
    }
    if (live) w.shade[i] = sh;
    const uint32_t total = wave_max_u32(running);
    if (__lane_id() == 0u) w.wave_cnt[gw] = total;
}

struct BatchShadow {
    const BatchWave* w;
    uint32_t n;
    LdsU32* tally;
    __device__ uint32_t count() const { return n; }
    template <class St>
    __device__ bool begin(uint32_t i, lux::Walk& wk, St& st) {
        float4 r0 = w->shadow[2ull * i], r1 = w->shadow[2ull * i + 1];
        uint32_t slot = __float_as_uint(r1.w);
        uint32_t s = find_scene(w->item_base, w->n_scenes, slot % w->stride);  // the slot's item
        if (!admit_query_scene(*w, s, tally)) {
            w->vis[slot] = 2u;
            return false;
        }
        const BatchScene& S = w->sc[s];
        wk.sp = &S.a.scene;
        lux::walk_begin(S.a.scene, wk, make_float3(r0.x, r0.y, r0.z), make_float3(r1.x, r1.y, r1.z), r0.w,
                        lux::kWalkOccluded, st);
        if (wk.phase != lux::kPhaseDone) return true;
        finish(i, wk);
        return false;
    }
    __device__ void finish(uint32_t i, const lux::Walk& wk) {
        uint32_t slot = __float_as_uint(w->shadow[2ull * i + 1].w);
        w->vis[slot] = (wk.h.err & lux::kHitOccluded) ? 1u : ((wk.h.err & lux::kErrUnsupported) ? 2u : 0u);
    }
};

template <int Depth>
__global__ __launch_bounds__(kWg, lux::walk_min_waves(Depth, kWg, 4 * kTallyScenes)) void k_b_shadow(BatchWave w) {
    __shared__ uint2 lds[Depth][kWg];
    lux::LaneStack<Depth, kWg> stk;
    stk.bind(&lds[0][threadIdx.x]);
    lux::NoStats st;
    __shared__ uint32_t tally[kTallyScenes];
    tally_clear(tally);
    BatchShadow src{&w, w.counters[1], (LdsU32*)tally};
    lux::regroup_walks_multi(src, &w.counters[2], w.regroup, w.batch, stk, st);
    tally_flush(w, tally);
}

__global__ __launch_bounds__(kWg) void k_b_accum(BatchWave w) {
    uint32_t j = blockIdx.x * kWg + threadIdx.x;
    if (j >= w.interval_base[w.n_scenes]) return;
    uint32_t s = find_scene(w.interval_base, w.n_scenes, j);
    const BatchScene& S = w.sc[s];
    const GiArgs& a = S.a;
    uint64_t g = j - w.interval_base[s];
    float3 o, d;
    float lim;
    bool terminal;
    item_ray(a, g, 0u, false, o, d, lim, terminal);
    uint32_t count = a.p.emissive_samples / 2u;
    if (count < 1u) count = 1u;
    float3 jsum = make_float3(0.0f, 0.0f, 0.0f);
    float bsum = 0.0f;
    uint32_t n = 0;
    const bool track = a.dep != nullptr && w.dep_item != nullptr;
    Dep dep = dep_empty();
    for (uint32_t r = 0; r < S.R; ++r) {
        uint32_t i = (uint32_t)(w.item_base[s] + g * S.R + r);
        uint4 pr = w.prim[i];
        if (pr.w & lux::kErrUnsupported) continue;
        float3 rad = make_float3(0.0f, 0.0f, 0.0f);
        float beta;
        bool nee_ran = false;
        if (pr.y != lux::kEmpty) {
            float4 sh = w.shade[i];
            float3 rr = make_float3(sh.x, sh.y, sh.z);
            if (sh.w != 0.0f) {
                nee_ran = true;
                float3 sum = make_float3(0.0f, 0.0f, 0.0f);
                bool ok = true;
                for (uint32_t smp = 0; smp < count; ++smp) {
                    uint32_t slot = smp * w.stride + i;
                    float4 s4 = w.slots[slot];
                    if (s4.w == 0.0f) continue;
                    uint32_t v = w.vis[slot];
                    if (v == 1u) continue;
                    if (v == 2u) {
                        ok = false;
                        break;
                    }
                    sum.x = sum.x + s4.x;
                    sum.y = sum.y + s4.y;
                    sum.z = sum.z + s4.z;
                }
                if (!ok) continue;
                float fc = (float)count;
                rr = make_float3(rr.x + sum.x / fc, rr.y + sum.y / fc, rr.z + sum.z / fc);
            }
            rad = shade_end(rr);
            beta = 0.0f;
        } else {
            beta = terminal ? 0.0f : 1.0f;
        }
        if (track) {
            // the ray's segment, then its NEE occlusion segments
            float3 ro, rdir;
            float rlim;
            bool rterm;
            item_ray(a, g, r, false, ro, rdir, rlim, rterm);
            Dep rd = dep_empty();
            dep_ray(rd, a.scene, ro, rdir, rlim, pr.y != lux::kEmpty, __uint_as_float(pr.x));
            if (nee_ran) dep_union(rd, dep_load(w.dep_item + 2ull * i));
            dep_union(dep, rd);
        }
        jsum = make_float3(jsum.x + rad.x, jsum.y + rad.y, jsum.z + rad.z);
        bsum = bsum + beta;
        n++;
    }
    float4 out;
    if (n == 0) {
        out = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    } else {
        float fn = (float)n;
        out = make_float4(jsum.x / fn, jsum.y / fn, jsum.z / fn, bsum / fn);
    }
    a.intervals[g] = out;
    if (a.samples) a.samples[g] = n;
    if (track) dep_store(a.dep + 2 * g, dep);
}

// Temporal history of every traced interval of the scenes that keep one.
__global__ __launch_bounds__(kWg) void k_b_history(BatchWave w) {
    uint32_t j = blockIdx.x * kWg + threadIdx.x;
    if (j >= w.interval_base[w.n_scenes]) return;
    uint32_t s = find_scene(w.interval_base, w.n_scenes, j);
    const GiArgs& a = w.sc[s].a;
    if (!a.history) return;
    uint64_t g = j - w.interval_base[s];
    uint32_t L = 0;
    while (L + 1 < a.p.levels && g >= a.offset[L + 1]) ++L;
    history_blend(a, g, L, g - a.offset[L]);
}

__global__ void k_b_history_done(const BatchScene* sc, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && sc[i].a.hist_state) {
        sc[i].a.hist_state[0] = 1u;
        sc[i].a.hist_state[1] = 0u;
        sc[i].a.hist_state[3] = 0u;
    }
}

__global__ __launch_bounds__(kWg) void k_b_merge(BatchWave w, uint32_t L) {
    const uint64_t* base = w.merge_base + (size_t)L * (w.n_scenes + 1);
    uint64_t x = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (x >= base[w.n_scenes]) return;
    uint32_t s = find_scene(base, w.n_scenes, x);
    merge_interval(w.sc[s].a, L, x - base[s]);
}

// Each block packs up to kPackBlockProbes level-0 probes of one scene with
// that scene's cosine table (k_gi_pack's layout).
__global__ __launch_bounds__(256) void k_b_pack(BatchWave w) {
    extern __shared__ float pack_lds[];
    const uint32_t s = find_scene(w.pack_block_base, w.n_scenes, blockIdx.x);
    const GiArgs& a = w.sc[s].a;
    const uint32_t D = a.p.base_phi * a.p.base_y;
    float* ctab = pack_lds;
    float* nom = ctab + 36u * D;
    float3* dirs = reinterpret_cast<float3*>(nom + 36u);
    float4* smp = reinterpret_cast<float4*>(pack_lds + pack_smp_offset(D));
    pack_table(a.p, ctab, nom, dirs);
    const uint32_t wv = threadIdx.x >> 6;
    const uint64_t first = (blockIdx.x - w.pack_block_base[s]) * kPackBlockProbes;
    const uint64_t probes = w.probe_base[s + 1] - w.probe_base[s];
    for (uint32_t j = 0; j < kPackBlockProbes / 4u; ++j) {
        const uint64_t probe = first + wv * (kPackBlockProbes / 4u) + j;
        if (probe < probes) pack_probe_table(a, (uint32_t)probe, ctab, nom, smp + wv * D);
    }
}

__global__ void k_b_zero(BatchScene* sc, uint32_t n, uint32_t* counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        sc[i].a.counters[0] = 0u;
        sc[i].a.counters[1] = 0u;
    }
    if (i < 4) counters[i] = 0u;
}

__global__ void k_b_set_frame(BatchScene* sc, uint32_t n, uint32_t frame, uint32_t seed) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        sc[i].a.p.frame = frame;
        sc[i].a.p.seed = seed;
    }
}

__global__ __launch_bounds__(kWg) void k_b_sur_queries(const BatchScene* sc, uint32_t s, uint32_t L,
                                                       uint64_t count, float* out) {
    uint64_t i = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (i >= count) return;
    level_query(sc[s].a.p, L, i, out + 9ull * i);
}

__global__ __launch_bounds__(kWg) void k_b_sur_apply(const float* Y, uint64_t count, float4* merged) {
    uint64_t i = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (i >= count) return;
    const float* y = Y + 16ull * i;
    merged[i] = make_float4(smax(y[0], 0.0f), smax(y[1], 0.0f), smax(y[2], 0.0f), 1.0f);
}

}  // namespace

// ======================================================================== host
struct lux_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string error;
    std::vector<lux_ctx*> scenes;
    std::vector<lux_surrogate*> sur;
    std::vector<uint32_t> sur_level;
    std::vector<float> sur_bound;  // per scene: the fallback's error bound (+inf: none)
    std::vector<lux_buffer> sur_fb;  // per scene: the batch's own fallback words
    std::vector<BatchScene> host_table;
    std::vector<uint64_t> item_base, interval_base, probe_base, merge_base, pack_block_base;
    uint32_t pack_dirs = 1;  // most level-0 directions of a scene (k_b_pack's LDS)
    uint32_t nee_max = 1, max_levels = 1;
    bool any_history = false;
    uint64_t query_max = 0;
    uint64_t setup_waves = 0;  // waves of the NEE setup launch (occlusion slabs)
    lux_ctx* owner = nullptr;  // allocations go through the first scene's context
    lux_buffer table, bases, prim, shade, slots, vis, shadow, counters, queries, sur_scratch, dep_item;
    uint32_t wpc = 24, depth = 12, regroup = 48, batch = 16;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool ready = false;
};

namespace {

int bfail(lux_batch* b, int code, const std::string& msg) {
    b->error = msg;
    return code;
}

#define LUX_BHIP(b, call)                                                            \
    do {                                                                             \
        hipError_t _e = (call);                                                      \
        if (_e != hipSuccess) return bfail((b), LUX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

void drop_graph(lux_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int ensure(lux_batch* b, lux_buffer& buf, size_t bytes) {
    if (int rc = lux_ensure(b->owner, buf, bytes)) return bfail(b, rc, lux_last_error(b->owner));
    return LUX_OK;
}

// Build the device table, the index spaces and the scratch for the current
// scenes and surrogate assignments.
int prepare(lux_batch* b) {
    drop_graph(b);
    b->ready = false;
    // scratch may be reallocated below: no queued update may still use it
    LUX_BHIP(b, hipStreamSynchronize(b->stream));
    const uint32_t n = (uint32_t)b->scenes.size();
    b->host_table.assign(n, BatchScene{});
    b->item_base.assign(n + 1, 0);
    b->interval_base.assign(n + 1, 0);
    b->probe_base.assign(n + 1, 0);
    b->merge_base.assign((size_t)kMaxBatchLevels * (n + 1), 0);
    b->pack_block_base.assign(n + 1, 0);
    b->pack_dirs = 1;
    b->nee_max = 1;
    b->max_levels = 1;
    b->query_max = 0;
    b->any_history = false;
    for (uint32_t s = 0; s < n; ++s) {
        lux_ctx* c = b->scenes[s];
        if (int rc = lux_gi_check_internal(c)) return bfail(b, rc, std::string("scene ") + std::to_string(s) + ": " + lux_last_error(c));
        BatchScene& bs = b->host_table[s];
        bs.a = gi_args(c);
        if (bs.a.history) b->any_history = true;
        bs.R = c->gi.rays_per_interval ? c->gi.rays_per_interval : 1u;
        uint32_t levels = c->gi.levels;
        uint32_t S = b->sur[s] ? b->sur_level[s] : levels;
        bs.trace_levels = S;
        uint32_t nee = c->gi.emissive_samples / 2u > 1u ? c->gi.emissive_samples / 2u : 1u;
        b->nee_max = std::max(b->nee_max, nee);
        b->max_levels = std::max(b->max_levels, levels);
        uint64_t traced = c->gi_offset[S];
        b->interval_base[s + 1] = b->interval_base[s] + traced;
        b->item_base[s + 1] = b->item_base[s] + traced * bs.R;
        b->probe_base[s + 1] = b->probe_base[s] + c->gi_probes0;
        b->pack_block_base[s + 1] = b->pack_block_base[s] + (c->gi_probes0 + kPackBlockProbes - 1) / kPackBlockProbes;
        b->pack_dirs = std::max(b->pack_dirs, c->gi.base_phi * c->gi.base_y);
        for (uint32_t L = 0; L < kMaxBatchLevels; ++L) {
            uint64_t cnt = (L < S) ? c->gi_offset[L + 1] - c->gi_offset[L] : 0;
            size_t row = (size_t)L * (n + 1);
            b->merge_base[row + s + 1] = b->merge_base[row + s] + cnt;
        }
        if (b->sur[s]) {
            // the scene may have been reconfigured since the surrogate was set
            if (S < 1 || S >= levels)
                return bfail(b, LUX_ERR_STATE, std::string("scene ") + std::to_string(s) + ": surrogate level " +
                                                   std::to_string(S) + " is no longer in 1..levels-1 (" +
                                                   std::to_string(levels) + " levels); set the surrogate again");
            uint64_t q = c->gi_offset[S + 1] - c->gi_offset[S];
            b->query_max = std::max(b->query_max, q);
            if (q >= (1ull << 31)) return bfail(b, LUX_ERR_INVALID, "surrogate level too large");
            // the fallback's probe list, count and errors follow the scene's current grid
            if (b->sur_bound[s] < INFINITY)
                if (int rc = lux_ensure(c, b->sur_fb[s], 4ull * lux_sur_fallback_words(c, S)))
                    return bfail(b, rc, lux_last_error(c));
        }
    }
    const uint64_t items = b->item_base[n];
    if (items * b->nee_max >= (1ull << 31) || b->probe_base[n] >= (1ull << 31))
        return bfail(b, LUX_ERR_INVALID, "batch too large for one launch (2^31 NEE slots or probes)");
    if (int rc = ensure(b, b->table, sizeof(BatchScene) * n)) return rc;
    size_t nb = 4 * (n + 1) + (size_t)kMaxBatchLevels * (n + 1);
    if (int rc = ensure(b, b->bases, 8 * nb)) return rc;
    if (int rc = ensure(b, b->prim, 16 * items + 16)) return rc;
    if (int rc = ensure(b, b->shade, 16 * items + 16)) return rc;
    if (int rc = ensure(b, b->slots, 16 * items * b->nee_max + 16)) return rc;
    if (int rc = ensure(b, b->vis, 4 * items * b->nee_max + 16)) return rc;
    if (b->any_history)
        if (int rc = ensure(b, b->dep_item, 32 * items + 32)) return rc;
    // occlusion rays: per setup wave a slab of 64 * nee_max, then the packed list
    b->setup_waves = ((items + kWg - 1) / kWg) * (kWg / 64u);
    if (b->setup_waves * 64u * b->nee_max >= (1ull << 31))
        return bfail(b, LUX_ERR_INVALID, "batch too large for one launch (occlusion list)");
    if (int rc = ensure(b, b->shadow, 64ull * b->setup_waves * 64u * b->nee_max + 32)) return rc;
    if (int rc = ensure(b, b->counters, 4ull * (4u + b->setup_waves + 1u + lux_scan_sums_words((uint32_t)b->setup_waves + 1u))))
        return rc;
    if (int rc = ensure(b, b->queries, 36 * b->query_max + 36)) return rc;
    if (int rc = ensure(b, b->sur_scratch, lux_surrogate_scratch_bytes((uint32_t)b->query_max) + 256)) return rc;
    std::vector<uint64_t> all;
    all.insert(all.end(), b->item_base.begin(), b->item_base.end());
    all.insert(all.end(), b->interval_base.begin(), b->interval_base.end());
    all.insert(all.end(), b->probe_base.begin(), b->probe_base.end());
    all.insert(all.end(), b->merge_base.begin(), b->merge_base.end());
    all.insert(all.end(), b->pack_block_base.begin(), b->pack_block_base.end());
    LUX_BHIP(b, hipMemcpy(b->table.ptr, b->host_table.data(), sizeof(BatchScene) * n, hipMemcpyHostToDevice));
    LUX_BHIP(b, hipMemcpy(b->bases.ptr, all.data(), 8 * all.size(), hipMemcpyHostToDevice));
    b->ready = true;
    return LUX_OK;
}

BatchWave wave_of(lux_batch* b) {
    const uint32_t n = (uint32_t)b->scenes.size();
    BatchWave w;
    w.sc = static_cast<const BatchScene*>(b->table.ptr);
    w.n_scenes = n;
    const uint64_t* bases = static_cast<const uint64_t*>(b->bases.ptr);
    w.item_base = bases;
    w.interval_base = bases + (n + 1);
    w.probe_base = bases + 2 * (n + 1);
    w.merge_base = bases + 3 * (n + 1);
    w.pack_block_base = bases + 3 * (n + 1) + (size_t)kMaxBatchLevels * (n + 1);
    w.nee_max = b->nee_max;
    w.stride = (uint32_t)b->item_base[n];
    w.prim = static_cast<uint4*>(b->prim.ptr);
    w.shade = static_cast<float4*>(b->shade.ptr);
    w.slots = static_cast<float4*>(b->slots.ptr);
    w.vis = static_cast<uint32_t*>(b->vis.ptr);
    w.shadow = static_cast<float4*>(b->shadow.ptr);
    w.counters = static_cast<uint32_t*>(b->counters.ptr);
    w.wave_cnt = w.counters + 4;
    w.slab = w.shadow + 2ull * b->setup_waves * 64u * b->nee_max;
    w.regroup = b->regroup ? b->regroup : 1u;
    w.batch = b->batch ? b->batch : 8u;
    w.dep_item = b->any_history ? static_cast<float4*>(b->dep_item.ptr) : nullptr;
    return w;
}

// The launch sequence of one update of every scene (also what is captured).
int enqueue(lux_batch* b) {
    const uint32_t n = (uint32_t)b->scenes.size();
    hipStream_t st = b->stream;
    BatchWave w = wave_of(b);
    const uint64_t items = b->item_base[n], intervals = b->interval_base[n], probes = b->probe_base[n];
    int cu = b->owner->cu_count;
    uint32_t pblocks = (uint32_t)cu * (b->wpc ? b->wpc : 24u) / (kWg / 64);
    hipLaunchKernelGGL(k_b_zero, dim3((n + 4 + 255) / 256), dim3(256), 0, st, static_cast<BatchScene*>(b->table.ptr),
                       n, w.counters);
    if (items) {
        uint32_t need = (uint32_t)((items + kWg - 1) / kWg);
        uint32_t pb = pblocks < need ? pblocks : need;
        switch (b->depth) {
        case 4: hipLaunchKernelGGL(k_b_primary<4>, dim3(pb), dim3(kWg), 0, st, w); break;
        case 8: hipLaunchKernelGGL(k_b_primary<8>, dim3(pb), dim3(kWg), 0, st, w); break;
        case 12: hipLaunchKernelGGL(k_b_primary<12>, dim3(pb), dim3(kWg), 0, st, w); break;
        default: hipLaunchKernelGGL(k_b_primary<16>, dim3(pb), dim3(kWg), 0, st, w); break;
        }
        hipLaunchKernelGGL(k_b_setup, dim3(need), dim3(kWg), 0, st, w);
        const uint32_t nw = need * (kWg / 64u);
        hipLaunchKernelGGL(k_fill_u32_one, dim3(1), dim3(1), 0, st, w.wave_cnt + nw, 0u);
        uint32_t* scan_sums = w.wave_cnt + b->setup_waves + 1u;
        if (int rc = lux_scan_exclusive_on(b->owner, st, w.wave_cnt, nw + 1u, scan_sums, &w.counters[1]))
            return bfail(b, rc, lux_last_error(b->owner));
        hipLaunchKernelGGL(k_shadow_compact, dim3((nw + 3u) / 4u), dim3(256), 0, st, w.slab, w.wave_cnt, nw,
                           64u * w.nee_max, w.shadow);
        switch (b->depth) {
        case 4: hipLaunchKernelGGL(k_b_shadow<4>, dim3(pblocks), dim3(kWg), 0, st, w); break;
        case 8: hipLaunchKernelGGL(k_b_shadow<8>, dim3(pblocks), dim3(kWg), 0, st, w); break;
        case 12: hipLaunchKernelGGL(k_b_shadow<12>, dim3(pblocks), dim3(kWg), 0, st, w); break;
        default: hipLaunchKernelGGL(k_b_shadow<16>, dim3(pblocks), dim3(kWg), 0, st, w); break;
        }
        hipLaunchKernelGGL(k_b_accum, dim3((uint32_t)((intervals + kWg - 1) / kWg)), dim3(kWg), 0, st, w);
        if (b->any_history) {
            hipLaunchKernelGGL(k_b_history, dim3((uint32_t)((intervals + kWg - 1) / kWg)), dim3(kWg), 0, st, w);
            hipLaunchKernelGGL(k_b_history_done, dim3((n + 255) / 256), dim3(256), 0, st,
                               static_cast<const BatchScene*>(b->table.ptr), n);
            for (uint32_t s = 0; s < n; ++s)
                if (b->scenes[s]->gi_hist_on) b->scenes[s]->gi_hist_valid = true;
        }
    }
    for (uint32_t s = 0; s < n; ++s) {
        if (!b->sur[s]) continue;
        lux_ctx* c = b->scenes[s];
        uint32_t S = b->sur_level[s];
        uint64_t q = c->gi_offset[S + 1] - c->gi_offset[S];
        uint32_t qb = (uint32_t)((q + kWg - 1) / kWg);
        float* qbuf = static_cast<float*>(b->queries.ptr);
        hipLaunchKernelGGL(k_b_sur_queries, dim3(qb), dim3(kWg), 0, st, w.sc, s, S, q, qbuf);
        const float* Y = nullptr;
        if (int rc = lux_surrogate_forward_into(b->sur[s], st, qbuf, (uint32_t)q, b->sur_scratch.ptr, &Y))
            return bfail(b, rc, lux_last_error(c));
        hipLaunchKernelGGL(k_b_sur_apply, dim3(qb), dim3(kWg), 0, st, Y, q,
                           static_cast<float4*>(c->gi_merged.ptr) + c->gi_offset[S]);
        if (b->sur_bound[s] < INFINITY)
            if (int rc = lux_sur_fallback_on(c, st, S, Y, b->sur_bound[s], static_cast<uint32_t*>(b->sur_fb[s].ptr)))
                return bfail(b, rc, lux_last_error(c));
    }
    for (int L = (int)b->max_levels - 1; L >= 0; --L) {
        uint64_t cnt = b->merge_base[(size_t)L * (n + 1) + n];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(k_b_merge, dim3((uint32_t)((cnt + kWg - 1) / kWg)), dim3(kWg), 0, st, w, (uint32_t)L);
    }
    if (probes)
        hipLaunchKernelGGL(k_b_pack, dim3((uint32_t)b->pack_block_base[n]), dim3(256), pack_lds_bytes(b->pack_dirs), st,
                           w);
    LUX_BHIP(b, hipGetLastError());
    return LUX_OK;
}

}  // namespace

extern "C" int lux_batch_create(lux_ctx* const* scenes, uint32_t n, lux_batch** out) {
    if (!out) return LUX_ERR_INVALID;
    *out = nullptr;
    if (!scenes || n == 0) return LUX_ERR_INVALID;
    for (uint32_t s = 0; s < n; ++s)
        if (!scenes[s] || scenes[s]->device != scenes[0]->device) return LUX_ERR_INVALID;
    lux_batch* b = new lux_batch();
    b->device = scenes[0]->device;
    b->owner = scenes[0];
    b->scenes.assign(scenes, scenes + n);
    b->sur.assign(n, nullptr);
    b->sur_level.assign(n, 0);
    b->sur_bound.assign(n, INFINITY);
    b->sur_fb.resize(n);
    if (hipSetDevice(b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        lux_batch_destroy(b);
        return LUX_ERR_HIP;
    }
    for (lux_ctx* c : b->scenes) (void)hipStreamSynchronize(c->stream);
    if (int rc = prepare(b)) {
        // keep the message on the first scene's context for the caller
        lux_fail(scenes[0], rc, b->error);
        lux_batch_destroy(b);
        return rc;
    }
    *out = b;
    return LUX_OK;
}

extern "C" void lux_batch_destroy(lux_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    lux_buffer* bufs[] = {&b->table, &b->bases, &b->prim, &b->shade, &b->slots,
                          &b->vis, &b->shadow, &b->counters, &b->queries, &b->sur_scratch, &b->dep_item};
    for (lux_buffer* x : bufs) lux_release(*x);
    for (lux_buffer& x : b->sur_fb) lux_release(x);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" const char* lux_batch_last_error(const lux_batch* b) {
    return b ? b->error.c_str() : "null batch";
}

extern "C" int lux_batch_set_surrogate(lux_batch* b, uint32_t scene, lux_surrogate* s, uint32_t level) {
    if (!b) return LUX_ERR_INVALID;
    if (scene >= b->scenes.size()) return bfail(b, LUX_ERR_INVALID, "lux_batch_set_surrogate: scene index");
    lux_ctx* c = b->scenes[scene];
    if (s) {
        if (lux_surrogate_context(s) != c)
            return bfail(b, LUX_ERR_INVALID, "lux_batch_set_surrogate: surrogate of another context");
        if (level < 1 || level >= c->gi.levels)
            return bfail(b, LUX_ERR_INVALID, "lux_batch_set_surrogate: level must be in 1..levels-1");
    }
    drop_graph(b);  // a captured graph would keep the old prediction set
    b->sur[scene] = s;
    b->sur_bound[scene] = INFINITY;  // a new network or level starts without fallback
    b->sur_level[scene] = s ? level : 0u;
    return prepare(b);
}

extern "C" int lux_batch_set_surrogate_bound(lux_batch* b, uint32_t scene, float error_bound) {
    if (!b) return LUX_ERR_INVALID;
    if (scene >= b->scenes.size()) return bfail(b, LUX_ERR_INVALID, "lux_batch_set_surrogate_bound: scene index");
    if (!b->sur[scene])
        return bfail(b, LUX_ERR_STATE, "lux_batch_set_surrogate_bound: the scene has no surrogate");
    if (std::isnan(error_bound)) return bfail(b, LUX_ERR_INVALID, "lux_batch_set_surrogate_bound: NaN bound");
    // a captured graph holds the bound and the words' address: drop it
    drop_graph(b);
    LUX_BHIP(b, hipStreamSynchronize(b->stream));
    if (error_bound < INFINITY) {
        lux_ctx* c = b->scenes[scene];
        if (int rc = lux_ensure(c, b->sur_fb[scene], 4ull * lux_sur_fallback_words(c, b->sur_level[scene])))
            return bfail(b, rc, lux_last_error(c));
    }
    b->sur_bound[scene] = error_bound;
    return LUX_OK;
}

extern "C" int lux_batch_refresh(lux_batch* b) {
    if (!b) return LUX_ERR_INVALID;
    for (lux_ctx* c : b->scenes) (void)hipStreamSynchronize(c->stream);
    return prepare(b);
}

extern "C" int lux_batch_set_frame(lux_batch* b, uint32_t frame, uint32_t seed) {
    if (!b || !b->ready) return LUX_ERR_INVALID;
    uint32_t n = (uint32_t)b->scenes.size();
    for (BatchScene& bs : b->host_table) {
        bs.a.p.frame = frame;
        bs.a.p.seed = seed;
    }
    // the scene contexts follow, so a later per-scene call sees the same frame
    for (lux_ctx* c : b->scenes) {
        c->gi.frame = frame;
        c->gi.seed = seed;
    }
    hipLaunchKernelGGL(k_b_set_frame, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                       static_cast<BatchScene*>(b->table.ptr), n, frame, seed);
    LUX_BHIP(b, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_batch_update(lux_batch* b, float* ms) {
    if (!b) return LUX_ERR_INVALID;
    if (!b->ready) return bfail(b, LUX_ERR_STATE, "lux_batch_update: batch not prepared");
    LUX_BHIP(b, hipEventRecord(b->ev0, b->stream));
    if (int rc = enqueue(b)) return rc;
    for (lux_ctx* c : b->scenes) c->gi_dirty_n = 0;
    LUX_BHIP(b, hipEventRecord(b->ev1, b->stream));
    if (ms) {
        LUX_BHIP(b, hipEventSynchronize(b->ev1));
        LUX_BHIP(b, hipEventElapsedTime(ms, b->ev0, b->ev1));
    }
    return LUX_OK;
}

extern "C" int lux_batch_capture(lux_batch* b) {
    if (!b) return LUX_ERR_INVALID;
    if (!b->ready) return bfail(b, LUX_ERR_STATE, "lux_batch_capture: batch not prepared");
    drop_graph(b);
    LUX_BHIP(b, hipStreamSynchronize(b->stream));
    LUX_BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, LUX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    LUX_BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return LUX_OK;
}

extern "C" int lux_batch_launch(lux_batch* b, float* ms) {
    if (!b) return LUX_ERR_INVALID;
    if (!b->exec) return bfail(b, LUX_ERR_STATE, "lux_batch_launch: no captured graph (lux_batch_capture)");
    LUX_BHIP(b, hipEventRecord(b->ev0, b->stream));
    LUX_BHIP(b, hipGraphLaunch(b->exec, b->stream));
    for (lux_ctx* c : b->scenes) c->gi_dirty_n = 0;
    LUX_BHIP(b, hipEventRecord(b->ev1, b->stream));
    if (ms) {
        LUX_BHIP(b, hipEventSynchronize(b->ev1));
        LUX_BHIP(b, hipEventElapsedTime(ms, b->ev0, b->ev1));
    }
    return LUX_OK;
}

extern "C" int lux_batch_synchronize(lux_batch* b) {
    if (!b) return LUX_ERR_INVALID;
    LUX_BHIP(b, hipStreamSynchronize(b->stream));
    return LUX_OK;
}

extern "C" int lux_batch_sizes(lux_batch* b, uint64_t* items, uint64_t* intervals, uint64_t* probes,
                               uint32_t* scenes) {
    if (!b) return LUX_ERR_INVALID;
    uint32_t n = (uint32_t)b->scenes.size();
    if (items) *items = b->item_base[n];
    if (intervals) *intervals = b->interval_base[n];
    if (probes) *probes = b->probe_base[n];
    if (scenes) *scenes = n;
    return LUX_OK;
}

extern "C" int lux_batch_set_config(lux_batch* b, const lux_trace_config* cfg) {
    if (!b || !cfg) return LUX_ERR_INVALID;
    if (cfg->waves_per_cu > 32 || cfg->regroup_threshold > 64 || cfg->steps_per_refill > 256 ||
        (cfg->lds_stack_depth != 0 && cfg->lds_stack_depth != 4 && cfg->lds_stack_depth != 8 &&
         cfg->lds_stack_depth != 12 && cfg->lds_stack_depth != 16))
        return bfail(b, LUX_ERR_INVALID, "lux_batch_set_config: value out of range");
    b->wpc = cfg->waves_per_cu;
    b->depth = cfg->lds_stack_depth;
    b->regroup = cfg->regroup_threshold;
    b->batch = cfg->steps_per_refill;
    drop_graph(b);
    return LUX_OK;
}

extern "C" int lux_batch_get_config(lux_batch* b, lux_trace_config* cfg) {
    if (!b || !cfg) return LUX_ERR_INVALID;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->waves_per_cu = b->wpc;
    cfg->lds_stack_depth = b->depth;
    cfg->regroup_threshold = b->regroup;
    cfg->steps_per_refill = b->batch;
    return LUX_OK;
}

// Autotuner of the batched launches: coordinate descent over the same knobs,
// each candidate timed over whole batch updates (median of `repeats` after a
// warm-up).  The timed updates skip the temporal history pass (it does not
// depend on the knobs), so every scene's history, its validity and its
// pending dirty regions are as they were before tuning.  Drops a captured
// graph (capture again afterwards).
extern "C" int lux_batch_autotune(lux_batch* b, uint32_t repeats, lux_trace_config* best, float* best_ms,
                                  uint32_t* evaluations) {
    if (!b) return LUX_ERR_INVALID;
    if (!b->ready) return bfail(b, LUX_ERR_STATE, "lux_batch_autotune: batch not prepared");
    if (repeats == 0) repeats = 2;
    if (repeats > 15) repeats = 15;
    drop_graph(b);
    const uint32_t saved[4] = {b->wpc, b->depth, b->regroup, b->batch};
    // history state restored on every return path
    struct HistoryHold {
        lux_batch* b;
        bool any;
        std::vector<uint32_t> dirty;
        std::vector<char> valid;
        explicit HistoryHold(lux_batch* bb) : b(bb), any(bb->any_history) {
            for (lux_ctx* c : b->scenes) {
                dirty.push_back(c->gi_dirty_n);
                valid.push_back(c->gi_hist_valid ? 1 : 0);
            }
            b->any_history = false;
        }
        ~HistoryHold() {
            b->any_history = any;
            for (size_t s = 0; s < b->scenes.size(); ++s) {
                b->scenes[s]->gi_dirty_n = dirty[s];
                b->scenes[s]->gi_hist_valid = valid[s] != 0;
            }
        }
    } hold(b);
    auto measure = [&](const uint32_t c[4], float& ms) -> int {
        b->wpc = c[0];
        b->depth = c[1];
        b->regroup = c[2];
        b->batch = c[3];
        float t[16];
        for (uint32_t r = 0; r <= repeats; ++r) {
            float e = 0.0f;
            if (int rc = lux_batch_update(b, &e)) return rc;
            if (r > 0) t[r - 1] = e;
        }
        std::sort(t, t + repeats);
        ms = t[repeats / 2];
        return LUX_OK;
    };
    static const uint32_t cand_wpc[] = {16, 20, 24, 32};
    static const uint32_t cand_depth[] = {4, 8, 12, 16};
    static const uint32_t cand_regroup[] = {4, 8, 16, 32, 48};
    static const uint32_t cand_batch[] = {4, 8, 16};
    const uint32_t* cands[4] = {cand_wpc, cand_depth, cand_regroup, cand_batch};
    const uint32_t ncand[4] = {4, 4, 5, 3};
    uint32_t cur[4] = {24, 12, 48, 16};
    float cur_ms = 0.0f;
    uint32_t evals = 1;
    int rc = measure(cur, cur_ms);
    for (int pass = 0; pass < 2 && rc == LUX_OK; ++pass)
        for (int k = 0; k < 4 && rc == LUX_OK; ++k)
            for (uint32_t j = 0; j < ncand[k] && rc == LUX_OK; ++j) {
                if (cands[k][j] == cur[k]) continue;
                uint32_t c[4] = {cur[0], cur[1], cur[2], cur[3]};
                c[k] = cands[k][j];
                float ms = 0.0f;
                rc = measure(c, ms);
                ++evals;
                // a candidate must beat the incumbent by more than the timing noise
                if (rc == LUX_OK && ms < cur_ms * kTuneMargin) {
                    cur_ms = ms;
                    cur[k] = c[k];
                }
            }
    const uint32_t* fin = rc == LUX_OK ? cur : saved;
    b->wpc = fin[0];
    b->depth = fin[1];
    b->regroup = fin[2];
    b->batch = fin[3];
    if (rc != LUX_OK) return rc;
    if (best) {
        std::memset(best, 0, sizeof *best);
        best->waves_per_cu = cur[0];
        best->lds_stack_depth = cur[1];
        best->regroup_threshold = cur[2];
        best->steps_per_refill = cur[3];
    }
    if (best_ms) *best_ms = cur_ms;
    if (evaluations) *evaluations = evals;
    return LUX_OK;
}

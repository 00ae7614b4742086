// lux_receivers.hip -- receiver-driven probe selection and the receiver
// trace on gfx950: the reference's sparse Split update (seed / rank /
// parents / prefix / generateDirections / trace / deposit / reduce,
// shaders/split.slang:567-1476) with its hash-keyed intervals addressed in
// the dense grids of a LOD pyramid (grid k holds LOD k).
//
// One update, on the first LOD context's stream:
//   k_rx_seed      one thread per receiver: eligibility, LOD pair and the
//                  overlap's LOD phase, fair-window selection, the keys of
//                  2 variants x levels descriptors
//   k_rx_insert    one thread per descriptor: open-addressing probe directory
//                  per level (CAS on the owner word); canonical = the lowest
//                  descriptor index of the key (atomicMin)
//   ordinals       stable radix sort of the (variant, receiver) ray markers by
//                  level-0 directory slot: ordinal = rank in the run, level-0
//                  ray count = run length
//   k_rx_parents   levels 1.. in order: ray count from the eight children
//   roots          root counts scattered by canonical index, exclusive scan
//   k_rx_prefix    levels top-1 .. 0: parent offset + earlier siblings' counts
//   k_rx_generate  sequence = ordinal + probe offset; tickets by a scan
//   k_rx_trace     one thread per (ticket, variant): R2-jittered direction,
//                  closest hit, shading with emissive NEE, deposits per level
//                  up to the terminating one into dense slots
//   reduce         stable radix sort of the deposits by slot, per run the
//                  mean in ticket order
// then each LOD grid: history blend (when on), merge, pack.
//
// The reference leaves three orders to its atomics (canonical descriptor,
// ordinals, root ranges and tickets); here each is the order of a
// sequential pass over ascending descriptor / receiver index, so the
// results are a deterministic function of the input, equal to the oracle's.
#include <algorithm>
#include <cstring>
#include "lux_gi_device.h"

namespace {

constexpr uint32_t kRxEmpty = 0xffffffffu;
constexpr uint32_t kRxValid = 1u, kRxRay = 2u;
constexpr uint32_t kRxLodMask = 31u << 18;
constexpr uint32_t kRxMaxLods = 8;
constexpr uint32_t kScanBlock = 1024;

// counters: [0] eligible receivers, [1] tickets, [2] rays, [3] unsupported,
// [4] deposits kept, [5] deposits outside, [6] root total, [7] overflow
__device__ __forceinline__ void wave_count(uint32_t* c, bool pred) {
    uint64_t m = __ballot(pred);
    if (m && __lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1)) atomicAdd(c, (uint32_t)__popcll(m));
}

// ---------------- candidate permutation (split.slang:567-663) ----------------
__device__ __forceinline__ uint32_t perm(uint32_t v, uint32_t mask, uint32_t seed) {
    v = (v ^ seed) & mask;
    v = (v ^ (v >> 16)) & mask;
    v = (v * 0x7feb352du) & mask;
    v = (v ^ (v >> 15)) & mask;
    v = (v * 0x846ca68bu) & mask;
    v = (v ^ (v >> 16)) & mask;
    return v;
}

__device__ uint32_t candidate_rank(uint32_t sample, uint32_t w, uint32_t h, uint32_t count) {
    if (count <= 1u) return 0u;
    uint32_t mask = 0;
    while (mask < count - 1u) mask = (mask << 1) | 1u;
    uint32_t seed = ((w * 0x9e3779b9u) ^ (h * 0x85ebca6bu) ^ count) & mask;
    uint32_t r = perm(sample, mask, seed);
    while (r >= count) r = perm(r, mask, seed);
    return r;
}

struct RxSeed {
    uint32_t count, width, height, budget, cursor, quota, frame, flags;
    uint32_t levels, has_view;
    float spacing;
    lux_lod_view view;
};

__device__ bool candidate_selected(const RxSeed& s, uint32_t sample) {
    if (s.flags & 1u) return false;
    if (s.count == 0u || s.budget == 0u || s.budget >= s.count) return true;
    uint32_t start = (s.cursor & 0x03ffffffu) % s.count;
    uint32_t rank = candidate_rank(sample, s.width, s.height, s.count);
    if (s.quota == 0u) return false;
    return rank >= start ? rank - start < s.quota : s.count - start + rank < s.quota;
}

// Eligible receiver (finite position and normal, |n|^2 > 0.01), unit normal.
__device__ __forceinline__ bool rx_surface(const float* r, float3& pos, float3& n) {
    float x = r[3], y = r[4], z = r[5];
    float nn = (x * x + y * y) + z * z;
    if (!(lux::isfinite_bits(r[0]) && lux::isfinite_bits(r[1]) && lux::isfinite_bits(r[2]))) return false;
    if (!(lux::isfinite_bits(x) && lux::isfinite_bits(y) && lux::isfinite_bits(z)) || !(nn > 0.01f)) return false;
    float nl = sqrtf(nn);
    pos = make_float3(r[0], r[1], r[2]);
    n = make_float3(x / nl, y / nl, z / nl);
    return true;
}

// cacheKey (common.slang:418-423), direction 0.
__device__ __forceinline__ int4 rx_cell(float3 p, float base, uint32_t level, uint32_t lod) {
    uint32_t l = lod < 24u ? lod : 24u;
    float s = base * pow2i((int)(level + l));
    return make_int4((int32_t)floorf(p.x / s), (int32_t)floorf(p.y / s), (int32_t)floorf(p.z / s),
                     (int32_t)((level << 12) | (l << 18)));
}

__global__ __launch_bounds__(kWg) void k_rx_seed(RxSeed s, const float* recv, int4* keys, uint32_t* flags,
                                                 uint32_t* lods, uint32_t* counters) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    bool ok = false;
    if (i < s.count) {
        float3 pos, n;
        ok = rx_surface(recv + 6ull * i, pos, n);
        uint32_t low = 0, high = 0;
        float blend = 0.0f;
        float pp[3] = {pos.x, pos.y, pos.z};
        if (ok && s.has_view) lod_select(s.view, pp, low, high, blend);
        uint32_t sel_lod = 0;
        if (high != low) {
            uint32_t bits = transport_hash(i ^ (s.frame * 0x9e3779b9u));
            float phase = (float)(bits >> 8) * (1.0f / 16777216.0f);
            sel_lod = phase < blend ? 1u : 0u;
        }
        const bool ray = ok && candidate_selected(s, i);
        for (uint32_t v = 0; v < 2; ++v) {
            lods[(uint64_t)v * s.count + i] = v ? high : low;
            for (uint32_t L = 0; L < s.levels; ++L) {
                uint64_t d = ((uint64_t)v * s.count + i) * s.levels + L;
                if (!ok || (v == 1 && high == low)) {
                    keys[d] = make_int4(0, 0, 0, 0);
                    flags[d] = 0;
                    continue;
                }
                keys[d] = rx_cell(pos, s.spacing, L, v ? high : low);
                flags[d] = kRxValid | ((L == 0 && ray && v == sel_lod) ? kRxRay : 0u);
            }
        }
    }
    wave_count(&counters[0], ok);
}

// ---------------- probe directory ----------------
__device__ __forceinline__ uint32_t dir_bucket(int4 k, uint32_t cap) {
    uint32_t w = (uint32_t)k.w & (kRxLodMask | (3u << 12));
    uint32_t h = (uint32_t)k.x * 0x8da6b343u ^ (uint32_t)k.y * 0xd8163841u ^ (uint32_t)k.z * 0xcb1ab31fu ^
                 w * 0x165667b1u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    return h & (cap - 1u);
}

__device__ __forceinline__ bool same_probe(int4 a, int4 b) {
    const uint32_t m = kRxLodMask | (0xfu << 12);
    return a.x == b.x && a.y == b.y && a.z == b.z && (((uint32_t)a.w ^ (uint32_t)b.w) & m) == 0u;
}

// Slot of a key in one level's table, or empty.
__device__ uint32_t dir_find(const int4* keys, const uint32_t* owner, uint32_t cap, int4 k) {
    uint32_t b = dir_bucket(k, cap);
    for (uint32_t a = 0; a < cap; ++a) {
        uint32_t o = owner[b];
        if (o == kRxEmpty) return kRxEmpty;
        if (same_probe(keys[o], k)) return b;
        b = (b + 1u) & (cap - 1u);
    }
    return kRxEmpty;
}

// Directory words: owner[levels][cap], canon[levels][cap], count[levels][cap],
// offset[levels][cap].
struct RxDir {
    uint32_t* owner;
    uint32_t* canon;
    uint32_t* count;
    uint32_t* offset;
    uint32_t cap, levels;
};

__global__ __launch_bounds__(kWg) void k_rx_dir_clear(RxDir D) {
    const uint64_t n = (uint64_t)D.levels * D.cap;
    for (uint64_t i = (uint64_t)blockIdx.x * kWg + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kWg) {
        D.owner[i] = kRxEmpty;
        D.canon[i] = kRxEmpty;
        D.count[i] = 0u;
        D.offset[i] = 0u;
    }
}

__global__ __launch_bounds__(kWg) void k_rx_insert(RxDir D, const int4* keys, const uint32_t* flags, uint64_t nd,
                                                   uint32_t* slot_out, uint32_t* counters) {
    const uint64_t d = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (d >= nd) return;
    if (!(flags[d] & kRxValid)) {
        slot_out[d] = kRxEmpty;
        return;
    }
    const uint32_t L = (uint32_t)(d % D.levels);
    const int4 k = keys[d];
    uint32_t* ow = D.owner + (uint64_t)L * D.cap;
    uint32_t b = dir_bucket(k, D.cap);
    for (uint32_t a = 0; a < D.cap; ++a) {
        // keys are all published by the seed launch; the CAS publishes an index
        uint32_t cur = atomicCAS(&ow[b], kRxEmpty, (uint32_t)d);
        if (cur == kRxEmpty || same_probe(keys[cur], k)) {
            atomicMin(&D.canon[(uint64_t)L * D.cap + b], (uint32_t)d);
            slot_out[d] = b;
            return;
        }
        b = (b + 1u) & (D.cap - 1u);
    }
    slot_out[d] = kRxEmpty;
    atomicAdd(&counters[7], 1u);
}

// Sort keys of the (variant, receiver) ray markers: level-0 slot, or empty.
__global__ __launch_bounds__(kWg) void k_rx_ray_keys(const uint32_t* flags, const uint32_t* slot, uint32_t m,
                                                     uint32_t levels, uint32_t* out) {
    const uint32_t j = blockIdx.x * kWg + threadIdx.x;
    if (j >= m) return;
    const uint64_t d0 = (uint64_t)j * levels;
    out[j] = (flags[d0] & kRxRay) && slot[d0] != kRxEmpty ? slot[d0] : kRxEmpty;
}

__global__ __launch_bounds__(kWg) void k_rx_run_start(const uint32_t* sk, uint32_t m, uint32_t* start) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= m) return;
    const uint32_t k = sk[i];
    if (k != kRxEmpty && (i == 0 || sk[i - 1] != k)) start[k] = i;
}

__global__ __launch_bounds__(kWg) void k_rx_ordinals(const uint32_t* sk, const uint32_t* sv, uint32_t m,
                                                     const uint32_t* start, uint32_t* count0, uint32_t* ord) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= m) return;
    const uint32_t k = sk[i];
    if (k == kRxEmpty) return;
    const uint32_t s = start[k];
    ord[sv[i]] = i - s;
    if (i + 1 == m || sk[i + 1] != k) count0[k] = i - s + 1u;
}

// parents (split.slang:898-922): level L's ray count from its eight children.
__global__ __launch_bounds__(kWg) void k_rx_parents(RxDir D, const int4* keys, uint32_t L) {
    const uint32_t b = blockIdx.x * kWg + threadIdx.x;
    if (b >= D.cap) return;
    const uint32_t o = D.owner[(uint64_t)L * D.cap + b];
    if (o == kRxEmpty) return;
    const int4 k = keys[o];
    const uint32_t lodbits = (uint32_t)k.w & kRxLodMask;
    const uint32_t* ow = D.owner + (uint64_t)(L - 1u) * D.cap;
    const uint32_t* cn = D.count + (uint64_t)(L - 1u) * D.cap;
    uint32_t sum = 0;
    for (uint32_t c = 0; c < 8u; ++c) {
        int4 ck = make_int4(k.x * 2 + (int32_t)(c & 1u), k.y * 2 + (int32_t)((c >> 1) & 1u),
                            k.z * 2 + (int32_t)((c >> 2) & 1u), (int32_t)(((L - 1u) << 12) | lodbits));
        uint32_t cs = dir_find(keys, ow, D.cap, ck);
        if (cs != kRxEmpty) sum += cn[cs];
    }
    D.count[(uint64_t)L * D.cap + b] = sum;
}

__global__ __launch_bounds__(kWg) void k_rx_root_scatter(RxDir D, uint32_t* root_arr) {
    const uint32_t b = blockIdx.x * kWg + threadIdx.x;
    if (b >= D.cap) return;
    const uint64_t top = (uint64_t)(D.levels - 1u) * D.cap + b;
    if (D.owner[top] == kRxEmpty) return;
    root_arr[D.canon[top] / D.levels] = D.count[top];
}

__global__ __launch_bounds__(kWg) void k_rx_root_offsets(RxDir D, const uint32_t* root_scan) {
    const uint32_t b = blockIdx.x * kWg + threadIdx.x;
    if (b >= D.cap) return;
    const uint64_t top = (uint64_t)(D.levels - 1u) * D.cap + b;
    if (D.owner[top] == kRxEmpty) return;
    D.offset[top] = root_scan[D.canon[top] / D.levels];
}

// prefix (split.slang:924-961) below the roots.
__global__ __launch_bounds__(kWg) void k_rx_prefix(RxDir D, const int4* keys, uint32_t L) {
    const uint32_t b = blockIdx.x * kWg + threadIdx.x;
    if (b >= D.cap) return;
    const uint32_t o = D.owner[(uint64_t)L * D.cap + b];
    if (o == kRxEmpty) return;
    const int4 k = keys[o];
    const uint32_t lodbits = (uint32_t)k.w & kRxLodMask;
    const int4 parent = make_int4(k.x >> 1, k.y >> 1, k.z >> 1, (int32_t)(((L + 1u) << 12) | lodbits));
    const uint32_t ps = dir_find(keys, D.owner + (uint64_t)(L + 1u) * D.cap, D.cap, parent);
    uint32_t off = ps != kRxEmpty ? D.offset[(uint64_t)(L + 1u) * D.cap + ps] : 0u;
    const uint32_t own = (uint32_t)((k.x - parent.x * 2) + (k.y - parent.y * 2) * 2 + (k.z - parent.z * 2) * 4);
    const uint32_t* ow = D.owner + (uint64_t)L * D.cap;
    const uint32_t* cn = D.count + (uint64_t)L * D.cap;
    for (uint32_t sib = 0; sib < own; ++sib) {
        int4 sk = make_int4(parent.x * 2 + (int32_t)(sib & 1u), parent.y * 2 + (int32_t)((sib >> 1) & 1u),
                            parent.z * 2 + (int32_t)((sib >> 2) & 1u), (int32_t)((L << 12) | lodbits));
        uint32_t ss = dir_find(keys, ow, D.cap, sk);
        if (ss != kRxEmpty) off += cn[ss];
    }
    D.offset[(uint64_t)L * D.cap + b] = off;
}

// generateDirections (split.slang:963-983): sequences and ticket flags.
__global__ __launch_bounds__(kWg) void k_rx_generate(const uint32_t* flags, const uint32_t* slot, const uint32_t* ord,
                                                     const uint32_t* offset0, uint32_t count, uint32_t levels,
                                                     uint32_t* seq, uint32_t* sel) {
    const uint32_t s = blockIdx.x * kWg + threadIdx.x;
    if (s >= count) return;
    uint32_t any = 0;
    for (uint32_t v = 0; v < 2; ++v) {
        const uint64_t j = (uint64_t)v * count + s;
        const uint64_t d0 = j * levels;
        uint32_t q = kRxEmpty;
        if ((flags[d0] & kRxRay) && slot[d0] != kRxEmpty) {
            q = ord[j] + offset0[slot[d0]];
            any = 1u;
        }
        seq[j] = q;
    }
    sel[s] = any;
}

__global__ __launch_bounds__(kWg) void k_rx_compact(const uint32_t* seq, const uint32_t* pos, uint32_t count,
                                                    uint32_t* tickets) {
    const uint32_t s = blockIdx.x * kWg + threadIdx.x;
    if (s >= count) return;
    if (seq[s] != kRxEmpty || seq[(uint64_t)count + s] != kRxEmpty) tickets[pos[s]] = s;
}

// ---------------- exclusive scan (multi-block) ----------------
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kScanBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63u) wsum[wave] = x;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < kScanBlock / 64 ? wsum[lane] : 0u;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            uint32_t y = __shfl_up(w, off, 64);
            if (lane >= (uint32_t)off) w += y;
        }
        if (lane < kScanBlock / 64) wsum[lane] = w;
    }
    __syncthreads();
    const uint32_t before = wave ? wsum[wave - 1] : 0u;
    *total = wsum[kScanBlock / 64 - 1];
    return before + x - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_blocks(uint32_t* data, uint32_t n, uint32_t* sums) {
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    uint32_t total;
    uint32_t e = block_scan_excl(i < n ? data[i] : 0u, &total);
    if (i < n) data[i] = e;
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Scan of the block sums in one block (looping); the grand total goes to *out.
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(uint32_t* sums, uint32_t nb, uint32_t* out) {
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += kScanBlock) {
        const uint32_t i = base + threadIdx.x;
        uint32_t total;
        const uint32_t c = carry;
        uint32_t e = block_scan_excl(i < nb ? sums[i] : 0u, &total);
        if (i < nb) sums[i] = c + e;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(uint32_t* data, uint32_t n, const uint32_t* sums) {
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    if (i < n) data[i] += sums[blockIdx.x];
}

// ---------------- receiver trace ----------------
// atan of |t| <= tan(pi/8): Taylor series to t^47 in f64.
__device__ __forceinline__ double atan_small(double t) {
    double t2 = t * t;
    double s = 1.0 / 47.0;
#pragma unroll
    for (int k = 22; k >= 0; --k) s = s * t2 + ((k & 1) ? -1.0 : 1.0) / (double)(2 * k + 1);
    return s * t;
}

// atan2 from the fixed series (deterministic, IEEE signed-zero conventions).
__device__ __forceinline__ double atan2_fixed(double y, double x) {
    double ax = fabs(x), ay = fabs(y);
    double mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    double a = mx > 0.0 ? mn / mx : 0.0;
    double r = a > 0.41421356237309503 ? 0.7853981633974483 + atan_small((a - 1.0) / (a + 1.0)) : atan_small(a);
    if (ay > ax) r = 1.5707963267948966 - r;
    if (signbit(x)) r = 3.141592653589793 - r;
    if (signbit(y)) r = -r;
    return r;
}

// directionIndex (common.slang:384-394) over nphi x ny bins.
__device__ __forceinline__ uint32_t direction_index(float3 d, uint32_t nphi, uint32_t ny) {
    float u = (float)(atan2_fixed((double)d.z, (double)d.x) * 0.15915494309189535);
    u = fracf(u + 1.0f);
    float v = d.y * 0.5f + 0.5f;
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    uint32_t x = (uint32_t)(u * (float)nphi), y = (uint32_t)(v * (float)ny);
    if (x > nphi - 1u) x = nphi - 1u;
    if (y > ny - 1u) y = ny - 1u;
    return x + y * nphi;
}

struct RxGrid {
    float origin[3];
    float spacing;
    uint32_t dims[3];
    uint32_t base_phi, base_y;
    uint32_t base;  // first global slot of this grid
    uint64_t offset[9];
};

struct RxGrids {
    uint32_t count;
    RxGrid g[kRxMaxLods];
};

__global__ __launch_bounds__(kWg) void k_rx_trace(GiArgs a, RxGrids grids, const float* recv, uint32_t count,
                                                  const uint32_t* tickets, uint32_t n_items, const uint32_t* seq,
                                                  const uint32_t* lods, uint32_t frame, uint32_t* dep_key,
                                                  float4* dep_val, uint32_t* counters) {
    __shared__ uint2 lds[kLdsDepth][kWg];
    Stack stk;
    stk.col = &lds[0][threadIdx.x];
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    const lux_gi_params& p = a.p;
    const uint32_t levels = p.levels;
    bool traced = false, unsupported = false;
    uint32_t kept = 0, outside = 0;
    if (i < n_items) {
        uint32_t* kout = dep_key + (uint64_t)i * levels;
        float4* vout = dep_val + (uint64_t)i * levels;
        for (uint32_t L = 0; L < levels; ++L) kout[L] = kRxEmpty;
        const uint32_t smp = tickets[i >> 1];
        const uint64_t j = (uint64_t)(i & 1u) * count + smp;
        const uint32_t sq = seq[j];
        float3 pos, n;
        if (sq != kRxEmpty && rx_surface(recv + 6ull * smp, pos, n)) {
            traced = true;
            const uint32_t lod = lods[j];
            const float2 jit = r2(transport_hash(frame ^ 0x5a17c9e3u));
            const float2 q = r2(sq + 1u);
            const float ux = fracf((0.5f + q.x) + jit.x), uy = fracf((0.5f + q.y) + jit.y);
            float c, s;
            cos_sin_2pi(ux, c, s);
            const float y = uy * 2.0f - 1.0f;
            const float qq = 1.0f - y * y;
            const float r = sqrtf(qq > 0.0f ? qq : 0.0f);
            float3 d = make_float3(r * c, y, r * s);
            if (dot3(d, n) < 0.0f) d = make_float3(-d.x, -d.y, -d.z);
            const float3 o = ray_offset(pos, n, p.normal_bias);
            lux::HitState h;
            lux::NoStats ns;
            lux::trace_closest(a.scene, o, d, 1e20f, h, stk, ns, false);
            float3 rad = make_float3(0.0f, 0.0f, 0.0f);
            const bool hit = h.tri != lux::kEmpty;
            if ((h.err & lux::kErrUnsupported) || (hit && !shade(a, o, d, h, stk, rad))) {
                unsupported = true;
            } else {
                const RxGrid* g = lod < grids.count ? &grids.g[lod] : nullptr;
                const float lod_scale = pow2i((int)(lod < 24u ? lod : 24u));
                for (uint32_t L = 0; L < levels; ++L) {
                    const float start = L == 0 ? 0.0f : interval_end(p, L - 1u) * lod_scale;
                    const float end = interval_end(p, L) * lod_scale;
                    const bool terminal = p.terminal_tail != 0u && L + 1u == levels;
                    const bool term = terminal || (hit && h.t > start && h.t <= end);
                    const int4 key = rx_cell(pos, p.spacing, L, lod);
                    bool in = false;
                    if (g) {
                        const float sl = g->spacing * pow2i((int)L);
                        const uint32_t dx = level_dim(g->dims[0], L), dy = level_dim(g->dims[1], L),
                                       dz = level_dim(g->dims[2], L);
                        const int64_t ix = (int64_t)key.x - (int64_t)floorf(g->origin[0] / sl);
                        const int64_t iy = (int64_t)key.y - (int64_t)floorf(g->origin[1] / sl);
                        const int64_t iz = (int64_t)key.z - (int64_t)floorf(g->origin[2] / sl);
                        in = ix >= 0 && iy >= 0 && iz >= 0 && ix < dx && iy < dy && iz < dz;
                        if (in) {
                            const uint32_t nphi = g->base_phi << L, ny = g->base_y << L;
                            const uint64_t probe = (uint64_t)ix + (uint64_t)dx * ((uint64_t)iy + (uint64_t)dy * (uint64_t)iz);
                            kout[L] = g->base + (uint32_t)(g->offset[L] + probe * nphi * ny + direction_index(d, nphi, ny));
                            vout[L] = term ? make_float4(rad.x, rad.y, rad.z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
                        }
                    }
                    if (in) ++kept;
                    else ++outside;
                    if (term) break;
                }
            }
        }
    }
    wave_count(&counters[2], traced);
    wave_count(&counters[3], unsupported);
    if (kept) atomicAdd(&counters[4], kept);
    if (outside) atomicAdd(&counters[5], outside);
}

struct RxOut {
    uint32_t count;
    uint32_t base[kRxMaxLods + 1];
    float4* iv[kRxMaxLods];
    uint32_t* sm[kRxMaxLods];
};

__global__ __launch_bounds__(kWg) void k_rx_clear(float4* iv, uint32_t* sm, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * kWg + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kWg) {
        iv[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        sm[i] = 0u;
    }
}

// reduce (split.slang:1386-1476): per slot, the mean of its deposits in
// ticket order.
__global__ __launch_bounds__(kWg) void k_rx_reduce(const uint32_t* sk, const uint32_t* sv, const float4* val,
                                                   uint32_t n, RxOut out) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = sk[i];
    if (k == kRxEmpty || (i > 0 && sk[i - 1] == k)) return;
    float jx = 0.0f, jy = 0.0f, jz = 0.0f, b = 0.0f;
    uint32_t c = 0;
    for (uint32_t t = i; t < n && sk[t] == k; ++t) {
        const float4 v = val[sv[t]];
        jx = jx + v.x;
        jy = jy + v.y;
        jz = jz + v.z;
        b = b + v.w;
        ++c;
    }
    uint32_t g = 0;
    while (g + 1 < out.count && k >= out.base[g + 1]) ++g;
    const uint32_t local = k - out.base[g];
    const float fc = (float)c;
    out.iv[g][local] = make_float4(jx / fc, jy / fc, jz / fc, b / fc);
    out.sm[g][local] = c;
}

// Descriptor view for downloads: canonical, count and offset of each
// descriptor's probe.
__global__ __launch_bounds__(kWg) void k_rx_describe(RxDir D, const uint32_t* slot, uint64_t nd, uint32_t* canon,
                                                     uint32_t* cnt, uint32_t* off) {
    const uint64_t d = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (d >= nd) return;
    const uint32_t b = slot[d];
    const uint64_t L = d % D.levels;
    canon[d] = b == kRxEmpty ? kRxEmpty : D.canon[L * D.cap + b];
    cnt[d] = b == kRxEmpty ? 0u : D.count[L * D.cap + b];
    off[d] = b == kRxEmpty ? kRxEmpty : D.offset[L * D.cap + b];
}

uint32_t directory_capacity(uint32_t count) {
    uint32_t v = (count * 4u > 1u ? count * 4u : 1u) - 1u;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1u;
}

uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kWg - 1) / kWg); }

RxDir dir_of(lux_ctx* c) {
    RxDir D;
    uint32_t* w = static_cast<uint32_t*>(c->rx_dir.ptr);
    const uint64_t lc = (uint64_t)c->rx_levels * c->rx_cap;
    D.owner = w;
    D.canon = w + lc;
    D.count = w + 2 * lc;
    D.offset = w + 3 * lc;
    D.cap = c->rx_cap;
    D.levels = c->rx_levels;
    return D;
}

// Exclusive scan of n words in place; the total is written to *total (device).
int scan_exclusive(lux_ctx* c, uint32_t* data, uint32_t n, uint32_t* total) {
    const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (int rc = lux_ensure(c, c->rx_scan, 4ull * (nb ? nb : 1))) return rc;
    uint32_t* sums = static_cast<uint32_t*>(c->rx_scan.ptr);
    if (nb == 0) {
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, c->stream, sums, 0u, total);
    } else {
        hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanBlock), 0, c->stream, data, n, sums);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, c->stream, sums, nb, total);
        hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanBlock), 0, c->stream, data, n, sums);
    }
    LUX_HIP(c, hipGetLastError());
    return LUX_OK;
}

// Stable sort of n u32 keys (values = input positions) in the sort scratch:
// keys_in at words [0, n), results at [n, 2n) keys and [2n, 3n) values.
int sort_keys(lux_ctx* c, uint32_t n, uint32_t** keys_in, uint32_t** keys_out, uint32_t** vals_out) {
    const size_t hw = lux_sort_hist_words(n ? n : 1);
    if (int rc = lux_ensure(c, c->rx_sort, 4ull * (5ull * (n ? n : 1) + hw))) return rc;
    uint32_t* w = static_cast<uint32_t*>(c->rx_sort.ptr);
    *keys_in = w;
    *keys_out = w + n;
    *vals_out = w + 2ull * n;
    return LUX_OK;
}

int run_sort(lux_ctx* c, uint32_t n) {
    uint32_t* w = static_cast<uint32_t*>(c->rx_sort.ptr);
    return lux_sort_pairs_impl(c, w, n, w + n, w + 2ull * n, w + 3ull * n, w + 4ull * n, w + 5ull * n);
}

int check_pyramid(lux_ctx* const* lods, uint32_t lod_count) {
    lux_ctx* c0 = lods[0];
    const lux_gi_params& p0 = c0->gi;
    for (uint32_t k = 0; k < lod_count; ++k) {
        lux_ctx* c = lods[k];
        if (!c || c->device != c0->device)
            return lux_fail(c0, LUX_ERR_INVALID, "lux_gi_update_receivers: LOD grid of another device");
        if (int rc = lux_gi_check_internal(c)) return lux_fail(c0, rc, c->error);
        const lux_gi_params& p = c->gi;
        if (p.levels != p0.levels || p.base_phi != p0.base_phi || p.base_y != p0.base_y ||
            p.interval_mode != p0.interval_mode || p.terminal_tail != p0.terminal_tail)
            return lux_fail(c0, LUX_ERR_INVALID,
                            "lux_gi_update_receivers: LOD grids differ in levels, directions or intervals");
        const float sk = p0.spacing * (float)(1u << k);
        if (p.spacing != sk || p.interval_length != p0.interval_length * (float)(1u << k))
            return lux_fail(c0, LUX_ERR_INVALID,
                            "lux_gi_update_receivers: grid k needs spacing and interval length x 2^k");
        const float top = p.spacing * (float)(1u << (p.levels - 1u));
        for (int q = 0; q < 3; ++q) {
            float cells = p.origin[q] / top;
            if (!(cells == std::floor(cells)) || std::fabs(cells) > 1e7f)
                return lux_fail(c0, LUX_ERR_INVALID,
                                "lux_gi_update_receivers: grid origin not on the coarsest level's cells");
        }
    }
    return LUX_OK;
}

}  // namespace

extern "C" int lux_gi_update_receivers(lux_ctx* const* lods, uint32_t lod_count, const lux_lod_view* view,
                                       const float* receivers, const lux_receiver_params* rp,
                                       lux_receiver_stats* stats) {
    if (!lods || lod_count == 0 || !lods[0]) return LUX_ERR_INVALID;
    lux_ctx* c = lods[0];
    if (lod_count > kRxMaxLods) return lux_fail(c, LUX_ERR_INVALID, "lux_gi_update_receivers: at most 8 LOD grids");
    if (!rp) return lux_fail(c, LUX_ERR_INVALID, "lux_gi_update_receivers: null parameters");
    const uint64_t count64 = (uint64_t)rp->width * rp->height;
    if (count64 == 0) return lux_fail(c, LUX_ERR_INVALID, "lux_gi_update_receivers: empty receiver grid");
    if (!receivers) return lux_fail(c, LUX_ERR_INVALID, "lux_gi_update_receivers: null receivers");
    if (int rc = check_pyramid(lods, lod_count)) return rc;
    const uint32_t levels = c->gi.levels;
    if (2ull * count64 * levels >= 0xffffffffull || count64 > (1ull << 28))
        return lux_fail(c, LUX_ERR_INVALID, "lux_gi_update_receivers: too many receivers");
    const uint32_t count = (uint32_t)count64;
    uint64_t slots = 0;
    for (uint32_t k = 0; k < lod_count; ++k) slots += lods[k]->gi_total;
    if (slots >= 0xffffffffull)
        return lux_fail(c, LUX_ERR_INVALID, "lux_gi_update_receivers: LOD grids hold more than 2^32 intervals");
    LUX_HIP(c, hipSetDevice(c->device));
    // the other grids' previous work runs on their own streams
    for (uint32_t k = 1; k < lod_count; ++k) LUX_HIP(c, hipStreamSynchronize(lods[k]->stream));
    for (uint32_t k = 0; k < lod_count; ++k) lods[k]->gi.frame = rp->frame;
    const uint64_t nd = 2ull * count * levels;
    const uint32_t m = 2u * count;
    const uint32_t cap = directory_capacity(count);
    if (int rc = lux_ensure(c, c->rx_keys, 16ull * nd)) return rc;
    if (int rc = lux_ensure(c, c->rx_flags, 4ull * nd)) return rc;
    if (int rc = lux_ensure(c, c->rx_slot, 4ull * nd)) return rc;
    if (int rc = lux_ensure(c, c->rx_lod, 4ull * m)) return rc;
    if (int rc = lux_ensure(c, c->rx_ord, 4ull * m)) return rc;
    if (int rc = lux_ensure(c, c->rx_seq, 4ull * m)) return rc;
    if (int rc = lux_ensure(c, c->rx_dir, 16ull * levels * cap)) return rc;
    if (int rc = lux_ensure(c, c->rx_tickets, 8ull * count)) return rc;
    if (int rc = lux_ensure(c, c->rx_counters, 64)) return rc;
    for (uint32_t k = 0; k < lod_count; ++k)
        if (int rc = lux_ensure(lods[k], lods[k]->gi_samples, 4ull * lods[k]->gi_total)) return rc;
    c->rx_count = count;
    c->rx_levels = levels;
    c->rx_cap = cap;
    c->rx_tickets_n = 0;
    hipStream_t st = c->stream;
    LUX_HIP(c, hipEventRecord(c->ev0, st));
    uint32_t* counters = static_cast<uint32_t*>(c->rx_counters.ptr);
    if (int rc = lux_device_zero(c, counters, 64)) return rc;
    int4* keys = static_cast<int4*>(c->rx_keys.ptr);
    uint32_t* flags = static_cast<uint32_t*>(c->rx_flags.ptr);
    uint32_t* slot = static_cast<uint32_t*>(c->rx_slot.ptr);
    uint32_t* lodv = static_cast<uint32_t*>(c->rx_lod.ptr);
    uint32_t* ord = static_cast<uint32_t*>(c->rx_ord.ptr);
    uint32_t* seq = static_cast<uint32_t*>(c->rx_seq.ptr);
    uint32_t* tickets = static_cast<uint32_t*>(c->rx_tickets.ptr);
    uint32_t* sel = tickets + count;  // ticket flags, scanned into positions
    // seed
    RxSeed sd;
    std::memset(&sd, 0, sizeof sd);
    sd.count = count;
    sd.width = rp->width;
    sd.height = rp->height;
    sd.budget = rp->budget;
    sd.cursor = rp->fair_cursor;
    sd.quota = rp->fair_quota;
    sd.frame = rp->frame;
    sd.flags = rp->flags;
    sd.levels = levels;
    sd.has_view = view != nullptr;
    sd.spacing = c->gi.spacing;
    if (view) sd.view = *view;
    hipLaunchKernelGGL(k_rx_seed, dim3(blocks_for(count)), dim3(kWg), 0, st, sd, receivers, keys, flags, lodv,
                       counters);
    // rank: directory, canonical descriptors, ordinals and level-0 counts
    RxDir D = dir_of(c);
    const uint32_t clear_blocks = std::min<uint32_t>(blocks_for((uint64_t)levels * cap), 4096u);
    hipLaunchKernelGGL(k_rx_dir_clear, dim3(clear_blocks), dim3(kWg), 0, st, D);
    hipLaunchKernelGGL(k_rx_insert, dim3(blocks_for(nd)), dim3(kWg), 0, st, D, keys, flags, nd, slot, counters);
    uint32_t *sk_in, *sk, *sv;
    if (int rc = sort_keys(c, std::max<uint32_t>(m, 1u), &sk_in, &sk, &sv)) return rc;
    hipLaunchKernelGGL(k_rx_ray_keys, dim3(blocks_for(m)), dim3(kWg), 0, st, flags, slot, m, levels, sk_in);
    if (int rc = run_sort(c, m)) return rc;
    hipLaunchKernelGGL(k_rx_run_start, dim3(blocks_for(m)), dim3(kWg), 0, st, sk, m, D.offset);
    hipLaunchKernelGGL(k_rx_ordinals, dim3(blocks_for(m)), dim3(kWg), 0, st, sk, sv, m, D.offset, D.count, ord);
    // parents, roots, prefix
    for (uint32_t L = 1; L < levels; ++L)
        hipLaunchKernelGGL(k_rx_parents, dim3(blocks_for(cap)), dim3(kWg), 0, st, D, keys, L);
    uint32_t* root_arr = seq;  // root counts by canonical index; the sequences come later
    if (int rc = lux_device_zero(c, root_arr, 4ull * m)) return rc;
    hipLaunchKernelGGL(k_rx_root_scatter, dim3(blocks_for(cap)), dim3(kWg), 0, st, D, root_arr);
    if (int rc = scan_exclusive(c, root_arr, m, counters + 6)) return rc;
    hipLaunchKernelGGL(k_rx_root_offsets, dim3(blocks_for(cap)), dim3(kWg), 0, st, D, root_arr);
    for (uint32_t L = levels - 1u; L-- > 0;)
        hipLaunchKernelGGL(k_rx_prefix, dim3(blocks_for(cap)), dim3(kWg), 0, st, D, keys, L);
    // generateDirections: sequences, tickets in receiver order
    hipLaunchKernelGGL(k_rx_generate, dim3(blocks_for(count)), dim3(kWg), 0, st, flags, slot, ord, D.offset, count,
                       levels, seq, sel);
    if (int rc = scan_exclusive(c, sel, count, counters + 1)) return rc;
    hipLaunchKernelGGL(k_rx_compact, dim3(blocks_for(count)), dim3(kWg), 0, st, seq, sel, count, tickets);
    LUX_HIP(c, hipGetLastError());
    uint32_t nt = 0;
    LUX_HIP(c, hipMemcpyAsync(&nt, counters + 1, 4, hipMemcpyDeviceToHost, st));
    LUX_HIP(c, hipStreamSynchronize(st));
    c->rx_tickets_n = nt;
    // trace and deposit
    const uint32_t items = 2u * nt;
    const uint64_t n_dep = (uint64_t)items * levels;
    if (n_dep >= 0xffffffffull) return lux_fail(c, LUX_ERR_INVALID, "lux_gi_update_receivers: too many deposits");
    RxGrids grids;
    std::memset(&grids, 0, sizeof grids);
    RxOut out;
    std::memset(&out, 0, sizeof out);
    grids.count = lod_count;
    out.count = lod_count;
    uint32_t base = 0;
    for (uint32_t k = 0; k < lod_count; ++k) {
        lux_ctx* g = lods[k];
        RxGrid& rg = grids.g[k];
        for (int q = 0; q < 3; ++q) {
            rg.origin[q] = g->gi.origin[q];
            rg.dims[q] = g->gi.dims[q];
        }
        rg.spacing = g->gi.spacing;
        rg.base_phi = g->gi.base_phi;
        rg.base_y = g->gi.base_y;
        rg.base = base;
        for (uint32_t L = 0; L <= levels && L < 9; ++L) rg.offset[L] = g->gi_offset[L];
        out.base[k] = base;
        out.iv[k] = static_cast<float4*>(g->gi_intervals.ptr);
        out.sm[k] = static_cast<uint32_t*>(g->gi_samples.ptr);
        base += (uint32_t)g->gi_total;
        hipLaunchKernelGGL(k_rx_clear, dim3(std::min<uint32_t>(blocks_for(g->gi_total), 8192u)), dim3(kWg), 0, st,
                           out.iv[k], out.sm[k], (uint64_t)g->gi_total);
    }
    out.base[lod_count] = base;
    if (items > 0) {
        if (int rc = lux_ensure(c, c->rx_dep_val, 16ull * n_dep)) return rc;
        if (int rc = sort_keys(c, (uint32_t)n_dep, &sk_in, &sk, &sv)) return rc;
        GiArgs a = gi_args(c);
        a.admit_limit = 0;  // the fair window bounds the work
        hipLaunchKernelGGL(k_rx_trace, dim3(blocks_for(items)), dim3(kWg), 0, st, a, grids, receivers, count, tickets,
                           items, seq, lodv, rp->frame, sk_in, static_cast<float4*>(c->rx_dep_val.ptr), counters);
        if (int rc = run_sort(c, (uint32_t)n_dep)) return rc;
        hipLaunchKernelGGL(k_rx_reduce, dim3(blocks_for(n_dep)), dim3(kWg), 0, st, sk, sv,
                           static_cast<const float4*>(c->rx_dep_val.ptr), (uint32_t)n_dep, out);
        LUX_HIP(c, hipGetLastError());
    }
    // history, merge and pack of every grid
    for (uint32_t k = 0; k < lod_count; ++k)
        if (int rc = lux_gi_finish_internal(lods[k], st)) return rc;
    LUX_HIP(c, hipEventRecord(c->ev1, st));
    if (stats) {
        uint32_t cv[8];
        LUX_HIP(c, hipMemcpyAsync(cv, counters, 32, hipMemcpyDeviceToHost, st));
        LUX_HIP(c, hipStreamSynchronize(st));
        std::memset(stats, 0, sizeof *stats);
        stats->receivers = cv[0];
        stats->tickets = cv[1];
        stats->rays = cv[2];
        stats->unsupported = cv[3];
        stats->deposits = cv[4];
        stats->outside = cv[5];
        stats->root_rays = cv[6];
        stats->overflow = cv[7];
        LUX_HIP(c, hipEventElapsedTime(&stats->ms, c->ev0, c->ev1));
    }
    return LUX_OK;
}

extern "C" int lux_receivers_download(lux_ctx* ctx, int32_t* keys, uint32_t* flags, uint32_t* canonical,
                                      uint32_t* probe_count, uint32_t* probe_offset, uint32_t* lods,
                                      uint32_t* sequence, uint32_t* tickets, uint32_t* n_tickets) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->rx_count) return lux_fail(ctx, LUX_ERR_STATE, "lux_receivers_download: no receiver update yet");
    const uint64_t nd = 2ull * ctx->rx_count * ctx->rx_levels;
    const uint64_t m = 2ull * ctx->rx_count;
    LUX_HIP(ctx, hipSetDevice(ctx->device));
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (keys) LUX_HIP(ctx, hipMemcpy(keys, ctx->rx_keys.ptr, 16ull * nd, hipMemcpyDeviceToHost));
    if (flags) LUX_HIP(ctx, hipMemcpy(flags, ctx->rx_flags.ptr, 4ull * nd, hipMemcpyDeviceToHost));
    if (canonical || probe_count || probe_offset) {
        lux_buffer tmp;
        if (int rc = lux_ensure(ctx, tmp, 12ull * nd)) return rc;
        uint32_t* t = static_cast<uint32_t*>(tmp.ptr);
        hipLaunchKernelGGL(k_rx_describe, dim3(blocks_for(nd)), dim3(kWg), 0, ctx->stream, dir_of(ctx),
                           static_cast<const uint32_t*>(ctx->rx_slot.ptr), nd, t, t + nd, t + 2 * nd);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess && canonical) e = hipMemcpy(canonical, t, 4ull * nd, hipMemcpyDeviceToHost);
        if (e == hipSuccess && probe_count) e = hipMemcpy(probe_count, t + nd, 4ull * nd, hipMemcpyDeviceToHost);
        if (e == hipSuccess && probe_offset) e = hipMemcpy(probe_offset, t + 2 * nd, 4ull * nd, hipMemcpyDeviceToHost);
        lux_release(tmp);
        if (e != hipSuccess) return lux_hip_check(ctx, e, "lux_receivers_download");
    }
    if (lods) LUX_HIP(ctx, hipMemcpy(lods, ctx->rx_lod.ptr, 4ull * m, hipMemcpyDeviceToHost));
    if (sequence) LUX_HIP(ctx, hipMemcpy(sequence, ctx->rx_seq.ptr, 4ull * m, hipMemcpyDeviceToHost));
    if (tickets && ctx->rx_tickets_n)
        LUX_HIP(ctx, hipMemcpy(tickets, ctx->rx_tickets.ptr, 4ull * ctx->rx_tickets_n, hipMemcpyDeviceToHost));
    if (n_tickets) *n_tickets = ctx->rx_tickets_n;
    return LUX_OK;
}

extern "C" int lux_gi_samples_download(lux_ctx* ctx, uint32_t* samples) {
    if (!ctx) return LUX_ERR_INVALID;
    if (!ctx->gi_ready) return lux_fail(ctx, LUX_ERR_STATE, "GI not configured (lux_gi_configure)");
    if (!samples) return lux_fail(ctx, LUX_ERR_INVALID, "lux_gi_samples_download: null");
    if (!ctx->gi_samples.ptr || ctx->gi_samples.bytes < 4ull * ctx->gi_total)
        return lux_fail(ctx, LUX_ERR_STATE, "lux_gi_samples_download: no per-interval sample counts yet");
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(samples, ctx->gi_samples.ptr, 4ull * ctx->gi_total, hipMemcpyDeviceToHost));
    return LUX_OK;
}

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
//   cold priority  (first update of an invalid persistent history) every
//                  receiver is eligible; per probe a nominee (lowest
//                  frame-seeded rank), populations binned by log2, nominees
//                  admitted by bin within the budget (scan for the boundary
//                  bin), fair-window receivers fill the rest (scan)
//   ordinals       stable radix sort of the (variant, receiver) ray markers by
//                  level-0 directory slot (a cold nominee first): ordinal =
//                  rank in the run, level-0 ray count = run length
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
// Cold admission and fill tickets go in ascending fair rank (variant 0
// first), the order of the reference's rank-mapped dispatch.
#include <algorithm>
#include <cstring>
#include "lux_gi_device.h"

namespace {

constexpr uint32_t kRxEmpty = 0xffffffffu;
constexpr uint32_t kRxValid = 1u, kRxRay = 2u, kRxNominee = 4u, kRxFill = 8u;
constexpr uint32_t kRxLodMask = 31u << 18;
constexpr uint32_t kRxMaxLods = 8;
constexpr uint32_t kScanBlock = 1024;

// Descriptors (variant, receiver, level) are stored level-major: descriptor
// j = variant * count + receiver at level L is d = L * m + j (m = 2 * count),
// so the seed's stores and the directory insert's loads of one level are
// contiguous across a wave.  Canonical indices compare the same way as in
// the receiver-major order of the reference's dispatch within one level
// (lower j, lower d).  lux_receivers_download returns the receiver-major
// order, (j * levels + L).
__device__ __forceinline__ uint64_t rx_desc(uint64_t j, uint32_t L, uint64_t m) { return (uint64_t)L * m + j; }

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

__device__ __forceinline__ uint32_t inv_xorshift(uint32_t v, uint32_t shift) {
    uint32_t r = v;
    for (uint32_t off = shift; off < 32u; off += shift) r ^= v >> off;
    return r;
}

__device__ __forceinline__ uint32_t inv_perm(uint32_t v, uint32_t mask, uint32_t seed) {
    v = inv_xorshift(v, 16u) & mask;
    v = (v * 0x43021123u) & mask;
    v = inv_xorshift(v, 15u) & mask;
    v = (v * 0x1d69e2a5u) & mask;
    v = inv_xorshift(v, 16u) & mask;
    return (v ^ seed) & mask;
}

__device__ __forceinline__ uint32_t perm_mask(uint32_t count) {
    uint32_t mask = 0;
    while (mask < count - 1u) mask = (mask << 1) | 1u;
    return mask;
}

// Rank of a receiver in the cycle-walked permutation whose seed is the grid
// term xor `extra` (0: transportCandidateRank; frame * phi:
// coldCandidateRank, split.slang:607-645).
__device__ uint32_t rank_of(uint32_t sample, uint32_t w, uint32_t h, uint32_t count, uint32_t extra) {
    if (count <= 1u) return 0u;
    const uint32_t mask = perm_mask(count);
    const uint32_t seed = ((w * 0x9e3779b9u) ^ (h * 0x85ebca6bu) ^ count ^ extra) & mask;
    uint32_t r = perm(sample, mask, seed);
    while (r >= count) r = perm(r, mask, seed);
    return r;
}

// Its inverse (transportCandidateFromRank / coldCandidateFromRank, :622-657).
__device__ uint32_t sample_of(uint32_t rank, uint32_t w, uint32_t h, uint32_t count, uint32_t extra) {
    if (count <= 1u) return 0u;
    const uint32_t mask = perm_mask(count);
    const uint32_t seed = ((w * 0x9e3779b9u) ^ (h * 0x85ebca6bu) ^ count ^ extra) & mask;
    uint32_t s = inv_perm(rank, mask, seed);
    while (s >= count) s = inv_perm(s, mask, seed);
    return s;
}

struct RxSeed {
    uint32_t count, width, height, budget, cursor, quota, frame, flags;
    uint32_t levels, has_view, cold;
    float spacing;
    lux_lod_view view;
};

// transportCandidateSelected (split.slang:659-683) of the receiver with a
// given fair rank.
__device__ bool rank_selected(const RxSeed& s, uint32_t rank) {
    if (s.flags & 1u) return false;
    if (s.count == 0u || s.budget == 0u || s.budget >= s.count) return true;
    uint32_t start = (s.cursor & 0x03ffffffu) % s.count;
    if (s.quota == 0u) return false;
    return rank >= start ? rank - start < s.quota : s.count - start + rank < s.quota;
}

__device__ bool candidate_selected(const RxSeed& s, uint32_t sample) {
    if (s.flags & 1u) return false;
    if (s.count == 0u || s.budget == 0u || s.budget >= s.count) return true;
    return rank_selected(s, rank_of(sample, s.width, s.height, s.count, 0u));
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
    // grid-stride over the receivers; the eligible count goes to the global
    // counter once per block (one counter word taking an atomic from every
    // wave of a 2M-receiver grid costs more than the kernel's stores)
    __shared__ uint32_t block_eligible;
    if (threadIdx.x == 0) block_eligible = 0u;
    __syncthreads();
    uint32_t eligible = 0;
    for (uint32_t i = blockIdx.x * kWg + threadIdx.x; i < s.count; i += gridDim.x * kWg) {
        float3 pos, n;
        const bool ok = rx_surface(recv + 6ull * i, pos, n);
        eligible += ok ? 1u : 0u;
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
        // cold priority: every receiver's selected LOD is eligible (eligibleRay)
        const bool ray = ok && (s.cold || candidate_selected(s, i));
        for (uint32_t v = 0; v < 2; ++v) {
            lods[(uint64_t)v * s.count + i] = v ? high : low;
            for (uint32_t L = 0; L < s.levels; ++L) {
                const uint64_t d = rx_desc((uint64_t)v * s.count + i, L, 2ull * s.count);
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
    if (eligible) atomicAdd(&block_eligible, eligible);
    __syncthreads();
    if (threadIdx.x == 0 && block_eligible) atomicAdd(&counters[0], block_eligible);
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
    uint32_t m;  // descriptors per level, 2 * count
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

// Insert one key at level L (the owner word goes from empty to a descriptor
// index once; the canonical word only decreases, so stale plain loads are
// safe: they cost at most an extra atomic).
__device__ uint32_t dir_insert(const RxDir& D, const int4* keys, uint32_t L, int4 k, uint32_t d) {
    uint32_t* ow = D.owner + (uint64_t)L * D.cap;
    uint32_t* cn = D.canon + (uint64_t)L * D.cap;
    uint32_t b = dir_bucket(k, D.cap);
    for (uint32_t a = 0; a < D.cap; ++a) {
        uint32_t cur = ow[b];
        if (cur == kRxEmpty) {
            // keys are all published by the seed launch; the CAS publishes an index
            cur = atomicCAS(&ow[b], kRxEmpty, d);
            if (cur == kRxEmpty) cur = d;
        }
        if (cur == d || same_probe(keys[cur], k)) {
            if (d < cn[b]) atomicMin(&cn[b], d);
            return b;
        }
        b = (b + 1u) & (D.cap - 1u);
    }
    return kRxEmpty;
}

// Level-major over the descriptors (thread t: level t / m, (variant,
// receiver) t % m, m = 2 * count), so a wave holds 64 neighbouring receivers
// of one level; lanes with the same probe key elect their lowest lane, which
// inserts once for all of them.
__global__ __launch_bounds__(kWg) void k_rx_insert(RxDir D, const int4* keys, const uint32_t* flags, uint32_t m,
                                                   uint32_t* slot_out, uint32_t* counters) {
    const uint64_t t = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    const uint32_t lane = __lane_id();
    uint32_t L = 0, d = 0;
    bool valid = false;
    int4 k = make_int4(0, 0, 0, 0);
    if (t < (uint64_t)m * D.levels) {
        L = (uint32_t)(t / m);
        d = (uint32_t)t;  // rx_desc(t % m, L, m)
        valid = (flags[d] & kRxValid) != 0u;
        if (valid) k = keys[d];
        else slot_out[d] = kRxEmpty;
    }
    uint64_t pending = __ballot(valid);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int4 lk = make_int4(__shfl(k.x, leader, 64), __shfl(k.y, leader, 64), __shfl(k.z, leader, 64),
                                  __shfl(k.w, leader, 64));
        const uint32_t lL = __shfl(L, leader, 64);
        const bool peer = valid && ((pending >> lane) & 1ull) && L == lL && same_probe(k, lk);
        const uint64_t peers = __ballot(peer);
        uint32_t slot = kRxEmpty;
        if ((int)lane == leader) slot = dir_insert(D, keys, L, k, d);
        slot = __shfl(slot, leader, 64);
        if (peer) slot_out[d] = slot;
        if ((int)lane == leader && slot == kRxEmpty) atomicAdd(&counters[7], (uint32_t)__popcll(peers));
        pending &= ~peers;
    }
}

// Sort keys of the (variant, receiver) ray markers: level-0 slot, or `none`
// (above every key).  With `cold` a ray is traced when it is an admitted
// nominee or a kept fill (every other loses its ray flag here), and the key
// is slot * 2 + (not the probe's nominee), so a probe's nominee sorts first.
__global__ __launch_bounds__(kWg) void k_rx_ray_keys(uint32_t* flags, const uint32_t* slot, uint32_t m,
                                                     uint32_t levels, uint32_t none, uint32_t cold, uint32_t* out) {
    const uint32_t j = blockIdx.x * kWg + threadIdx.x;
    if (j >= m) return;
    const uint64_t d0 = rx_desc(j, 0u, m);
    uint32_t f = flags[d0];
    const uint32_t b = slot[d0];
    if (cold && (f & kRxRay)) {
        const uint32_t g = (f & (kRxNominee | kRxFill)) ? f & ~kRxFill : f & ~kRxRay;
        if (g != f) flags[d0] = g;
        f = g;
    }
    uint32_t k = none;
    if ((f & kRxRay) && b != kRxEmpty) k = cold ? (b << 1) | ((f & kRxNominee) ? 0u : 1u) : b;
    out[j] = k;
}

// Runs of equal slots (key >> shift) in the sorted keys.
__global__ __launch_bounds__(kWg) void k_rx_run_start(const uint32_t* sk, uint32_t m, uint32_t none, uint32_t shift,
                                                      uint32_t* start) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= m) return;
    const uint32_t k = sk[i];
    if (k != none && (i == 0 || (sk[i - 1] >> shift) != (k >> shift))) start[k >> shift] = i;
}

__global__ __launch_bounds__(kWg) void k_rx_ordinals(const uint32_t* sk, const uint32_t* sv, uint32_t m,
                                                     uint32_t none, uint32_t shift, const uint32_t* start,
                                                     uint32_t* count0, uint32_t* ord) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= m) return;
    const uint32_t k = sk[i];
    if (k == none) return;
    const uint32_t b = k >> shift;
    const uint32_t s = start[b];
    ord[sv[i]] = i - s;
    if (i + 1 == m || (sk[i + 1] >> shift) != b) count0[b] = i - s + 1u;
}

// ---------------- cold priority (split.slang:685-705, rank :836-888) ----------------
// Words: per level-0 slot the nominee priority (0xffffffff - lowest cold
// rank) and the population (eligible receivers); a table of 32 population
// bins, 32 admission quotas, the boundary bin, admitted nominees and the fill
// quota; m words of flags scanned into positions.
struct RxCold {
    uint32_t* prio;
    uint32_t* pop;
    uint32_t* table;
    uint32_t* pos;
};
constexpr uint32_t kColdTable = 80;

__device__ __forceinline__ uint32_t cold_bin(uint32_t population) {
    return population > 1u ? 31u - (uint32_t)__clz(population) : 0u;
}

// Phase 0: per eligible (variant, receiver) its probe's nominee priority and
// population; lanes of one probe reduce to one atomic pair.
__global__ __launch_bounds__(kWg) void k_rx_cold_nominate(RxSeed sd, const uint32_t* flags, const uint32_t* slot,
                                                          uint32_t m, RxCold cw) {
    const uint32_t j = blockIdx.x * kWg + threadIdx.x;
    const uint32_t lane = __lane_id();
    bool e = false;
    uint32_t b = kRxEmpty, pr = 0u;
    if (j < m) {
        const uint64_t d0 = rx_desc(j, 0u, m);
        b = slot[d0];
        e = (flags[d0] & kRxRay) && b != kRxEmpty;
        if (e) pr = 0xffffffffu - rank_of(j % sd.count, sd.width, sd.height, sd.count, sd.frame * 0x9e3779b9u);
    }
    uint64_t pending = __ballot(e);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lb = __shfl(b, leader, 64);
        const bool peer = e && b == lb;
        const uint64_t peers = __ballot(peer);
        const uint32_t mx = wave_max_u32(peer ? pr : 0u);
        if ((int)lane == leader) {
            atomicMax(&cw.prio[lb], mx);
            atomicAdd(&cw.pop[lb], (uint32_t)__popcll(peers));
        }
        pending &= ~peers;
    }
}

// Phase 1: probes per population bin.
__global__ __launch_bounds__(kWg) void k_rx_cold_hist(uint32_t cap, RxCold cw) {
    __shared__ uint32_t h[32];
    if (threadIdx.x < 32) h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t b = blockIdx.x * kWg + threadIdx.x;
    if (b < cap && cw.pop[b] != 0u) atomicAdd(&h[cold_bin(cw.pop[b])], 1u);
    __syncthreads();
    if (threadIdx.x < 32 && h[threadIdx.x]) atomicAdd(&cw.table[threadIdx.x], h[threadIdx.x]);
}

// Admission per bin: a probe of bin b is admitted while fewer than
// budget - (probes in higher bins) of its bin came before it; at most one bin
// is admitted in part.
__global__ void k_rx_cold_bounds(RxCold cw, uint32_t budget, uint32_t* counters) {
    if (threadIdx.x != 0) return;
    uint32_t higher = 0, admitted = 0, boundary = 32u;
    for (int b = 31; b >= 0; --b) {
        const uint32_t hb = cw.table[b];
        const uint32_t q = higher < budget ? budget - higher : 0u;
        cw.table[32 + b] = q;
        if (hb > q && q > 0u) boundary = (uint32_t)b;
        admitted += hb < q ? hb : q;
        higher += hb;
    }
    cw.table[64] = boundary;
    cw.table[65] = admitted;
    cw.table[66] = budget - admitted;
    counters[10] = admitted;
}

// Position p = fair rank * 2 + variant (the rank-mapped dispatch order) to
// its (variant, receiver) index.
__device__ __forceinline__ uint32_t rank_mapped(const RxSeed& sd, uint32_t p) {
    return (p & 1u) * sd.count + sample_of(p >> 1, sd.width, sd.height, sd.count, 0u);
}

// Canonical level-0 descriptors of the boundary bin, flagged in rank-mapped
// order (scanned: their admission tickets).
__global__ __launch_bounds__(kWg) void k_rx_cold_order(RxSeed sd, RxDir D, const uint32_t* slot, uint32_t m,
                                                       RxCold cw) {
    const uint32_t p = blockIdx.x * kWg + threadIdx.x;
    if (p >= m) return;
    const uint32_t boundary = cw.table[64];
    uint32_t f = 0u;
    if (boundary < 32u) {
        const uint64_t d0 = rx_desc(rank_mapped(sd, p), 0u, m);
        const uint32_t b = slot[d0];
        if (b != kRxEmpty && D.canon[b] == (uint32_t)d0 && cw.pop[b] != 0u && cold_bin(cw.pop[b]) == boundary) f = 1u;
    }
    cw.pos[p] = f;
}

// Phase 2: each probe's nominee (the receiver of its lowest cold rank, on
// the variant that holds the probe) is admitted or loses its ray.
__global__ __launch_bounds__(kWg) void k_rx_cold_admit(RxSeed sd, RxDir D, const uint32_t* slot, uint32_t* flags,
                                                       RxCold cw) {
    const uint32_t b = blockIdx.x * kWg + threadIdx.x;
    if (b >= D.cap || cw.prio[b] == 0u) return;
    const uint32_t ns = sample_of(0xffffffffu - cw.prio[b], sd.width, sd.height, sd.count, sd.frame * 0x9e3779b9u);
    const uint64_t m = 2ull * sd.count;
    uint64_t nom = rx_desc(ns, 0u, m);
    if (slot[nom] != b) nom = rx_desc((uint64_t)sd.count + ns, 0u, m);
    const uint32_t bin = cold_bin(cw.pop[b]);
    const uint32_t q = cw.table[32 + bin];
    bool admitted = q > 0u;
    if (bin == cw.table[64]) {
        const uint32_t jc = D.canon[b];  // a level-0 descriptor: rx_desc(jc, 0, m) = jc
        const uint32_t sc = jc % sd.count, vc = jc / sd.count;
        admitted = cw.pos[rank_of(sc, sd.width, sd.height, sd.count, 0u) * 2u + vc] < q;
    }
    const uint32_t f = flags[nom];
    flags[nom] = admitted ? (f | kRxNominee) : (f & ~kRxRay);
}

// Phase 3 over the fair window only (its ranks [start, start + len) mod
// count; every fill candidate is in it): position t enumerates the window's
// (rank, variant) pairs in ascending rank -- the wrapped part [0, start + len
// - count) first -- so a scan of the candidate flags gives the fill tickets.
struct RxWindow {
    uint32_t start, len, wrapped;
};

__device__ __forceinline__ uint64_t window_descriptor(const RxSeed& sd, const RxWindow& w, uint32_t t) {
    const uint32_t i = t >> 1;
    const uint32_t r = i < w.wrapped ? i : w.start + (i - w.wrapped);
    return rx_desc((uint64_t)(t & 1u) * sd.count + sample_of(r, sd.width, sd.height, sd.count, 0u), 0u,
                   2ull * sd.count);
}

__device__ __forceinline__ bool cold_fill_candidate(uint32_t f, uint32_t b) {
    return (f & kRxRay) && !(f & kRxNominee) && b != kRxEmpty;
}

__global__ __launch_bounds__(kWg) void k_rx_cold_fill_flags(RxSeed sd, RxWindow w, const uint32_t* flags,
                                                            const uint32_t* slot, RxCold cw) {
    const uint32_t t = blockIdx.x * kWg + threadIdx.x;
    if (t >= 2u * w.len) return;
    const uint64_t d0 = window_descriptor(sd, w, t);
    cw.pos[t] = cold_fill_candidate(flags[d0], slot[d0]) ? 1u : 0u;
}

// The first fill-quota candidates are marked kept (kRxFill).
__global__ __launch_bounds__(kWg) void k_rx_cold_fill(RxSeed sd, RxWindow w, uint32_t* flags, const uint32_t* slot,
                                                      RxCold cw) {
    const uint32_t t = blockIdx.x * kWg + threadIdx.x;
    if (t >= 2u * w.len) return;
    const uint64_t d0 = window_descriptor(sd, w, t);
    const uint32_t f = flags[d0];
    if (cold_fill_candidate(f, slot[d0]) && cw.pos[t] < cw.table[66]) flags[d0] = f | kRxFill;
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
    root_arr[D.canon[top] - (uint64_t)(D.levels - 1u) * D.m] = D.count[top];
}

__global__ __launch_bounds__(kWg) void k_rx_root_offsets(RxDir D, const uint32_t* root_scan) {
    const uint32_t b = blockIdx.x * kWg + threadIdx.x;
    if (b >= D.cap) return;
    const uint64_t top = (uint64_t)(D.levels - 1u) * D.cap + b;
    if (D.owner[top] == kRxEmpty) return;
    D.offset[top] = root_scan[D.canon[top] - (uint64_t)(D.levels - 1u) * D.m];
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

// generateDirections (split.slang:963-983): sequences, ticket flags and the
// number of rays of each receiver.
__global__ __launch_bounds__(kWg) void k_rx_generate(const uint32_t* flags, const uint32_t* slot, const uint32_t* ord,
                                                     const uint32_t* offset0, uint32_t count, uint32_t levels,
                                                     uint32_t* seq, uint32_t* sel, uint32_t* nray) {
    const uint32_t s = blockIdx.x * kWg + threadIdx.x;
    if (s >= count) return;
    uint32_t rays = 0;
    for (uint32_t v = 0; v < 2; ++v) {
        const uint64_t j = (uint64_t)v * count + s;
        const uint64_t d0 = rx_desc(j, 0u, 2ull * count);
        uint32_t q = kRxEmpty;
        if ((flags[d0] & kRxRay) && slot[d0] != kRxEmpty) {
            q = ord[j] + offset0[slot[d0]];
            ++rays;
        }
        seq[j] = q;
    }
    sel[s] = rays ? 1u : 0u;
    nray[s] = rays;
}

// Tickets (receivers with rays) and the ray list ((variant, receiver) index
// j = variant * count + receiver), both in receiver order.
__global__ __launch_bounds__(kWg) void k_rx_compact(const uint32_t* seq, const uint32_t* tpos, const uint32_t* rpos,
                                                    uint32_t count, uint32_t* tickets, uint32_t* rays) {
    const uint32_t s = blockIdx.x * kWg + threadIdx.x;
    if (s >= count) return;
    const bool a = seq[s] != kRxEmpty, b = seq[(uint64_t)count + s] != kRxEmpty;
    if (a || b) tickets[tpos[s]] = s;
    uint32_t p = rpos[s];
    if (a) rays[p++] = s;
    if (b) rays[p] = count + s;
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

// deposit (split.slang:1056-1118) of one ray: the levels up to the
// terminating one, those inside a grid written contiguously to kout / vout.
__device__ void rx_deposits(const lux_gi_params& p, const RxGrids& grids, float3 pos, float3 d, uint32_t lod,
                            bool hit, float t, float3 rad, uint32_t* kout, float4* vout, uint32_t& kept,
                            uint32_t& outside) {
    const uint32_t levels = p.levels;
    const RxGrid* g = lod < grids.count ? &grids.g[lod] : nullptr;
    const float lod_scale = pow2i((int)(lod < 24u ? lod : 24u));
    for (uint32_t L = 0; L < levels; ++L) {
        const float start = L == 0 ? 0.0f : interval_end(p, L - 1u) * lod_scale;
        const float end = interval_end(p, L) * lod_scale;
        const bool terminal = p.terminal_tail != 0u && L + 1u == levels;
        const bool term = terminal || (hit && t > start && t <= end);
        const int4 key = rx_cell(pos, p.spacing, L, lod);
        bool in = false;
        if (g) {
            const float sl = g->spacing * pow2i((int)L);
            const uint32_t dx = level_dim(g->dims[0], L), dy = level_dim(g->dims[1], L), dz = level_dim(g->dims[2], L);
            const int64_t ix = (int64_t)key.x - (int64_t)floorf(g->origin[0] / sl);
            const int64_t iy = (int64_t)key.y - (int64_t)floorf(g->origin[1] / sl);
            const int64_t iz = (int64_t)key.z - (int64_t)floorf(g->origin[2] / sl);
            in = ix >= 0 && iy >= 0 && iz >= 0 && ix < dx && iy < dy && iz < dz;
            if (in) {
                const uint32_t nphi = g->base_phi << L, ny = g->base_y << L;
                const uint64_t probe = (uint64_t)ix + (uint64_t)dx * ((uint64_t)iy + (uint64_t)dy * (uint64_t)iz);
                kout[kept] = g->base + (uint32_t)(g->offset[L] + probe * nphi * ny + direction_index(d, nphi, ny));
                vout[kept] = term ? make_float4(rad.x, rad.y, rad.z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            }
        }
        if (in) ++kept;
        else ++outside;
        if (term) break;
    }
}

// Fused form: one thread per ray of the ray list traces, shades (emissive
// NEE with its occlusion rays) and deposits; dcount[ray] = deposits kept.
__global__ __launch_bounds__(kWg) void k_rx_trace(GiArgs a, RxGrids grids, const float* recv, uint32_t count,
                                                  const uint32_t* rays, uint32_t n_rays, const uint32_t* seq,
                                                  const uint32_t* lods, uint32_t frame, uint32_t* dkey,
                                                  float4* dval, uint32_t* dcount, uint32_t* counters) {
    __shared__ uint2 lds[kLdsDepth][kWg];
    Stack stk;
    stk.bind(&lds[0][threadIdx.x]);
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    bool traced = false, unsupported = false;
    uint32_t kept = 0, outside = 0;
    if (i < n_rays) {
        const uint32_t j = rays[i];
        float3 pos, o, d;
        if (rx_ray(recv, count, j, seq[j], frame, a.p.normal_bias, pos, o, d)) {
            traced = true;
            lux::HitState h;
            lux::NoStats ns;
            lux::trace_closest(a.scene, o, d, 1e20f, h, stk, ns, false);
            float3 rad = make_float3(0.0f, 0.0f, 0.0f);
            const bool hit = h.tri != lux::kEmpty;
            if ((h.err & lux::kErrUnsupported) || (hit && !shade(a, o, d, h, stk, rad)))
                unsupported = true;
            else
                rx_deposits(a.p, grids, pos, d, lods[j], hit, h.t, rad, dkey + (uint64_t)i * a.p.levels,
                            dval + (uint64_t)i * a.p.levels, kept, outside);
        }
        dcount[i] = kept;
    }
    wave_count(&counters[2], traced);
    wave_count(&counters[3], unsupported);
    if (kept) atomicAdd(&counters[4], kept);
    if (outside) atomicAdd(&counters[5], outside);
}

// Wavefront form: deposits of each ray from its traced (rad, t) and status.
__global__ __launch_bounds__(kWg) void k_rx_deposit(lux_gi_params p, RxGrids grids, const float* recv,
                                                    uint32_t count, const uint32_t* rays, uint32_t n_rays,
                                                    const uint32_t* seq, const uint32_t* lods, uint32_t frame,
                                                    const float4* traced_rt, const uint32_t* status,
                                                    uint32_t* dkey, float4* dval, uint32_t* dcount,
                                                    uint32_t* counters) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    bool traced = false, unsupported = false;
    uint32_t kept = 0, outside = 0;
    if (i < n_rays) {
        const uint32_t j = rays[i];
        float3 pos, o, d;
        if (rx_ray(recv, count, j, seq[j], frame, p.normal_bias, pos, o, d)) {
            traced = true;
            const uint32_t stt = status[i];
            const float4 rt = traced_rt[i];
            if (stt == 2u)
                unsupported = true;
            else
                rx_deposits(p, grids, pos, d, lods[j], stt == 1u, rt.w, make_float3(rt.x, rt.y, rt.z),
                            dkey + (uint64_t)i * p.levels, dval + (uint64_t)i * p.levels, kept, outside);
        }
        dcount[i] = kept;
    }
    wave_count(&counters[2], traced);
    wave_count(&counters[3], unsupported);
    if (kept) atomicAdd(&counters[4], kept);
    if (outside) atomicAdd(&counters[5], outside);
}

// Deposits of ray i from its fixed slots to [dpos[i], dpos[i + 1]).
__global__ __launch_bounds__(kWg) void k_rx_pack(const uint32_t* dkey, const float4* dval, const uint32_t* dpos,
                                                 uint32_t n_rays, uint32_t levels, uint32_t* key_out,
                                                 float4* val_out) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n_rays) return;
    const uint32_t p = dpos[i], n = dpos[i + 1] - p;
    for (uint32_t k = 0; k < n; ++k) {
        key_out[p + k] = dkey[(uint64_t)i * levels + k];
        val_out[p + k] = dval[(uint64_t)i * levels + k];
    }
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

// reduce (split.slang:1386-1476): per slot, the mean of its deposits.  The
// sum of a run of deposits (in ray order) is fixed as: chunks of 64
// consecutive deposits, each summed by a pairwise tree (x[i] += x[i + s] for
// i a multiple of 2s, s = 1, 2, 4, ..., 32), the chunk sums added in order.
// Runs of at most 8 deposits are summed by one thread, longer ones by a wave.
__device__ __forceinline__ float4 add4(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

__device__ __forceinline__ void rx_write(const RxOut& out, uint32_t k, float4 sum, uint32_t c) {
    uint32_t g = 0;
    while (g + 1 < out.count && k >= out.base[g + 1]) ++g;
    const uint32_t local = k - out.base[g];
    const float fc = (float)c;
    out.iv[g][local] = make_float4(sum.x / fc, sum.y / fc, sum.z / fc, sum.w / fc);
    out.sm[g][local] = c;
}

__global__ __launch_bounds__(kWg) void k_rx_reduce_short(const uint32_t* sk, const uint32_t* sv, const float4* val,
                                                         uint32_t n, RxOut out, uint32_t* long_runs,
                                                         uint32_t* n_long) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = sk[i];
    if (i > 0 && sk[i - 1] == k) return;
    uint32_t c = 1;
    while (c <= 8u && i + c < n && sk[i + c] == k) ++c;
    if (c > 8u) {
        long_runs[atomicAdd(n_long, 1u)] = i;
        return;
    }
    float4 x[8];
#pragma unroll
    for (uint32_t t = 0; t < 8u; ++t) x[t] = t < c ? val[sv[i + t]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (uint32_t st = 1; st < 8u; st <<= 1)
#pragma unroll
        for (uint32_t t = 0; t + st < 8u; t += 2u * st)
            if (t + st < c) x[t] = add4(x[t], x[t + st]);
    rx_write(out, k, x[0], c);
}

__global__ __launch_bounds__(kWg) void k_rx_reduce_long(const uint32_t* sk, const uint32_t* sv, const float4* val,
                                                        uint32_t n, RxOut out, const uint32_t* long_runs,
                                                        const uint32_t* n_long) {
    const uint32_t lane = __lane_id();
    const uint32_t waves = gridDim.x * (kWg / 64u);
    const uint32_t total = *n_long;
    for (uint32_t r = blockIdx.x * (kWg / 64u) + (threadIdx.x >> 6); r < total; r += waves) {
        const uint32_t start = long_runs[r];
        const uint32_t k = sk[start];
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uint32_t c = 0;
        for (uint32_t base = start;; base += 64u) {
            const uint32_t pos = base + lane;
            const bool in = pos < n && sk[pos] == k;
            const uint32_t nc = (uint32_t)__popcll(__ballot(in));
            if (nc == 0) break;
            float4 x = in ? val[sv[pos]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
            for (uint32_t st = 1; st < 64u; st <<= 1) {
                float4 y = make_float4(__shfl_down(x.x, st, 64), __shfl_down(x.y, st, 64), __shfl_down(x.z, st, 64),
                                       __shfl_down(x.w, st, 64));
                if ((lane & (2u * st - 1u)) == 0u && lane + st < nc) x = add4(x, y);
            }
            acc = c == 0 ? x : add4(acc, x);
            c += nc;
            if (nc < 64u) break;
        }
        if (lane == 0) rx_write(out, k, acc, c);
    }
}

// Descriptor view for downloads, in the receiver-major order p = j * levels
// + L: key, flags, and the canonical descriptor (in the same numbering),
// count and offset of each descriptor's probe.
__global__ __launch_bounds__(kWg) void k_rx_describe(RxDir D, const int4* keys, const uint32_t* flags,
                                                     const uint32_t* slot, uint64_t nd, int4* key_out,
                                                     uint32_t* flag_out, uint32_t* canon, uint32_t* cnt,
                                                     uint32_t* off) {
    const uint64_t p = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (p >= nd) return;
    const uint64_t j = p / D.levels;
    const uint32_t L = (uint32_t)(p % D.levels);
    const uint64_t d = rx_desc(j, L, D.m);
    key_out[p] = keys[d];
    flag_out[p] = flags[d];
    const uint32_t b = slot[d];
    const uint64_t e = (uint64_t)L * D.cap + b;
    uint32_t c = kRxEmpty;
    if (b != kRxEmpty && D.canon[e] != kRxEmpty)
        c = (uint32_t)((D.canon[e] - (uint64_t)L * D.m) * D.levels + L);
    canon[p] = c;
    cnt[p] = b == kRxEmpty ? 0u : D.count[e];
    off[p] = b == kRxEmpty ? kRxEmpty : D.offset[e];
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
    D.m = 2u * c->rx_count;
    return D;
}

RxCold cold_of(lux_ctx* c, uint32_t cap) {
    uint32_t* w = static_cast<uint32_t*>(c->rx_cold.ptr);
    RxCold cw;
    cw.prio = w;
    cw.pop = w + cap;
    cw.table = w + 2ull * cap;
    cw.pos = cw.table + kColdTable;
    return cw;
}

// Exclusive scan of n words in place; the total is written to *total (device).
int scan_exclusive(lux_ctx* c, uint32_t* data, uint32_t n, uint32_t* total) {
    if (int rc = lux_ensure(c, c->rx_scan, 4ull * lux_scan_sums_words(n))) return rc;
    return lux_scan_exclusive_impl(c, data, n, static_cast<uint32_t*>(c->rx_scan.ptr), total);
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

// Sort of keys below 2^bits.
int run_sort(lux_ctx* c, uint32_t n, uint32_t bits) {
    uint32_t* w = static_cast<uint32_t*>(c->rx_sort.ptr);
    return lux_sort_pairs_impl(c, w, n, w + n, w + 2ull * n, w + 3ull * n, w + 4ull * n, w + 5ull * n, bits);
}

uint32_t bits_for(uint64_t max_key) {
    uint32_t b = 1;
    while (b < 32u && (max_key >> b) != 0u) ++b;
    return b;
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
    if (int rc = lux_ensure(c, c->rx_tickets, 20ull * count + 16)) return rc;
    if (int rc = lux_ensure(c, c->rx_counters, 64)) return rc;
    for (uint32_t k = 0; k < lod_count; ++k)
        if (int rc = lux_ensure(lods[k], lods[k]->gi_samples, 4ull * lods[k]->gi_total)) return rc;
    // cold priority (coldPriorityEnabled, split.slang:685-689; src/render.rs:
    // 128-131): the first update of a persistent history that is not valid
    const bool cold = !(rp->flags & 1u) && rp->budget > 0u && c->gi_hist_on && c->gi_hist.history_frames > 0u &&
                      c->gi_hist.persistent != 0u && !c->gi_hist_valid;
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
    // receiver words: tickets, ticket flags (scanned into positions), ray
    // counts (scanned into positions), the ray list (up to 2 per receiver)
    uint32_t* tickets = static_cast<uint32_t*>(c->rx_tickets.ptr);
    uint32_t* tsel = tickets + count;
    uint32_t* rpos = tsel + count;
    uint32_t* rays = rpos + count;
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
    sd.cold = cold ? 1u : 0u;
    sd.spacing = c->gi.spacing;
    if (view) sd.view = *view;
    hipLaunchKernelGGL(k_rx_seed, dim3(std::min<uint32_t>(blocks_for(count), 2048u)), dim3(kWg), 0, st, sd, receivers,
                       keys, flags, lodv, counters);
    // rank: directory, canonical descriptors, ordinals and level-0 counts
    RxDir D = dir_of(c);
    const uint32_t clear_blocks = std::min<uint32_t>(blocks_for((uint64_t)levels * cap), 4096u);
    hipLaunchKernelGGL(k_rx_dir_clear, dim3(clear_blocks), dim3(kWg), 0, st, D);
    hipLaunchKernelGGL(k_rx_insert, dim3(blocks_for(nd)), dim3(kWg), 0, st, D, keys, flags, m, slot, counters);
    if (cold) {
        // nominate, histogram the populations, admit nominees by bin, fill
        // from the fair window up to the budget
        if (int rc = lux_ensure(c, c->rx_cold, 4ull * (2ull * cap + kColdTable + m))) return rc;
        const RxCold cw = cold_of(c, cap);
        if (int rc = lux_device_zero(c, cw.prio, 4ull * (2ull * cap + kColdTable))) return rc;
        hipLaunchKernelGGL(k_rx_cold_nominate, dim3(blocks_for(m)), dim3(kWg), 0, st, sd, flags, slot, m, cw);
        hipLaunchKernelGGL(k_rx_cold_hist, dim3(blocks_for(cap)), dim3(kWg), 0, st, cap, cw);
        hipLaunchKernelGGL(k_rx_cold_bounds, dim3(1), dim3(64), 0, st, cw, rp->budget, counters);
        hipLaunchKernelGGL(k_rx_cold_order, dim3(blocks_for(m)), dim3(kWg), 0, st, sd, D, slot, m, cw);
        if (int rc = scan_exclusive(c, cw.pos, m, counters + 11)) return rc;
        hipLaunchKernelGGL(k_rx_cold_admit, dim3(blocks_for(cap)), dim3(kWg), 0, st, sd, D, slot, flags, cw);
        // the fair window (transportCandidateSelected's ranks)
        RxWindow win;
        if (rp->budget >= count || rp->fair_quota >= count) {
            win.start = 0;
            win.len = count;
        } else {
            win.start = (rp->fair_cursor & 0x03ffffffu) % count;
            win.len = rp->fair_quota;
        }
        win.wrapped = win.start + win.len > count ? win.start + win.len - count : 0u;
        if (win.len > 0) {
            hipLaunchKernelGGL(k_rx_cold_fill_flags, dim3(blocks_for(2ull * win.len)), dim3(kWg), 0, st, sd, win,
                               flags, slot, cw);
            if (int rc = scan_exclusive(c, cw.pos, 2u * win.len, counters + 12)) return rc;
            hipLaunchKernelGGL(k_rx_cold_fill, dim3(blocks_for(2ull * win.len)), dim3(kWg), 0, st, sd, win, flags,
                               slot, cw);
        }
    }
    const uint32_t none = cold ? cap << 1 : cap, shift = cold ? 1u : 0u;
    uint32_t *sk_in, *sk, *sv;
    if (int rc = sort_keys(c, std::max<uint32_t>(m, 1u), &sk_in, &sk, &sv)) return rc;
    hipLaunchKernelGGL(k_rx_ray_keys, dim3(blocks_for(m)), dim3(kWg), 0, st, flags, slot, m, levels, none,
                       cold ? 1u : 0u, sk_in);
    if (int rc = run_sort(c, m, bits_for(none))) return rc;
    hipLaunchKernelGGL(k_rx_run_start, dim3(blocks_for(m)), dim3(kWg), 0, st, sk, m, none, shift, D.offset);
    hipLaunchKernelGGL(k_rx_ordinals, dim3(blocks_for(m)), dim3(kWg), 0, st, sk, sv, m, none, shift, D.offset,
                       D.count, ord);
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
    // generateDirections: sequences, tickets and the ray list in receiver order
    hipLaunchKernelGGL(k_rx_generate, dim3(blocks_for(count)), dim3(kWg), 0, st, flags, slot, ord, D.offset, count,
                       levels, seq, tsel, rpos);
    if (int rc = scan_exclusive(c, tsel, count, counters + 1)) return rc;
    if (int rc = scan_exclusive(c, rpos, count, counters + 8)) return rc;
    hipLaunchKernelGGL(k_rx_compact, dim3(blocks_for(count)), dim3(kWg), 0, st, seq, tsel, rpos, count, tickets,
                       rays);
    LUX_HIP(c, hipGetLastError());
    uint32_t nts[2] = {0, 0};
    LUX_HIP(c, hipMemcpyAsync(&nts[0], counters + 1, 4, hipMemcpyDeviceToHost, st));
    LUX_HIP(c, hipMemcpyAsync(&nts[1], counters + 8, 4, hipMemcpyDeviceToHost, st));
    LUX_HIP(c, hipStreamSynchronize(st));
    c->rx_tickets_n = nts[0];
    const uint32_t n_rays = nts[1];
    // trace and deposit
    const uint64_t n_fixed = (uint64_t)n_rays * levels;
    if (n_fixed >= 0xffffffffull) return lux_fail(c, LUX_ERR_INVALID, "lux_gi_update_receivers: too many deposits");
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
    if (n_rays > 0) {
        // fixed deposit slots (key, value) per ray, then the compacted copies
        if (int rc = lux_ensure(c, c->rx_dep_val, 20ull * n_fixed + 4ull * (n_rays + 1u))) return rc;
        float4* dval = static_cast<float4*>(c->rx_dep_val.ptr);
        uint32_t* dkey = reinterpret_cast<uint32_t*>(dval + n_fixed);
        uint32_t* dpos = dkey + n_fixed;
        if (int rc = lux_device_zero(c, dpos + n_rays, 4)) return rc;
        if (c->gi_trace_mode == 1u) {
            GiArgs a = gi_args(c);
            a.admit_limit = 0;  // the fair window bounds the work
            hipLaunchKernelGGL(k_rx_trace, dim3(blocks_for(n_rays)), dim3(kWg), 0, st, a, grids, receivers, count,
                               rays, n_rays, seq, lodv, rp->frame, dkey, dval, dpos, counters);
        } else {
            // regrouped primary and occlusion walks; (rad, t) and status per ray
            if (int rc = lux_ensure(c, c->rx_traced, 20ull * n_rays)) return rc;
            float4* rt = static_cast<float4*>(c->rx_traced.ptr);
            uint32_t* stt = reinterpret_cast<uint32_t*>(rt + n_rays);
            if (int rc = lux_rx_wave_trace(c, receivers, count, rays, n_rays, seq, rp->frame, rt, stt)) return rc;
            hipLaunchKernelGGL(k_rx_deposit, dim3(blocks_for(n_rays)), dim3(kWg), 0, st, c->gi, grids, receivers,
                               count, rays, n_rays, seq, lodv, rp->frame, rt, stt, dkey, dval, dpos, counters);
        }
        if (int rc = scan_exclusive(c, dpos, n_rays + 1u, counters + 9)) return rc;
        uint32_t n_dep = 0;
        LUX_HIP(c, hipMemcpyAsync(&n_dep, counters + 9, 4, hipMemcpyDeviceToHost, st));
        LUX_HIP(c, hipStreamSynchronize(st));
        if (n_dep > 0) {
            if (int rc = lux_ensure(c, c->rx_pack, 16ull * n_dep + 4ull * (n_dep + 2u))) return rc;
            float4* cval = static_cast<float4*>(c->rx_pack.ptr);
            uint32_t* long_runs = reinterpret_cast<uint32_t*>(cval + n_dep);
            uint32_t* n_long = long_runs + n_dep;
            if (int rc = sort_keys(c, n_dep, &sk_in, &sk, &sv)) return rc;
            hipLaunchKernelGGL(k_rx_pack, dim3(blocks_for(n_rays)), dim3(kWg), 0, st, dkey, dval, dpos, n_rays, levels,
                               sk_in, cval);
            if (int rc = run_sort(c, n_dep, bits_for(base ? base - 1u : 0u))) return rc;
            if (int rc = lux_device_zero(c, n_long, 4)) return rc;
            hipLaunchKernelGGL(k_rx_reduce_short, dim3(blocks_for(n_dep)), dim3(kWg), 0, st, sk, sv, cval, n_dep, out,
                               long_runs, n_long);
            hipLaunchKernelGGL(k_rx_reduce_long, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(blocks_for(n_dep), 4096u))),
                               dim3(kWg), 0, st, sk, sv, cval, n_dep, out, long_runs, n_long);
        }
        LUX_HIP(c, hipGetLastError());
    }
    // history, merge and pack of every grid
    for (uint32_t k = 0; k < lod_count; ++k)
        if (int rc = lux_gi_finish_internal(lods[k], st)) return rc;
    LUX_HIP(c, hipEventRecord(c->ev1, st));
    if (stats) {
        uint32_t cv[11];
        LUX_HIP(c, hipMemcpyAsync(cv, counters, 44, hipMemcpyDeviceToHost, st));
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
        stats->cold = cold ? 1u : 0u;
        stats->nominees = cold ? cv[10] : 0u;
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
    if (keys || flags || canonical || probe_count || probe_offset) {
        // the device holds descriptors level-major; the view is receiver-major
        lux_buffer tmp;
        if (int rc = lux_ensure(ctx, tmp, 32ull * nd)) return rc;
        int4* tk = static_cast<int4*>(tmp.ptr);
        uint32_t* t = reinterpret_cast<uint32_t*>(tk + nd);
        hipLaunchKernelGGL(k_rx_describe, dim3(blocks_for(nd)), dim3(kWg), 0, ctx->stream, dir_of(ctx),
                           static_cast<const int4*>(ctx->rx_keys.ptr), static_cast<const uint32_t*>(ctx->rx_flags.ptr),
                           static_cast<const uint32_t*>(ctx->rx_slot.ptr), nd, tk, t, t + nd, t + 2 * nd, t + 3 * nd);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess && keys) e = hipMemcpy(keys, tk, 16ull * nd, hipMemcpyDeviceToHost);
        if (e == hipSuccess && flags) e = hipMemcpy(flags, t, 4ull * nd, hipMemcpyDeviceToHost);
        if (e == hipSuccess && canonical) e = hipMemcpy(canonical, t + nd, 4ull * nd, hipMemcpyDeviceToHost);
        if (e == hipSuccess && probe_count) e = hipMemcpy(probe_count, t + 2 * nd, 4ull * nd, hipMemcpyDeviceToHost);
        if (e == hipSuccess && probe_offset)
            e = hipMemcpy(probe_offset, t + 3 * nd, 4ull * nd, hipMemcpyDeviceToHost);
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

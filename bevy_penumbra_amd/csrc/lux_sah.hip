// lux_sah.hip -- binned-SAH BLAS build on gfx950.
//
// The reference's build_tree (src/scene.rs:2398-2554) for every mesh at
// once, node for node and element for element the tree its sequential
// recursion writes:
//   * node box: min/max fold over the node's elements in their current
//     order, each step a per-component select (a < b ? a : b, as glam's
//     Vec3::min/max), so among equal values the LAST element's wins (this
//     only shows on signed zeros; positions are finite);
//   * 12 bins per axis over the centroid span, f64 bin index, f64 area x
//     count costs, the first strictly cheaper split with both sides
//     non-empty in axis order x, y, z;
//   * the in-place partition `if left { l += 1 } else { r -= 1; swap(l, r) }`;
//   * coincident centroids: stable order by the total-order key of the
//     centre on the widest centroid axis, split at count / 2;
//   * leaves of at most 4 elements and at depth 48;
//   * nodes in the recursion's pre-order with children pairs appended.
//
// Nodes of more than kSmall elements are built level by level over all
// meshes: bounds and bins by atomics (ordered-integer images; the node box
// keys carry the element position so the last equal element wins), the
// split sweep one thread per node, and the partition in closed form: with
// mid left elements, f = mid + (element mid goes right), the first f
// positions are the front; a front left element keeps its offset, the k-th
// front right element moves to (offset of the (k-1)-th back left element,
// counted from the end) - 1 (the first to count - 1), the k-th back left
// element to the offset of the k-th front right element, and a back right
// element one offset down -- exactly where the sequential swap loop leaves
// them.  Subtrees of at most kSmall elements run the sequential recursion,
// one thread each.  Node indices come last: subtree sizes bottom-up, then
// pre-order positions top-down.
#include "lux_internal.h"

namespace {

constexpr uint32_t kBins = 12;
constexpr uint32_t kLeaf = 4;
constexpr uint32_t kMaxDepth = 48;
constexpr uint32_t kSmall = 64;
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kBinWords = 7;                    // count, min xyz, max xyz
constexpr uint32_t kNodeBinWords = 3 * kBins * kBinWords;

enum : uint32_t { kLarge = 0, kSmallRoot = 1, kLeafNode = 2, kSplit = 3, kMedian = 4 };

struct SahRec {
    uint32_t first, count, depth, mesh;   // global first position
    uint32_t kind, child, axis, bin;      // child: record of the left child
    uint32_t mid, front, desc, final_idx; // desc: nodes below; final: index in the mesh
    uint32_t pair, pad0, pad1, pad2;      // pair: index of the children pair
    float box[6];
    float cmin[3], cmax[3];
    double span;
};

// One node of a small subtree (local pre-order): box, first (local pair
// index, or global first position for a leaf), count.
struct SmallNode {
    float box[6];
    uint32_t first, count;
};

__device__ __forceinline__ float gmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float gmax(float a, float b) { return a > b ? a : b; }

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
// both zeros as one key (ties between them resolve by position)
__device__ __forceinline__ uint32_t f2ord_z(float f) { return f2ord(f == 0.0f ? 0.0f : f); }

constexpr uint32_t kOrdInf = 0xff800000u;     // f2ord(+inf)
constexpr uint32_t kOrdNegInf = 0x007fffffu;  // f2ord(-inf)

__device__ __forceinline__ float comp(const float4& v, uint32_t a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ __forceinline__ float centre(const float4& l, const float4& h, uint32_t a) {
    return comp(l, a) * 0.5f + comp(h, a) * 0.5f;
}

__device__ __forceinline__ double area(const float* mn, const float* mx) {
    const double sx = (double)mx[0] - (double)mn[0], sy = (double)mx[1] - (double)mn[1];
    const double sz = (double)mx[2] - (double)mn[2];
    return 2.0 * (sx * sy + sy * sz + sz * sx);
}

__device__ __forceinline__ uint32_t bin_of(float c, float cmin, double span) {
    const double q = ((double)c - (double)cmin) / span * 12.0;
    // a float-to-unsigned cast saturates at 0 (q >= 0 for a centre inside
    // the node's centroid bounds)
    const uint64_t b = q > 0.0 ? (uint64_t)q : 0ull;
    return b < 11 ? (uint32_t)b : 11u;
}

inline uint32_t nblocks(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

__device__ __forceinline__ uint64_t lanes_below(uint32_t k) { return k >= 64 ? ~0ull : ((1ull << k) - 1ull); }

__global__ void k_sah_finite(const float* pos, uint64_t words, uint32_t* bad) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words && !(fabsf(pos[i]) <= 3.402823466e38f)) *bad = 1u;
}

__global__ void k_sah_roots(uint32_t n_seg, const uint32_t* tri_base, const uint32_t* count, SahRec* recs,
                            uint32_t* small_list, uint32_t* small_n) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_seg) return;
    SahRec r = {};
    r.first = tri_base[m];
    r.count = count[m];
    r.mesh = m;
    r.kind = r.count <= kSmall ? kSmallRoot : kLarge;
    r.final_idx = 0;
    r.pair = 1;
    recs[m] = r;
    if (r.kind == kSmallRoot) small_list[atomicAdd(small_n, 1u)] = m;
}

__global__ void k_sah_posmap_init(uint32_t n, const uint32_t* seg_of, const uint32_t* count, uint32_t* pos_node,
                                  uint32_t* perm) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t m = seg_of[p];
    pos_node[p] = count[m] > kSmall ? m : kNone;
    perm[p] = p;
}

__global__ void k_sah_level_init(const SahRec* lv, uint32_t L, unsigned long long* keys, uint32_t* cb,
                                 uint32_t* bins) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per = 6 + 6 + kNodeBinWords;
    if (w >= per * L) return;
    const uint32_t i = (uint32_t)(w / per), k = (uint32_t)(w % per);
    if (lv[i].kind != kLarge) return;
    if (k < 6) {
        keys[6ull * i + k] = k < 3 ? ~0ull : 0ull;
    } else if (k < 12) {
        cb[6ull * i + k - 6] = k - 6 < 3 ? kOrdInf : kOrdNegInf;
    } else {
        const uint32_t b = k - 12, f = b % kBinWords;
        bins[(uint64_t)kNodeBinWords * i + b] = f == 0 ? 0u : (f < 4 ? kOrdInf : kOrdNegInf);
    }
}

// Runs of one node in a block of 256 positions: a node's positions are
// contiguous and a large node holds more than kSmall of them, so a block
// meets at most kRunSlots nodes.  slot: the run of this position (kNone
// for positions outside large nodes); run_v[k]: the node of run k.
constexpr uint32_t kRunSlots = 256 / (kSmall + 1) + 2;

__device__ __forceinline__ uint32_t block_runs(uint32_t v, uint32_t* s_v, uint32_t* s_heads, uint32_t* run_v,
                                               uint32_t* n_runs) {
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    s_v[t] = v;
    __syncthreads();
    const bool head = v != kNone && (t == 0 || s_v[t - 1] != v);
    const uint64_t hm = __ballot(head);
    if (lane == 0) s_heads[wave] = (uint32_t)__popcll(hm);
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += s_heads[w];
    const uint32_t upto = base + (uint32_t)__popcll(hm & lanes_below(lane + 1));
    // more runs than slots cannot happen (see above); such positions would
    // take kNone here and be counted in *n_runs no further
    const uint32_t slot = v != kNone && upto - 1 < kRunSlots ? upto - 1 : kNone;
    if (head && slot != kNone) run_v[slot] = v;
    if (t == 255) *n_runs = upto < kRunSlots ? upto : kRunSlots;
    __syncthreads();
    return slot;
}

// Per active position: the node box keys (value, position) and the
// centroid bounds, gathered per node run in LDS, then one atomic per word
// and run.
__global__ __launch_bounds__(256) void k_sah_bounds(uint32_t n, const uint32_t* pos_node, const uint32_t* perm,
                                                    const float4* lo, const float4* hi,
                                                    unsigned long long* keys, uint32_t* cb) {
    __shared__ uint32_t s_v[256], s_heads[4], s_run_v[kRunSlots], s_nruns;
    __shared__ unsigned long long s_key[kRunSlots][6];
    __shared__ uint32_t s_cb[kRunSlots][6];
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    const uint32_t v = p < n ? pos_node[p] : kNone;
    if (threadIdx.x < 6 * kRunSlots) {
        const uint32_t k = threadIdx.x / 6, a = threadIdx.x % 6;
        s_key[k][a] = a < 3 ? ~0ull : 0ull;
        s_cb[k][a] = a < 3 ? kOrdInf : kOrdNegInf;
    }
    const uint32_t slot = block_runs(v, s_v, s_heads, s_run_v, &s_nruns);
    if (slot != kNone) {
        const uint32_t e = perm[p];
        const float4 l = lo[e], h = hi[e];
#pragma unroll
        for (uint32_t a = 0; a < 3; ++a) {
            atomicMin(&s_key[slot][a], ((unsigned long long)f2ord_z(comp(l, a)) << 32) | (0xffffffffu - p));
            atomicMax(&s_key[slot][3 + a], ((unsigned long long)f2ord_z(comp(h, a)) << 32) | p);
            const uint32_t c = f2ord(centre(l, h, a));
            atomicMin(&s_cb[slot][a], c);
            atomicMax(&s_cb[slot][3 + a], c);
        }
    }
    __syncthreads();
    const uint32_t runs = s_nruns;
    if (threadIdx.x < 6 * runs) {
        const uint32_t k = threadIdx.x / 6, a = threadIdx.x % 6, v0 = s_run_v[k];
        if (a < 3) {
            atomicMin(&keys[6ull * v0 + a], s_key[k][a]);
            atomicMin(&cb[6ull * v0 + a], s_cb[k][a]);
        } else {
            atomicMax(&keys[6ull * v0 + a], s_key[k][a]);
            atomicMax(&cb[6ull * v0 + a], s_cb[k][a]);
        }
    }
}

// Per large node of the level: box from the winning elements, centroid
// bounds, leaf at depth 48.
__global__ void k_sah_node(SahRec* lv, uint32_t L, const unsigned long long* keys, const uint32_t* cb,
                           const uint32_t* perm, const float4* lo, const float4* hi) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L || lv[i].kind != kLarge) return;
    SahRec& r = lv[i];
#pragma unroll
    for (uint32_t a = 0; a < 3; ++a) {
        const uint32_t pmin = 0xffffffffu - (uint32_t)keys[6ull * i + a];
        const uint32_t pmax = (uint32_t)keys[6ull * i + 3 + a];
        r.box[a] = comp(lo[perm[pmin]], a);
        r.box[3 + a] = comp(hi[perm[pmax]], a);
        r.cmin[a] = ord2f(cb[6ull * i + a]);
        r.cmax[a] = ord2f(cb[6ull * i + 3 + a]);
    }
    if (r.depth == kMaxDepth) r.kind = kLeafNode;
}

// Per active position of a large inner node: the element into its bin on
// every axis with a positive span, gathered per node run in LDS, then
// flushed (non-empty bins only) to the node's bins.
__global__ __launch_bounds__(256) void k_sah_bin(uint32_t n, const uint32_t* pos_node, const SahRec* lv,
                                                 const uint32_t* perm, const float4* lo, const float4* hi,
                                                 uint32_t* bins) {
    __shared__ uint32_t s_v[256], s_heads[4], s_run_v[kRunSlots], s_nruns;
    __shared__ uint32_t s_bins[kRunSlots * kNodeBinWords];
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    uint32_t v = p < n ? pos_node[p] : kNone;
    if (v != kNone && lv[v].kind != kLarge) v = kNone;
    for (uint32_t w = threadIdx.x; w < kRunSlots * kNodeBinWords; w += 256u) {
        const uint32_t f = (w % kNodeBinWords) % kBinWords;
        s_bins[w] = f == 0 ? 0u : (f < 4 ? kOrdInf : kOrdNegInf);
    }
    const uint32_t slot = block_runs(v, s_v, s_heads, s_run_v, &s_nruns);
    if (slot != kNone) {
        const SahRec& r = lv[v];
        const uint32_t e = perm[p];
        const float4 l = lo[e], h = hi[e];
        const uint32_t ol[3] = {f2ord(l.x), f2ord(l.y), f2ord(l.z)};
        const uint32_t oh[3] = {f2ord(h.x), f2ord(h.y), f2ord(h.z)};
        for (uint32_t axis = 0; axis < 3; ++axis) {
            const double span = (double)r.cmax[axis] - (double)r.cmin[axis];
            if (!(span > 0.0)) continue;
            const uint32_t b = bin_of(centre(l, h, axis), r.cmin[axis], span);
            uint32_t* bw = s_bins + slot * kNodeBinWords + (axis * kBins + b) * kBinWords;
            atomicAdd(&bw[0], 1u);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                atomicMin(&bw[1 + a], ol[a]);
                atomicMax(&bw[4 + a], oh[a]);
            }
        }
    }
    __syncthreads();
    const uint32_t runs = s_nruns;
    for (uint32_t w = threadIdx.x; w < runs * kNodeBinWords; w += 256u) {
        const uint32_t k = w / kNodeBinWords, bw = w % kNodeBinWords, f = bw % kBinWords;
        if (s_bins[k * kNodeBinWords + bw - f] == 0u) continue;  // an empty bin
        uint32_t* g = bins + (uint64_t)kNodeBinWords * s_run_v[k] + bw;
        const uint32_t val = s_bins[w];
        if (f == 0) atomicAdd(g, val);
        else if (f < 4) atomicMin(g, val);
        else atomicMax(g, val);
    }
}

struct Bin {
    float mn[3], mx[3];
    uint32_t count;
    __device__ void clear() {
        for (int a = 0; a < 3; ++a) {
            mn[a] = INFINITY;
            mx[a] = -INFINITY;
        }
        count = 0;
    }
    __device__ void join(const Bin& o) {
        for (int a = 0; a < 3; ++a) {
            mn[a] = gmin(mn[a], o.mn[a]);
            mx[a] = gmax(mx[a], o.mx[a]);
        }
        count += o.count;
    }
    __device__ double cost() const { return count == 0 ? 0.0 : area(mn, mx) * (double)count; }
};

// Per large inner node: the cheapest split over the bins, or the median.
__global__ void k_sah_split(SahRec* lv, uint32_t L, const uint32_t* bins, uint32_t* any_median) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L || lv[i].kind != kLarge) return;
    SahRec& r = lv[i];
    double best = INFINITY, best_span = 0.0;
    int best_axis = -1;
    uint32_t best_bin = 0;
    for (uint32_t axis = 0; axis < 3; ++axis) {
        const double span = (double)r.cmax[axis] - (double)r.cmin[axis];
        if (span <= 0.0) continue;
        Bin b[kBins], right[kBins];
        const uint32_t* src = bins + (uint64_t)kNodeBinWords * i + axis * kBins * kBinWords;
        for (uint32_t k = 0; k < kBins; ++k) {
            b[k].count = src[k * kBinWords];
            for (int a = 0; a < 3; ++a) {
                b[k].mn[a] = ord2f(src[k * kBinWords + 1 + a]);
                b[k].mx[a] = ord2f(src[k * kBinWords + 4 + a]);
            }
        }
        Bin sum;
        sum.clear();
        for (int k = (int)kBins - 1; k >= 0; --k) {
            sum.join(b[k]);
            right[k] = sum;
        }
        Bin left;
        left.clear();
        for (uint32_t k = 0; k + 1 < kBins; ++k) {
            left.join(b[k]);
            const double cost = left.cost() + right[k + 1].cost();
            if (left.count > 0 && right[k + 1].count > 0 && cost < best) {
                best = cost;
                best_axis = (int)axis;
                best_bin = k;
                best_span = span;
            }
        }
    }
    if (best_axis >= 0) {
        r.kind = kSplit;
        r.axis = (uint32_t)best_axis;
        r.bin = best_bin;
        r.span = best_span;
    } else {
        const float sx = r.cmax[0] - r.cmin[0], sy = r.cmax[1] - r.cmin[1], sz = r.cmax[2] - r.cmin[2];
        r.kind = kMedian;
        r.axis = (sx >= sy && sx >= sz) ? 0u : (sy >= sz ? 1u : 2u);
        r.mid = r.count / 2;
        *any_median = 1u;
    }
}

__device__ __forceinline__ bool goes_left(const SahRec& r, const float4& l, const float4& h) {
    return bin_of(centre(l, h, r.axis), r.cmin[r.axis], r.span) <= r.bin;
}

// Left/right flags of every position (zero outside split nodes), n + 1
// words each for the exclusive scans.
__global__ void k_sah_flags(uint32_t n, const uint32_t* pos_node, const SahRec* lv, const uint32_t* perm,
                            const float4* lo, const float4* hi, uint32_t* fl, uint32_t* fr) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > n) return;
    uint32_t a = 0, b = 0;
    const uint32_t v = p < n ? pos_node[p] : kNone;
    if (v != kNone && lv[v].kind == kSplit) {
        const uint32_t e = perm[p];
        const bool left = goes_left(lv[v], lo[e], hi[e]);
        a = left ? 1u : 0u;
        b = 1u - a;
    }
    fl[p] = a;
    fr[p] = b;
}

__global__ void k_sah_mid(SahRec* lv, uint32_t L, const uint32_t* sl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L || lv[i].kind != kSplit) return;
    SahRec& r = lv[i];
    const uint32_t mid = sl[r.first + r.count] - sl[r.first];
    r.mid = mid;
    r.front = mid == r.count ? mid : mid + (1u - (sl[r.first + mid + 1] - sl[r.first + mid]));
}

__global__ void k_sah_tables(uint32_t n, const uint32_t* pos_node, const SahRec* lv, const uint32_t* sl,
                             const uint32_t* sr, uint32_t* tab_l, uint32_t* tab_r) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t v = pos_node[p];
    if (v == kNone || lv[v].kind != kSplit) return;
    const SahRec& r = lv[v];
    const uint32_t i = p - r.first;
    const bool left = sl[p + 1] - sl[p] != 0u;
    if (i < r.front) {
        if (!left) tab_r[r.first + (sr[p + 1] - sr[r.first]) - 1] = i;
    } else if (left) {
        tab_l[r.first + (sl[r.first + r.count] - sl[p]) - 1] = i;
    }
}

__global__ void k_sah_scatter(uint32_t n, const uint32_t* pos_node, const SahRec* lv, const uint32_t* sl,
                              const uint32_t* sr, const uint32_t* tab_l, const uint32_t* tab_r,
                              const uint32_t* perm, uint32_t* tmp, uint32_t* err) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t v = pos_node[p];
    if (v == kNone || lv[v].kind != kSplit) return;
    const SahRec& r = lv[v];
    const uint32_t i = p - r.first;
    const bool left = sl[p + 1] - sl[p] != 0u;
    uint32_t dst;
    if (i < r.front) {
        if (left) {
            dst = i;
        } else {
            const uint32_t k = sr[p + 1] - sr[r.first];
            dst = k == 1 ? r.count - 1 : tab_l[r.first + k - 2] - 1;
        }
    } else if (left) {
        dst = tab_r[r.first + (sl[r.first + r.count] - sl[p]) - 1];
    } else {
        dst = i - 1;
    }
    if (dst < r.count) tmp[r.first + dst] = perm[p];
    else *err = 1u;  // no destination inside the node: reported, nothing written
}

__global__ void k_sah_copy_back(uint32_t n, const uint32_t* pos_node, const SahRec* lv, const uint32_t* tmp,
                                uint32_t* perm, uint32_t kind) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t v = pos_node[p];
    if (v == kNone || lv[v].kind != kind) return;
    perm[p] = tmp[p];
}

// Median nodes: sort keys (total order of the centre on the node's axis)
// and segment words (the node's first position; every other position its
// own).
__global__ void k_sah_median_keys(uint32_t n, const uint32_t* pos_node, const SahRec* lv, const uint32_t* perm,
                                  const float4* lo, const float4* hi, uint32_t* key, uint32_t* seg) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t v = pos_node[p];
    if (v != kNone && lv[v].kind == kMedian) {
        const uint32_t e = perm[p];
        key[p] = f2ord(centre(lo[e], hi[e], lv[v].axis));
        seg[p] = lv[v].first;
    } else {
        key[p] = 0u;
        seg[p] = p;
    }
}

__global__ void k_sah_gather(const uint32_t* src, const uint32_t* idx, uint32_t n, uint32_t* dst) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dst[j] = src[idx[j]];
}

__global__ void k_sah_children(SahRec* recs, uint32_t lv_begin, uint32_t L, uint32_t next_begin,
                               uint32_t* next_count, uint32_t* small_list, uint32_t* small_n, uint32_t* err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    SahRec& r = recs[lv_begin + i];
    if (r.kind != kSplit && r.kind != kMedian) return;
    if (r.mid == 0 || r.mid >= r.count) {  // both sides hold elements by construction
        *err = 1u;
        r.kind = kLeafNode;
        return;
    }
    const uint32_t c = next_begin + atomicAdd(next_count, 2u);
    r.child = c;
    const uint32_t cnt[2] = {r.mid, r.count - r.mid}, first[2] = {r.first, r.first + r.mid};
    for (uint32_t s = 0; s < 2; ++s) {
        SahRec k = {};
        k.first = first[s];
        k.count = cnt[s];
        k.depth = r.depth + 1;
        k.mesh = r.mesh;
        k.kind = cnt[s] <= kSmall ? kSmallRoot : kLarge;
        recs[c + s] = k;
        if (k.kind == kSmallRoot) small_list[atomicAdd(small_n, 1u)] = c + s;
    }
}

__global__ void k_sah_posmap(uint32_t n, uint32_t* pos_node, const SahRec* lv, const SahRec* recs,
                             uint32_t next_begin) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t v = pos_node[p];
    if (v == kNone) return;
    const SahRec& r = lv[v];
    if (r.kind == kLeafNode) {
        pos_node[p] = kNone;
        return;
    }
    const uint32_t c = p < r.first + r.mid ? r.child : r.child + 1;
    pos_node[p] = recs[c].kind == kLarge ? c - next_begin : kNone;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}
static_assert(kSmall <= 64, "a small subtree holds one element per lane");

// The recursion over one subtree of at most kSmall elements, one wave: lane
// j holds the element at offset j, a node is a run of lanes.  Each node
// computes what the sequential recursion computes -- box by the select fold
// (the wave minimum, its value taken from the last lane holding it), bins in
// LDS, the sweep one lane per axis (the first strictly cheapest (axis, bin)
// in order), the partition's closed form from ballots, the median as a
// stable rank by key -- and nodes go out in local pre-order; the subtree's
// element order returns to perm.
__global__ __launch_bounds__(64) void k_sah_small(SahRec* recs, const uint32_t* small_list, uint32_t small_n,
                                                  uint32_t* perm, const float4* lo, const float4* hi,
                                                  SmallNode* snodes) {
    struct Task {
        uint32_t index, a, cnt, depth;
    };
    __shared__ uint32_t s_e[64];
    __shared__ float4 s_l[64], s_h[64];
    __shared__ uint32_t s_bins[kNodeBinWords];
    __shared__ uint32_t s_tl[64], s_tr[64];
    __shared__ Task s_stack[kSmall + 2];
    if (blockIdx.x >= small_n) return;
    SahRec& root = recs[small_list[blockIdx.x]];
    const uint32_t j = threadIdx.x, first = root.first, count = root.count;
    SmallNode* out = snodes + 2ull * first;
    uint32_t e = 0;
    float4 l = make_float4(0.0f, 0.0f, 0.0f, 0.0f), h = l;
    if (j < count) {
        e = perm[first + j];
        l = lo[e];
        h = hi[e];
    }
    if (j == 0) s_stack[0] = Task{0u, 0u, count, root.depth};
    uint32_t top = 1, next = 1;
    __syncthreads();
    while (top > 0) {
        --top;
        const Task t = s_stack[top];
        const bool act = j >= t.a && j < t.a + t.cnt;
        const uint64_t node_mask = lanes_below(t.a + t.cnt) & ~lanes_below(t.a);
        float box[6];
#pragma unroll
        for (uint32_t a = 0; a < 3; ++a) {
            const uint32_t on = act ? f2ord_z(comp(l, a)) : 0xffffffffu;
            const uint32_t mn = wave_min(on);
            const uint64_t wn = __ballot(act && on == mn);
            box[a] = __shfl(comp(l, a), 63 - __clzll(wn), 64);
            const uint32_t ox = act ? f2ord_z(comp(h, a)) : 0u;
            const uint32_t mx = wave_max(ox);
            const uint64_t wx = __ballot(act && ox == mx);
            box[3 + a] = __shfl(comp(h, a), 63 - __clzll(wx), 64);
        }
        const bool leaf = t.cnt <= kLeaf || t.depth == kMaxDepth;
        if (j == 0) {
            SmallNode& nd = out[t.index];
            for (int a = 0; a < 6; ++a) nd.box[a] = box[a];
            nd.first = leaf ? first + t.a : next;
            nd.count = leaf ? t.cnt : 0u;
        }
        if (leaf) {
            __syncthreads();
            continue;
        }
        float c[3], cmin[3], cmax[3];
#pragma unroll
        for (uint32_t a = 0; a < 3; ++a) {
            c[a] = centre(l, h, a);
            cmin[a] = ord2f(wave_min(act ? f2ord(c[a]) : kOrdInf));
            cmax[a] = ord2f(wave_max(act ? f2ord(c[a]) : kOrdNegInf));
        }
        for (uint32_t w = j; w < kNodeBinWords; w += 64) {
            const uint32_t f = w % kBinWords;
            s_bins[w] = f == 0 ? 0u : (f < 4 ? kOrdInf : kOrdNegInf);
        }
        __syncthreads();
        if (act) {
            const uint32_t ol[3] = {f2ord(l.x), f2ord(l.y), f2ord(l.z)};
            const uint32_t oh[3] = {f2ord(h.x), f2ord(h.y), f2ord(h.z)};
            for (uint32_t axis = 0; axis < 3; ++axis) {
                const double span = (double)cmax[axis] - (double)cmin[axis];
                if (!(span > 0.0)) continue;
                uint32_t* bw = s_bins + (axis * kBins + bin_of(c[axis], cmin[axis], span)) * kBinWords;
                atomicAdd(&bw[0], 1u);
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    atomicMin(&bw[1 + q], ol[q]);
                    atomicMax(&bw[4 + q], oh[q]);
                }
            }
        }
        __syncthreads();
        // lane `axis` sweeps its axis: the cheapest bin, the first of equal cost
        double my_cost = INFINITY;
        uint32_t my_bin = 0;
        if (j < 3) {
            const double span = (double)cmax[j] - (double)cmin[j];
            if (span > 0.0) {
                Bin b[kBins], right[kBins];
                for (uint32_t k = 0; k < kBins; ++k) {
                    const uint32_t* src = s_bins + (j * kBins + k) * kBinWords;
                    b[k].count = src[0];
                    for (int q = 0; q < 3; ++q) {
                        b[k].mn[q] = ord2f(src[1 + q]);
                        b[k].mx[q] = ord2f(src[4 + q]);
                    }
                }
                Bin sum;
                sum.clear();
                for (int k = (int)kBins - 1; k >= 0; --k) {
                    sum.join(b[k]);
                    right[k] = sum;
                }
                Bin left;
                left.clear();
                for (uint32_t k = 0; k + 1 < kBins; ++k) {
                    left.join(b[k]);
                    const double cost = left.cost() + right[k + 1].cost();
                    if (left.count > 0 && right[k + 1].count > 0 && cost < my_cost) {
                        my_cost = cost;
                        my_bin = k;
                    }
                }
            }
        }
        double best = INFINITY;
        int best_axis = -1;
        uint32_t best_bin = 0;
#pragma unroll
        for (int axis = 0; axis < 3; ++axis) {
            const double ca = __shfl(my_cost, axis, 64);
            const uint32_t ba = __shfl(my_bin, axis, 64);
            if (ca < best) {
                best = ca;
                best_axis = axis;
                best_bin = ba;
            }
        }
        const uint32_t i = j - t.a;
        uint32_t mid, dst = i;
        if (best_axis >= 0) {
            const double span = (double)cmax[best_axis] - (double)cmin[best_axis];
            const bool left = act && bin_of(c[best_axis], cmin[best_axis], span) <= best_bin;
            const uint64_t lm = __ballot(left), rm = node_mask & ~lm;
            mid = (uint32_t)__popcll(lm);
            const uint32_t front = mid == t.cnt ? mid : mid + (uint32_t)(((lm >> (t.a + mid)) & 1ull) ^ 1ull);
            const uint64_t front_mask = node_mask & lanes_below(t.a + front);
            uint32_t kr = 0, kl = 0;
            if (act) {
                if (i < front && !left) {
                    kr = (uint32_t)__popcll(rm & front_mask & lanes_below(j + 1));
                    s_tr[kr - 1] = i;
                } else if (i >= front && left) {
                    kl = (uint32_t)__popcll(lm & ~front_mask & ~lanes_below(j));
                    s_tl[kl - 1] = i;
                }
            }
            __syncthreads();
            if (act) {
                if (i < front) dst = left ? i : (kr == 1 ? t.cnt - 1 : s_tl[kr - 2] - 1);
                else dst = left ? s_tr[kl - 1] : i - 1;
            }
        } else {
            const float sx = cmax[0] - cmin[0], sy = cmax[1] - cmin[1], sz = cmax[2] - cmin[2];
            const uint32_t axis = (sx >= sy && sx >= sz) ? 0u : (sy >= sz ? 1u : 2u);
            const uint32_t key = f2ord(c[axis]);
            s_tl[j] = key;
            __syncthreads();
            if (act) {
                uint32_t r = 0;
                for (uint32_t q = t.a; q < t.a + t.cnt; ++q) {
                    const uint32_t kq = s_tl[q];
                    r += (kq < key || (kq == key && q < j)) ? 1u : 0u;
                }
                dst = r;
            }
            mid = t.cnt / 2;
        }
        // move the node's elements to their places
        if (act) {
            s_e[t.a + dst] = e;
            s_l[t.a + dst] = l;
            s_h[t.a + dst] = h;
        }
        __syncthreads();
        if (act) {
            e = s_e[j];
            l = s_l[j];
            h = s_h[j];
        }
        const uint32_t children = next;
        next += 2;
        // the left subtree runs first, as in the recursion
        if (j == 0) {
            s_stack[top] = Task{children + 1, t.a + mid, t.cnt - mid, t.depth + 1};
            s_stack[top + 1] = Task{children, t.a, mid, t.depth + 1};
        }
        top += 2;
        __syncthreads();
    }
    if (j < count) perm[first + j] = e;
    if (j == 0) root.desc = next - 1;
}

__global__ void k_sah_desc(SahRec* recs, uint32_t lv_begin, uint32_t L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    SahRec& r = recs[lv_begin + i];
    if (r.kind == kLeafNode) r.desc = 0;
    else if (r.kind == kSplit || r.kind == kMedian) r.desc = 2 + recs[r.child].desc + recs[r.child + 1].desc;
}

__global__ void k_sah_number(SahRec* recs, uint32_t lv_begin, uint32_t L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    const SahRec& r = recs[lv_begin + i];
    if (r.kind != kSplit && r.kind != kMedian) return;
    SahRec& a = recs[r.child];
    SahRec& b = recs[r.child + 1];
    a.final_idx = r.pair;
    a.pair = r.pair + 2;
    b.final_idx = r.pair + 1;
    b.pair = r.pair + 2 + a.desc;
}

__device__ __forceinline__ void put(float4* nodes, uint64_t slot, const float* box, uint32_t first,
                                    uint32_t count) {
    nodes[2 * slot] = make_float4(box[0], box[1], box[2], __uint_as_float(first));
    nodes[2 * slot + 1] = make_float4(box[3], box[4], box[5], __uint_as_float(count));
}

// Node slots are checked against the mesh's 2n - 1 region (a numbering
// error is reported, not written).
__global__ void k_sah_write(const SahRec* recs, uint32_t n_recs, const uint32_t* blas_base, const uint32_t* count,
                            float4* nodes, uint32_t* err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_recs) return;
    const SahRec& r = recs[i];
    if (r.kind == kSmallRoot) return;
    const uint32_t bb = blas_base[r.mesh], region = 2 * count[r.mesh] - 1;
    if (r.final_idx >= region || (r.kind != kLeafNode && r.pair + 1 >= region)) {
        *err = 1u;
        return;
    }
    if (r.kind == kLeafNode) put(nodes, bb + r.final_idx, r.box, r.first, r.count);
    else put(nodes, bb + r.final_idx, r.box, bb + r.pair, 0u);
}

__global__ void k_sah_write_small(const SahRec* recs, const uint32_t* small_list, uint32_t small_n,
                                  const SmallNode* snodes, const uint32_t* blas_base, const uint32_t* count,
                                  float4* nodes, uint32_t* err) {
    const uint32_t s = blockIdx.x;
    if (s >= small_n) return;
    const SahRec& r = recs[small_list[s]];
    const SmallNode* in = snodes + 2ull * r.first;
    const uint32_t bb = blas_base[r.mesh], region = 2 * count[r.mesh] - 1;
    if (r.desc >= 2 * r.count) {  // at most 2 count - 1 nodes
        if (threadIdx.x == 0) *err = 1u;
        return;
    }
    for (uint32_t k = threadIdx.x; k <= r.desc; k += blockDim.x) {
        const SmallNode& nd = in[k];
        const uint32_t slot = k == 0 ? r.final_idx : r.pair + k - 1;
        if (slot >= region || (nd.count == 0 && r.pair + nd.first >= region)) {
            *err = 1u;
            continue;
        }
        if (nd.count) put(nodes, bb + slot, nd.box, nd.first, nd.count);
        else put(nodes, bb + slot, nd.box, bb + r.pair + nd.first - 1, 0u);
    }
}

__global__ void k_sah_zero(float4* nodes, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) nodes[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace

int lux_blas_build_sah(lux_ctx* ctx, const float* pos, const uint32_t* mat, uint32_t n, uint32_t n_seg,
                       const uint32_t* seg_info, uint32_t n_blas, float4* nodes, float4* tris, uint32_t* order) {
    if (n == 0 || n_seg == 0) return LUX_OK;
    hipStream_t st = ctx->stream;
    const uint32_t* blas_base = seg_info;
    const uint32_t* tri_base = seg_info + n_seg;
    const uint32_t* count = seg_info + 2ull * n_seg;
    // records: the mesh roots and two per inner large node.  Large nodes
    // of one level are disjoint with more than kSmall elements, so at most
    // n / (kSmall + 1) of them are a tree's lowest large nodes, and each has
    // at most kMaxDepth large ancestors.
    const uint64_t rec_cap = (uint64_t)n_seg + 2ull * (kMaxDepth + 1) * (n / (kSmall + 1)) + 64;
    // nodes of one level: at most two per large node of the level above
    const uint64_t lv_cap = (uint64_t)n_seg + 2ull * (n / (kSmall + 1)) + 2;
    const size_t hw = lux_sort_hist_words(n);
    const size_t sw = lux_scan_sums_words(n + 1);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t b_recs = al(sizeof(SahRec) * rec_cap);
    const size_t b_keys = al(8ull * 6 * lv_cap);
    const size_t b_cb = al(4ull * 6 * lv_cap);
    const size_t b_bins = al(4ull * kNodeBinWords * lv_cap);
    const size_t b_f4 = al(16ull * n);
    const size_t b_u = al(4ull * (n + 1));
    const size_t b_snodes = al(sizeof(SmallNode) * 2ull * n);
    const size_t b_sums = al(4ull * sw);
    const size_t b_hist = al(4ull * hw);
    const size_t total = b_recs + b_keys + b_cb + b_bins + 2 * b_f4 + 16 * b_u + b_snodes + 2 * b_sums + b_hist + 256;
    if (int rc = lux_ensure(ctx, ctx->scratch, total)) return rc;
    char* w = static_cast<char*>(ctx->scratch.ptr);
    auto take = [&](size_t b) {
        char* p = w;
        w += b;
        return p;
    };
    SahRec* recs = reinterpret_cast<SahRec*>(take(b_recs));
    auto* keys = reinterpret_cast<unsigned long long*>(take(b_keys));
    auto* cb = reinterpret_cast<uint32_t*>(take(b_cb));
    auto* bins = reinterpret_cast<uint32_t*>(take(b_bins));
    auto* lo = reinterpret_cast<float4*>(take(b_f4));
    auto* hi = reinterpret_cast<float4*>(take(b_f4));
    auto* seg_of = reinterpret_cast<uint32_t*>(take(b_u));
    auto* pos_node = reinterpret_cast<uint32_t*>(take(b_u));
    auto* perm = reinterpret_cast<uint32_t*>(take(b_u));
    auto* tmp = reinterpret_cast<uint32_t*>(take(b_u));
    auto* fl = reinterpret_cast<uint32_t*>(take(b_u));
    auto* fr = reinterpret_cast<uint32_t*>(take(b_u));
    auto* tab_l = reinterpret_cast<uint32_t*>(take(b_u));
    auto* tab_r = reinterpret_cast<uint32_t*>(take(b_u));
    auto* key = reinterpret_cast<uint32_t*>(take(b_u));
    auto* segw = reinterpret_cast<uint32_t*>(take(b_u));
    auto* sk = reinterpret_cast<uint32_t*>(take(b_u));
    auto* idx1 = reinterpret_cast<uint32_t*>(take(b_u));
    auto* idx2 = reinterpret_cast<uint32_t*>(take(b_u));
    auto* tk = reinterpret_cast<uint32_t*>(take(b_u));
    auto* tv = reinterpret_cast<uint32_t*>(take(b_u));
    auto* small_list = reinterpret_cast<uint32_t*>(take(b_u));
    auto* snodes = reinterpret_cast<SmallNode*>(take(b_snodes));
    auto* sums_l = reinterpret_cast<uint32_t*>(take(b_sums));
    auto* sums_r = reinterpret_cast<uint32_t*>(take(b_sums));
    auto* hist = reinterpret_cast<uint32_t*>(take(b_hist));
    auto* ctl = reinterpret_cast<uint32_t*>(take(64));  // 0 next count, 1 small count, 2 any median, 3 non-finite

    // per-segment centre bounds are not used here; the LBVH bounds words go
    // to fl (6 per segment fit in n + 1 words only when n_seg is small, so
    // keep them in the bins area, which the first level re-initialises)
    if (6ull * n_seg > (uint64_t)kNodeBinWords * lv_cap)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_bvh_build: SAH scratch layout");
    LUX_HIP(ctx, hipMemsetAsync(ctl, 0, 64, st));
    hipLaunchKernelGGL(k_sah_finite, dim3(nblocks(9ull * n, 256)), dim3(256), 0, st, pos, 9ull * n, ctl + 3);
    if (int rc = lux_seg_prepare(ctx, pos, n, tri_base, n_seg, lo, hi, seg_of, bins)) return rc;
    hipLaunchKernelGGL(k_sah_roots, dim3(nblocks(n_seg, 256)), dim3(256), 0, st, n_seg, tri_base, count, recs,
                       small_list, ctl + 1);
    hipLaunchKernelGGL(k_sah_posmap_init, dim3(nblocks(n, 256)), dim3(256), 0, st, n, seg_of, count, pos_node, perm);
    LUX_HIP(ctx, hipGetLastError());
    uint32_t h_ctl[4];
    LUX_HIP(ctx, hipMemcpyAsync(h_ctl, ctl, 16, hipMemcpyDeviceToHost, st));
    LUX_HIP(ctx, hipStreamSynchronize(st));
    if (h_ctl[3]) return lux_fail(ctx, LUX_ERR_INVALID, "lux_bvh_build: SAH needs finite triangle positions");

    std::vector<uint32_t> lv_begin{0}, lv_size{n_seg};
    uint32_t lv = 0;
    const uint32_t nb = nblocks(n, 256), nb1 = nblocks(n + 1ull, 256);
    while (true) {
        const uint32_t b0 = lv_begin[lv], L = lv_size[lv];
        SahRec* lvr = recs + b0;
        const uint32_t next_begin = b0 + L;
        if (next_begin + 2ull * L > rec_cap || L > lv_cap)
            return lux_fail(ctx, LUX_ERR_INVALID, "lux_bvh_build: SAH node records exceed their bound");
        const uint64_t init_words = (uint64_t)(12 + kNodeBinWords) * L;
        hipLaunchKernelGGL(k_sah_level_init, dim3(nblocks(init_words, 256)), dim3(256), 0, st, lvr, L, keys, cb, bins);
        hipLaunchKernelGGL(k_sah_bounds, dim3(nb), dim3(256), 0, st, n, pos_node, perm, lo, hi, keys, cb);
        hipLaunchKernelGGL(k_sah_node, dim3(nblocks(L, 128)), dim3(128), 0, st, lvr, L, keys, cb, perm, lo, hi);
        hipLaunchKernelGGL(k_sah_bin, dim3(nb), dim3(256), 0, st, n, pos_node, lvr, perm, lo, hi, bins);
        hipLaunchKernelGGL(k_sah_split, dim3(nblocks(L, 64)), dim3(64), 0, st, lvr, L, bins, ctl + 2);
        // split nodes: the closed-form partition
        hipLaunchKernelGGL(k_sah_flags, dim3(nb1), dim3(256), 0, st, n, pos_node, lvr, perm, lo, hi, fl, fr);
        if (int rc = lux_scan_exclusive_on(ctx, st, fl, n + 1, sums_l, sums_l + sw - 1)) return rc;
        if (int rc = lux_scan_exclusive_on(ctx, st, fr, n + 1, sums_r, sums_r + sw - 1)) return rc;
        hipLaunchKernelGGL(k_sah_mid, dim3(nblocks(L, 128)), dim3(128), 0, st, lvr, L, fl);
        hipLaunchKernelGGL(k_sah_tables, dim3(nb), dim3(256), 0, st, n, pos_node, lvr, fl, fr, tab_l, tab_r);
        hipLaunchKernelGGL(k_sah_scatter, dim3(nb), dim3(256), 0, st, n, pos_node, lvr, fl, fr, tab_l, tab_r, perm,
                           tmp, ctl + 4);
        hipLaunchKernelGGL(k_sah_copy_back, dim3(nb), dim3(256), 0, st, n, pos_node, lvr, tmp, perm, (uint32_t)kSplit);
        LUX_HIP(ctx, hipGetLastError());
        LUX_HIP(ctx, hipMemcpyAsync(h_ctl, ctl, 16, hipMemcpyDeviceToHost, st));
        LUX_HIP(ctx, hipStreamSynchronize(st));
        if (h_ctl[2]) {
            // median nodes: stable by key, then stable by node (every other
            // position is a node of its own and stays)
            hipLaunchKernelGGL(k_sah_median_keys, dim3(nb), dim3(256), 0, st, n, pos_node, lvr, perm, lo, hi, key,
                               segw);
            if (int rc = lux_sort_pairs_impl(ctx, key, n, sk, idx1, tk, tv, hist, 32)) return rc;
            hipLaunchKernelGGL(k_sah_gather, dim3(nb), dim3(256), 0, st, segw, idx1, n, key);
            const uint32_t bits = n > 1 ? 32u - (uint32_t)__builtin_clz(n - 1u) : 1u;
            if (int rc = lux_sort_pairs_vals_impl(ctx, key, idx1, n, sk, idx2, tk, tv, hist, bits)) return rc;
            hipLaunchKernelGGL(k_sah_gather, dim3(nb), dim3(256), 0, st, perm, idx2, n, tmp);
            hipLaunchKernelGGL(k_sah_copy_back, dim3(nb), dim3(256), 0, st, n, pos_node, lvr, tmp, perm,
                               (uint32_t)kMedian);
            LUX_HIP(ctx, hipMemsetAsync(ctl + 2, 0, 4, st));
        }
        LUX_HIP(ctx, hipMemsetAsync(ctl, 0, 4, st));
        hipLaunchKernelGGL(k_sah_children, dim3(nblocks(L, 128)), dim3(128), 0, st, recs, b0, L, next_begin, ctl,
                           small_list, ctl + 1, ctl + 4);
        hipLaunchKernelGGL(k_sah_posmap, dim3(nb), dim3(256), 0, st, n, pos_node, lvr, recs, next_begin);
        LUX_HIP(ctx, hipGetLastError());
        LUX_HIP(ctx, hipMemcpyAsync(h_ctl, ctl, 16, hipMemcpyDeviceToHost, st));
        LUX_HIP(ctx, hipStreamSynchronize(st));
        if (h_ctl[0] == 0) break;
        lv_begin.push_back(next_begin);
        lv_size.push_back(h_ctl[0]);
        ++lv;
    }
    const uint32_t n_recs = lv_begin.back() + lv_size.back();
    const uint32_t small_n = h_ctl[1];
    if (small_n)
        hipLaunchKernelGGL(k_sah_small, dim3(small_n), dim3(64), 0, st, recs, small_list, small_n, perm, lo, hi,
                           snodes);
    for (int k = (int)lv; k >= 0; --k)
        hipLaunchKernelGGL(k_sah_desc, dim3(nblocks(lv_size[k], 128)), dim3(128), 0, st, recs, lv_begin[k],
                           lv_size[k]);
    for (uint32_t k = 0; k <= lv; ++k)
        hipLaunchKernelGGL(k_sah_number, dim3(nblocks(lv_size[k], 128)), dim3(128), 0, st, recs, lv_begin[k],
                           lv_size[k]);
    hipLaunchKernelGGL(k_sah_zero, dim3(nblocks(2ull * n_blas, 256)), dim3(256), 0, st, nodes, 2ull * n_blas);
    hipLaunchKernelGGL(k_sah_write, dim3(nblocks(n_recs, 128)), dim3(128), 0, st, recs, n_recs, blas_base, count,
                       nodes, ctl + 4);
    if (small_n)
        hipLaunchKernelGGL(k_sah_write_small, dim3(small_n), dim3(64), 0, st, recs, small_list, small_n, snodes,
                           blas_base, count, nodes, ctl + 4);
    LUX_HIP(ctx, hipGetLastError());
    if (int rc = lux_seg_gather(ctx, pos, mat, perm, seg_of, tri_base, n, tris, order)) return rc;
    LUX_HIP(ctx, hipMemcpyAsync(h_ctl, ctl + 4, 4, hipMemcpyDeviceToHost, st));
    LUX_HIP(ctx, hipStreamSynchronize(st));
    if (h_ctl[0]) return lux_fail(ctx, LUX_ERR_HIP, "lux_bvh_build: SAH build produced an inconsistent tree");
    return LUX_OK;
}

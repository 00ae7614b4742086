// lux_bvh.hip -- Morton-30 LBVH build on gfx950.
//
// Replaces the reference's CPU binned-SAH build (src/scene.rs:2374-2554)
// with a GPU build whose primitive order obeys the reference's sort
// contract: stable by key, ties by original index (gpuSortLess,
// shaders/split.slang:3-10).  Pipeline per primitive set:
//   1. centre bounds       atomic min/max on order-preserving integer images
//   2. Morton codes        10 bits/axis of c = lo*0.5 + hi*0.5 (scene.rs:2461)
//   3. stable LSD radix    8-bit digits (4 passes for 30-bit codes), wave-ballot
//                          ranking (wave64), multi-block histogram scan
//   4. Karras radix tree   one thread per internal node
//   5. bottom-up boxes     one thread per leaf, atomic arrival counters
//   6. node emission       reference BvhNode layout; internal node k keeps its
//                          children in slots 1+2k, 2+2k; subtrees of at most
//                          leaf_max primitives become leaves
// Box unions use explicit compare-selects so that results never depend on
// the sign-of-zero behaviour of min/max instructions.
#include "lux_internal.h"

namespace {

constexpr int kRsThreads = 256;
constexpr int kRsChunks = 8;
constexpr int kRsTile = kRsThreads * kRsChunks;
constexpr int kRsBins = 256;

__device__ __forceinline__ float sel_min(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float sel_max(float a, float b) { return b > a ? b : a; }

// Order-preserving map float -> uint32 (for atomic min/max).
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// bounds[0..2] = min (ordered), bounds[3..5] = max (ordered); preset by caller.
__global__ __launch_bounds__(256) void k_centre_bounds(const float4* lo, const float4* hi,
                                                       uint32_t n, uint32_t* bounds) {
    __shared__ uint32_t smin[3], smax[3];
    if (threadIdx.x < 3) {
        smin[threadIdx.x] = 0xffffffffu;
        smax[threadIdx.x] = 0u;
    }
    __syncthreads();
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 l = lo[i], h = hi[i];
        float c[3] = {l.x * 0.5f + h.x * 0.5f, l.y * 0.5f + h.y * 0.5f, l.z * 0.5f + h.z * 0.5f};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (c[a] != c[a]) continue;  // NaN centres do not widen the bounds
            uint32_t o = f2ord(c[a]);
            mn[a] = o < mn[a] ? o : mn[a];
            mx[a] = o > mx[a] ? o : mx[a];
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        atomicMin(&smin[a], mn[a]);
        atomicMax(&smax[a], mx[a]);
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        atomicMin(&bounds[threadIdx.x], smin[threadIdx.x]);
        atomicMax(&bounds[3 + threadIdx.x], smax[threadIdx.x]);
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ uint32_t quantize10(float c, float lo, float ext) {
    if (!(ext > 0.0f)) return 0u;
    float q = (c - lo) / ext * 1024.0f;
    if (!(q > 0.0f)) return 0u;
    if (q >= 1023.0f) return 1023u;
    return (uint32_t)q;
}

// Morton code of the box centre inside bounds[0..5] (ordered images).
__device__ __forceinline__ uint32_t morton_code(float4 l, float4 h, const uint32_t* bounds) {
    float cmin[3], ext[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float mnv = bounds[a] == 0xffffffffu ? 0.0f : ord2f(bounds[a]);
        float mxv = bounds[3 + a] == 0u ? 0.0f : ord2f(bounds[3 + a]);
        cmin[a] = mnv;
        ext[a] = mxv - mnv;
    }
    float c0 = l.x * 0.5f + h.x * 0.5f, c1 = l.y * 0.5f + h.y * 0.5f, c2 = l.z * 0.5f + h.z * 0.5f;
    uint32_t q0 = quantize10(c0, cmin[0], ext[0]);
    uint32_t q1 = quantize10(c1, cmin[1], ext[1]);
    uint32_t q2 = quantize10(c2, cmin[2], ext[2]);
    return (spread10(q0) << 2) | (spread10(q1) << 1) | spread10(q2);
}

__global__ __launch_bounds__(256) void k_morton(const float4* lo, const float4* hi, uint32_t n,
                                                const uint32_t* bounds, uint32_t* codes) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    codes[i] = morton_code(lo[i], hi[i], bounds);
}

// ---------------- stable LSD radix sort, 8-bit digits ----------------
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t* keys, uint32_t n,
                                                        uint32_t shift, uint32_t nblocks,
                                                        uint32_t* hist) {
    __shared__ uint32_t h[kRsBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kRsTile;
#pragma unroll
    for (int c = 0; c < kRsChunks; ++c) {
        uint32_t i = base + c * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// ---------------- exclusive scan (multi-block) ----------------
constexpr uint32_t kScanBlock = 1024;

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

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint32_t* keys_in,
                                                           const uint32_t* vals_in, uint32_t* keys_out,
                                                           uint32_t* vals_out, uint32_t n,
                                                           uint32_t shift, uint32_t nblocks,
                                                           const uint32_t* hist) {
    __shared__ uint32_t base[kRsBins];
    __shared__ uint32_t wcount[kRsThreads / 64][kRsBins];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    base[tid] = hist[tid * nblocks + blockIdx.x];
    const uint64_t lt_mask = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    for (int c = 0; c < kRsChunks; ++c) {
        for (int w = 0; w < kRsThreads / 64; ++w) wcount[w][tid] = 0;
        __syncthreads();
        uint32_t i = blockIdx.x * kRsTile + c * kRsThreads + tid;
        bool valid = i < n;
        uint32_t key = valid ? keys_in[i] : 0u;
        uint32_t val = valid ? (vals_in ? vals_in[i] : i) : 0u;
        uint32_t digit = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bits = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? bits : ~bits;
        }
        uint32_t rank = __popcll(peers & lt_mask);
        uint32_t cnt = __popcll(peers);
        bool leader = valid && rank == 0;
        if (leader) wcount[wave][digit] = cnt;
        __syncthreads();
        if (valid) {
            uint32_t prefix = 0;
            for (uint32_t w = 0; w < wave; ++w) prefix += wcount[w][digit];
            uint32_t pos = base[digit] + prefix + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int w = 0; w < kRsThreads / 64; ++w) tot += wcount[w][tid];
        base[tid] += tot;
        __syncthreads();
    }
}

// ---------------- Karras tree ----------------
__device__ __forceinline__ int kdelta(const uint32_t* k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

// Internal node ii of one primitive set's radix tree (keys sk[0..n), ties
// broken by index): gamma, range lo/hi, and the parents of its children.
// Every pointer is relative to the set, so one set of a segmented build
// runs exactly the code of a single build.
__device__ __forceinline__ void karras_node(const uint32_t* sk, uint32_t n, uint32_t ii, uint32_t* gamma_out,
                                            uint32_t* rlo, uint32_t* rhi, uint32_t* parent_leaf,
                                            uint32_t* parent_int) {
    int64_t i = ii, N = n;
    int d = kdelta(sk, N, i, i + 1) - kdelta(sk, N, i, i - 1) > 0 ? 1 : -1;
    int dmin = kdelta(sk, N, i, i - d);
    int64_t lmax = 2;
    while (kdelta(sk, N, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(sk, N, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = kdelta(sk, N, i, j);
    int64_t s = 0, div = 2, t;
    do {
        t = (l + div - 1) / div;
        if (kdelta(sk, N, i, i + (s + t) * d) > dnode) s += t;
        div *= 2;
    } while (t > 1);
    int64_t g = i + s * d + (d < 0 ? -1 : 0);
    uint32_t lo = (uint32_t)(i < j ? i : j), hi = (uint32_t)(i < j ? j : i);
    gamma_out[ii] = (uint32_t)g;
    rlo[ii] = lo;
    rhi[ii] = hi;
    if ((uint32_t)g == lo) parent_leaf[g] = ii;
    else parent_int[g] = ii;
    if ((uint32_t)g + 1 == hi) parent_leaf[g + 1] = ii;
    else parent_int[g + 1] = ii;
}

// children[k] = gamma; range lo/hi; parents of leaves/internal nodes.
__global__ __launch_bounds__(256) void k_karras(const uint32_t* sk, uint32_t n, uint32_t* gamma_out,
                                                uint32_t* rlo, uint32_t* rhi, uint32_t* parent_leaf,
                                                uint32_t* parent_int) {
    uint32_t ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii + 1 >= n) return;
    karras_node(sk, n, ii, gamma_out, rlo, rhi, parent_leaf, parent_int);
}

__device__ __forceinline__ float4 ld_relaxed(const float4* p) {
    const uint32_t* u = reinterpret_cast<const uint32_t*>(p);
    float4 r;
    r.x = __uint_as_float(__hip_atomic_load(u + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    r.y = __uint_as_float(__hip_atomic_load(u + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    r.z = __uint_as_float(__hip_atomic_load(u + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    r.w = 0.0f;
    return r;
}
__device__ __forceinline__ void st_relaxed(float4* p, float4 v) {
    uint32_t* u = reinterpret_cast<uint32_t*>(p);
    __hip_atomic_store(u + 0, __float_as_uint(v.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(u + 1, __float_as_uint(v.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(u + 2, __float_as_uint(v.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bottom-up boxes: every sorted leaf walks up from its parent; the second
// arrival at a node unions its children (left first) and continues.  The
// sorted positions are cut into tiles of at most kBuTile leaves, one
// workgroup each (two leaves per thread).  A node whose leaf range lies
// inside its tile only ever sees arrivals from that workgroup: its counter,
// records and boxes live in LDS and order at workgroup scope.  A node
// reaching outside the tile (and every node above it) counts in global
// memory with agent-scope acquire/release RMWs, its boxes stored and loaded
// as agent-coherent atomics.  Every box is also stored to plo/phi/ilo/ihi
// for node emission.  Seg: sets are segments of the sorted positions
// (seg_of, tri_base, count; sets of at most leaf_max are skipped) and tile k
// covers [tiles[k], tiles[k+1]) with no tile boundary other than at a
// segment start or kBuTile multiples into one segment; otherwise one set of
// n_all in tiles of kBuTile.  Indices inside a set are local, as in a single
// build of that set.
constexpr uint32_t kBuThreads = 1024;
constexpr uint32_t kBuTile = 2 * kBuThreads;

struct Box3 {
    float x, y, z;
};
__device__ __forceinline__ Box3 box3(float4 v) { return Box3{v.x, v.y, v.z}; }
__device__ __forceinline__ float4 box4(Box3 v) { return make_float4(v.x, v.y, v.z, 0.0f); }

struct BuTile {
    Box3 plo[kBuTile], phi[kBuTile], ilo[kBuTile], ihi[kBuTile];
    uint32_t cnt[kBuTile], gamma[kBuTile], rlo[kBuTile], rhi[kBuTile], pint[kBuTile];
};

// The walk of sorted leaf j (tile slot j - start) of a set at tb of n.
__device__ __forceinline__ void bu_walk(BuTile& T, uint32_t j, uint32_t start, uint32_t len, uint32_t tb,
                                        const uint32_t* gamma, const uint32_t* rlo, const uint32_t* rhi,
                                        const uint32_t* parent_leaf, const uint32_t* parent_int, float4* plo,
                                        float4* phi, float4* ilo, float4* ihi, uint32_t* arrivals) {
    // the set's arrays, indexed by local node / leaf index
    const uint32_t* g_gamma = gamma + tb;
    const uint32_t* g_rlo = rlo + tb;
    const uint32_t* g_rhi = rhi + tb;
    const uint32_t* g_pint = parent_int + tb;
    float4* g_plo = plo + tb;
    float4* g_phi = phi + tb;
    float4* g_ilo = ilo + tb;
    float4* g_ihi = ihi + tb;
    // local index x of the set sits at tile slot tb + x - start
    const int64_t off = (int64_t)tb - (int64_t)start, L = (int64_t)len;
    uint32_t p = parent_leaf[j];
    bool in_tile = true;
    while (true) {
        // p lies in its own range, so a node inside the tile has its records
        // in the tile
        const int64_t sp = off + (int64_t)p;
        const bool p_here = in_tile && sp >= 0 && sp < L;
        const uint32_t r0 = p_here ? T.rlo[sp] : g_rlo[p];
        const uint32_t r1 = p_here ? T.rhi[sp] : g_rhi[p];
        const uint32_t g = p_here ? T.gamma[sp] : g_gamma[p];
        in_tile = p_here && off + (int64_t)r0 >= 0 && off + (int64_t)r1 < L;
        float4 alo, ahi, blo, bhi;
        if (in_tile) {
            const uint32_t old = __hip_atomic_fetch_add(&T.cnt[sp], 1u, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0u) return;
            if (g == r0) { alo = box4(T.plo[off + g]); ahi = box4(T.phi[off + g]); }
            else { alo = box4(T.ilo[off + g]); ahi = box4(T.ihi[off + g]); }
            if (g + 1 == r1) { blo = box4(T.plo[off + g + 1]); bhi = box4(T.phi[off + g + 1]); }
            else { blo = box4(T.ilo[off + g + 1]); bhi = box4(T.ihi[off + g + 1]); }
        } else {
            const uint32_t old = __hip_atomic_fetch_add(&arrivals[tb + p], 1u, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            if (old == 0u) return;
            if (g == r0) { alo = ld_relaxed(&g_plo[g]); ahi = ld_relaxed(&g_phi[g]); }
            else { alo = ld_relaxed(&g_ilo[g]); ahi = ld_relaxed(&g_ihi[g]); }
            if (g + 1 == r1) { blo = ld_relaxed(&g_plo[g + 1]); bhi = ld_relaxed(&g_phi[g + 1]); }
            else { blo = ld_relaxed(&g_ilo[g + 1]); bhi = ld_relaxed(&g_ihi[g + 1]); }
        }
        float4 olo = make_float4(sel_min(alo.x, blo.x), sel_min(alo.y, blo.y), sel_min(alo.z, blo.z), 0.0f);
        float4 ohi = make_float4(sel_max(ahi.x, bhi.x), sel_max(ahi.y, bhi.y), sel_max(ahi.z, bhi.z), 0.0f);
        if (in_tile) {
            T.ilo[sp] = box3(olo);
            T.ihi[sp] = box3(ohi);
        }
        st_relaxed(&g_ilo[p], olo);
        st_relaxed(&g_ihi[p], ohi);
        if (p == 0) return;
        p = p_here ? T.pint[sp] : g_pint[p];
    }
}

template <bool Seg>
__global__ __launch_bounds__(kBuThreads) void k_bottom_up_t(const float4* lo, const float4* hi, const uint32_t* perm,
                                                            uint32_t n_all, const uint32_t* tiles,
                                                            const uint32_t* seg_of, const uint32_t* tri_base,
                                                            const uint32_t* count, uint32_t leaf_max,
                                                            const uint32_t* gamma, const uint32_t* rlo,
                                                            const uint32_t* rhi, const uint32_t* parent_leaf,
                                                            const uint32_t* parent_int, float4* plo, float4* phi,
                                                            float4* ilo, float4* ihi, uint32_t* arrivals) {
    __shared__ BuTile T;
    const uint32_t t = threadIdx.x;
    uint32_t start, end;
    if (Seg) {
        start = tiles[blockIdx.x];
        end = tiles[blockIdx.x + 1];
    } else {
        start = blockIdx.x * kBuTile;
        end = n_all - start < kBuTile ? n_all : start + kBuTile;
    }
    const uint32_t len = end - start;
    // the tile's counters, tree records and leaf boxes (slot x = position
    // start + x; positions that are no internal node of a set hold records
    // that no in-tile lookup reads)
#pragma unroll
    for (uint32_t r = 0; r < 2; ++r) {
        const uint32_t x = t + r * kBuThreads;
        T.cnt[x] = 0u;
        if (x >= len) continue;
        const uint32_t j = start + x;
        T.gamma[x] = gamma[j];
        T.rlo[x] = rlo[j];
        T.rhi[x] = rhi[j];
        T.pint[x] = parent_int[j];
        const uint32_t src = perm[j];
        float4 l = lo[src], h = hi[src];
        l.w = 0.0f;
        h.w = 0.0f;
        T.plo[x] = box3(l);
        T.phi[x] = box3(h);
        st_relaxed(&plo[j], l);
        st_relaxed(&phi[j], h);
    }
    __syncthreads();
#pragma unroll 1
    for (uint32_t r = 0; r < 2; ++r) {
        const uint32_t x = t + r * kBuThreads;
        if (x >= len) break;
        const uint32_t j = start + x;
        uint32_t tb = 0, n = n_all;
        if (Seg) {
            const uint32_t m = seg_of[j];
            tb = tri_base[m];
            n = count[m];
            if (n <= leaf_max) continue;
        }
        if (n < 2) continue;
        bu_walk(T, j, start, len, tb, gamma, rlo, rhi, parent_leaf, parent_int, plo, phi, ilo, ihi, arrivals);
    }
}

__device__ __forceinline__ void put_node(float4* nodes, uint32_t slot, float4 mn, float4 mx,
                                         uint32_t first, uint32_t count) {
    nodes[2u * slot] = make_float4(mn.x, mn.y, mn.z, __uint_as_float(first));
    nodes[2u * slot + 1u] = make_float4(mx.x, mx.y, mx.z, __uint_as_float(count));
}

// Internal node k of one set writes its two children into slots 1+2k,
// 2+2k (node 0 also writes the root); pointers are relative to the set.
__device__ __forceinline__ void emit_node(uint32_t n, uint32_t leaf_max, uint32_t k, const uint32_t* gamma,
                                          const uint32_t* rlo, const uint32_t* rhi, const float4* plo,
                                          const float4* phi, const float4* ilo, const float4* ihi,
                                          float4* nodes, uint32_t node_off, uint32_t leaf_off) {
    if (k == 0) put_node(nodes, 0, ilo[0], ihi[0], 1u + node_off, 0u);
    uint32_t g = gamma[k];
    uint32_t cl[2] = {rlo[k], g + 1}, ch[2] = {g, rhi[k]};
    uint32_t cidx[2] = {g, g + 1};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        uint32_t size = ch[c] - cl[c] + 1;
        float4 mn = size == 1 ? plo[cl[c]] : ilo[cidx[c]];
        float4 mx = size == 1 ? phi[cl[c]] : ihi[cidx[c]];
        uint32_t first, count;
        if (size <= leaf_max) {
            first = cl[c] + leaf_off;
            count = size;
        } else {
            first = 1u + 2u * cidx[c] + node_off;
            count = 0u;
        }
        put_node(nodes, 1u + 2u * k + (uint32_t)c, mn, mx, first, count);
    }
}

__global__ __launch_bounds__(256) void k_emit(uint32_t n, uint32_t leaf_max, const uint32_t* gamma,
                                              const uint32_t* rlo, const uint32_t* rhi,
                                              const float4* plo, const float4* phi,
                                              const float4* ilo, const float4* ihi, float4* nodes,
                                              uint32_t node_off, uint32_t leaf_off) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k + 1 >= n) return;
    emit_node(n, leaf_max, k, gamma, rlo, rhi, plo, phi, ilo, ihi, nodes, node_off, leaf_off);
}

// A set of at most leaf_max primitives: the root is a leaf over all of
// them, its box the union in sorted order.
__device__ __forceinline__ void small_root(const float4* lo, const float4* hi, const uint32_t* perm, uint32_t n,
                                           float4* nodes, uint32_t leaf_off) {
    float4 mn = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
    float4 mx = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
    for (uint32_t i = 0; i < n; ++i) {
        float4 l = lo[perm[i]], h = hi[perm[i]];
        mn = make_float4(sel_min(mn.x, l.x), sel_min(mn.y, l.y), sel_min(mn.z, l.z), 0.0f);
        mx = make_float4(sel_max(mx.x, h.x), sel_max(mx.y, h.y), sel_max(mx.z, h.z), 0.0f);
    }
    put_node(nodes, 0, mn, mx, leaf_off, n);
}

__global__ void k_small_root(const float4* lo, const float4* hi, const uint32_t* perm, uint32_t n,
                             float4* nodes, uint32_t leaf_off) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    small_root(lo, hi, perm, n, nodes, leaf_off);
}

// ---------------- segmented build: every mesh's tree in one pass ----------------
// Segment m holds primitives [tri_base[m], tri_base[m] + count[m]) (every
// segment non-empty, so tri_base is strictly increasing); its nodes start at
// slot blas_base[m].  Each segment's tree is the tree a single build of that
// segment writes: Morton codes inside its own centre bounds, order by
// (code, index within the segment), the same Karras/bottom-up/emit code on
// pointers offset to the segment.

// Largest m with tri_base[m] <= i.
__device__ __forceinline__ uint32_t seg_find(const uint32_t* tri_base, uint32_t n_seg, uint32_t i) {
    uint32_t a = 0, b = n_seg;
    while (b - a > 1) {
        const uint32_t mid = (a + b) >> 1;
        if (tri_base[mid] <= i) a = mid;
        else b = mid;
    }
    return a;
}

__global__ void k_seg_bounds_init(uint32_t* bounds, uint32_t n_seg) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 6u * n_seg) bounds[i] = (i % 6u) < 3u ? 0xffffffffu : 0u;
}

// Triangle boxes (a.min(b).min(c) as compare-selects, as the per-mesh
// build), segment ids, and per-segment centre bounds: a wave whose lanes all
// lie in one segment reduces in registers and issues one atomic per word.
__global__ __launch_bounds__(256) void k_seg_boxes(const float* pos, uint32_t n, const uint32_t* tri_base,
                                                   uint32_t n_seg, float4* lo, float4* hi, uint32_t* seg_of,
                                                   uint32_t* bounds) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    uint32_t m = 0;
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    if (valid) {
        const float* p = pos + 9ull * i;
        float4 l, h;
        l.x = sel_min(sel_min(p[0], p[3]), p[6]);
        l.y = sel_min(sel_min(p[1], p[4]), p[7]);
        l.z = sel_min(sel_min(p[2], p[5]), p[8]);
        h.x = sel_max(sel_max(p[0], p[3]), p[6]);
        h.y = sel_max(sel_max(p[1], p[4]), p[7]);
        h.z = sel_max(sel_max(p[2], p[5]), p[8]);
        l.w = h.w = 0.0f;
        lo[i] = l;
        hi[i] = h;
        m = seg_find(tri_base, n_seg, i);
        seg_of[i] = m;
        const float c[3] = {l.x * 0.5f + h.x * 0.5f, l.y * 0.5f + h.y * 0.5f, l.z * 0.5f + h.z * 0.5f};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (c[a] != c[a]) continue;  // NaN centres do not widen the bounds
            mn[a] = f2ord(c[a]);
            mx[a] = mn[a];
        }
    }
    // a block whose valid triangles all lie in one segment reduces through
    // LDS and issues one atomic per word; otherwise each wave of one segment
    // reduces in registers, and lanes of mixed waves go one by one
    __shared__ uint32_t s_m0;
    __shared__ uint32_t s_b[6];
    if (threadIdx.x == 0) s_m0 = m;  // the block's first triangle is valid
    if (threadIdx.x < 6) s_b[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
    __syncthreads();
    const uint32_t mb = s_m0;
    const bool block_uniform = !__syncthreads_or(valid && m != mb);
    const uint32_t m0 = __shfl(m, 0, 64);
    const uint64_t act = __ballot(valid);
    const bool uniform = act != 0ull && __ballot(valid && m == m0) == act;
    if (uniform) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const uint32_t om = __shfl_xor(mn[a], off, 64), ox = __shfl_xor(mx[a], off, 64);
                mn[a] = om < mn[a] ? om : mn[a];
                mx[a] = ox > mx[a] ? ox : mx[a];
            }
        }
        if ((threadIdx.x & 63u) == 0u) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (block_uniform) {
                    atomicMin(&s_b[a], mn[a]);
                    atomicMax(&s_b[3 + a], mx[a]);
                } else {
                    if (mn[a] != 0xffffffffu) atomicMin(&bounds[6u * m0 + a], mn[a]);
                    if (mx[a] != 0u) atomicMax(&bounds[6u * m0 + 3u + a], mx[a]);
                }
            }
        }
    } else if (valid) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (mn[a] != 0xffffffffu) atomicMin(&bounds[6u * m + a], mn[a]);
            if (mx[a] != 0u) atomicMax(&bounds[6u * m + 3u + a], mx[a]);
        }
    }
    if (!block_uniform) return;
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t v = s_b[threadIdx.x];
        if (threadIdx.x < 3) {
            if (v != 0xffffffffu) atomicMin(&bounds[6u * mb + threadIdx.x], v);
        } else if (v != 0u) {
            atomicMax(&bounds[6u * mb + threadIdx.x], v);
        }
    }
}

__global__ __launch_bounds__(256) void k_seg_morton(const float4* lo, const float4* hi, uint32_t n,
                                                    const uint32_t* seg_of, const uint32_t* bounds,
                                                    uint32_t* codes) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    codes[i] = morton_code(lo[i], hi[i], bounds + 6u * seg_of[i]);
}

__global__ void k_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t n, uint32_t* dst) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dst[j] = src[idx[j]];
}

__global__ __launch_bounds__(256) void k_seg_karras(const uint32_t* sk, uint32_t n_all, const uint32_t* seg_of,
                                                    const uint32_t* tri_base, const uint32_t* count,
                                                    uint32_t leaf_max, uint32_t* gamma, uint32_t* rlo,
                                                    uint32_t* rhi, uint32_t* parent_leaf, uint32_t* parent_int) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_all) return;
    const uint32_t m = seg_of[j], tb = tri_base[m], n = count[m], ii = j - tb;
    if (n <= leaf_max || ii + 1 >= n) return;
    karras_node(sk + tb, n, ii, gamma + tb, rlo + tb, rhi + tb, parent_leaf + tb, parent_int + tb);
}

// Segments above leaf_max emit as a single build; smaller ones write a leaf
// root and zero the rest of their 2n - 1 slots.
__global__ __launch_bounds__(256) void k_seg_emit(uint32_t n_all, uint32_t leaf_max, const uint32_t* seg_of,
                                                  const uint32_t* blas_base, const uint32_t* tri_base,
                                                  const uint32_t* count, const uint32_t* gamma,
                                                  const uint32_t* rlo, const uint32_t* rhi, const float4* plo,
                                                  const float4* phi, const float4* ilo, const float4* ihi,
                                                  const float4* lo, const float4* hi, const uint32_t* perm,
                                                  float4* nodes) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_all) return;
    const uint32_t m = seg_of[j], tb = tri_base[m], n = count[m], bb = blas_base[m], k = j - tb;
    float4* nm = nodes + 2ull * bb;
    if (n <= leaf_max) {
        if (k == 0) {
            small_root(lo, hi, perm + tb, n, nm, tb);
        } else {
            const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            nm[2u * (2u * k - 1u)] = z;
            nm[2u * (2u * k - 1u) + 1u] = z;
            nm[2u * (2u * k)] = z;
            nm[2u * (2u * k) + 1u] = z;
        }
        return;
    }
    if (k + 1 >= n) return;
    emit_node(n, leaf_max, k, gamma + tb, rlo + tb, rhi + tb, plo + tb, phi + tb, ilo + tb, ihi + tb, nm, bb,
              tb);
}

// Triangles in leaf order: w of the first vertex holds the index within
// the mesh, of the second the material; order[j] is the global index.
__global__ __launch_bounds__(256) void k_seg_gather(const float* pos, const uint32_t* mat, const uint32_t* perm,
                                                    const uint32_t* seg_of, const uint32_t* tri_base, uint32_t n,
                                                    float4* tri, uint32_t* order) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t s = perm[j];
    const float* p = pos + 9ull * s;
    float4* t = tri + 3ull * j;
    t[0] = make_float4(p[0], p[1], p[2], __uint_as_float(s - tri_base[seg_of[j]]));
    t[1] = make_float4(p[3], p[4], p[5], __uint_as_float(mat ? mat[s] : 0u));
    t[2] = make_float4(p[6], p[7], p[8], 0.0f);
    order[j] = s;
}

inline uint32_t blocks_for(uint32_t n, uint32_t t) { return (n + t - 1) / t; }

}  // namespace

size_t lux_scan_sums_words(uint32_t n) { return (size_t)(n + kScanBlock - 1) / kScanBlock + 2; }

int lux_scan_exclusive_impl(lux_ctx* ctx, uint32_t* data, uint32_t n, uint32_t* sums, uint32_t* total) {
    return lux_scan_exclusive_on(ctx, ctx->stream, data, n, sums, total);
}

int lux_scan_exclusive_on(lux_ctx* ctx, hipStream_t st, uint32_t* data, uint32_t n, uint32_t* sums,
                          uint32_t* total) {
    const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb) hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanBlock), 0, st, data, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, st, sums, nb, total);
    if (nb) hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanBlock), 0, st, data, n, sums);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

// Histogram words (256 digits x tiles) followed by the histogram scan's
// block sums and total.
size_t lux_sort_hist_words(uint32_t n) {
    uint32_t nb = (n + kRsTile - 1) / kRsTile;
    if (nb == 0) nb = 1;
    return (size_t)kRsBins * nb + lux_scan_sums_words(kRsBins * nb);
}

int lux_sort_pairs_impl(lux_ctx* ctx, const uint32_t* keys_in, uint32_t n, uint32_t* keys_out,
                        uint32_t* vals_out, uint32_t* tmp_keys, uint32_t* tmp_vals,
                        uint32_t* hist, uint32_t bits) {
    return lux_sort_pairs_vals_impl(ctx, keys_in, nullptr, n, keys_out, vals_out, tmp_keys, tmp_vals, hist, bits);
}

int lux_sort_pairs_vals_impl(lux_ctx* ctx, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t n,
                             uint32_t* keys_out, uint32_t* vals_out, uint32_t* tmp_keys, uint32_t* tmp_vals,
                             uint32_t* hist, uint32_t bits) {
    if (n == 0) return LUX_OK;
    hipStream_t st = ctx->stream;
    uint32_t nb = (n + kRsTile - 1) / kRsTile;
    const uint32_t words = (uint32_t)kRsBins * nb;
    uint32_t* sums = hist + words;
    const int passes = bits == 0 ? 1 : (int)((bits + 7u) / 8u);
    // ping-pong so that the last pass writes keys_out / vals_out
    const uint32_t* kin = keys_in;
    const uint32_t* vin = vals_in;
    for (int pass = 0; pass < passes; ++pass) {
        uint32_t shift = 8u * (uint32_t)pass;
        const bool last_parity = ((passes - 1 - pass) & 1) == 0;
        uint32_t* kout = last_parity ? keys_out : tmp_keys;
        uint32_t* vout = last_parity ? vals_out : tmp_vals;
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kRsThreads), 0, st, kin, n, shift, nb, hist);
        if (int rc = lux_scan_exclusive_impl(ctx, hist, words, sums, sums + lux_scan_sums_words(words) - 1))
            return rc;
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kRsThreads), 0, st, kin, vin, kout, vout, n,
                           shift, nb, hist);
        kin = kout;
        vin = vout;
    }
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

int lux_lbvh_build_device(lux_ctx* ctx, const float4* lo, const float4* hi, uint32_t n,
                          uint32_t leaf_max, float4* nodes, uint32_t* perm,
                          uint32_t node_first_offset, uint32_t leaf_first_offset) {
    if (n == 0) return LUX_OK;
    hipStream_t st = ctx->stream;
    // scratch layout (uint32 words)
    size_t hw = lux_sort_hist_words(n);
    size_t words = 6 + 5ull * n + hw + 5ull * n + 4ull * 4 * n + 4ull * 4 * n + n;
    if (int rc = lux_ensure(ctx, ctx->scratch, words * 4 + 1024)) return rc;
    uint32_t* w = static_cast<uint32_t*>(ctx->scratch.ptr);
    uint32_t* bounds = w;
    uint32_t* codes = bounds + 8;
    uint32_t* sk = codes + n;
    uint32_t* tk = sk + n;
    uint32_t* tv = tk + n;
    uint32_t* hist = tv + n;
    uint32_t* gamma = hist + hw;
    uint32_t* rlo = gamma + n;
    uint32_t* rhi = rlo + n;
    uint32_t* parent_leaf = rhi + n;
    uint32_t* parent_int = parent_leaf + n;
    uint32_t* arrivals = parent_int + n;
    // align float4 arrays to 16 bytes
    uintptr_t fa = reinterpret_cast<uintptr_t>(arrivals + n);
    fa = (fa + 15) & ~uintptr_t(15);
    float4* plo = reinterpret_cast<float4*>(fa);
    float4* phi = plo + n;
    float4* ilo = phi + n;
    float4* ihi = ilo + n;

    uint32_t init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    LUX_HIP(ctx, hipMemcpyAsync(bounds, init, sizeof init, hipMemcpyHostToDevice, st));
    uint32_t cb = blocks_for(n, 256);
    if (cb > 2048) cb = 2048;
    hipLaunchKernelGGL(k_centre_bounds, dim3(cb), dim3(256), 0, st, lo, hi, n, bounds);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, 256)), dim3(256), 0, st, lo, hi, n, bounds, codes);
    if (int rc = lux_sort_pairs_impl(ctx, codes, n, sk, perm, tk, tv, hist, 32)) return rc;
    if (n <= leaf_max) {
        // the root is a leaf; the remaining 2n-2 slots of this range are unused
        // and kept zero, as in every other unused slot
        uint32_t words = 8u * (2u * n - 1u);
        hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(words, 256)), dim3(256), 0, st,
                           reinterpret_cast<uint32_t*>(nodes), 0u, words);
        hipLaunchKernelGGL(k_small_root, dim3(1), dim3(64), 0, st, lo, hi, perm, n, nodes,
                           leaf_first_offset);
        LUX_HIP(ctx, hipGetLastError());
        return LUX_OK;
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, st, arrivals, 0u, n);
    hipLaunchKernelGGL(k_karras, dim3(blocks_for(n, 256)), dim3(256), 0, st, sk, n, gamma, rlo, rhi,
                       parent_leaf, parent_int);
    hipLaunchKernelGGL(k_bottom_up_t<false>, dim3(blocks_for(n, kBuTile)), dim3(kBuThreads), 0, st, lo, hi, perm,
                       n, nullptr, nullptr, nullptr, nullptr, leaf_max, gamma, rlo, rhi, parent_leaf, parent_int, plo,
                       phi, ilo, ihi, arrivals);
    hipLaunchKernelGGL(k_emit, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, leaf_max, gamma, rlo,
                       rhi, plo, phi, ilo, ihi, nodes, node_first_offset, leaf_first_offset);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

void lux_segment_tiles(const uint32_t* count, uint32_t n_seg, std::vector<uint32_t>& tiles) {
    // whole segments pack into a tile while they fit; a larger segment
    // starts a tile and is cut every kBuTile leaves, its last piece left
    // open for the segments that follow
    tiles.clear();
    uint32_t pos = 0, fill = 0;
    for (uint32_t m = 0; m < n_seg; ++m) {
        const uint32_t c = count[m];
        if (!tiles.empty() && fill + c <= kBuTile) {
            fill += c;
        } else {
            uint32_t k = 0;
            while (c - k * kBuTile > kBuTile) tiles.push_back(pos + kBuTile * k++);
            tiles.push_back(pos + kBuTile * k);
            fill = c - k * kBuTile;
        }
        pos += c;
    }
    tiles.push_back(pos);
}

int lux_blas_build_segmented(lux_ctx* ctx, const float* pos, const uint32_t* mat, uint32_t n, uint32_t n_seg,
                             const uint32_t* seg_info, const uint32_t* tiles, uint32_t n_tiles, uint32_t leaf_max,
                             float4* nodes, float4* tris, uint32_t* order) {
    if (n == 0 || n_seg == 0) return LUX_OK;
    hipStream_t st = ctx->stream;
    const uint32_t* blas_base = seg_info;
    const uint32_t* tri_base = seg_info + n_seg;
    const uint32_t* count = seg_info + 2ull * n_seg;
    const size_t hw = lux_sort_hist_words(n);
    const size_t words = 6ull * n_seg + 16ull * n + hw + 4 + 4ull * 6 * n;
    if (int rc = lux_ensure(ctx, ctx->scratch, words * 4 + 1024)) return rc;
    uint32_t* w = static_cast<uint32_t*>(ctx->scratch.ptr);
    uint32_t* bounds = w;
    uint32_t* seg_of = bounds + 6ull * n_seg;
    uint32_t* codes = seg_of + n;
    uint32_t* sk = codes + n;
    uint32_t* perm1 = sk + n;
    uint32_t* key2 = perm1 + n;
    uint32_t* perm = key2 + n;
    uint32_t* tk = perm + n;
    uint32_t* tv = tk + n;
    uint32_t* gamma = tv + n;
    uint32_t* rlo = gamma + n;
    uint32_t* rhi = rlo + n;
    uint32_t* parent_leaf = rhi + n;
    uint32_t* parent_int = parent_leaf + n;
    uint32_t* arrivals = parent_int + n;
    uint32_t* hist = arrivals + n;
    uintptr_t fa = reinterpret_cast<uintptr_t>(hist + hw);
    fa = (fa + 15) & ~uintptr_t(15);
    float4* lo = reinterpret_cast<float4*>(fa);
    float4* hi = lo + n;
    float4* plo = hi + n;
    float4* phi = plo + n;
    float4* ilo = phi + n;
    float4* ihi = ilo + n;
    const uint32_t nb = blocks_for(n, 256);

    hipLaunchKernelGGL(k_seg_bounds_init, dim3(blocks_for(6u * n_seg, 256)), dim3(256), 0, st, bounds, n_seg);
    hipLaunchKernelGGL(k_seg_boxes, dim3(nb), dim3(256), 0, st, pos, n, tri_base, n_seg, lo, hi, seg_of, bounds);
    hipLaunchKernelGGL(k_seg_morton, dim3(nb), dim3(256), 0, st, lo, hi, n, seg_of, bounds, codes);
    // stable by code (ties by global index), then stable by segment: within
    // a segment the order is by (code, index within the segment)
    if (n_seg == 1) {
        if (int rc = lux_sort_pairs_impl(ctx, codes, n, sk, perm, tk, tv, hist, 32)) return rc;
    } else {
        if (int rc = lux_sort_pairs_impl(ctx, codes, n, sk, perm1, tk, tv, hist, 32)) return rc;
        hipLaunchKernelGGL(k_gather_u32, dim3(nb), dim3(256), 0, st, seg_of, perm1, n, key2);
        const uint32_t seg_bits = 32u - (uint32_t)__builtin_clz(n_seg - 1u);
        if (int rc = lux_sort_pairs_vals_impl(ctx, key2, perm1, n, sk, perm, tk, tv, hist, seg_bits)) return rc;
        hipLaunchKernelGGL(k_gather_u32, dim3(nb), dim3(256), 0, st, codes, perm, n, sk);
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(nb), dim3(256), 0, st, arrivals, 0u, n);
    hipLaunchKernelGGL(k_seg_karras, dim3(nb), dim3(256), 0, st, sk, n, seg_of, tri_base, count, leaf_max, gamma,
                       rlo, rhi, parent_leaf, parent_int);
    hipLaunchKernelGGL(k_bottom_up_t<true>, dim3(n_tiles), dim3(kBuThreads), 0, st, lo, hi, perm, n, tiles, seg_of,
                       tri_base, count, leaf_max, gamma, rlo, rhi, parent_leaf, parent_int, plo, phi, ilo, ihi,
                       arrivals);
    hipLaunchKernelGGL(k_seg_emit, dim3(nb), dim3(256), 0, st, n, leaf_max, seg_of, blas_base, tri_base, count,
                       gamma, rlo, rhi, plo, phi, ilo, ihi, lo, hi, perm, nodes);
    hipLaunchKernelGGL(k_seg_gather, dim3(nb), dim3(256), 0, st, pos, mat, perm, seg_of, tri_base, n, tris, order);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

int lux_seg_prepare(lux_ctx* ctx, const float* pos, uint32_t n, const uint32_t* tri_base, uint32_t n_seg,
                    float4* lo, float4* hi, uint32_t* seg_of, uint32_t* bounds) {
    if (n == 0 || n_seg == 0) return LUX_OK;
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_seg_bounds_init, dim3(blocks_for(6u * n_seg, 256)), dim3(256), 0, st, bounds, n_seg);
    hipLaunchKernelGGL(k_seg_boxes, dim3(blocks_for(n, 256)), dim3(256), 0, st, pos, n, tri_base, n_seg, lo, hi,
                       seg_of, bounds);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

int lux_seg_gather(lux_ctx* ctx, const float* pos, const uint32_t* mat, const uint32_t* perm, const uint32_t* seg_of,
                   const uint32_t* tri_base, uint32_t n, float4* tris, uint32_t* order) {
    if (n == 0) return LUX_OK;
    hipLaunchKernelGGL(k_seg_gather, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, pos, mat, perm, seg_of,
                       tri_base, n, tris, order);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

extern "C" int lux_morton_codes_device(lux_ctx* ctx, const float* lo4, const float* hi4, uint32_t n,
                                       uint32_t* codes) {
    if (!ctx) return LUX_ERR_INVALID;
    if (n == 0) return LUX_OK;
    if (!lo4 || !hi4 || !codes)
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_morton_codes_device: null buffer");
    if (int rc = lux_ensure(ctx, ctx->scratch, 64)) return rc;
    uint32_t* bounds = static_cast<uint32_t*>(ctx->scratch.ptr);
    hipStream_t st = ctx->stream;
    uint32_t init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    LUX_HIP(ctx, hipMemcpyAsync(bounds, init, sizeof init, hipMemcpyHostToDevice, st));
    const float4* lo = reinterpret_cast<const float4*>(lo4);
    const float4* hi = reinterpret_cast<const float4*>(hi4);
    uint32_t cb = blocks_for(n, 256);
    if (cb > 2048) cb = 2048;
    hipLaunchKernelGGL(k_centre_bounds, dim3(cb), dim3(256), 0, st, lo, hi, n, bounds);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, 256)), dim3(256), 0, st, lo, hi, n, bounds, codes);
    LUX_HIP(ctx, hipGetLastError());
    LUX_HIP(ctx, hipStreamSynchronize(st));
    return LUX_OK;
}

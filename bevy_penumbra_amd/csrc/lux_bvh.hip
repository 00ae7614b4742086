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

__global__ __launch_bounds__(256) void k_morton(const float4* lo, const float4* hi, uint32_t n,
                                                const uint32_t* bounds, uint32_t* codes) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float cmin[3], ext[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float mnv = bounds[a] == 0xffffffffu ? 0.0f : ord2f(bounds[a]);
        float mxv = bounds[3 + a] == 0u ? 0.0f : ord2f(bounds[3 + a]);
        cmin[a] = mnv;
        ext[a] = mxv - mnv;
    }
    float4 l = lo[i], h = hi[i];
    float c0 = l.x * 0.5f + h.x * 0.5f, c1 = l.y * 0.5f + h.y * 0.5f, c2 = l.z * 0.5f + h.z * 0.5f;
    uint32_t q0 = quantize10(c0, cmin[0], ext[0]);
    uint32_t q1 = quantize10(c1, cmin[1], ext[1]);
    uint32_t q2 = quantize10(c2, cmin[2], ext[2]);
    codes[i] = (spread10(q0) << 2) | (spread10(q1) << 1) | spread10(q2);
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

// children[k] = gamma; range lo/hi; parents of leaves/internal nodes.
__global__ __launch_bounds__(256) void k_karras(const uint32_t* sk, uint32_t n, uint32_t* gamma_out,
                                                uint32_t* rlo, uint32_t* rhi, uint32_t* parent_leaf,
                                                uint32_t* parent_int) {
    uint32_t ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii + 1 >= n) return;
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

// Sorted primitive boxes, then walk up: the second arrival at a node unions
// its children (left first) and continues.
__global__ __launch_bounds__(256) void k_bottom_up(const float4* lo, const float4* hi,
                                                   const uint32_t* perm, uint32_t n,
                                                   const uint32_t* gamma, const uint32_t* rlo,
                                                   const uint32_t* rhi, const uint32_t* parent_leaf,
                                                   const uint32_t* parent_int, float4* plo,
                                                   float4* phi, float4* ilo, float4* ihi,
                                                   uint32_t* arrivals) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t src = perm[j];
    float4 l = lo[src], h = hi[src];
    l.w = 0.0f;
    h.w = 0.0f;
    st_relaxed(&plo[j], l);
    st_relaxed(&phi[j], h);
    if (n < 2) return;
    uint32_t p = parent_leaf[j];
    while (true) {
        __atomic_thread_fence(__ATOMIC_RELEASE);  // publish this thread's boxes
        uint32_t old = __hip_atomic_fetch_add(&arrivals[p], 1u, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0u) return;
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
        uint32_t g = gamma[p];
        float4 alo, ahi, blo, bhi;
        if (g == rlo[p]) { alo = ld_relaxed(&plo[g]); ahi = ld_relaxed(&phi[g]); }
        else { alo = ld_relaxed(&ilo[g]); ahi = ld_relaxed(&ihi[g]); }
        if (g + 1 == rhi[p]) { blo = ld_relaxed(&plo[g + 1]); bhi = ld_relaxed(&phi[g + 1]); }
        else { blo = ld_relaxed(&ilo[g + 1]); bhi = ld_relaxed(&ihi[g + 1]); }
        float4 olo = make_float4(sel_min(alo.x, blo.x), sel_min(alo.y, blo.y), sel_min(alo.z, blo.z), 0.0f);
        float4 ohi = make_float4(sel_max(ahi.x, bhi.x), sel_max(ahi.y, bhi.y), sel_max(ahi.z, bhi.z), 0.0f);
        st_relaxed(&ilo[p], olo);
        st_relaxed(&ihi[p], ohi);
        if (p == 0) return;
        p = parent_int[p];
    }
}

__device__ __forceinline__ void put_node(float4* nodes, uint32_t slot, float4 mn, float4 mx,
                                         uint32_t first, uint32_t count) {
    nodes[2u * slot] = make_float4(mn.x, mn.y, mn.z, __uint_as_float(first));
    nodes[2u * slot + 1u] = make_float4(mx.x, mx.y, mx.z, __uint_as_float(count));
}

__global__ __launch_bounds__(256) void k_emit(uint32_t n, uint32_t leaf_max, const uint32_t* gamma,
                                              const uint32_t* rlo, const uint32_t* rhi,
                                              const float4* plo, const float4* phi,
                                              const float4* ilo, const float4* ihi, float4* nodes,
                                              uint32_t node_off, uint32_t leaf_off) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k + 1 >= n) return;
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

__global__ void k_small_root(const float4* lo, const float4* hi, const uint32_t* perm, uint32_t n,
                             float4* nodes, uint32_t leaf_off) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float4 mn = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
    float4 mx = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
    for (uint32_t i = 0; i < n; ++i) {
        float4 l = lo[perm[i]], h = hi[perm[i]];
        mn = make_float4(sel_min(mn.x, l.x), sel_min(mn.y, l.y), sel_min(mn.z, l.z), 0.0f);
        mx = make_float4(sel_max(mx.x, h.x), sel_max(mx.y, h.y), sel_max(mx.z, h.z), 0.0f);
    }
    put_node(nodes, 0, mn, mx, leaf_off, n);
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
    if (n == 0) return LUX_OK;
    hipStream_t st = ctx->stream;
    uint32_t nb = (n + kRsTile - 1) / kRsTile;
    const uint32_t words = (uint32_t)kRsBins * nb;
    uint32_t* sums = hist + words;
    const int passes = bits == 0 ? 1 : (int)((bits + 7u) / 8u);
    // ping-pong so that the last pass writes keys_out / vals_out
    const uint32_t* kin = keys_in;
    const uint32_t* vin = nullptr;
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
    hipLaunchKernelGGL(k_bottom_up, dim3(blocks_for(n, 256)), dim3(256), 0, st, lo, hi, perm, n,
                       gamma, rlo, rhi, parent_leaf, parent_int, plo, phi, ilo, ihi, arrivals);
    hipLaunchKernelGGL(k_emit, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, leaf_max, gamma, rlo,
                       rhi, plo, phi, ilo, ihi, nodes, node_first_offset, leaf_first_offset);
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

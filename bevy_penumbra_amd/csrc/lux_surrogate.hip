// lux_surrogate.hip -- learned far-cascade radiance surrogate on gfx950.
//
// A multiresolution hash-grid encoder (16 levels x 2 FP16 features, 2^19
// entries per level, base resolution 16, growth 1.5; levels whose (res+1)^3
// grid fits the table are indexed densely) plus an 8-band real
// spherical-harmonic direction encoding and the normal feed an MLP
// 128 -> 64 -> 64 -> 16 (ReLU, no biases; input 99 is a constant 1) whose
// weights are BF16 and whose products run on MFMA
// (__builtin_amdgcn_mfma_f32_16x16x16bf16_1k) with FP32 accumulation.
// Outputs 0-2: radiance; output 3: log of the predicted squared error.
//
// One training step over S samples, all on the library stream:
//   k_sur_encode      X (S x 128 bf16) from (position, direction, normal)
//   k_sur_forward     H1, H2 (bf16, post-ReLU) and Y (f32), MFMA
//   k_sur_loss        dY: relative L2 on RGB, L2 on the error head
//   k_sur_backward    dH2, dH1 (bf16) and dX[:, 0:32] (f32), MFMA; the error
//                     head's gradient stops at its column of W3
//   k_sur_wgrad       dW = A^T B over samples, MFMA, FP32 atomics per tile
//   k_sur_hash_bwd    trilinear scatter of dX into the FP32 table gradient
//   k_sur_adam        Adam on the FP32 masters; refreshes the BF16 weight and
//                     FP16 table copies the training pass reads and, with a
//                     nonzero EMA decay, the exponential moving average of the
//                     parameters whose copies inference reads
// The reference has no learned component (README.md:384-385 defers neural
// radiance caching); the FP64 oracle of this network is oracle/surrogate.py.
//
// MFMA 16x16x16 operand layout (lane l of a 64-wide wave), checked on gfx950:
//   A: A[l % 16][4 (l / 16) + i], B: B[4 (l / 16) + i][l % 16], i = 0..3
//   D: D[4 (l / 16) + j][l % 16], j = 0..3
#include <cmath>
#include <cstring>
#include <hip/hip_fp16.h>
#include "lux_internal.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 128, kHid = 64, kOut = 16;
constexpr int kLevels = 16, kFeat = 2;
constexpr uint32_t kTableLog2 = 19, kTable = 1u << kTableLog2;
constexpr int kSH = 64;

__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) {
    return __uint_as_float(((uint32_t)(uint16_t)b) << 16);
}
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bool finite3(const float* t) {
    return isfinite(t[0]) && isfinite(t[1]) && isfinite(t[2]);
}

struct SurDev {
    const __half* table;       // [kLevels][kTable][kFeat] FP16 copy
    const short* w1;           // [kIn][kHid]  BF16 copies, row-major (in, out)
    const short* w2;           // [kHid][kHid]
    const short* w3;           // [kHid][kOut]
    const float* sh_const;     // [64] real-SH normalisation
    uint32_t res[kLevels];     // grid resolution per level
    uint32_t dense[kLevels];   // 1 when (res+1)^3 <= table size
    float box_min[3], box_inv[3];
};

// ---------------------------------------------------------------- encoding
__device__ __forceinline__ uint32_t grid_index(uint32_t x, uint32_t y, uint32_t z, uint32_t res,
                                               uint32_t dense) {
    if (dense) return x + (res + 1u) * (y + (res + 1u) * z);
    return (x * 1u ^ y * 2654435761u ^ z * 805459861u) & (kTable - 1u);
}

__device__ __forceinline__ void unit_position(const SurDev& d, const float* r, float* p) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // NaN and -inf go to 0, +inf to 1 - 2^-24: every corner index stays in the grid
        float q = (r[a] - d.box_min[a]) * d.box_inv[a];
        p[a] = q > 0.99999994f ? 0.99999994f : (q >= 0.0f ? q : 0.0f);
    }
}

// Real SH of a unit direction, bands 0..7, index l*l + l + m, polar axis z,
// no Condon-Shortley phase: (x + i y)^m by complex products and the
// associated-Legendre recurrence in z.
__device__ void sh_encode(float x, float y, float z, const float* c, float* out) {
    float re[8], im[8];
    re[0] = 1.0f;
    im[0] = 0.0f;
    for (int m = 1; m < 8; ++m) {
        re[m] = re[m - 1] * x - im[m - 1] * y;
        im[m] = re[m - 1] * y + im[m - 1] * x;
    }
    for (int m = 0; m < 8; ++m) {
        float pmm = 1.0f;
        for (int k = 1; k <= m; ++k) pmm = pmm * (float)(2 * k - 1);
        float p_prev = 0.0f, p_cur = pmm;
        for (int l = m; l < 8; ++l) {
            float p;
            if (l == m) p = pmm;
            else if (l == m + 1) p = z * (float)(2 * m + 1) * pmm;
            else p = ((float)(2 * l - 1) * z * p_cur - (float)(l + m - 1) * p_prev) / (float)(l - m);
            if (l > m) { p_prev = p_cur; p_cur = p; }
            int base = l * l + l;
            if (m == 0) out[base] = c[base] * p;
            else {
                out[base + m] = c[base + m] * re[m] * p;
                out[base - m] = c[base - m] * im[m] * p;
            }
        }
    }
}

// rec: `stride` floats per sample, the first 9 = position, direction, normal.
__global__ __launch_bounds__(256) void k_sur_encode(SurDev d, const float* rec, uint32_t stride,
                                                    uint32_t n, short* X) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float* r = rec + (size_t)stride * s;
    short* row = X + (size_t)s * kIn;
    float p[3];
    unit_position(d, r, p);
    for (int l = 0; l < kLevels; ++l) {
        float res = (float)d.res[l];
        float gx = p[0] * res, gy = p[1] * res, gz = p[2] * res;
        float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
        uint32_t cx = (uint32_t)fx, cy = (uint32_t)fy, cz = (uint32_t)fz;
        float tx = gx - fx, ty = gy - fy, tz = gz - fz;
        float f0 = 0.0f, f1 = 0.0f;
        const __half* tab = d.table + (size_t)l * kTable * kFeat;
        for (int c = 0; c < 8; ++c) {
            uint32_t bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
            float w = (bx ? tx : 1.0f - tx) * (by ? ty : 1.0f - ty) * (bz ? tz : 1.0f - tz);
            uint32_t idx = grid_index(cx + bx, cy + by, cz + bz, d.res[l], d.dense[l]);
            __half2 v = *reinterpret_cast<const __half2*>(tab + 2ull * idx);
            f0 = f0 + w * __low2float(v);
            f1 = f1 + w * __high2float(v);
        }
        row[2 * l] = f2bf(f0);
        row[2 * l + 1] = f2bf(f1);
    }
    float sh[kSH];
    sh_encode(r[3], r[4], r[5], d.sh_const, sh);
    for (int k = 0; k < kSH; ++k) row[32 + k] = f2bf(sh[k]);
    row[96] = f2bf(r[6]);
    row[97] = f2bf(r[7]);
    row[98] = f2bf(r[8]);
    row[99] = f2bf(1.0f);
    for (int k = 100; k < kIn; ++k) row[k] = 0;
}

// ---------------------------------------------------------------- MLP forward
// Weights staged transposed in LDS ([out][in]) so a lane's B operand is four
// consecutive BF16.  One 16-sample tile per wave per iteration; the trip count
// depends on blockIdx only, so block barriers are reached by every wave.
__global__ __launch_bounds__(256) void k_sur_forward(SurDev d, const short* X, uint32_t n, short* H1,
                                                     short* H2, float* Y, uint32_t ntiles) {
    __shared__ short w1t[kHid * kIn];
    __shared__ short w2t[kHid * kHid];
    __shared__ short w3t[kOut * kHid];
    __shared__ short h[4][2][16 * kHid];
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    for (int i = tid; i < kHid * kIn; i += 256) w1t[i] = d.w1[(i % kIn) * kHid + i / kIn];
    for (int i = tid; i < kHid * kHid; i += 256) w2t[i] = d.w2[(i % kHid) * kHid + i / kHid];
    for (int i = tid; i < kOut * kHid; i += 256) w3t[i] = d.w3[(i % kHid) * kOut + i / kHid];
    __syncthreads();
    const int r16 = l & 15, q = l >> 4;
    short* h1 = h[wave][0];
    short* h2 = h[wave][1];
    for (uint32_t base = blockIdx.x * 4u; base < ntiles; base += gridDim.x * 4u) {
        uint32_t tile = base + wave;
        uint32_t s0 = tile * 16u;
        uint32_t srow = s0 + r16;
        bool valid_row = tile < ntiles && srow < n;
        bf16x4 a[8];
#pragma unroll
        for (int ks = 0; ks < 8; ++ks)
            a[ks] = valid_row ? *reinterpret_cast<const bf16x4*>(X + (size_t)srow * kIn + ks * 16 + 4 * q)
                              : bf16x4{0, 0, 0, 0};
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 8; ++ks) {
                bf16x4 b = *reinterpret_cast<const bf16x4*>(w1t + (nt * 16 + r16) * kIn + ks * 16 + 4 * q);
                acc = mfma16(a[ks], b, acc);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                h1[(4 * q + j) * kHid + nt * 16 + r16] = f2bf(acc[j] > 0.0f ? acc[j] : 0.0f);
        }
        __syncthreads();
        bf16x4 a2[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) a2[ks] = *reinterpret_cast<const bf16x4*>(h1 + r16 * kHid + ks * 16 + 4 * q);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                bf16x4 b = *reinterpret_cast<const bf16x4*>(w2t + (nt * 16 + r16) * kHid + ks * 16 + 4 * q);
                acc = mfma16(a2[ks], b, acc);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                h2[(4 * q + j) * kHid + nt * 16 + r16] = f2bf(acc[j] > 0.0f ? acc[j] : 0.0f);
        }
        __syncthreads();
        bf16x4 a3[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) a3[ks] = *reinterpret_cast<const bf16x4*>(h2 + r16 * kHid + ks * 16 + 4 * q);
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            bf16x4 b = *reinterpret_cast<const bf16x4*>(w3t + r16 * kHid + ks * 16 + 4 * q);
            acc = mfma16(a3[ks], b, acc);
        }
        if (tile < ntiles) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                uint32_t s = s0 + 4 * q + j;
                if (s < n) Y[(size_t)s * kOut + r16] = acc[j];
            }
            if (H1 && srow < n)
                for (int c = 0; c < 4; ++c) {
                    *reinterpret_cast<bf16x4*>(H1 + (size_t)srow * kHid + c * 16 + 4 * q) =
                        *reinterpret_cast<const bf16x4*>(h1 + r16 * kHid + c * 16 + 4 * q);
                    *reinterpret_cast<bf16x4*>(H2 + (size_t)srow * kHid + c * 16 + 4 * q) =
                        *reinterpret_cast<const bf16x4*>(h2 + r16 * kHid + c * 16 + 4 * q);
                }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- loss
// Relative L2 with the weight 1/(y^2 + 0.01) held constant, plus
// err_weight * (y3 - log(mse + 1e-4))^2 with its target held constant;
// averaged over n samples.  A sample whose target is not finite contributes
// nothing.  Target RGB sits at floats 9..11 of each 12-float sample.
__global__ __launch_bounds__(256) void k_sur_loss(const float* Y, const float* samples, uint32_t n,
                                                  float err_weight, short* dYb, float* loss) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    float lsum = 0.0f;
    if (s < n) {
        const float* y = Y + (size_t)s * kOut;
        const float* t = samples + 12ull * s + 9;
        float dv[kOut];
        for (int c = 0; c < kOut; ++c) dv[c] = 0.0f;
        if (finite3(t)) {
            float inv = 1.0f / (3.0f * (float)n);
            float se = 0.0f;
            for (int c = 0; c < 3; ++c) {
                float diff = y[c] - t[c];
                float w = 1.0f / (y[c] * y[c] + 0.01f);
                lsum = lsum + diff * diff * w * inv;
                dv[c] = 2.0f * diff * w * inv;
                se = se + diff * diff;
            }
            float te = logf(se / 3.0f + 1e-4f);
            float de = y[3] - te;
            lsum = lsum + err_weight * de * de / (float)n;
            dv[3] = 2.0f * err_weight * de / (float)n;
        }
        for (int c = 0; c < kOut; ++c) dYb[(size_t)s * kOut + c] = f2bf(dv[c]);
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, lsum);
}

// ---------------------------------------------------------------- backward (data)
// dH2 = (dY' W3^T) [H2 > 0], dH1 = (dH2 W2^T) [H1 > 0], dX[:, 0:32] = dH1 W1^T,
// where dY' is dY with the error-head column zeroed: the error head learns
// through its own column of W3 only and leaves the shared layers to radiance.
// Weights staged row-major ([in][out]) so B = W^T reads four consecutive BF16.
__global__ __launch_bounds__(256) void k_sur_backward(SurDev d, const short* dYb, const short* H1,
                                                      const short* H2, uint32_t n, short* dH1,
                                                      short* dH2, float* dXh, uint32_t ntiles) {
    __shared__ short w1s[kIn * kHid];
    __shared__ short w2s[kHid * kHid];
    __shared__ short w3s[kHid * kOut];
    __shared__ short g[4][16 * kHid];
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    for (int i = tid; i < kIn * kHid; i += 256) w1s[i] = d.w1[i];
    for (int i = tid; i < kHid * kHid; i += 256) w2s[i] = d.w2[i];
    for (int i = tid; i < kHid * kOut; i += 256) w3s[i] = d.w3[i];
    __syncthreads();
    const int r16 = l & 15, q = l >> 4;
    short* gw = g[wave];
    for (uint32_t base = blockIdx.x * 4u; base < ntiles; base += gridDim.x * 4u) {
        uint32_t tile = base + wave;
        uint32_t s0 = tile * 16u, srow = s0 + r16;
        bool valid = tile < ntiles && srow < n;
        bf16x4 a = valid ? *reinterpret_cast<const bf16x4*>(dYb + (size_t)srow * kOut + 4 * q)
                         : bf16x4{0, 0, 0, 0};
        if (q == 0) a[3] = 0;  // the error head (output 3) does not train the trunk
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            int hid = nt * 16 + r16;
            bf16x4 b = *reinterpret_cast<const bf16x4*>(w3s + hid * kOut + 4 * q);
            f32x4 acc = mfma16(a, b, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                uint32_t s = s0 + 4 * q + j;
                float m = (tile < ntiles && s < n) ? bf2f(H2[(size_t)s * kHid + hid]) : 0.0f;
                gw[(4 * q + j) * kHid + hid] = f2bf(m > 0.0f ? acc[j] : 0.0f);
            }
        }
        __syncthreads();
        if (valid)
            for (int c = 0; c < 4; ++c)
                *reinterpret_cast<bf16x4*>(dH2 + (size_t)srow * kHid + c * 16 + 4 * q) =
                    *reinterpret_cast<const bf16x4*>(gw + r16 * kHid + c * 16 + 4 * q);
        bf16x4 a2[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) a2[ks] = *reinterpret_cast<const bf16x4*>(gw + r16 * kHid + ks * 16 + 4 * q);
        f32x4 acc1[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            int hid = nt * 16 + r16;
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                bf16x4 b = *reinterpret_cast<const bf16x4*>(w2s + hid * kHid + ks * 16 + 4 * q);
                acc = mfma16(a2[ks], b, acc);
            }
            acc1[nt] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            int hid = nt * 16 + r16;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                uint32_t s = s0 + 4 * q + j;
                float m = (tile < ntiles && s < n) ? bf2f(H1[(size_t)s * kHid + hid]) : 0.0f;
                gw[(4 * q + j) * kHid + hid] = f2bf(m > 0.0f ? acc1[nt][j] : 0.0f);
            }
        }
        __syncthreads();
        if (valid)
            for (int c = 0; c < 4; ++c)
                *reinterpret_cast<bf16x4*>(dH1 + (size_t)srow * kHid + c * 16 + 4 * q) =
                    *reinterpret_cast<const bf16x4*>(gw + r16 * kHid + c * 16 + 4 * q);
        bf16x4 a3[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) a3[ks] = *reinterpret_cast<const bf16x4*>(gw + r16 * kHid + ks * 16 + 4 * q);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            int in = nt * 16 + r16;
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                bf16x4 b = *reinterpret_cast<const bf16x4*>(w1s + in * kHid + ks * 16 + 4 * q);
                acc = mfma16(a3[ks], b, acc);
            }
            if (tile < ntiles)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    uint32_t s = s0 + 4 * q + j;
                    if (s < n) dXh[(size_t)s * 32 + in] = acc[j];
                }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- weight gradients
// C[M x N] += A^T B over n samples (A: n x lda BF16, B: n x ldb BF16), the
// sample axis as the MFMA K dimension.  One wave per (16x16 output tile,
// sample chunk); partial tiles meet in FP32 atomics.
__global__ __launch_bounds__(64) void k_sur_wgrad(const short* A, uint32_t lda, const short* B,
                                                  uint32_t ldb, uint32_t n, uint32_t mt_count,
                                                  uint32_t nt_count, uint32_t chunk, float* C,
                                                  uint32_t ldc) {
    uint32_t tile = blockIdx.x % (mt_count * nt_count);
    uint32_t ci = blockIdx.x / (mt_count * nt_count);
    uint32_t mt = tile / nt_count, nt = tile % nt_count;
    const int l = threadIdx.x, r16 = l & 15, q = l >> 4;
    uint32_t sb = ci * chunk;
    uint32_t se = sb + chunk < n ? sb + chunk : n;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t s0 = sb; s0 < se; s0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint32_t s = s0 + 4 * q + i;
            bool ok = s < se;
            a[i] = ok ? A[(size_t)s * lda + mt * 16 + r16] : (short)0;
            b[i] = ok ? B[(size_t)s * ldb + nt * 16 + r16] : (short)0;
        }
        acc = mfma16(a, b, acc);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
        atomicAdd(&C[(size_t)(mt * 16 + 4 * q + j) * ldc + nt * 16 + r16], acc[j]);
}

// ---------------------------------------------------------------- hash backward
// Sort key of a sample: a hash of its unit-cube position bits, so samples
// at one position sort next to each other.
__global__ __launch_bounds__(256) void k_sur_pos_key(SurDev d, const float* rec, uint32_t stride, uint32_t n,
                                                     uint32_t* key) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float p[3];
    unit_position(d, rec + (size_t)stride * s, p);
    uint32_t h = __float_as_uint(p[0]) * 0x9e3779b1u;
    h = (h ^ (h >> 15) ^ __float_as_uint(p[1])) * 0x85ebca77u;
    h = (h ^ (h >> 13) ^ __float_as_uint(p[2])) * 0xc2b2ae3du;
    key[s] = h ^ (h >> 16);
}

// Trilinear scatter of dX into the FP32 table gradient, samples taken in
// the order `order` (samples at one position adjacent).  Lanes of a wave
// whose positions are bitwise equal form a run: the corner weights are the
// run's, so a segmented scan sums the run's 32 gradient values and only its
// last lane adds them to the table (one atomic per word and run instead of
// one per sample; training samples concentrate on few positions).
__global__ __launch_bounds__(256) void k_sur_hash_bwd(SurDev d, const float* rec, uint32_t stride,
                                                      uint32_t n, const uint32_t* order, const float* dXh,
                                                      float* grad) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63u;
    const bool valid = j < n;
    const uint32_t s = valid ? (order ? order[j] : j) : 0u;
    float p[3] = {0.0f, 0.0f, 0.0f};
    if (valid) unit_position(d, rec + (size_t)stride * s, p);
    // run heads: the first lane, an invalid neighbour, or a new position
    const uint32_t px = __float_as_uint(p[0]), py = __float_as_uint(p[1]), pz = __float_as_uint(p[2]);
    const uint32_t qx = __shfl_up(px, 1, 64), qy = __shfl_up(py, 1, 64), qz = __shfl_up(pz, 1, 64);
    const bool prev_valid = __shfl_up(valid ? 1u : 0u, 1, 64) != 0u;
    const bool head = lane == 0 || !valid || !prev_valid || px != qx || py != qy || pz != qz;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const uint32_t my_head = 63u - (uint32_t)__clzll(heads & upto);
    const bool tail = valid && (lane == 63 || ((heads >> (lane + 1)) & 1ull) != 0ull);
    float g[32];
    const float* gs = dXh + (size_t)s * 32;
#pragma unroll
    for (int k = 0; k < 32; ++k) g[k] = valid ? gs[k] : 0.0f;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const bool take = lane >= (uint32_t)off && lane - (uint32_t)off >= my_head;
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const float v = __shfl_up(g[k], off, 64);
            if (take) g[k] += v;
        }
    }
    if (!tail) return;
    for (int l = 0; l < kLevels; ++l) {
        const float g0 = g[2 * l], g1 = g[2 * l + 1];
        if (g0 == 0.0f && g1 == 0.0f) continue;
        float res = (float)d.res[l];
        float gx = p[0] * res, gy = p[1] * res, gz = p[2] * res;
        float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
        uint32_t cx = (uint32_t)fx, cy = (uint32_t)fy, cz = (uint32_t)fz;
        float tx = gx - fx, ty = gy - fy, tz = gz - fz;
        float* gt = grad + (size_t)l * kTable * kFeat;
        for (int c = 0; c < 8; ++c) {
            uint32_t bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
            float w = (bx ? tx : 1.0f - tx) * (by ? ty : 1.0f - ty) * (bz ? tz : 1.0f - tz);
            uint32_t idx = grid_index(cx + bx, cy + by, cz + bz, d.res[l], d.dense[l]);
            atomicAdd(&gt[2ull * idx], w * g0);
            atomicAdd(&gt[2ull * idx + 1], w * g1);
        }
    }
}

// ---------------------------------------------------------------- Adam
__global__ __launch_bounds__(256) void k_sur_adam(float* param, float* grad, float* m, float* v,
                                                  size_t count, float lr, float b1, float b2,
                                                  float eps, float c1, float c2, short* bf_copy,
                                                  __half* h_copy, float* ema, float decay,
                                                  float ce, short* ema_bf, __half* ema_h) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) {
        float g = grad[i];
        float mi = b1 * m[i] + (1.0f - b1) * g;
        float vi = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        float mh = mi / c1, vh = vi / c2;
        float p = param[i] - lr * mh / (sqrtf(vh) + eps);
        param[i] = p;
        grad[i] = 0.0f;
        if (bf_copy) bf_copy[i] = f2bf(p);
        if (h_copy) h_copy[i] = __float2half(p);
        if (ema) {
            // zero-started average with bias correction ce = 1 - decay^t
            float e = decay * ema[i] + (1.0f - decay) * p;
            ema[i] = e;
            float ec = e / ce;
            if (ema_bf) ema_bf[i] = f2bf(ec);
            if (ema_h) ema_h[i] = __float2half(ec);
        }
    }
}

__global__ void k_sur_copies(const float* src, size_t count, short* bf, __half* hf) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) {
        if (bf) bf[i] = f2bf(src[i]);
        if (hf) hf[i] = __float2half(src[i]);
    }
}

__global__ void k_sur_fill(float* p, size_t count, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

constexpr size_t kW1 = (size_t)kIn * kHid, kW2 = (size_t)kHid * kHid, kW3 = (size_t)kHid * kOut;
constexpr size_t kMlp = kW1 + kW2 + kW3;
constexpr size_t kTableElems = (size_t)kLevels * kTable * kFeat;

uint32_t mix32(uint32_t v) {
    v ^= v >> 16; v *= 0x7feb352du; v ^= v >> 15; v *= 0x846ca68bu; v ^= v >> 16;
    return v;
}
float uniform01(uint32_t seed, uint32_t i) {
    return ((float)(mix32(seed * 0x9e3779b9u + mix32(i)) >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

void sh_constants(float* c) {
    const double pi = 3.14159265358979323846;
    for (int l = 0; l < 8; ++l)
        for (int m = 0; m <= l; ++m) {
            double num = 1.0, den = 1.0;
            for (int k = 1; k <= l - m; ++k) num *= k;
            for (int k = 1; k <= l + m; ++k) den *= k;
            double K = std::sqrt((2.0 * l + 1.0) / (4.0 * pi) * num / den);
            if (m == 0) c[l * l + l] = (float)K;
            else c[l * l + l + m] = c[l * l + l - m] = (float)(std::sqrt(2.0) * K);
        }
}

uint32_t grid_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

}  // namespace

// ======================================================================== host
struct lux_surrogate {
    lux_ctx* ctx = nullptr;
    lux_surrogate_params prm{};
    SurDev dev_train{}, dev_infer{};
    // parameters: FP32 masters, Adam moments, gradients, low-precision copies
    lux_buffer table_f32, table_m, table_v, table_g, table_h;
    lux_buffer mlp_f32, mlp_m, mlp_v, mlp_g, mlp_bf;  // W1 | W2 | W3
    // exponential moving average of the parameters (zero-started, bias
    // corrected over ema_steps updates) and its copies (inference)
    bool ema = false;
    uint32_t ema_steps = 0;
    lux_buffer table_ema, table_ema_h, mlp_ema, mlp_ema_bf;
    lux_buffer sh_const;
    // per-call scratch sized for the largest batch seen
    lux_buffer X, H1, H2, Y, dYb, dH1, dH2, dXh, loss;
    lux_buffer bwd_sort;  // position keys, sorted keys, order, sort temporaries and histogram
    uint32_t cap = 0;
    uint32_t step = 0;
};

namespace {

int sur_ensure(lux_surrogate* s, uint32_t n) {
    lux_ctx* ctx = s->ctx;
    if (n <= s->cap) return LUX_OK;
    size_t N = n;
    if (int rc = lux_ensure(ctx, s->X, 2 * kIn * N)) return rc;
    if (int rc = lux_ensure(ctx, s->H1, 2 * kHid * N)) return rc;
    if (int rc = lux_ensure(ctx, s->H2, 2 * kHid * N)) return rc;
    if (int rc = lux_ensure(ctx, s->Y, 4 * kOut * N)) return rc;
    if (int rc = lux_ensure(ctx, s->dYb, 2 * kOut * N)) return rc;
    if (int rc = lux_ensure(ctx, s->dH1, 2 * kHid * N)) return rc;
    if (int rc = lux_ensure(ctx, s->dH2, 2 * kHid * N)) return rc;
    if (int rc = lux_ensure(ctx, s->dXh, 4 * 32 * N)) return rc;
    if (int rc = lux_ensure(ctx, s->loss, 16)) return rc;
    s->cap = n;
    return LUX_OK;
}

void fill_dev(lux_surrogate* s, SurDev& d, const lux_buffer& table_h, const lux_buffer& mlp_bf) {
    d.table = static_cast<const __half*>(table_h.ptr);
    const short* bf = static_cast<const short*>(mlp_bf.ptr);
    d.w1 = bf;
    d.w2 = bf + kW1;
    d.w3 = bf + kW1 + kW2;
    d.sh_const = static_cast<const float*>(s->sh_const.ptr);
    for (int l = 0; l < kLevels; ++l) {
        double r = std::floor((double)s->prm.base_resolution * std::pow((double)s->prm.growth, (double)l));
        d.res[l] = (uint32_t)r;
        double cells = (r + 1.0) * (r + 1.0) * (r + 1.0);
        d.dense[l] = cells <= (double)kTable ? 1u : 0u;
    }
    for (int a = 0; a < 3; ++a) {
        d.box_min[a] = s->prm.box_min[a];
        float ext = s->prm.box_max[a] - s->prm.box_min[a];
        d.box_inv[a] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
}

void refresh_dev(lux_surrogate* s) {
    fill_dev(s, s->dev_train, s->table_h, s->mlp_bf);
    if (s->ema) fill_dev(s, s->dev_infer, s->table_ema_h, s->mlp_ema_bf);
    else s->dev_infer = s->dev_train;
}

int fill(lux_ctx* ctx, lux_buffer& b, size_t count, float v) {
    uint32_t blocks = grid_for(count, 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_sur_fill, dim3(blocks ? blocks : 1), dim3(256), 0, ctx->stream,
                       static_cast<float*>(b.ptr), count, v);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

// Low-precision copies of the masters; with EMA, the average restarts (its
// copies equal the masters' until the next step).
int refresh_copies(lux_surrogate* s) {
    lux_ctx* ctx = s->ctx;
    hipLaunchKernelGGL(k_sur_copies, dim3(8192), dim3(256), 0, ctx->stream,
                       static_cast<const float*>(s->table_f32.ptr), kTableElems, (short*)nullptr,
                       static_cast<__half*>(s->table_h.ptr));
    hipLaunchKernelGGL(k_sur_copies, dim3(64), dim3(256), 0, ctx->stream,
                       static_cast<const float*>(s->mlp_f32.ptr), kMlp,
                       static_cast<short*>(s->mlp_bf.ptr), (__half*)nullptr);
    if (s->ema) {
        s->ema_steps = 0;
        if (int rc = fill(ctx, s->table_ema, kTableElems, 0.0f)) return rc;
        if (int rc = fill(ctx, s->mlp_ema, kMlp, 0.0f)) return rc;
        hipLaunchKernelGGL(k_sur_copies, dim3(8192), dim3(256), 0, ctx->stream,
                           static_cast<const float*>(s->table_f32.ptr), kTableElems, (short*)nullptr,
                           static_cast<__half*>(s->table_ema_h.ptr));
        hipLaunchKernelGGL(k_sur_copies, dim3(64), dim3(256), 0, ctx->stream,
                           static_cast<const float*>(s->mlp_f32.ptr), kMlp,
                           static_cast<short*>(s->mlp_ema_bf.ptr), (__half*)nullptr);
    }
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

// Deterministic initialisation: table U(-1e-4, 1e-4), weights Xavier-uniform;
// Adam state and gradients zero.
int sur_init(lux_surrogate* s) {
    lux_ctx* ctx = s->ctx;
    const uint32_t seed = s->prm.seed;
    std::vector<float> table(kTableElems), mlp(kMlp);
    for (size_t i = 0; i < kTableElems; ++i)
        table[i] = (uniform01(seed, (uint32_t)i) * 2.0f - 1.0f) * 1e-4f;
    const size_t sizes[3] = {kW1, kW2, kW3};
    const int fan[3][2] = {{kIn, kHid}, {kHid, kHid}, {kHid, kOut}};
    size_t off = 0;
    for (int m = 0; m < 3; ++m) {
        float lim = std::sqrt(6.0f / (float)(fan[m][0] + fan[m][1]));
        for (size_t i = 0; i < sizes[m]; ++i)
            mlp[off + i] = (uniform01(seed ^ (0x51u + (uint32_t)m), (uint32_t)(off + i)) * 2.0f - 1.0f) * lim;
        off += sizes[m];
    }
    float shc[64];
    sh_constants(shc);
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUX_HIP(ctx, hipMemcpy(s->table_f32.ptr, table.data(), 4 * kTableElems, hipMemcpyHostToDevice));
    LUX_HIP(ctx, hipMemcpy(s->mlp_f32.ptr, mlp.data(), 4 * kMlp, hipMemcpyHostToDevice));
    LUX_HIP(ctx, hipMemcpy(s->sh_const.ptr, shc, sizeof shc, hipMemcpyHostToDevice));
    lux_buffer* tz[] = {&s->table_m, &s->table_v, &s->table_g};
    for (lux_buffer* b : tz)
        if (int rc = fill(ctx, *b, kTableElems, 0.0f)) return rc;
    lux_buffer* mz[] = {&s->mlp_m, &s->mlp_v, &s->mlp_g};
    for (lux_buffer* b : mz)
        if (int rc = fill(ctx, *b, kMlp, 0.0f)) return rc;
    if (int rc = refresh_copies(s)) return rc;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->step = 0;
    refresh_dev(s);
    return LUX_OK;
}

int sur_encode(lux_surrogate* s, const SurDev& dev, const float* rec, uint32_t stride, uint32_t n) {
    lux_ctx* ctx = s->ctx;
    hipLaunchKernelGGL(k_sur_encode, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, dev, rec,
                       stride, n, static_cast<short*>(s->X.ptr));
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

uint32_t tile_blocks(lux_ctx* ctx, uint32_t n) {
    uint32_t ntiles = (n + 15) / 16;
    uint32_t blocks = (ntiles + 3) / 4;
    uint32_t maxb = (uint32_t)ctx->cu_count * 4u;
    if (blocks > maxb) blocks = maxb;
    return blocks ? blocks : 1u;
}

// keep: training pass (current parameters, activations kept for backward);
// otherwise inference (EMA parameters when enabled).
int sur_forward(lux_surrogate* s, const float* rec, uint32_t stride, uint32_t n, bool keep) {
    lux_ctx* ctx = s->ctx;
    const SurDev& dev = keep ? s->dev_train : s->dev_infer;
    if (int rc = sur_encode(s, dev, rec, stride, n)) return rc;
    uint32_t ntiles = (n + 15) / 16;
    hipLaunchKernelGGL(k_sur_forward, dim3(tile_blocks(ctx, n)), dim3(256), 0, ctx->stream, dev,
                       static_cast<const short*>(s->X.ptr), n,
                       keep ? static_cast<short*>(s->H1.ptr) : (short*)nullptr,
                       static_cast<short*>(s->H2.ptr), static_cast<float*>(s->Y.ptr), ntiles);
    LUX_HIP(ctx, hipGetLastError());
    return LUX_OK;
}

}  // namespace

int lux_surrogate_forward_internal(lux_surrogate* s, const float* queries, uint32_t n, const float** Y) {
    if (int rc = sur_ensure(s, n)) return rc;
    if (int rc = sur_forward(s, queries, 9, n, false)) return rc;
    *Y = static_cast<const float*>(s->Y.ptr);
    return LUX_OK;
}

lux_ctx* lux_surrogate_context(lux_surrogate* s) { return s->ctx; }

size_t lux_surrogate_scratch_bytes(uint32_t n) {
    size_t N = ((size_t)n + 15) / 16 * 16;
    return 2 * kIn * N + 4 * kOut * N;
}

int lux_surrogate_forward_into(lux_surrogate* s, hipStream_t st, const float* queries, uint32_t n, void* scratch,
                               const float** Y) {
    lux_ctx* ctx = s->ctx;
    size_t N = ((size_t)n + 15) / 16 * 16;
    short* X = static_cast<short*>(scratch);
    float* out = reinterpret_cast<float*>(X + kIn * N);
    const SurDev& dev = s->dev_infer;
    // inference: no activations are kept (H1 == nullptr leaves H2 unused)
    hipLaunchKernelGGL(k_sur_encode, dim3(grid_for(n, 256)), dim3(256), 0, st, dev, queries, 9u, n, X);
    hipLaunchKernelGGL(k_sur_forward, dim3(tile_blocks(ctx, n)), dim3(256), 0, st, dev,
                       static_cast<const short*>(X), n, (short*)nullptr, (short*)nullptr, out, (n + 15) / 16);
    LUX_HIP(ctx, hipGetLastError());
    *Y = out;
    return LUX_OK;
}

extern "C" void lux_surrogate_default_params(lux_surrogate_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->levels = kLevels;
    p->features = kFeat;
    p->log2_table = kTableLog2;
    p->base_resolution = 16.0f;
    p->growth = 1.5f;
    p->box_max[0] = p->box_max[1] = p->box_max[2] = 1.0f;
    p->learning_rate = 1e-2f;
    p->error_weight = 0.1f;
    p->seed = 1u;
    p->ema_decay = 0.95f;
}

extern "C" int lux_surrogate_create(lux_ctx* ctx, const lux_surrogate_params* params,
                                    lux_surrogate** out) {
    if (!ctx || !out) return LUX_ERR_INVALID;
    *out = nullptr;
    if (!params) return lux_fail(ctx, LUX_ERR_INVALID, "lux_surrogate_create: null params");
    if (params->levels != (uint32_t)kLevels || params->features != (uint32_t)kFeat ||
        params->log2_table != kTableLog2)
        return lux_fail(ctx, LUX_ERR_INVALID,
                        "lux_surrogate_create: the encoder is built for 16 levels x 2 features, 2^19 entries");
    double top = (double)params->base_resolution * std::pow((double)params->growth, 15.0);
    if (!(params->base_resolution >= 1.0f) || !(params->growth >= 1.0f) || !(top <= 1.0e9))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_surrogate_create: base resolution / growth out of range");
    if (!(params->learning_rate > 0.0f) || !std::isfinite(params->learning_rate))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_surrogate_create: learning rate must be positive");
    if (!(params->table_learning_rate >= 0.0f) || !std::isfinite(params->table_learning_rate))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_surrogate_create: table learning rate must be >= 0");
    if (!(params->ema_decay >= 0.0f && params->ema_decay < 1.0f))
        return lux_fail(ctx, LUX_ERR_INVALID, "lux_surrogate_create: EMA decay must be in [0, 1)");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(params->box_min[a]) || !std::isfinite(params->box_max[a]) ||
            !(params->box_max[a] > params->box_min[a]))
            return lux_fail(ctx, LUX_ERR_INVALID, "lux_surrogate_create: empty encoder box");
    lux_surrogate* s = new lux_surrogate();
    s->ctx = ctx;
    s->prm = *params;
    s->ema = params->ema_decay > 0.0f;
    int rc = LUX_OK;
    lux_buffer* big[] = {&s->table_f32, &s->table_m, &s->table_v, &s->table_g};
    for (lux_buffer* b : big)
        if (!rc) rc = lux_ensure(ctx, *b, 4 * kTableElems);
    if (!rc) rc = lux_ensure(ctx, s->table_h, 2 * kTableElems);
    lux_buffer* small[] = {&s->mlp_f32, &s->mlp_m, &s->mlp_v, &s->mlp_g};
    for (lux_buffer* b : small)
        if (!rc) rc = lux_ensure(ctx, *b, 4 * kMlp);
    if (!rc) rc = lux_ensure(ctx, s->mlp_bf, 2 * kMlp);
    if (!rc) rc = lux_ensure(ctx, s->sh_const, 4 * 64);
    if (s->ema) {
        if (!rc) rc = lux_ensure(ctx, s->table_ema, 4 * kTableElems);
        if (!rc) rc = lux_ensure(ctx, s->table_ema_h, 2 * kTableElems);
        if (!rc) rc = lux_ensure(ctx, s->mlp_ema, 4 * kMlp);
        if (!rc) rc = lux_ensure(ctx, s->mlp_ema_bf, 2 * kMlp);
    }
    if (!rc) rc = sur_init(s);
    if (rc) {
        lux_surrogate_destroy(s);
        return rc;
    }
    *out = s;
    return LUX_OK;
}

extern "C" void lux_surrogate_destroy(lux_surrogate* s) {
    if (!s) return;
    if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
    lux_buffer* bufs[] = {&s->table_f32, &s->table_m, &s->table_v, &s->table_g, &s->table_h,
                          &s->mlp_f32,   &s->mlp_m,   &s->mlp_v,   &s->mlp_g,   &s->mlp_bf,
                          &s->sh_const,  &s->X,       &s->H1,      &s->H2,      &s->Y,
                          &s->dYb,       &s->dH1,     &s->dH2,     &s->dXh,     &s->loss,
                          &s->table_ema, &s->table_ema_h, &s->mlp_ema, &s->mlp_ema_bf, &s->bwd_sort};
    for (lux_buffer* b : bufs) lux_release(*b);
    delete s;
}

extern "C" int lux_surrogate_reset(lux_surrogate* s) {
    if (!s) return LUX_ERR_INVALID;
    return sur_init(s);
}

extern "C" int lux_surrogate_encode(lux_surrogate* s, const float* queries, uint32_t n, uint16_t* X) {
    if (!s) return LUX_ERR_INVALID;
    lux_ctx* ctx = s->ctx;
    if (n == 0) return LUX_OK;
    if (!queries || !X) return lux_fail(ctx, LUX_ERR_INVALID, "lux_surrogate_encode: null buffer");
    if (int rc = sur_ensure(s, n)) return rc;
    if (int rc = sur_encode(s, s->dev_infer, queries, 9, n)) return rc;
    LUX_HIP(ctx, hipMemcpyAsync(X, s->X.ptr, 2ull * kIn * n, hipMemcpyDeviceToDevice, ctx->stream));
    return LUX_OK;
}

extern "C" int lux_surrogate_infer(lux_surrogate* s, const float* queries, uint32_t n, float* out) {
    if (!s) return LUX_ERR_INVALID;
    lux_ctx* ctx = s->ctx;
    if (n == 0) return LUX_OK;
    if (!queries || !out) return lux_fail(ctx, LUX_ERR_INVALID, "lux_surrogate_infer: null buffer");
    const float* Y = nullptr;
    if (int rc = lux_surrogate_forward_internal(s, queries, n, &Y)) return rc;
    LUX_HIP(ctx, hipMemcpyAsync(out, Y, 4ull * kOut * n, hipMemcpyDeviceToDevice, ctx->stream));
    return LUX_OK;
}

extern "C" int lux_surrogate_train_step(lux_surrogate* s, const float* samples, uint32_t n,
                                        float* loss_out) {
    if (!s) return LUX_ERR_INVALID;
    lux_ctx* ctx = s->ctx;
    if (n == 0) return LUX_OK;
    if (!samples) return lux_fail(ctx, LUX_ERR_INVALID, "lux_surrogate_train_step: null buffer");
    if (int rc = sur_ensure(s, n)) return rc;
    hipStream_t st = ctx->stream;
    if (int rc = sur_forward(s, samples, 12, n, true)) return rc;
    if (int rc = fill(ctx, s->loss, 1, 0.0f)) return rc;
    hipLaunchKernelGGL(k_sur_loss, dim3(grid_for(n, 256)), dim3(256), 0, st,
                       static_cast<const float*>(s->Y.ptr), samples, n, s->prm.error_weight,
                       static_cast<short*>(s->dYb.ptr), static_cast<float*>(s->loss.ptr));
    uint32_t ntiles = (n + 15) / 16;
    hipLaunchKernelGGL(k_sur_backward, dim3(tile_blocks(ctx, n)), dim3(256), 0, st, s->dev_train,
                       static_cast<const short*>(s->dYb.ptr), static_cast<const short*>(s->H1.ptr),
                       static_cast<const short*>(s->H2.ptr), n, static_cast<short*>(s->dH1.ptr),
                       static_cast<short*>(s->dH2.ptr), static_cast<float*>(s->dXh.ptr), ntiles);
    float* g = static_cast<float*>(s->mlp_g.ptr);
    const uint32_t chunk = 2048;
    uint32_t chunks = (n + chunk - 1) / chunk;
    // dW1 = X^T dH1, dW2 = H1^T dH2, dW3 = H2^T dY
    hipLaunchKernelGGL(k_sur_wgrad, dim3((kIn / 16) * (kHid / 16) * chunks), dim3(64), 0, st,
                       static_cast<const short*>(s->X.ptr), (uint32_t)kIn,
                       static_cast<const short*>(s->dH1.ptr), (uint32_t)kHid, n, (uint32_t)(kIn / 16),
                       (uint32_t)(kHid / 16), chunk, g, (uint32_t)kHid);
    hipLaunchKernelGGL(k_sur_wgrad, dim3((kHid / 16) * (kHid / 16) * chunks), dim3(64), 0, st,
                       static_cast<const short*>(s->H1.ptr), (uint32_t)kHid,
                       static_cast<const short*>(s->dH2.ptr), (uint32_t)kHid, n, (uint32_t)(kHid / 16),
                       (uint32_t)(kHid / 16), chunk, g + kW1, (uint32_t)kHid);
    hipLaunchKernelGGL(k_sur_wgrad, dim3((kHid / 16) * (kOut / 16) * chunks), dim3(64), 0, st,
                       static_cast<const short*>(s->H2.ptr), (uint32_t)kHid,
                       static_cast<const short*>(s->dYb.ptr), (uint32_t)kOut, n, (uint32_t)(kHid / 16),
                       (uint32_t)(kOut / 16), chunk, g + kW1 + kW2, (uint32_t)kOut);
    // the scatter in position order (a stable sort of position hashes)
    const size_t hw = lux_sort_hist_words(n);
    if (int rc = lux_ensure(ctx, s->bwd_sort, 4ull * (5ull * n + hw) + 64)) return rc;
    uint32_t* key = static_cast<uint32_t*>(s->bwd_sort.ptr);
    uint32_t* skey = key + n;
    uint32_t* order = skey + n;
    uint32_t* tk = order + n;
    uint32_t* tv = tk + n;
    uint32_t* hist = tv + n;
    hipLaunchKernelGGL(k_sur_pos_key, dim3(grid_for(n, 256)), dim3(256), 0, st, s->dev_train, samples, 12u, n, key);
    if (int rc = lux_sort_pairs_impl(ctx, key, n, skey, order, tk, tv, hist, 32)) return rc;
    hipLaunchKernelGGL(k_sur_hash_bwd, dim3(grid_for(n, 256)), dim3(256), 0, st, s->dev_train, samples, 12u, n,
                       static_cast<const uint32_t*>(order), static_cast<const float*>(s->dXh.ptr),
                       static_cast<float*>(s->table_g.ptr));
    LUX_HIP(ctx, hipGetLastError());
    if (s->prm.freeze_updates == 0u) {
        s->step += 1;
        const float b1 = 0.9f, b2 = 0.99f, eps = 1e-15f;
        float c1 = (float)(1.0 - std::pow((double)b1, (double)s->step));
        float c2 = (float)(1.0 - std::pow((double)b2, (double)s->step));
        if (s->ema) s->ema_steps += 1;
        float ce = (float)(1.0 - std::pow((double)s->prm.ema_decay, (double)s->ema_steps));
        hipLaunchKernelGGL(k_sur_adam, dim3(64), dim3(256), 0, st, static_cast<float*>(s->mlp_f32.ptr), g,
                           static_cast<float*>(s->mlp_m.ptr), static_cast<float*>(s->mlp_v.ptr), kMlp,
                           s->prm.learning_rate, b1, b2, eps, c1, c2, static_cast<short*>(s->mlp_bf.ptr),
                           (__half*)nullptr, s->ema ? static_cast<float*>(s->mlp_ema.ptr) : nullptr,
                           s->prm.ema_decay, ce, static_cast<short*>(s->mlp_ema_bf.ptr), (__half*)nullptr);
        hipLaunchKernelGGL(k_sur_adam, dim3(8192), dim3(256), 0, st, static_cast<float*>(s->table_f32.ptr),
                           static_cast<float*>(s->table_g.ptr), static_cast<float*>(s->table_m.ptr),
                           static_cast<float*>(s->table_v.ptr), kTableElems,
                           s->prm.table_learning_rate > 0.0f ? s->prm.table_learning_rate : s->prm.learning_rate, b1, b2,
                           eps, c1, c2, (short*)nullptr, static_cast<__half*>(s->table_h.ptr),
                           s->ema ? static_cast<float*>(s->table_ema.ptr) : nullptr, s->prm.ema_decay, ce,
                           (short*)nullptr, static_cast<__half*>(s->table_ema_h.ptr));
        LUX_HIP(ctx, hipGetLastError());
    }
    if (loss_out) {
        LUX_HIP(ctx, hipMemcpyAsync(loss_out, s->loss.ptr, 4, hipMemcpyDeviceToHost, st));
        LUX_HIP(ctx, hipStreamSynchronize(st));
    }
    return LUX_OK;
}

extern "C" int lux_surrogate_sizes(lux_surrogate* s, uint64_t* mlp_params, uint64_t* table_params,
                                   uint32_t* steps) {
    if (!s) return LUX_ERR_INVALID;
    if (mlp_params) *mlp_params = kMlp;
    if (table_params) *table_params = kTableElems;
    if (steps) *steps = s->step;
    return LUX_OK;
}

extern "C" int lux_surrogate_download(lux_surrogate* s, float* mlp, float* mlp_grad, float* table,
                                      float* table_grad) {
    if (!s) return LUX_ERR_INVALID;
    lux_ctx* ctx = s->ctx;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (mlp) LUX_HIP(ctx, hipMemcpy(mlp, s->mlp_f32.ptr, 4 * kMlp, hipMemcpyDeviceToHost));
    if (mlp_grad) LUX_HIP(ctx, hipMemcpy(mlp_grad, s->mlp_g.ptr, 4 * kMlp, hipMemcpyDeviceToHost));
    if (table) LUX_HIP(ctx, hipMemcpy(table, s->table_f32.ptr, 4 * kTableElems, hipMemcpyDeviceToHost));
    if (table_grad)
        LUX_HIP(ctx, hipMemcpy(table_grad, s->table_g.ptr, 4 * kTableElems, hipMemcpyDeviceToHost));
    return LUX_OK;
}

extern "C" int lux_surrogate_upload(lux_surrogate* s, const float* mlp, const float* table) {
    if (!s) return LUX_ERR_INVALID;
    lux_ctx* ctx = s->ctx;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (mlp) LUX_HIP(ctx, hipMemcpy(s->mlp_f32.ptr, mlp, 4 * kMlp, hipMemcpyHostToDevice));
    if (table) LUX_HIP(ctx, hipMemcpy(s->table_f32.ptr, table, 4 * kTableElems, hipMemcpyHostToDevice));
    if (int rc = refresh_copies(s)) return rc;
    LUX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LUX_OK;
}

extern "C" int lux_surrogate_zero_grad(lux_surrogate* s) {
    if (!s) return LUX_ERR_INVALID;
    if (int rc = fill(s->ctx, s->mlp_g, kMlp, 0.0f)) return rc;
    return fill(s->ctx, s->table_g, kTableElems, 0.0f);
}

// cnn.hip - the CU-split predictor: a small BF16 CNN on MFMA with FP32
// accumulation, forward and backward (training) kernels for gfx950.
//
// Network (per 32x32 luma block b, with normalised QP q = QP / 51):
//   x   = (b - 128) / 64                              1 x 32 x 32
//   a1  = relu(conv3x3(x,  W1) + b1)                  32 x 32 x 32
//   a2  = relu(conv3x3(a1, W2) + b2)                  32 x 32 x 32
//   a3  = relu(conv3x3(a2, W3) + b3)                  32 x 32 x 32
//   p   = mean over pixels of a3                      32
//   h   = relu(W4 [p; q] + b4)                        64
//   z   = W5 h + b5                                   6 logits, CodingTreeSplit order
// Convolutions are zero-padded, stride 1. Activations, their gradients and
// the convolution weights are stored as bf16 (round to nearest even); every
// sum accumulates in fp32 (MFMA v_mfma_f32_32x32x8_bf16). Pooling, the MLP,
// the loss and Adam run in fp32.
//
// Layouts: activations NHWC [n][y][x][c]; conv weights [cout][ky][kx][cin]
// (k = (ky * 3 + kx) * cin_count + cin), padded with zeros to a multiple of 8
// in k. The input gradient of a convolution is the convolution of the output
// gradient with the flipped, channel-swapped weights, so it runs on the
// forward kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cnn.h"

namespace rmh {
namespace cnn {
namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));

constexpr int kRowsPerWave = 2;  // output rows per wave of conv_fwd_kernel
typedef float float16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Forward convolution, one 256-lane workgroup per 4 * RW output rows of one
// image; wave w computes rows y0 + RW w .. + RW - 1, each 32 pixels x 32
// output channels as one 32x32 MFMA tile, K = 9 * CIN (padded to KP) in steps
// of 8. The input rows (with a one-sample zero border) and the weights are
// staged in LDS with 16-byte loads; each weight operand read feeds RW MFMAs.
//   out = act(conv(in, w) + bias); act: 0 none, 1 relu, 2 multiply by (mask > 0)
template <int CIN, int RW>
__global__ __launch_bounds__(256) void conv_fwd_kernel(const uint16_t* __restrict__ in, const uint16_t* __restrict__ w,
                                                       const float* __restrict__ bias, const uint16_t* __restrict__ mask,
                                                       uint16_t* __restrict__ out, int act) {
    constexpr int K = 9 * CIN, KP = (K + 7) & ~7;
    constexpr int OROWS = 4 * RW, ROWS = OROWS + 2, COLS = 34;
    __shared__ __align__(16) uint16_t tile[ROWS * COLS * CIN];
    __shared__ __align__(16) uint16_t wl[32 * KP];
    const int n = blockIdx.x, y0 = blockIdx.y * OROWS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint16_t* img = in + (size_t)n * 1024 * CIN;
    // stage rows y0 - 1 .. y0 + OROWS, columns -1 .. 32 (zero outside the image)
    if constexpr (CIN % 8 == 0) {
        constexpr int V = CIN / 8;  // 16-byte units per pixel
        uint4* t4 = reinterpret_cast<uint4*>(tile);
        const uint4* i4 = reinterpret_cast<const uint4*>(img);
        for (int e = tid; e < ROWS * COLS * V; e += 256) {
            const int v = e % V, col = (e / V) % COLS, row = e / (V * COLS);
            const int y = y0 - 1 + row, x = col - 1;
            t4[e] = (y >= 0 && y < 32 && x >= 0 && x < 32) ? i4[(y * 32 + x) * V + v] : make_uint4(0, 0, 0, 0);
        }
    } else {
        for (int e = tid; e < ROWS * COLS * CIN; e += 256) {
            const int c = e % CIN, col = (e / CIN) % COLS, row = e / (CIN * COLS);
            const int y = y0 - 1 + row, x = col - 1;
            tile[e] = (y >= 0 && y < 32 && x >= 0 && x < 32) ? img[(y * 32 + x) * CIN + c] : (uint16_t)0;
        }
    }
    static_assert((32 * KP) % 8 == 0, "weights are staged in 16-byte units");
    for (int e = tid; e < 32 * KP / 8; e += 256)
        reinterpret_cast<uint4*>(wl)[e] = reinterpret_cast<const uint4*>(w)[e];
    __syncthreads();

    const int i = lane & 31, kq = (lane >> 5) * 4;  // A: pixel i; k offset within a step
    float16_t acc[RW];
#pragma unroll
    for (int r = 0; r < RW; r++) acc[r] = float16_t{};
#pragma unroll 4
    for (int s = 0; s < KP / 8; s++) {
        const int k0 = s * 8 + kq;
        short4_t b;
        const uint2 vb = *reinterpret_cast<const uint2*>(&wl[i * KP + k0]);  // B: output channel j = i
        b[0] = (short)(vb.x & 0xffff);
        b[1] = (short)(vb.x >> 16);
        b[2] = (short)(vb.y & 0xffff);
        b[3] = (short)(vb.y >> 16);
#pragma unroll
        for (int r = 0; r < RW; r++) {
            const int y = wave * RW + r;  // output row within the workgroup
            short4_t a;
            if (CIN % 4 == 0) {
                const int tap = k0 / CIN, c = k0 % CIN, ky = tap / 3, kx = tap % 3;
                const uint2 va = *reinterpret_cast<const uint2*>(&tile[((y + ky) * COLS + i + kx) * CIN + c]);
                a[0] = (short)(va.x & 0xffff);
                a[1] = (short)(va.x >> 16);
                a[2] = (short)(va.y & 0xffff);
                a[3] = (short)(va.y >> 16);
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int k = k0 + q;
                    short v = 0;
                    if (k < K) {
                        const int tap = k / CIN, c = k % CIN, ky = tap / 3, kx = tap % 3;
                        v = (short)tile[((y + ky) * COLS + i + kx) * CIN + c];
                    }
                    a[q] = v;
                }
            }
            acc[r] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[r], 0, 0, 0);
        }
    }
    // D: channel j = lane % 32; pixel = 8 * (q / 4) + 4 * (lane / 32) + q % 4
    const int j = lane & 31;
    const float bj = bias ? bias[j] : 0.0f;
#pragma unroll
    for (int r = 0; r < RW; r++) {
        const int oy = y0 + wave * RW + r;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int px = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            const size_t o = ((size_t)n * 1024 + oy * 32 + px) * 32 + j;
            float v = acc[r][q] + bj;
            if (act == 1) v = v > 0.0f ? v : 0.0f;
            if (act == 2 && !(bf2f(mask[o]) > 0.0f)) v = 0.0f;
            out[o] = f2bf(v);
        }
    }
}

// Weight gradient of a 32-channel-input convolution for one image and one
// tap (ky, kx): dW[cout][tap][cin] += sum_p dz[p][cout] * a[p + tap][cin],
// as a 32x32 MFMA tile over K = pixels. Wave w takes rows 8w .. 8w+7; the
// four partial tiles are summed in LDS and written to part[n][cout][tap][cin].
__global__ __launch_bounds__(256) void conv_wgrad32_kernel(const uint16_t* __restrict__ dz,
                                                           const uint16_t* __restrict__ a, float* __restrict__ part) {
    __shared__ float red[4][32 * 32];
    const int n = blockIdx.x, tap = blockIdx.y, ky = tap / 3, kx = tap % 3;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint16_t* dzi = dz + (size_t)n * 1024 * 32;
    const uint16_t* ai = a + (size_t)n * 1024 * 32;
    const int i = lane & 31, kq = (lane >> 5) * 4;
    float16_t acc = {};
    for (int s = 0; s < 32; s++) {  // 8 rows x 32 pixels = 256 pixels = 32 steps of 8
        const int p0 = wave * 256 + s * 8 + kq;
        short4_t va, vb;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int p = p0 + q, py = p >> 5, px = p & 31;
            va[q] = (short)dzi[p * 32 + i];  // A[cout = i][pixel]
            const int sy = py + ky - 1, sx = px + kx - 1;
            vb[q] = (sy >= 0 && sy < 32 && sx >= 0 && sx < 32) ? (short)ai[(sy * 32 + sx) * 32 + i] : (short)0;  // B[pixel][cin = i]
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(va, vb, acc, 0, 0, 0);
    }
    const int j = lane & 31;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int co = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        red[wave][co * 32 + j] = acc[r];
    }
    __syncthreads();
    for (int e = tid; e < 1024; e += 256) {
        const int co = e >> 5, ci = e & 31;
        part[(size_t)n * 9216 + (co * 9 + tap) * 32 + ci] = red[0][e] + red[1][e] + red[2][e] + red[3][e];
    }
}

// Weight gradient of the 1-channel first convolution (VALU): lane (cout, tap)
// sums over one quarter (8 rows) of the pixels of one image;
// part[n * 4 + quarter][cout][tap].
__global__ __launch_bounds__(288) void conv_wgrad1_kernel(const uint16_t* __restrict__ dz, const uint16_t* __restrict__ x,
                                                          float* __restrict__ part) {
    __shared__ float xs[34 * 34];
    const int n = blockIdx.x, quarter = blockIdx.y, tid = threadIdx.x;
    for (int e = tid; e < 34 * 34; e += 288) {
        const int r = e / 34 - 1, c = e % 34 - 1;
        xs[e] = (r >= 0 && r < 32 && c >= 0 && c < 32) ? bf2f(x[(size_t)n * 1024 + r * 32 + c]) : 0.0f;
    }
    __syncthreads();
    const int co = tid / 9, tap = tid % 9, ky = tap / 3, kx = tap % 3;
    const uint16_t* d = dz + (size_t)n * 1024 * 32;
    float s = 0.0f;
#pragma unroll 4
    for (int p = quarter * 256; p < quarter * 256 + 256; p++) s += bf2f(d[p * 32 + co]) * xs[((p >> 5) + ky) * 34 + (p & 31) + kx];
    part[((size_t)n * 4 + quarter) * 288 + co * 9 + tap] = s;
}

// Per-channel sum over the 1024 pixels of one image's [p][32] bf16 map, in a
// fixed order: 16-byte loads (8 channels of one pixel), lane t always holding
// channels 8 (t & 3) .. + 7, then the 64 lanes of each channel group added in
// lane order. Returns the sum for channel tid (< 32) of the workgroup.
__device__ __forceinline__ float channel_sum(const uint16_t* __restrict__ map, float (*red)[9]) {
    const int tid = threadIdx.x;
    const uint4* d = reinterpret_cast<const uint4*>(map);
    float acc[8] = {};
#pragma unroll 4
    for (int k = 0; k < 16; k++) {
        const uint4 v = d[tid + 256 * k];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acc[2 * j] += bf2f((uint16_t)(w[j] & 0xffff));
            acc[2 * j + 1] += bf2f((uint16_t)(w[j] >> 16));
        }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) red[tid][j] = acc[j];
    __syncthreads();
    float t = 0.0f;
    if (tid < 32)
        for (int q = tid >> 3; q < 256; q += 4) t += red[q][tid & 7];
    return t;
}

// Bias gradient partials: part[n][c] = sum_p dz[n][p][c].
__global__ __launch_bounds__(256) void bias_grad_kernel(const uint16_t* __restrict__ dz, float* __restrict__ part) {
    __shared__ float red[256][9];
    const int n = blockIdx.x;
    const float t = channel_sum(dz + (size_t)n * 1024 * 32, red);
    if (threadIdx.x < 32) part[(size_t)n * 32 + threadIdx.x] = t;
}

// Sums partials over images in a fixed order: out[e] = sum_k part[k][e],
// k = 0 .. n - 1. A workgroup takes 32 elements; its 8 lane groups sum
// images g, g + 8, ... and the 8 group sums are added in order.
__global__ __launch_bounds__(256) void reduce_kernel(const float* __restrict__ part, int n, int len, float* __restrict__ out) {
    __shared__ float red[8][32];
    const int tid = threadIdx.x, el = tid & 31, g = tid >> 5;
    const int e = blockIdx.x * 32 + el;
    float s = 0.0f;
    if (e < len) {
#pragma unroll 4
        for (int k = g; k < n; k += 8) s += part[(size_t)k * len + e];
    }
    red[g][el] = s;
    __syncthreads();
    if (g == 0 && e < len) {
        float t = 0.0f;
        for (int q = 0; q < 8; q++) t += red[q][el];
        out[e] = t;
    }
}

// Input of the first convolution: x = (b - 128) / 64 as bf16 (exact).
__global__ void input_kernel(const uint8_t* __restrict__ blocks, int total, uint16_t* __restrict__ x) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < total) x[e] = f2bf(((float)blocks[e] - 128.0f) * (1.0f / 64.0f));
}

// Global average pool: p[n][c] = mean_p a3[n][p][c].
__global__ __launch_bounds__(256) void pool_kernel(const uint16_t* __restrict__ a3, float* __restrict__ pooled) {
    __shared__ float red[256][9];
    const int n = blockIdx.x;
    const float t = channel_sum(a3 + (size_t)n * 1024 * 32, red);
    if (threadIdx.x < 32) pooled[n * 32 + threadIdx.x] = t * (1.0f / 1024.0f);
}

// MLP forward (and, with labels or targets, the softmax cross-entropy
// backward), one 64-lane workgroup per sample, lane o owning hidden unit o.
// Writes probs[n][6]; in training also loss[n], dpool[n][32] and the
// per-sample vectors the weight gradients are built from: hidden hv[n][64],
// its gradient dhv[n][64] and the logit gradient dzv[n][6]. Sums over hidden
// units run in unit order (LDS), so results do not depend on the lane layout.
__global__ __launch_bounds__(64) void mlp_kernel(const float* __restrict__ pooled, const float* __restrict__ qp,
                                                 const float* __restrict__ w4, const float* __restrict__ b4,
                                                 const float* __restrict__ w5, const float* __restrict__ b5, int n,
                                                 float* __restrict__ probs, const uint8_t* __restrict__ labels,
                                                 const float* __restrict__ targets, float inv_batch,
                                                 float* __restrict__ loss, float* __restrict__ dpool,
                                                 float* __restrict__ hv, float* __restrict__ dhv,
                                                 float* __restrict__ dzv) {
    __shared__ float in[33], h[64], dh[64], zs[6], dz[6];
    const int s = blockIdx.x, o = threadIdx.x;
    if (o < 32) in[o] = pooled[s * 32 + o];
    if (o == 32) in[32] = qp[s];
    __syncthreads();
    float t = b4[o];
    for (int c = 0; c < 33; c++) t += w4[o * 33 + c] * in[c];
    h[o] = t > 0.0f ? t : 0.0f;
    __syncthreads();
    if (o < 6) {
        float z = b5[o];
        for (int c = 0; c < 64; c++) z += w5[o * 64 + c] * h[c];
        zs[o] = z;
    }
    __syncthreads();
    float mx = -1e30f;
    for (int k = 0; k < 6; k++) mx = zs[k] > mx ? zs[k] : mx;
    float den = 0.0f, e[6];
    for (int k = 0; k < 6; k++) {
        e[k] = expf(zs[k] - mx);
        den += e[k];
    }
    if (o < 6) probs[s * 6 + o] = e[o] / den;
    if (!labels && !targets) return;
    // target distribution: one-hot label (split modes 0..5) or a given distribution
    float tg[6];
    if (targets) {
        for (int k = 0; k < 6; k++) tg[k] = targets[s * 6 + k];
    } else {
        const int lab = labels[s] < 6 ? labels[s] : 5;
        for (int k = 0; k < 6; k++) tg[k] = k == lab ? 1.0f : 0.0f;
    }
    float tsum = 0.0f;
    for (int k = 0; k < 6; k++) tsum += tg[k];
    if (o == 0) {
        const float lden = logf(den);
        float l = 0.0f;
        for (int k = 0; k < 6; k++) l -= tg[k] * (zs[k] - mx - lden);
        loss[s] = l;
    }
    if (o < 6) {
        const float d = (tsum * e[o] / den - tg[o]) * inv_batch;
        dz[o] = d;
        dzv[s * 6 + o] = d;
    }
    __syncthreads();
    float g = 0.0f;
    for (int k = 0; k < 6; k++) g += w5[k * 64 + o] * dz[k];
    const float dho = h[o] > 0.0f ? g : 0.0f;
    dh[o] = dho;
    hv[s * 64 + o] = h[o];
    dhv[s * 64 + o] = dho;
    __syncthreads();
    if (o < 32) {
        float d = 0.0f;
        for (int k = 0; k < 64; k++) d += w4[k * 33 + o] * dh[k];
        dpool[s * 32 + o] = d;
    }
}

// Weight gradients of the MLP summed over the batch (layout W4[64][33], b4[64],
// W5[6][64], b5[6]): dW4[o][c] = sum_s dh[s][o] in[s][c] with
// in = [pooled; qp], db4[o] = sum_s dh[s][o], dW5[k][o] = sum_s dz[s][k] h[s][o],
// db5[k] = sum_s dz[s][k]. A workgroup takes 32 parameters; its 8 lane groups
// sum samples g, g + 8, ... and the group sums are added in order.
__global__ __launch_bounds__(256) void mlp_wgrad_kernel(const float* __restrict__ pooled, const float* __restrict__ qp,
                                                        const float* __restrict__ hv, const float* __restrict__ dhv,
                                                        const float* __restrict__ dzv, int n, float* __restrict__ g) {
    __shared__ float red[8][32];
    const int tid = threadIdx.x, el = tid & 31, grp = tid >> 5;
    const int p = blockIdx.x * 32 + el;
    float s = 0.0f;
    if (p < 64 * 33) {
        const int o = p / 33, c = p % 33;
#pragma unroll 4
        for (int k = grp; k < n; k += 8) s += dhv[k * 64 + o] * (c < 32 ? pooled[k * 32 + c] : qp[k]);
    } else if (p < 64 * 33 + 64) {
        const int o = p - 64 * 33;
#pragma unroll 4
        for (int k = grp; k < n; k += 8) s += dhv[k * 64 + o];
    } else if (p < 64 * 33 + 64 + 6 * 64) {
        const int q = p - (64 * 33 + 64), kk = q / 64, o = q % 64;
#pragma unroll 4
        for (int k = grp; k < n; k += 8) s += dzv[k * 6 + kk] * hv[k * 64 + o];
    } else if (p < kMlpParams) {
        const int kk = p - (64 * 33 + 64 + 6 * 64);
#pragma unroll 4
        for (int k = grp; k < n; k += 8) s += dzv[k * 6 + kk];
    }
    red[grp][el] = s;
    __syncthreads();
    if (grp == 0 && p < kMlpParams) {
        float t = 0.0f;
        for (int q = 0; q < 8; q++) t += red[q][el];
        g[p] = t;
    }
}

// Gradient into the last convolution's output: dz3 = (a3 > 0) * dpool / 1024.
__global__ void pool_back_kernel(const uint16_t* __restrict__ a3, const float* __restrict__ dpool, int total,
                                 uint16_t* __restrict__ dz3) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int n = e / (1024 * 32), c = e & 31;
    dz3[e] = bf2f(a3[e]) > 0.0f ? f2bf(dpool[n * 32 + c] * (1.0f / 1024.0f)) : (uint16_t)0;
}

// bf16 copies of the conv weights for the forward pass (padded k) and of the
// flipped, channel-swapped 32x32 weights for the input-gradient pass.
__global__ void pack_weights_kernel(const float* __restrict__ params, uint16_t* __restrict__ w1, uint16_t* __restrict__ w2,
                                    uint16_t* __restrict__ w3, uint16_t* __restrict__ w2t, uint16_t* __restrict__ w3t) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < 32 * 16) {
        const int co = e / 16, k = e % 16;
        w1[e] = k < 9 ? f2bf(params[kOffW1 + co * 9 + k]) : (uint16_t)0;
    }
    if (e < 9216) {
        w2[e] = f2bf(params[kOffW2 + e]);
        w3[e] = f2bf(params[kOffW3 + e]);
        // e indexes the transposed weight [a][ky][kx][b] = W[b][2-ky][2-kx][a]
        const int a = e / 288, tap = (e / 32) % 9, b = e % 32;
        const int src = (b * 9 + (8 - tap)) * 32 + a;
        w2t[e] = f2bf(params[kOffW2 + src]);
        w3t[e] = f2bf(params[kOffW3 + src]);
    }
}

__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, int len, float lr, float b1, float b2, float c1, float c2,
                            float eps) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.0f - b1) * gi;
    const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

namespace {

// Input, the three convolutions and the pool: b.pooled[n][32].
void features(const Buffers& b, const uint8_t* blocks, int n, hipStream_t s) {
    const size_t px = (size_t)n * 1024;
    hipLaunchKernelGGL(input_kernel, dim3(blocks_for(px, 256)), dim3(256), 0, s, blocks, (int)px, b.x);
    hipLaunchKernelGGL((conv_fwd_kernel<1, kRowsPerWave>), dim3(n, 32 / (4 * kRowsPerWave)), dim3(256), 0, s, b.x, b.w1, b.params + kOffB1, nullptr, b.a1, 1);
    hipLaunchKernelGGL((conv_fwd_kernel<32, kRowsPerWave>), dim3(n, 32 / (4 * kRowsPerWave)), dim3(256), 0, s, b.a1, b.w2, b.params + kOffB2, nullptr, b.a2, 1);
    hipLaunchKernelGGL((conv_fwd_kernel<32, kRowsPerWave>), dim3(n, 32 / (4 * kRowsPerWave)), dim3(256), 0, s, b.a2, b.w3, b.params + kOffB3, nullptr, b.a3, 1);
    hipLaunchKernelGGL(pool_kernel, dim3(n), dim3(256), 0, s, b.a3, b.pooled);
}

void reduce(const float* part, int n, int len, float* out, hipStream_t s) {
    hipLaunchKernelGGL(reduce_kernel, dim3((len + 31) / 32), dim3(256), 0, s, part, n, len, out);
}

}  // namespace

hipError_t forward(const Buffers& b, const uint8_t* blocks, const float* qp, int n, float* probs, hipStream_t s) {
    features(b, blocks, n, s);
    hipLaunchKernelGGL(mlp_kernel, dim3(n), dim3(64), 0, s, b.pooled, qp, b.params + kOffW4,
                       b.params + kOffB4, b.params + kOffW5, b.params + kOffB5, n, probs, nullptr, nullptr, 0.0f,
                       nullptr, nullptr, nullptr, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t gradients(const Buffers& b, const uint8_t* blocks, const float* qp, const uint8_t* labels,
                     const float* targets, int n, float* grads, float* loss, hipStream_t s) {
    features(b, blocks, n, s);
    const size_t act = (size_t)n * 1024 * 32;
    // the MLP's per-sample vectors live in b.part until the convolution gradients reuse it
    float* hv = b.part;
    float* dhv = hv + (size_t)n * 64;
    float* dzv = dhv + (size_t)n * 64;
    hipLaunchKernelGGL(mlp_kernel, dim3(n), dim3(64), 0, s, b.pooled, qp, b.params + kOffW4,
                       b.params + kOffB4, b.params + kOffW5, b.params + kOffB5, n, b.probs, labels, targets,
                       1.0f / (float)n, loss, b.dpool, hv, dhv, dzv);
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((kMlpParams + 31) / 32), dim3(256), 0, s, b.pooled, qp, hv, dhv, dzv, n,
                       grads + kOffW4);
    // conv3
    hipLaunchKernelGGL(pool_back_kernel, dim3(blocks_for(act, 256)), dim3(256), 0, s, b.a3, b.dpool, (int)act, b.dz);
    hipLaunchKernelGGL(conv_wgrad32_kernel, dim3(n, 9), dim3(256), 0, s, b.dz, b.a2, b.part);
    reduce(b.part, n, 9216, grads + kOffW3, s);
    hipLaunchKernelGGL(bias_grad_kernel, dim3(n), dim3(256), 0, s, b.dz, b.part);
    reduce(b.part, n, 32, grads + kOffB3, s);
    // into conv2's output: dz2 = (a2 > 0) * conv(dz3, W3 flipped)
    hipLaunchKernelGGL((conv_fwd_kernel<32, kRowsPerWave>), dim3(n, 32 / (4 * kRowsPerWave)), dim3(256), 0, s, b.dz, b.w3t, nullptr, b.a2, b.dz2, 2);
    hipLaunchKernelGGL(conv_wgrad32_kernel, dim3(n, 9), dim3(256), 0, s, b.dz2, b.a1, b.part);
    reduce(b.part, n, 9216, grads + kOffW2, s);
    hipLaunchKernelGGL(bias_grad_kernel, dim3(n), dim3(256), 0, s, b.dz2, b.part);
    reduce(b.part, n, 32, grads + kOffB2, s);
    // into conv1's output: dz1 = (a1 > 0) * conv(dz2, W2 flipped)
    hipLaunchKernelGGL((conv_fwd_kernel<32, kRowsPerWave>), dim3(n, 32 / (4 * kRowsPerWave)), dim3(256), 0, s, b.dz2, b.w2t, nullptr, b.a1, b.dz, 2);
    hipLaunchKernelGGL(conv_wgrad1_kernel, dim3(n, 4), dim3(288), 0, s, b.dz, b.x, b.part);
    reduce(b.part, 4 * n, 288, grads + kOffW1, s);
    hipLaunchKernelGGL(bias_grad_kernel, dim3(n), dim3(256), 0, s, b.dz, b.part);
    reduce(b.part, n, 32, grads + kOffB1, s);
    return hipGetLastError();
}

hipError_t pack_weights(const Buffers& b, hipStream_t s) {
    hipLaunchKernelGGL(pack_weights_kernel, dim3(blocks_for(9216, 256)), dim3(256), 0, s, b.params, b.w1, b.w2, b.w3,
                       b.w2t, b.w3t);
    return hipGetLastError();
}

hipError_t adam(const Buffers& b, int t, float lr, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, s, b.params, b.adam_m, b.adam_v,
                       b.grads, kParams, lr, b1, b2, c1, c2, 1e-8f);
    return pack_weights(b, s);
}

}  // namespace cnn
}  // namespace rmh

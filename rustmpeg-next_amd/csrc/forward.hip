// forward.hip - the encoder's forward path for gfx950: forward transform,
// forward LFNST and quantisation (flat, dependent, transform skip) of every
// TU of a batch, and the per-TU reconstruction SSE of a candidate batch.
//
// Semantics: forward_transform_rect_bd (crates/rustmpeg-video/src/vvc/dct.rs:
// 865-991), forward_lfnst_rect (vvc/lfnst.rs:131-204), quantize_rect_bd
// (dct.rs:1113-1147), quantize_dep_quant_bd (dct.rs:1252-1301) and
// quantize_ts_with_mode (vvc/transform.rs:415-427), in the order of the
// encoder's plan_primary (vvc/slice.rs:718-761).
//
// Mapping: one 64-lane workgroup per TU. The residual block is staged in LDS
// (int16); the first stage writes the column transforms of the retained
// coefficient columns to LDS (int32, exact: |sum| < 64 * 90 * 2^15), the
// second stage accumulates in 64-bit as the reference does; the coefficients
// that survive the zero-out (at most 32 x 32) stay in LDS for the LFNST and
// the quantiser. Dependent quantisation runs the reverse scan's 4-state
// automaton as a wave prefix of map compositions (each position's step is a
// map of the 4 states: the level, and so its parity, depends only on
// state >> 1), so all lanes quantise their chunks of the scan at once.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gmem.h"
#include "kernels.h"
#include "xcd.h"

namespace rmh {
namespace {

__device__ __forceinline__ int clip16(int64_t v) { return v < -32768 ? -32768 : v > 32767 ? 32767 : (int)v; }

// Entry (k, j) of the transform matrix of (type, n): the table holds
// (M[2kk][j], M[2kk + 1][j]) pairs at dword kk * n + j.
__device__ __forceinline__ int mat_at(const uint32_t* m, int n, int k, int j) {
    const uint32_t d = m[(k >> 1) * n + j];
    return (int)(int16_t)(k & 1 ? d >> 16 : d & 0xffffu);
}

constexpr int kQuantScales[6] = {26214, 23302, 20560, 18396, 16384, 14564};
constexpr int kQuantScalesSqrt2[6] = {18396, 16384, 14564, 13107, 11651, 10280};

// Maps of the 4 dependent-quantisation states packed as 4 x 2 bits (entry x
// = image of state x): g o f.
__device__ __forceinline__ uint32_t compose(uint32_t g, uint32_t f) {
    uint32_t h = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) h |= ((g >> (2 * ((f >> (2 * x)) & 3))) & 3) << (2 * x);
    return h;
}

__global__ __launch_bounds__(64) void forward_kernel(const FwdArgs a) {
    __shared__ int16_t R[64 * 64];   // residual, raster
    __shared__ int32_t T[32 * 64];   // first stage: [k][y]
    __shared__ int32_t C[32 * 32];   // retained coefficients: [y][x], pitch cwp
    __shared__ int32_t S[48];        // LFNST input / output
    const uint32_t tu = xcd_local_block(blockIdx.x, gridDim.x);
    if (tu >= a.n_tu) return;
    const int lane = threadIdx.x;
    const rmh_tu_desc d = a.tus[tu];
    const int lw = d.log2_w, lh = d.log2_h, w = 1 << lw, h = 1 << lh, n = w * h;
    const int bd = d.bit_depth, qp = d.qp < 4 ? 4 : d.qp;
    const bool intra = !(d.flags & RMH_TU_QUANT_INTER);
    const RMH_GLOBAL int16_t* res = gsrc<int16_t>(a.residual) + d.offset;
    RMH_GLOBAL int16_t* out = gdst<int16_t>(a.levels) + d.offset;

    if (d.flags & RMH_TU_TRANSFORM_SKIP) {
        // quantize_ts_with_mode per sample
        const int per = qp / 6, rem = qp % 6, qbits = 14 + per;
        const int64_t add = (int64_t)(intra ? 171 : 85) << (qbits - 9);
        for (int i = lane; i < n; i += 64) {
            const int r = res[i];
            const int64_t lv = ((int64_t)(r < 0 ? -r : r) * kQuantScales[rem] + add) >> qbits;
            out[i] = (int16_t)clip16(r < 0 ? -lv : lv);
        }
        return;
    }

    // zero-out: the default 32-point MTS and 64-point limits, the TU's own,
    // and the LFNST region (as the decoder's inverse_residual widens it)
    int sw = (d.tr_hor != RMH_TR_DCT2 && w == 32) ? 16 : (w > 32 ? w - 32 : 0);
    int sh = (d.tr_ver != RMH_TR_DCT2 && h == 32) ? 16 : (h > 32 ? h - 32 : 0);
    sw = max(sw, (int)d.skip_w);
    sh = max(sh, (int)d.skip_h);
    const bool lfnst = d.lfnst_idx != 0;
    const int sb = (w >= 8 && h >= 8) ? 8 : 4;
    if (lfnst) {
        const int nz = (w == 4 || h == 4) ? 4 : 8;
        sw = max(sw, w - nz);
        sh = max(sh, h - nz);
    }
    const int cw = w > 1 ? w - sw : 1, ch = h > 1 ? h - sh : 1;
    // the coefficient array also covers the LFNST output region
    const int cwp = lfnst ? max(cw, sb) : cw, chp = lfnst ? max(ch, sb) : ch;

    for (int i = lane; i < n; i += 64) R[i] = res[i];
    for (int i = lane; i < cwp * chp; i += 64) C[i] = 0;
    __syncthreads();

    if (w == 1 || h == 1) {
        // one-dimensional transform: one stage with shift log2(n) + bd - 9
        const int len = w == 1 ? h : w, llen = w == 1 ? lh : lw;
        const int type = w == 1 ? d.tr_ver : d.tr_hor;
        const uint32_t* m = a.mat + a.mat_off[type][llen];
        const int shift = llen + bd + 6 - 15;
        const int64_t add = shift > 0 ? (int64_t)1 << (shift - 1) : 0;
        const int cut = w == 1 ? ch : cw;
        for (int k = lane; k < cut; k += 64) {
            int64_t s = 0;
            for (int j = 0; j < len; j++) s += (int64_t)mat_at(m, len, k, j) * R[j];
            C[k] = clip16(shift >= 0 ? (s + add) >> shift : s << (-shift));
        }
    } else {
        const uint32_t* mh = a.mat + a.mat_off[d.tr_hor][lw];
        const uint32_t* mv = a.mat + a.mat_off[d.tr_ver][lh];
        const int shift1 = lw + bd + 6 - 15, shift2 = lh + 6;
        const int add1 = shift1 > 0 ? 1 << (shift1 - 1) : 0;
        const int64_t add2 = (int64_t)1 << (shift2 - 1);
        // first stage: T[k][y] = (sum_n M_h[k][n] r[y][n] + add1) >> shift1, k < cw
        for (int e = lane; e < cw * h; e += 64) {
            const int k = e >> lh, y = e & (h - 1);
            const int16_t* row = R + y * w;
            int s = 0;
            for (int j = 0; j < w; j++) s += mat_at(mh, w, k, j) * row[j];
            T[k * h + y] = (s + add1) >> shift1;
        }
        __syncthreads();
        // second stage: C[v][k] = clip((sum_y M_v[v][y] T[k][y] + add2) >> shift2)
        for (int e = lane; e < ch * cw; e += 64) {
            const int v = e / cw, k = e - v * cw;
            const int32_t* col = T + k * h;
            int64_t s = 0;
            for (int y = 0; y < h; y++) s += (int64_t)mat_at(mv, h, v, y) * col[y];
            C[v * cwp + k] = clip16((s + add2) >> shift2);
        }
    }
    __syncthreads();

    if (lfnst) {
        // forward_lfnst_rect: gather the sb x sb region (4x4, or 8x8 without
        // its bottom-right 4x4) in the reference's order, transposed for the
        // modes that transpose, multiply by the kernel, write the outputs
        // along the 8x8 diagonal scan (positions past zero_out get 0)
        const int tr_size = sb == 4 ? 16 : 48;
        const int zero_out = ((w == 4 && h == 4) || (w == 8 && h == 8)) ? 8 : 16;
        const int mode = d.lfnst_mode;
        const int set = mode <= 1 ? 0 : mode <= 12 ? 1 : mode <= 23 ? 2 : mode <= 44 ? 3 : mode <= 55 ? 2 : 1;
        const bool tp = (mode >= 35 && mode < 67) || mode >= 81;
        if (lane < tr_size) {
            const int i = lane;
            int x, y;
            if (sb == 4) {
                x = tp ? i >> 2 : i & 3;
                y = tp ? i & 3 : i >> 2;
            } else if (i < 32) {
                x = tp ? i >> 3 : i & 7;
                y = tp ? i & 7 : i >> 3;
            } else if (tp) {
                x = 4 + ((i - 32) >> 2);
                y = (i - 32) & 3;
            } else {
                x = (i - 32) & 3;
                y = 4 + ((i - 32) >> 2);
            }
            S[i] = C[y * cwp + x];
        }
        __syncthreads();
        int o = 0;
        if (lane < zero_out) {
            const int8_t* mt = sb == 8 ? a.lfnst8 + ((set * 2 + (d.lfnst_idx - 1)) * 16 + lane) * 48
                                       : a.lfnst4 + ((set * 2 + (d.lfnst_idx - 1)) * 16 + lane) * 16;
            int64_t s = 0;
            for (int i = 0; i < tr_size; i++) s += (int64_t)S[i] * mt[i];
            o = (int)((s + 64) >> 7);
        }
        __syncthreads();
        if (lane < tr_size) {
            const int r = a.scans[a.scan_off[3][3] + lane];  // 8x8 grouped diagonal scan, raster in an 8-wide block
            C[(r >> 3) * cwp + (r & 7)] = o;
        }
        __syncthreads();
    }

    const bool sqrt2 = ((lw + lh) & 1) != 0;
    const int tr_shift = 15 - bd - ((lw + lh) >> 1) - (sqrt2 ? 1 : 0);
    auto coef = [&](int x, int y) { return x < cwp && y < chp ? C[y * cwp + x] : 0; };
    if (!(d.flags & RMH_TU_DEP_QUANT)) {
        const int per = qp / 6, rem = qp % 6;
        const int qbits = 14 + per + tr_shift;
        const int64_t add = (int64_t)(intra ? 171 : 85) << max(qbits - 9, 0);
        const int64_t scale = sqrt2 ? kQuantScalesSqrt2[rem] : kQuantScales[rem];
        const int rs = max(qbits, 0);
        for (int i = lane; i < n; i += 64) {
            const int c = coef(i & (w - 1), i >> lw);
            const int64_t lv = ((int64_t)(c < 0 ? -c : c) * scale + add) >> rs;
            out[i] = (int16_t)clip16(c < 0 ? -lv : lv);
        }
        return;
    }

    // dependent quantisation over the reverse scan
    const int qdq = qp + 1, per = qdq / 6, rem = qdq % 6;
    const int qshift = 13 + per + tr_shift;
    const int64_t qscale = sqrt2 ? kQuantScalesSqrt2[rem] : kQuantScales[rem];
    const int64_t half = (int64_t)1 << max(qshift - 1, 0);
    const int rs = max(qshift, 0);
    const uint16_t* scan = a.scans + a.scan_off[lw][lh];
    const int len = a.scan_len[lw][lh];
    // the scan covers the top-left min(w, 32) x min(h, 32) region only: the
    // rest of a 64-sample side is zero (the reference clears the whole block)
    if (w > 32 || h > 32)
        for (int i = lane; i < n; i += 64)
            if ((i & (w - 1)) >= 32 || (i >> lw) >= 32) out[i] = 0;
    // the signed, clamped levels for the two reconstruction offsets (state >> 1 = 0, 1)
    auto levels_at = [&](int s, int* l0, int* l1) {
        const int r = scan[s];
        const int c = coef(r & (w - 1), r >> lw);
        const int64_t q = ((int64_t)(c < 0 ? -c : c) * qscale + half) >> rs;
        const int64_t m0 = q <= 0 ? 0 : (q + 1) >> 1, m1 = q <= 1 ? 0 : (q + 2) >> 1;
        *l0 = c < 0 ? clip16(-m0) : clip16(m0);
        *l1 = c < 0 ? clip16(-m1) : clip16(m1);
    };
    // step maps T0 = (0, 2, 1, 3), T1 = (2, 0, 3, 1) of 32040 >> (4 x + 2 parity); states
    // 0 and 1 step with the parity of l0, states 2 and 3 with that of l1
    constexpr uint32_t kT0 = 0u | 2u << 2 | 1u << 4 | 3u << 6, kT1 = 2u | 0u << 2 | 3u << 4 | 1u << 6;
    const int chunk = (len + 63) >> 6;
    const int t0 = lane * chunk, t1 = min(t0 + chunk, len);  // positions s = len - 1 - t
    uint32_t f = 0xE4u;  // identity
    for (int t = t0; t < t1; t++) {
        int l0, l1;
        levels_at(len - 1 - t, &l0, &l1);
        const uint32_t a0 = (l0 & 1) ? kT1 : kT0, a1 = (l1 & 1) ? kT1 : kT0;
        const uint32_t step = (a0 & 0x0Fu) | (a1 & 0xF0u);
        f = compose(step, f);
    }
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)f, dd);
        if (lane >= dd) f = compose(f, o);
    }
    const uint32_t prev = (uint32_t)__shfl_up((int)f, 1);
    int state = lane ? (int)(prev & 3) : 0;
    for (int t = t0; t < t1; t++) {
        const int s = len - 1 - t;
        int l0, l1;
        levels_at(s, &l0, &l1);
        const int level = (state >> 1) ? l1 : l0;
        out[scan[s]] = (int16_t)level;
        state = (32040 >> ((state << 2) + ((level & 1) << 1))) & 3;
    }
}

// Per-TU SSE of a candidate: sum over the block of
// (pred + residual - clip(pred + recon, 0, 2^bd - 1))^2, or of
// (residual - recon)^2 without a prediction.
__global__ __launch_bounds__(64) void sse_kernel(const SseArgs a) {
    const uint32_t tu = blockIdx.x;
    if (tu >= a.n_tu) return;
    const rmh_tu_desc d = a.tus[tu];
    const int n = 1 << (d.log2_w + d.log2_h), maxv = (1 << d.bit_depth) - 1;
    const RMH_GLOBAL int16_t* r = gsrc<int16_t>(a.residual) + d.offset;
    const RMH_GLOBAL int16_t* rr = gsrc<int16_t>(a.recon) + d.offset;
    uint64_t s = 0;
    if (a.pred) {
        const RMH_GLOBAL int16_t* p = gsrc<int16_t>(a.pred) + d.offset;
        for (int i = threadIdx.x; i < n; i += 64) {
            const int pv = p[i];
            int v = pv + rr[i];
            v = v < 0 ? 0 : v > maxv ? maxv : v;
            const int64_t e = (int64_t)(pv + r[i]) - v;
            s += (uint64_t)(e * e);
        }
    } else {
        for (int i = threadIdx.x; i < n; i += 64) {
            const int64_t e = (int64_t)r[i] - rr[i];
            s += (uint64_t)(e * e);
        }
    }
#pragma unroll
    for (int dd = 32; dd > 0; dd >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)s, dd), hi = (uint32_t)__shfl_xor((int)(s >> 32), dd);
        s += ((uint64_t)hi << 32) | lo;
    }
    if (threadIdx.x == 0) gdst<uint64_t>(a.sse)[tu] = s;
}

}  // namespace

hipError_t launch_forward(const FwdArgs& a, hipStream_t stream) {
    if (a.n_tu == 0) return hipSuccess;
    hipLaunchKernelGGL(forward_kernel, dim3(a.n_tu), dim3(64), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_sse(const SseArgs& a, hipStream_t stream) {
    if (a.n_tu == 0) return hipSuccess;
    hipLaunchKernelGGL(sse_kernel, dim3(a.n_tu), dim3(64), 0, stream, a);
    return hipGetLastError();
}

}  // namespace rmh

// mc_refine.hip - the decoder-side refinements of 8-bit bi-prediction for
// gfx950: DMVR (dmvr_kernel) and the bi-prediction of the PUs it refines or
// that use BDOF (mc_bi_refined_kernel); and the geometric-partitioning blend
// of two predictions (gpm_kernel, predict_gpm_block, inter.rs:1673-1730).
//
// Semantics: the bi branch of predict_inter_pu
// (crates/rustmpeg-video/src/vvc/inter.rs:1753-1801) with
// refine_dmvr_subblock / dmvr_subpel_1d (inter.rs:1341-1459) over the whole
// PU, interpolate_luma_block / interpolate_chroma_block_420 to 14-bit
// intermediates (inter.rs:886-1182), apply_bdof_luma (inter.rs:1204-1339) and
// weighted_average_bi (inter.rs:1189-1201).
//
// DMVR: one 256-lane workgroup per PU. The bilinear predictions of both lists
// over the PU plus the two-sample search margin are built once in LDS (8-bit
// values: the 2-tap filters of 8-bit samples stay within 0..255); every lane
// then accumulates, for the 25 integer offsets at once, the bilateral SAD of
// its even-row positions, and the sums are reduced across the workgroup. One
// lane takes the decision and writes the refined pair into the plan's MV
// array, which the prediction kernel reads.
//
// Prediction: one 256-lane workgroup per 16x16 luma tile of a PU (and its two
// 8x8 chroma tiles). The luma intermediates of both lists are computed into
// LDS over the tile plus a two-sample margin clamped to the PU (BDOF's
// replicated border: a position outside the PU reads the nearest PU sample,
// and so does its gradient); each 4x4 unit's optical-flow offset comes from
// its 6x6 window with sixteen lanes per unit, reduced with lane exchanges.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gmem.h"
#include "kernels.h"

namespace rmh {
namespace {

// LUMA_FILTER_8TAP, LUMA_ALT_HPEL_FILTER_8TAP and CHROMA_FILTER_4TAP
// (vvc/inter.rs:57-80, vvc/predict.rs:86-119; H.266 Tables 33, 34).
__constant__ int8_t kRLuma8[16][8] = {
    {0, 0, 0, 64, 0, 0, 0, 0},        {0, 1, -3, 63, 4, -2, 1, 0},      {-1, 2, -5, 62, 8, -3, 1, 0},
    {-1, 3, -8, 60, 13, -4, 1, 0},    {-1, 4, -10, 58, 17, -5, 1, 0},   {-1, 4, -11, 52, 26, -8, 3, -1},
    {-1, 3, -9, 47, 31, -10, 4, -1},  {-1, 4, -11, 45, 34, -10, 4, -1}, {-1, 4, -11, 40, 40, -11, 4, -1},
    {-1, 4, -10, 34, 45, -11, 4, -1}, {-1, 4, -10, 31, 47, -9, 3, -1},  {-1, 3, -8, 26, 52, -11, 4, -1},
    {0, 1, -5, 17, 58, -10, 4, -1},   {0, 1, -4, 13, 60, -8, 3, -1},    {0, 1, -3, 8, 62, -5, 2, -1},
    {0, 1, -2, 4, 63, -3, 1, 0}};
__constant__ int8_t kRLumaAltHpel[8] = {0, 3, 9, 20, 20, 9, 3, 0};
__constant__ int8_t kRChroma4[32][4] = {
    {0, 64, 0, 0},    {-1, 63, 2, 0},   {-2, 62, 4, 0},   {-2, 60, 7, -1},  {-2, 58, 10, -2}, {-3, 57, 12, -2},
    {-4, 56, 14, -2}, {-4, 55, 15, -2}, {-4, 54, 16, -2}, {-5, 53, 18, -2}, {-6, 52, 20, -2}, {-6, 49, 24, -3},
    {-6, 46, 28, -4}, {-5, 44, 29, -4}, {-4, 42, 30, -4}, {-4, 39, 33, -4}, {-4, 36, 36, -4}, {-4, 33, 39, -4},
    {-4, 30, 42, -4}, {-4, 29, 44, -5}, {-4, 28, 46, -6}, {-3, 24, 49, -6}, {-2, 20, 52, -6}, {-2, 18, 53, -5},
    {-2, 16, 54, -4}, {-2, 15, 55, -4}, {-2, 14, 56, -4}, {-2, 12, 57, -3}, {-2, 10, 58, -2}, {-1, 7, 60, -2},
    {0, 4, 62, -2},   {0, 2, 63, -1}};

constexpr int kOffs = 8192;  // IF_INTERNAL_OFFS

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

// sample_clamped (inter.rs:886-890) of an 8-bit plane.
__device__ __forceinline__ int smp8(const rmh_plane& p, int x, int y) {
    x = clampi(x, 0, (int)p.width - 1);
    y = clampi(y, 0, (int)p.height - 1);
    return (int)gsrc<uint8_t>(p.data)[(uint32_t)y * p.stride + (uint32_t)x];
}

// ------------------------------------------------------------------ DMVR

// dmvr_subpel_1d (inter.rs:1341-1363).
__device__ int dmvr_subpel_1d(int64_t s0, int64_t s1, int64_t s2) {
    const int64_t den = s1 + s2 - 2 * s0;
    if (den == 0) return 0;
    if (s1 != s0 && s2 != s0) {
        const int64_t num = (s1 - s2) * 8;  // ((s1 - s2) << 4) >> 1
        int64_t rem = num < 0 ? -num : num;
        int q = 0;
        for (int i = 2; i >= 0; i--) {
            const int64_t div = den * (1 << i);
            if (rem >= div) {
                rem -= div;
                q += 1 << i;
            }
        }
        return num < 0 ? -q : q;
    }
    return s1 == s0 ? -8 : 8;
}

__global__ __launch_bounds__(256) void dmvr_kernel(const rmh_mc_bi_frame* __restrict__ frames,
                                                   const DmvrJob* __restrict__ jobs, int32_t* __restrict__ mvs) {
    extern __shared__ uint8_t bil[];  // list 0 then list 1: (w + 4) x (h + 4) each
    __shared__ int red[4][25];
    const DmvrJob J = jobs[blockIdx.x];
    const rmh_mc_bi_frame& F = frames[J.frame];
    const int tid = threadIdx.x;
    const int lw = J.lw, lh = J.lh, w = 1 << lw, h = 1 << lh, ew = w + 4, eh = h + 4;
    // bilinear_pred at offsets -2 .. 2: B_l(u, v) is the prediction sample of
    // position (u - 2, v - 2) of the block at offset 0
#pragma unroll
    for (int l = 0; l < 2; l++) {
        const rmh_plane& p = l ? F.ref1[0] : F.ref0[0];
        const int mvx = J.mv[2 * l], mvy = J.mv[2 * l + 1];
        const int ix = (int)J.x + (mvx >> 4) - 2, iy = (int)J.y + (mvy >> 4) - 2;
        const int fx = mvx & 15, fy = mvy & 15;
        uint8_t* B = bil + l * ew * eh;
        for (int i = tid; i < ew * eh; i += 256) {
            const int v = i / ew, u = i - v * ew;
            const int px = ix + u, py = iy + v;
            const int h0 = (smp8(p, px, py) * (16 - fx) + smp8(p, px + 1, py) * fx + 8) >> 4;
            const int h1 = (smp8(p, px, py + 1) * (16 - fx) + smp8(p, px + 1, py + 1) * fx + 8) >> 4;
            B[i] = (uint8_t)((h0 * (16 - fy) + h1 * fy + 8) >> 4);
        }
    }
    __syncthreads();
    // bilateral costs: L0 at +(ox, oy), L1 at -(ox, oy), every second row
    const uint8_t* B0 = bil;
    const uint8_t* B1 = bil + ew * eh;
    int acc[25];
#pragma unroll
    for (int k = 0; k < 25; k++) acc[k] = 0;
    const int n_pos = w << (lh - 1);
    for (int q = tid; q < n_pos; q += 256) {
        const int x = q & (w - 1), y = (q >> lw) << 1;
#pragma unroll
        for (int oy = -2; oy <= 2; oy++)
#pragma unroll
            for (int ox = -2; ox <= 2; ox++) {
                const int a = B0[(y + oy + 2) * ew + x + ox + 2];
                const int b = B1[(y - oy + 2) * ew + x - ox + 2];
                acc[(oy + 2) * 5 + ox + 2] += abs(a - b);
            }
    }
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < 25; k++) {
        int v = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    // the 25 costs stay in LDS: a private array indexed at run time would be
    // given scratch memory for every lane of the launch
    __shared__ int64_t tot[25];
    if (tid < 25) tot[tid] = (int64_t)red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
    __syncthreads();
    if (tid != 0) return;
    int64_t (*sads)[5] = reinterpret_cast<int64_t (*)[5]>(tot);
    int32_t* out = mvs + 4 * (size_t)J.slot;
    int64_t min_cost = sads[2][2];
    min_cost -= min_cost >> 2;
    sads[2][2] = min_cost;
    int d[2] = {0, 0};
    if (min_cost >= (int64_t)w * h) {
        int bix = 0, biy = 0;
        for (int oy = -2; oy <= 2; oy++)
            for (int ox = -2; ox <= 2; ox++) {
                if (ox == 0 && oy == 0) continue;
                if (sads[oy + 2][ox + 2] < min_cost) {
                    min_cost = sads[oy + 2][ox + 2];
                    bix = ox;
                    biy = oy;
                }
            }
        d[0] = bix * 16;
        d[1] = biy * 16;
        if (bix > -2 && bix < 2 && biy > -2 && biy < 2) {
            const int ux = bix + 2, uy = biy + 2;
            d[0] += dmvr_subpel_1d(sads[uy][ux], sads[uy][ux - 1], sads[uy][ux + 1]);
            d[1] += dmvr_subpel_1d(sads[uy][ux], sads[uy - 1][ux], sads[uy + 1][ux]);
        }
    }
    constexpr int kLo = -(1 << 17), kHi = (1 << 17) - 1;
    out[0] = clampi(J.mv[0] + d[0], kLo, kHi);
    out[1] = clampi(J.mv[1] + d[1], kLo, kHi);
    out[2] = clampi(J.mv[2] - d[0], kLo, kHi);
    out[3] = clampi(J.mv[3] - d[1], kLo, kHi);
}

// ------------------------------------------------------------ prediction

constexpr int kM = 20;          // intermediates window: 16 + 2 x 2
constexpr int kTR = kM + 7;     // horizontal-pass rows

// right_shift_msb (inter.rs:1204-1211).
__device__ __forceinline__ int right_shift_msb(int numer, int denom) {
    return denom <= 0 ? 0 : numer >> (31 - __clz(denom));
}

__device__ __forceinline__ int sgn_mul(int s, int v) { return s < 0 ? -v : s == 0 ? 0 : v; }

__global__ __launch_bounds__(256) void mc_bi_refined_kernel(const rmh_mc_bi_frame* __restrict__ frames,
                                                            const McRefItem* __restrict__ items,
                                                            const int32_t* __restrict__ mvs) {
    __shared__ int32_t I[2][kM * kM];   // luma intermediates, [v - va][u - ua]
    __shared__ int32_t T[2][kTR * kM];  // horizontal pass of each list
    __shared__ int32_t flow[16][2];     // per 4x4 unit: (vx, vy)
    const McRefItem it = items[blockIdx.x];
    const rmh_mc_bi_frame& F = frames[it.frame];
    const int tid = threadIdx.x;
    const int pw = 1 << it.lw, ph = 1 << it.lh;
    const int tw = pw < 16 ? pw : 16, th = ph < 16 ? ph : 16;
    const int tx0 = 16 * it.tx, ty0 = 16 * it.ty;  // tile origin in the PU
    const bool bdof = (it.flags & 2) != 0;
    const int m = bdof ? 2 : 0;
    const int ua = max(tx0 - m, 0), ub = min(tx0 + tw - 1 + m, pw - 1);
    const int va = max(ty0 - m, 0), vb = min(ty0 + th - 1 + m, ph - 1);
    const int nu = ub - ua + 1, nv = vb - va + 1;
    const int bd = F.ref0[0].bit_depth, head_room = 14 - bd;
    int mv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) mv[k] = mvs[4 * (size_t)it.slot + k];

    // luma intermediates of both lists (interpolate_luma_block, for_bi_pred),
    // list 0 on threads 0-127 and list 1 on threads 128-255 side by side
    {
        const int l = tid >> 7, lt = tid & 127;
        const rmh_plane& p = l ? F.ref1[0] : F.ref0[0];
        const int mvx = mv[2 * l], mvy = mv[2 * l + 1];
        const int fx = mvx & 15, fy = mvy & 15;
        const int ix = (int)it.px + ua + (mvx >> 4), iy = (int)it.py + va + (mvy >> 4);
        const int8_t* cx = ((it.flags & 1) && fx == 8) ? kRLumaAltHpel : kRLuma8[fx];
        const int8_t* cy = ((it.flags & 1) && fy == 8) ? kRLumaAltHpel : kRLuma8[fy];
        int32_t* out = I[l];
        const bool two_d = fx != 0 && fy != 0;
        if (!two_d && fx == 0 && fy == 0) {
            for (int i = lt; i < nu * nv; i += 128) {
                const int v = i / nu, u = i - v * nu;
                out[v * kM + u] = smp8(p, ix + u, iy + v) * (1 << head_room) - kOffs;
            }
        } else if (!two_d) {
            const bool hor = fy == 0;
            const int8_t* c = hor ? cx : cy;
            const int shift = 6 - head_room, offset = -kOffs * (1 << shift);
            for (int i = lt; i < nu * nv; i += 128) {
                const int v = i / nu, u = i - v * nu;
                int s = 0;
#pragma unroll
                for (int k = 0; k < 8; k++)
                    s += smp8(p, ix + u + (hor ? k - 3 : 0), iy + v + (hor ? 0 : k - 3)) * c[k];
                out[v * kM + u] = (s + offset) >> shift;
            }
        } else {
            const int shift1 = 6 - head_room, offset1 = -kOffs * (1 << shift1);
            for (int i = lt; i < nu * (nv + 7); i += 128) {
                const int r = i / nu, u = i - r * nu;
                int s = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) s += smp8(p, ix + u + k - 3, iy + r - 3) * cx[k];
                T[l][r * kM + u] = (s + offset1) >> shift1;
            }
        }
        // every thread reaches both barriers; only a list with both
        // fractions non-zero has a vertical pass between them
        __syncthreads();
        if (two_d)
            for (int i = lt; i < nu * nv; i += 128) {
                const int v = i / nu, u = i - v * nu;
                int s = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) s += T[l][(v + k) * kM + u] * cy[k];
                out[v * kM + u] = s >> 6;
            }
        __syncthreads();
    }

    // PU-clamped access to the intermediates and their gradients
    auto P = [&](int l, int u, int v) {
        return I[l][(clampi(v, 0, ph - 1) - va) * kM + clampi(u, 0, pw - 1) - ua];
    };
    auto GX = [&](int l, int u, int v) {
        const int cu = clampi(u, 0, pw - 1), cv = clampi(v, 0, ph - 1);
        return (P(l, cu + 1, cv) >> 6) - (P(l, cu - 1, cv) >> 6);
    };
    auto GY = [&](int l, int u, int v) {
        const int cu = clampi(u, 0, pw - 1), cv = clampi(v, 0, ph - 1);
        return (P(l, cu, cv + 1) >> 6) - (P(l, cu, cv - 1) >> 6);
    };

    const int w1 = it.w1, w0 = 8 - w1;
    const int maxv = (1 << bd) - 1;
    if (bdof) {
        // the 6x6 window of 4x4 unit `unit`: sixteen lanes, up to three positions each
        const int unit = tid >> 4, sub = tid & 15;
        const int nux = tw >> 2, nuy = th >> 2;
        int sgx = 0, sgy = 0, sdx = 0, sdy = 0, sxy = 0;
        if (unit < nux * nuy) {
            const int bu = tx0 + 4 * (unit % nux) - 1, bv = ty0 + 4 * (unit / nux) - 1;
            for (int k = sub; k < 36; k += 16) {
                const int u = bu + k % 6, v = bv + k / 6;
                const int tgx = (GX(0, u, v) + GX(1, u, v)) >> 1;
                const int tgy = (GY(0, u, v) + GY(1, u, v)) >> 1;
                const int tdi = (P(1, u, v) >> 4) - (P(0, u, v) >> 4);
                sgx += abs(tgx);
                sgy += abs(tgy);
                sdx += sgn_mul(tgx, tdi);
                sdy += sgn_mul(tgy, tdi);
                sxy += sgn_mul(tgy, tgx);
            }
        }
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) {
            sgx += __shfl_xor(sgx, off);
            sgy += __shfl_xor(sgy, off);
            sdx += __shfl_xor(sdx, off);
            sdy += __shfl_xor(sdy, off);
            sxy += __shfl_xor(sxy, off);
        }
        if (sub == 0 && unit < nux * nuy) {
            const int vx = sgx == 0 ? 0 : clampi(right_shift_msb(4 * sdx, sgx), -15, 15);
            const int t = (sxy * vx) >> 1;
            const int vy = sgy == 0 ? 0 : clampi(right_shift_msb(4 * sdy - t, sgy), -15, 15);
            flow[unit][0] = vx;
            flow[unit][1] = vy;
        }
        __syncthreads();
    }
    const rmh_plane& Y = F.dst[0];
    {
        const int x = tid & 15, y = tid >> 4;
        const int u = tx0 + x, v = ty0 + y;
        const int X = (int)it.px + u, Yy = (int)it.py + v;
        if (x < tw && y < th && X < (int)Y.width && Yy < (int)Y.height) {
            const int p0 = P(0, u, v), p1 = P(1, u, v);
            int val;
            if (bdof) {
                const int unit = (y >> 2) * (tw >> 2) + (x >> 2);
                const int b = flow[unit][0] * (GX(0, u, v) - GX(1, u, v)) + flow[unit][1] * (GY(0, u, v) - GY(1, u, v));
                val = clampi((p0 + p1 + b + 64 + 2 * kOffs) >> 7, 0, 255);
            } else {
                const int sh = head_room + 3;
                val = clampi((p0 * w0 + p1 * w1 + (1 << (sh - 1)) + (kOffs << 3)) >> sh, 0, maxv);
            }
            gdst<uint8_t>(Y.data)[(uint32_t)Yy * Y.stride + (uint32_t)X] = (uint8_t)val;
        }
    }

    // chroma: both lists' 4-tap intermediates per sample, weighted average
    {
        const int cw = tw >> 1, chh = th >> 1;
        const int c = 1 + (tid >> 7), r = tid & 127;
        const int x = r & 7, y = r >> 3;
        const rmh_plane& D = F.dst[c];
        const int cx0 = ((int)it.px >> 1) + (tx0 >> 1), cy0 = ((int)it.py >> 1) + (ty0 >> 1);
        const int X = cx0 + x, Yc = cy0 + y;
        if (y < 8 && x < cw && y < chh && X < (int)D.width && Yc < (int)D.height) {
            int pv[2];
#pragma unroll
            for (int l = 0; l < 2; l++) {
                const rmh_plane& p = l ? F.ref1[c] : F.ref0[c];
                const int mvx = mv[2 * l], mvy = mv[2 * l + 1];
                const int fx = mvx & 31, fy = mvy & 31;
                const int ix = X + (mvx >> 5), iy = Yc + (mvy >> 5);
                if (fx == 0 && fy == 0) {
                    pv[l] = smp8(p, ix, iy) * (1 << head_room) - kOffs;
                } else if (fx == 0 || fy == 0) {
                    const bool hor = fy == 0;
                    const int8_t* cf = kRChroma4[hor ? fx : fy];
                    const int shift = 6 - head_room, offset = -kOffs * (1 << shift);
                    int s = 0;
#pragma unroll
                    for (int k = 0; k < 4; k++) s += smp8(p, ix + (hor ? k - 1 : 0), iy + (hor ? 0 : k - 1)) * cf[k];
                    pv[l] = (s + offset) >> shift;
                } else {
                    const int8_t* cxf = kRChroma4[fx];
                    const int8_t* cyf = kRChroma4[fy];
                    const int shift1 = 6 - head_room, offset1 = -kOffs * (1 << shift1);
                    int s2 = 0;
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) {
                        int s = 0;
#pragma unroll
                        for (int k = 0; k < 4; k++) s += smp8(p, ix - 1 + k, iy - 1 + rr) * cxf[k];
                        s2 += ((s + offset1) >> shift1) * cyf[rr];
                    }
                    pv[l] = s2 >> 6;
                }
            }
            const int sh = head_room + 3;
            const int val = clampi((pv[0] * w0 + pv[1] * w1 + (1 << (sh - 1)) + (kOffs << 3)) >> sh, 0, maxv);
            gdst<uint8_t>(D.data)[(uint32_t)Yc * D.stride + (uint32_t)X] = (uint8_t)val;
        }
    }
}

// ------------------------------------------------------------------- GPM

// GEO_DIS (inter.rs:131-134).
__constant__ int8_t kGeoDis[32] = {8, 8, 8, 8, 4, 4, 2, 1, 0, -1, -2, -4, -4, -8, -8, -8,
                                   -8, -8, -8, -8, -4, -4, -2, -1, 0, 1, 2, 4, 4, 8, 8, 8};

// predict_gpm_block's weight of prediction 0 at position (x, y) of a block
// whose split line lies at offsets (ox, oy) on the luma grid.
__device__ __forceinline__ int gpm_w0(int x, int y, int scale, int ox, int oy, int dx, int dy) {
    const int idx = (((x << scale) + ox) * 2 + 1) * kGeoDis[dx] + (((y << scale) + oy) * 2 + 1) * kGeoDis[dy];
    return clampi((32 + idx + 4) >> 3, 0, 8);
}

// One 256-lane workgroup per item: lane (x, y) of the 16x16 luma tile, then
// lanes 0-127 / 128-255 on the two 8x8 chroma tiles.
__global__ __launch_bounds__(256) void gpm_kernel(const rmh_gpm_frame* __restrict__ frames,
                                                  const GpmItem* __restrict__ items, int luma_scale, int chroma_scale) {
    const GpmItem it = items[blockIdx.x];
    const rmh_gpm_frame& F = frames[it.frame];
    const int tid = threadIdx.x;
    {
        const int x = tid & 15, y = tid >> 4;
        if (x < it.tw && y < it.th) {
            const int X = it.x0 + x, Y = it.y0 + y;
            const int w0 = gpm_w0(X - it.px, Y - it.py, luma_scale, it.lox, it.loy, it.dx, it.dy);
            const uint32_t o0 = (uint32_t)Y * F.pred0[0].stride + X, o1 = (uint32_t)Y * F.pred1[0].stride + X;
            const int p0 = gsrc<uint8_t>(F.pred0[0].data)[o0], p1 = gsrc<uint8_t>(F.pred1[0].data)[o1];
            gdst<uint8_t>(F.dst[0].data)[(uint32_t)Y * F.dst[0].stride + X] =
                (uint8_t)clampi((w0 * p0 + (8 - w0) * p1 + 4) >> 3, 0, 255);
        }
    }
    {
        const int c = 1 + (tid >> 7), r = tid & 127, x = r & 7, y = r >> 3;
        if (y < 8 && x < it.cw && y < it.ch) {
            const int X = (it.x0 >> 1) + x, Y = (it.y0 >> 1) + y;
            const int w0 = gpm_w0(X - (it.px >> 1), Y - (it.py >> 1), chroma_scale, it.cox, it.coy, it.dx, it.dy);
            const uint32_t o0 = (uint32_t)Y * F.pred0[c].stride + X, o1 = (uint32_t)Y * F.pred1[c].stride + X;
            const int p0 = gsrc<uint8_t>(F.pred0[c].data)[o0], p1 = gsrc<uint8_t>(F.pred1[c].data)[o1];
            gdst<uint8_t>(F.dst[c].data)[(uint32_t)Y * F.dst[c].stride + X] =
                (uint8_t)clampi((w0 * p0 + (8 - w0) * p1 + 4) >> 3, 0, 255);
        }
    }
}

}  // namespace

hipError_t launch_gpm(const rmh_gpm_frame* frames, const GpmItem* items, uint32_t n_items, uint32_t luma_scale,
                      uint32_t chroma_scale, hipStream_t stream) {
    if (n_items == 0) return hipSuccess;
    hipLaunchKernelGGL(gpm_kernel, dim3(n_items), dim3(256), 0, stream, frames, items, (int)luma_scale,
                       (int)chroma_scale);
    return hipGetLastError();
}

hipError_t launch_dmvr(const rmh_mc_bi_frame* frames, const DmvrJob* jobs, uint32_t n_jobs, uint32_t lds_bytes,
                       int32_t* mvs, hipStream_t stream) {
    if (n_jobs == 0) return hipSuccess;
    hipLaunchKernelGGL(dmvr_kernel, dim3(n_jobs), dim3(256), lds_bytes, stream, frames, jobs, mvs);
    return hipGetLastError();
}

hipError_t launch_mc_bi_refined(const rmh_mc_bi_frame* frames, const McRefItem* items, uint32_t n_items,
                                const int32_t* mvs, hipStream_t stream) {
    if (n_items == 0) return hipSuccess;
    hipLaunchKernelGGL(mc_bi_refined_kernel, dim3(n_items), dim3(256), 0, stream, frames, items, mvs);
    return hipGetLastError();
}

}  // namespace rmh

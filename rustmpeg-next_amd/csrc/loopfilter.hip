// loopfilter.hip - luma mapping (LMCS) and the adaptive loop filters (ALF,
// CCALF) over every plane of a batch of pictures.
//
// LMCS: apply_inverse_plane_16 / apply_forward_plane_16
// (crates/rustmpeg-video/src/vvc/filter.rs:2707-2722) map each sample of a
// luma plane through a 2^bd-entry table. One workgroup maps kLmcsRows rows
// of one plane with the table staged in LDS; 16-byte vector accesses where
// the rows allow. HBM-bound: one read and one write per sample.
//
// ALF: apply_alf (vvc/decode_general.rs:5225-5333) filters each CTB from an
// unmodified copy of the picture. Here the copy is the source plane and dst
// a separate plane that receives every sample once. A luma workgroup covers
// a 64 x 16 tile: it stages the tile with a 4-sample halo in LDS (reads
// clamped to the visible region, as get_src is), classifies the tile's 64
// 4x4 blocks with four lanes per block (one row of the 4 x 4 gradient grid
// each, summed with lane exchanges), fetches each block's transposed
// coefficients and clip values from the frame's filter bank, and filters the
// 7x7 diamond from LDS. The virtual-boundary row substitution of
// alf_luma_block (5505-5534) amounts to limiting the vertical tap offset to
// the distance from the boundary: d = vb_pos - 1 - yv above it, yv - vb_pos
// below it, and the stronger rounding shift where that distance is 0.
// Chroma workgroups cover 64 x 16 chroma samples and read their 5x5 diamond
// and the CCALF luma taps through the caches.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gmem.h"
#include "kernels.h"

namespace rmh {
namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------- LMCS

template <typename T>
__global__ __launch_bounds__(256) void lmcs_kernel(const LmcsJob* __restrict__ jobs,
                                                   const uint32_t* __restrict__ block_job, uint32_t block0) {
    constexpr int K = 16 / (int)sizeof(T);  // samples per 16-byte access
    __shared__ uint16_t lut[4096];
    const uint32_t b = blockIdx.x + block0;
    const LmcsJob& J = jobs[block_job[b]];
    const uint32_t cap = J.cap;
    for (uint32_t i = threadIdx.x; i <= cap; i += 256) lut[i] = J.lut[i];
    __syncthreads();
    const int width = (int)J.width;
    const uint32_t y0 = (b - J.first_block) * kLmcsRows;
    const uint32_t y1 = min(y0 + kLmcsRows, J.height);
    for (uint32_t y = y0; y < y1; y++) {
        const RMH_GLOBAL T* s = gsrc<T>(J.src) + (size_t)y * J.src_stride;
        RMH_GLOBAL T* d = gdst<T>(J.dst) + (size_t)y * J.dst_stride;
        for (int x = (int)threadIdx.x * K; x < width; x += 256 * K) {
            if (x + K <= width && (((uintptr_t)(s + x) | (uintptr_t)(d + x)) & 15) == 0) {
                T v[K];
                *reinterpret_cast<u32x4*>(v) = *reinterpret_cast<const RMH_GLOBAL u32x4*>(s + x);
#pragma unroll
                for (int k = 0; k < K; k++) v[k] = (T)lut[min((uint32_t)v[k], cap)];
                *reinterpret_cast<RMH_GLOBAL u32x4*>(d + x) = *reinterpret_cast<const u32x4*>(v);
            } else {
                for (int k = 0; k < K && x + k < width; k++) d[x + k] = (T)lut[min((uint32_t)s[x + k], cap)];
            }
        }
    }
}

// ---------------------------------------------------------------- ALF luma

constexpr int kHalo = 4;
constexpr int kLdsW = kAlfTileW + 2 * kHalo;  // 72
constexpr int kLdsH = kAlfTileH + 2 * kHalo;  // 24

// The 4x4 block's class index (H.266 8.8.5.3; derive_alf_class_4x4_vb_samples,
// vvc/filter.rs:1580-1691) from its gradient sums.
__device__ __forceinline__ void alf_class(int sv, int sh, int sd0, int sd1, int act_scale, int bd, int* cls, int* tr) {
    const int act = clampi(((sv + sh) * act_scale) >> (bd + 4), 0, 15);
    int c = act == 0 ? 0 : act == 1 ? 1 : act <= 6 ? 2 : act <= 14 ? 3 : 4;
    const bool v_major = sv > sh, d0_major = sd0 > sd1;
    const int hv1 = v_major ? sv : sh, hv0 = v_major ? sh : sv, dir_hv = v_major ? 1 : 3;
    const int d1 = d0_major ? sd0 : sd1, d0 = d0_major ? sd1 : sd0, dir_d = d0_major ? 0 : 2;
    const bool diag = (int64_t)d1 * hv0 > (int64_t)hv1 * d0;
    const int hi = diag ? d1 : hv1, lo = diag ? d0 : hv0;
    const int main_dir = diag ? dir_d : dir_hv, sec_dir = diag ? dir_hv : dir_d;
    const int strength = hi * 2 > 9 * lo ? 2 : (hi > 2 * lo ? 1 : 0);
    if (strength) c += (((main_dir & 1) << 1) + strength) * 5;
    *cls = c;
    // transposition by (main direction, secondary direction / 2): the table
    // {0, 1, 0, 2, 2, 3, 1, 3} of the standard
    const int k = main_dir * 2 + (sec_dir >> 1);
    *tr = (0x31322010u >> (4 * k)) & 3;
}

template <typename T>
__global__ __launch_bounds__(256) void alf_luma_kernel(const AlfLumaJob* __restrict__ jobs,
                                                       const uint32_t* __restrict__ block_job, uint32_t block0) {
    __shared__ int16_t tile[kLdsH][kLdsW];
    __shared__ int16_t prm[64][24];  // per 4x4 block: 12 coefficients, 12 clip values
    __shared__ uint8_t on[64];
    const uint32_t b = blockIdx.x + block0;
    const AlfLumaJob& J = jobs[block_job[b]];
    const uint32_t rel = b - J.first_block;
    const int tx0 = (int)(rel % J.tiles_x) * kAlfTileW, ty0 = (int)(rel / J.tiles_x) * kAlfTileH;
    const int bw = (int)J.bw, bh = (int)J.bh;
    const RMH_GLOBAL T* src = gsrc<T>(J.src);
    const size_t sstride = J.src_stride;
    const int tid = (int)threadIdx.x;

    for (int i = tid; i < kLdsH * kLdsW; i += 256) {
        const int r = i / kLdsW, c = i - r * kLdsW;
        const int y = clampi(ty0 - kHalo + r, 0, bh - 1), x = clampi(tx0 - kHalo + c, 0, bw - 1);
        tile[r][c] = (int16_t)src[(size_t)y * sstride + x];
    }
    __syncthreads();
#define L(xx, yy) ((int)tile[(yy) - ty0 + kHalo][(xx) - tx0 + kHalo])

    const int ctb_log2 = (int)J.ctb_log2, ctb_mask = (1 << ctb_log2) - 1, vb_pos = (1 << ctb_log2) - 4;
    const int bd = (int)J.bd;
    {
        // classification: lanes 4k .. 4k+3 of block k, lane q taking gradient row dy = 2q - 2
        const int bk = tid >> 2, q = tid & 3;
        const int bx = tx0 + (bk & 15) * 4, by = ty0 + (bk >> 4) * 4;
        int mode = 0;
        if (bx < (int)J.vis_w && by < (int)J.vis_h) {
            mode = J.ctbs[(size_t)(by >> ctb_log2) * J.ctb_cols + (bx >> ctb_log2)].luma;
            if (mode > RMH_ALF_LUMA_APS || (mode == RMH_ALF_LUMA_APS && !J.has_aps)) mode = 0;
        }
        const int byv = by & ctb_mask;
        const int dy = 2 * q - 2;
        int sv = 0, sh = 0, sd0 = 0, sd1 = 0;
        // rows next to the virtual boundary skip the gradient row across it
        if (mode && !(byv == vb_pos - 4 && dy == 4) && !(byv == vb_pos && dy == -2)) {
            const int y = by + dy;
            int down = y - 1, up2 = y + 2;
            const int up = y + 1;
            if (y > 0) {
                const int yy = y & ctb_mask;
                if (yy == vb_pos - 2)
                    up2 = y + 1;
                else if (yy == vb_pos)
                    down = y;
            }
#pragma unroll
            for (int dx = -2; dx < 6; dx += 2) {
                const int x = bx + dx;
                const int c0 = L(x, y) << 1, c1 = L(x + 1, y + 1) << 1;
                sv += abs(c0 - L(x, down) - L(x, up)) + abs(c1 - L(x + 1, y) - L(x + 1, up2));
                sh += abs(c0 - L(x + 1, y) - L(x - 1, y)) + abs(c1 - L(x + 2, y + 1) - L(x, y + 1));
                sd0 += abs(c0 - L(x - 1, down) - L(x + 1, up)) + abs(c1 - L(x, y) - L(x + 2, up2));
                sd1 += abs(c0 - L(x - 1, up) - L(x + 1, down)) + abs(c1 - L(x, up2) - L(x + 2, y));
            }
        }
        sv += __shfl_xor(sv, 1, 64);
        sh += __shfl_xor(sh, 1, 64);
        sd0 += __shfl_xor(sd0, 1, 64);
        sd1 += __shfl_xor(sd1, 1, 64);
        sv += __shfl_xor(sv, 2, 64);
        sh += __shfl_xor(sh, 2, 64);
        sd0 += __shfl_xor(sd0, 2, 64);
        sd1 += __shfl_xor(sd1, 2, 64);
        if (mode) {
            int cls, tr;
            alf_class(sv, sh, sd0, sd1, (byv == vb_pos - 4 || byv == vb_pos) ? 96 : 64, bd, &cls, &tr);
            const int bank = mode == RMH_ALF_LUMA_APS ? 16 : mode - 1;
            const int16_t* e = J.bank + ((bank * 25 + cls) * 4 + tr) * 24 + q * 6;
#pragma unroll
            for (int j = 0; j < 6; j++) prm[bk][q * 6 + j] = e[j];
        }
        if (q == 0) on[bk] = mode != 0;
    }
    __syncthreads();

    const int r = tid >> 4, c4 = (tid & 15) * 4;
    const int y = ty0 + r;
    if (y >= (int)J.height) return;
    const int kb = (r >> 2) * 16 + (tid & 15);
    const int width = (int)J.width;
    const int maxv = (1 << bd) - 1;
    // vertical tap limit at the virtual boundary
    const int yv = y & ctb_mask;
    int lim = 3;
    if (yv < vb_pos && yv + 4 >= vb_pos)
        lim = vb_pos - 1 - yv;
    else if (yv >= vb_pos && yv <= vb_pos + 3)
        lim = yv - vb_pos;
    const int m1 = min(1, lim), m2 = min(2, lim), m3 = lim;
    const int shift = lim == 0 ? 10 : 7;
    const bool filt = on[kb] && y < (int)J.vis_h;
    int fc[12], cl[12];
    if (filt) {
#pragma unroll
        for (int k = 0; k < 12; k++) {
            fc[k] = prm[kb][k];
            cl[k] = prm[kb][12 + k];
        }
    }
    T out[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int x = tx0 + c4 + i;
        int v;
        if (filt && x < (int)J.vis_w) {
            const int cur = L(x, y);
            int sum = 0;
#define TAP(k, dx, dy)                                                                        \
    sum += fc[k] * (clampi(L(x + (dx), y + (dy)) - cur, -cl[k], cl[k]) +                      \
                    clampi(L(x - (dx), y - (dy)) - cur, -cl[k], cl[k]))
            TAP(0, 0, m3);
            TAP(1, 1, m2);
            TAP(2, 0, m2);
            TAP(3, -1, m2);
            TAP(4, 2, m1);
            TAP(5, 1, m1);
            TAP(6, 0, m1);
            TAP(7, -1, m1);
            TAP(8, -2, m1);
            TAP(9, 3, 0);
            TAP(10, 2, 0);
            TAP(11, 1, 0);
#undef TAP
            v = clampi(cur + ((sum + (1 << (shift - 1))) >> shift), 0, maxv);
        } else if (x < bw && y < bh) {
            v = L(x, y);
        } else {
            v = x < width ? (int)src[(size_t)y * sstride + x] : 0;
        }
        out[i] = (T)v;
    }
#undef L
    RMH_GLOBAL T* d = gdst<T>(J.dst) + (size_t)y * J.dst_stride + tx0 + c4;
    const int x0 = tx0 + c4;
    if (x0 + 4 <= width && ((uintptr_t)d & (4 * sizeof(T) - 1)) == 0) {
        if (sizeof(T) == 2)
            *reinterpret_cast<RMH_GLOBAL u32x2*>(d) = *reinterpret_cast<const u32x2*>(out);
        else
            *reinterpret_cast<RMH_GLOBAL uint32_t*>(d) = *reinterpret_cast<const uint32_t*>(out);
    } else {
        for (int i = 0; i < 4 && x0 + i < width; i++) d[i] = out[i];
    }
}

// ---------------------------------------------------------------- ALF chroma + CCALF

template <typename T>
__global__ __launch_bounds__(256) void alf_chroma_kernel(const AlfChromaJob* __restrict__ jobs,
                                                         const uint32_t* __restrict__ block_job, uint32_t block0) {
    const uint32_t b = blockIdx.x + block0;
    const AlfChromaJob& J = jobs[block_job[b]];
    const uint32_t rel = b - J.first_block;
    const int tid = (int)threadIdx.x;
    const int x0 = (int)(rel % J.tiles_x) * kAlfTileW + (tid & 15) * 4;
    const int y = (int)(rel / J.tiles_x) * kAlfTileH + (tid >> 4);
    const int width = (int)J.width;
    if (y >= (int)J.height || x0 >= width) return;
    const RMH_GLOBAL T* src = gsrc<T>(J.src);
    const RMH_GLOBAL T* luma = gsrc<T>(J.luma);
    const size_t sstride = J.src_stride, lstride = J.luma_stride;
    const int ctb_log2 = (int)J.ctb_log2, ctb = 1 << ctb_log2;
    const int cl2 = ctb_log2 - 1, ctbc = 1 << cl2;
    const int bd = (int)J.bd, maxv = (1 << bd) - 1;
    const int cbw = (int)J.cbw, cbh = (int)J.cbh, lw = (int)J.lw, lh = (int)J.lh;
    const int gy = y >> cl2;
    const int hc = max(min(ctb, (int)J.vis_h - gy * ctb) >> 1, 1);
    const bool row_in = gy < (int)J.ctb_yn && y - gy * ctbc < hc;

    // chroma virtual boundary: tap limit 2, or the distance from ctbc - 2
    const int yv = y & (ctbc - 1), vbc = ctbc - 2;
    int lim = 2;
    if (yv < vbc && yv + 2 >= vbc)
        lim = vbc - 1 - yv;
    else if (yv >= vbc && yv <= vbc + 1)
        lim = yv - vbc;
    const int m1 = min(1, lim), m2 = lim;
    const int shift = lim == 0 ? 10 : 7;
    // CCALF luma rows around 2y at the luma virtual boundary ctb - 4
    const int ly = 2 * y, pos = ly & (ctb - 1), vbl = ctb - 4;
    int above = ly - 1, below = ly + 1, below2 = ly + 2;
    if (pos == vbl - 2 || pos == vbl + 1) {
        below2 = ly + 1;
    } else if (pos == vbl - 1 || pos == vbl) {
        above = ly;
        below = ly;
        below2 = ly;
    }
    const int yc = clampi(y, 0, cbh - 1);
    const int ra = clampi(above, 0, lh - 1), r0 = clampi(ly, 0, lh - 1), rb = clampi(below, 0, lh - 1),
              rb2 = clampi(below2, 0, lh - 1);
    const int cc_off = 1 << (bd - 1);

    T out[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int x = x0 + i;
        if (x >= width) {
            out[i] = 0;
            continue;
        }
        int v = src[(size_t)y * sstride + x];
        const int gx = x >> cl2;
        if (row_in && gx < (int)J.ctb_xn) {
            const int wc = max(min(ctb, (int)J.vis_w - gx * ctb) >> 1, 1);
            if (x - gx * ctbc < wc) {
                const rmh_alf_ctb s = J.ctbs[(size_t)gy * J.ctb_cols + gx];
                const int alt = J.comp == 1 ? s.cb : s.cr;
                const int idc = J.comp == 1 ? s.cc_cb : s.cc_cr;
                if (alt && J.chroma_bank) {
                    const int16_t* e = J.chroma_bank + min(alt - 1, (int)J.n_chroma - 1) * 12;
#define C(dx, dy) (int)src[(size_t)clampi(y + (dy), 0, cbh - 1) * sstride + clampi(x + (dx), 0, cbw - 1)]
                    const int cur = (int)src[(size_t)yc * sstride + clampi(x, 0, cbw - 1)];
                    int sum = 0;
#define TAP(k, dx, dy) \
    sum += e[k] * (clampi(C(dx, dy) - cur, -e[6 + k], e[6 + k]) + clampi(C(-(dx), -(dy)) - cur, -e[6 + k], e[6 + k]))
                    TAP(0, 0, m2);
                    TAP(1, 1, m1);
                    TAP(2, 0, m1);
                    TAP(3, -1, m1);
                    TAP(4, 2, 0);
                    TAP(5, 1, 0);
#undef TAP
#undef C
                    v = clampi(cur + ((sum + (1 << (shift - 1))) >> shift), 0, maxv);
                }
                if (idc && idc <= (int)J.n_cc && x < (int)J.cvis_w && y < (int)J.cvis_h) {
                    const int16_t* c = J.cc_bank + (idc - 1) * 7;
                    const int lx = 2 * x;
                    const int xl = clampi(lx - 1, 0, lw - 1), xm = clampi(lx, 0, lw - 1), xr = clampi(lx + 1, 0, lw - 1);
#define Y(xx, rr) (int)luma[(size_t)(rr) * lstride + (xx)]
                    const int cur = Y(xm, r0);
                    int sum = c[0] * (Y(xm, ra) - cur);
                    sum += c[1] * (Y(xl, r0) - cur);
                    sum += c[2] * (Y(xr, r0) - cur);
                    sum += c[3] * (Y(xl, rb) - cur);
                    sum += c[4] * (Y(xm, rb) - cur);
                    sum += c[5] * (Y(xr, rb) - cur);
                    sum += c[6] * (Y(xm, rb2) - cur);
#undef Y
                    v = clampi(v + clampi((sum + 64) >> 7, -cc_off, cc_off - 1), 0, maxv);
                }
            }
        }
        out[i] = (T)v;
    }
    RMH_GLOBAL T* d = gdst<T>(J.dst) + (size_t)y * J.dst_stride + x0;
    if (x0 + 4 <= width && ((uintptr_t)d & (4 * sizeof(T) - 1)) == 0) {
        if (sizeof(T) == 2)
            *reinterpret_cast<RMH_GLOBAL u32x2*>(d) = *reinterpret_cast<const u32x2*>(out);
        else
            *reinterpret_cast<RMH_GLOBAL uint32_t*>(d) = *reinterpret_cast<const uint32_t*>(out);
    } else {
        for (int i = 0; i < 4 && x0 + i < width; i++) d[i] = out[i];
    }
}

}  // namespace

hipError_t launch_lmcs(const LmcsJob* jobs, const uint32_t* block_job, const uint32_t (&range)[3], hipStream_t stream) {
    if (range[1] > range[0])
        hipLaunchKernelGGL(lmcs_kernel<uint8_t>, dim3(range[1] - range[0]), dim3(256), 0, stream, jobs, block_job, range[0]);
    if (range[2] > range[1])
        hipLaunchKernelGGL(lmcs_kernel<uint16_t>, dim3(range[2] - range[1]), dim3(256), 0, stream, jobs, block_job,
                           range[1]);
    return hipGetLastError();
}

hipError_t launch_alf(const AlfLumaJob* luma, const uint32_t* luma_block_job, const uint32_t (&luma_range)[3],
                      const AlfChromaJob* chroma, const uint32_t* chroma_block_job, const uint32_t (&chroma_range)[3],
                      hipStream_t stream) {
    if (luma_range[1] > luma_range[0])
        hipLaunchKernelGGL(alf_luma_kernel<uint8_t>, dim3(luma_range[1] - luma_range[0]), dim3(256), 0, stream, luma,
                           luma_block_job, luma_range[0]);
    if (luma_range[2] > luma_range[1])
        hipLaunchKernelGGL(alf_luma_kernel<uint16_t>, dim3(luma_range[2] - luma_range[1]), dim3(256), 0, stream, luma,
                           luma_block_job, luma_range[1]);
    if (chroma_range[1] > chroma_range[0])
        hipLaunchKernelGGL(alf_chroma_kernel<uint8_t>, dim3(chroma_range[1] - chroma_range[0]), dim3(256), 0, stream,
                           chroma, chroma_block_job, chroma_range[0]);
    if (chroma_range[2] > chroma_range[1])
        hipLaunchKernelGGL(alf_chroma_kernel<uint16_t>, dim3(chroma_range[2] - chroma_range[1]), dim3(256), 0, stream,
                           chroma, chroma_block_job, chroma_range[1]);
    return hipGetLastError();
}

}  // namespace rmh

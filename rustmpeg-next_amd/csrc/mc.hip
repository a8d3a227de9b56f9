// mc.hip - uni-prediction motion compensation for gfx950: 8-tap luma and
// 4-tap 4:2:0 chroma fractional-sample interpolation.
//
// Semantics follow interpolate_luma_block / interpolate_chroma_block_420
// (crates/rustmpeg-video/src/vvc/inter.rs:886-1182) at bit depth bd with
// head_room = 14 - bd (the spec-consistent second-stage shift 6 + head_room;
// see the note on inter.rs:2046 in the design notes), the uni-prediction path
// of predict_inter_pu (inter.rs:1734-1848) and write_prediction
// (decode_general.rs:2641-2667).
//
// Mapping: one 256-lane workgroup per 16x16 luma tile of a PU and the matching
// 8x8 tiles of both chroma planes. The reference window (tile + 7 / + 3
// samples of halo, edge-replicated) is staged once in LDS; the horizontal pass
// writes 14-bit intermediates to LDS; the vertical pass writes the prediction.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace rmh {
namespace {

__constant__ int8_t kLuma8[16][8] = {
    {0, 0, 0, 64, 0, 0, 0, 0},        {0, 1, -3, 63, 4, -2, 1, 0},      {-1, 2, -5, 62, 8, -3, 1, 0},
    {-1, 3, -8, 60, 13, -4, 1, 0},    {-1, 4, -10, 58, 17, -5, 1, 0},   {-1, 4, -11, 52, 26, -8, 3, -1},
    {-1, 3, -9, 47, 31, -10, 4, -1},  {-1, 4, -11, 45, 34, -10, 4, -1}, {-1, 4, -11, 40, 40, -11, 4, -1},
    {-1, 4, -10, 34, 45, -11, 4, -1}, {-1, 4, -10, 31, 47, -9, 3, -1},  {-1, 3, -8, 26, 52, -11, 4, -1},
    {0, 1, -5, 17, 58, -10, 4, -1},   {0, 1, -4, 13, 60, -8, 3, -1},    {0, 1, -3, 8, 62, -5, 2, -1},
    {0, 1, -2, 4, 63, -3, 1, 0}};
__constant__ int8_t kLumaAltHpel[8] = {0, 3, 9, 20, 20, 9, 3, 0};
__constant__ int8_t kChroma4[32][4] = {
    {0, 64, 0, 0},    {-1, 63, 2, 0},   {-2, 62, 4, 0},   {-2, 60, 7, -1},  {-2, 58, 10, -2}, {-3, 57, 12, -2},
    {-4, 56, 14, -2}, {-4, 55, 15, -2}, {-4, 54, 16, -2}, {-5, 53, 18, -2}, {-6, 52, 20, -2}, {-6, 49, 24, -3},
    {-6, 46, 28, -4}, {-5, 44, 29, -4}, {-4, 42, 30, -4}, {-4, 39, 33, -4}, {-4, 36, 36, -4}, {-4, 33, 39, -4},
    {-4, 30, 42, -4}, {-4, 29, 44, -5}, {-4, 28, 46, -6}, {-3, 24, 49, -6}, {-2, 20, 52, -6}, {-2, 18, 53, -5},
    {-2, 16, 54, -4}, {-2, 15, 55, -4}, {-2, 14, 56, -4}, {-2, 12, 57, -3}, {-2, 10, 58, -2}, {-1, 7, 60, -2},
    {0, 4, 62, -2},   {0, 2, 63, -1}};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

__device__ __forceinline__ int load_sample(const rmh_plane& p, int x, int y) {
    x = clampi(x, 0, (int)p.width - 1);
    y = clampi(y, 0, (int)p.height - 1);
    size_t o = (size_t)y * p.stride + x;
    return p.bytes_per_sample == 1 ? (int)((const uint8_t*)p.data)[o] : (int)((const uint16_t*)p.data)[o];
}

__device__ __forceinline__ void store_sample(const rmh_plane& p, int x, int y, int v) {
    size_t o = (size_t)y * p.stride + x;
    if (p.bytes_per_sample == 1)
        ((uint8_t*)p.data)[o] = (uint8_t)v;
    else
        ((uint16_t*)p.data)[o] = (uint16_t)v;
}

// Interpolates a tw x th tile (tw, th <= T) whose top-left output sample is
// (x0, y0) of `dst`; `ix`, `iy` are the integer reference positions of that
// sample and `fx`, `fy` the fractions. TAPS = 8 (luma) or 4 (chroma).
template <int TAPS, int T>
__device__ void interp_tile(const rmh_plane& ref, const rmh_plane& dst, int x0, int y0, int tw, int th, int ix,
                            int iy, const int8_t* cx, const int8_t* cy, bool hfrac, bool vfrac, int* win, int* tmp) {
    constexpr int HALO = TAPS - 1;
    constexpr int BEFORE = TAPS / 2 - 1;  // 3 for luma, 1 for chroma
    constexpr int WW = T + HALO;          // window width / height
    const int tid = threadIdx.x;
    const int bd = ref.bit_depth;
    const int head_room = 14 - bd;
    const int maxv = (1 << bd) - 1;

    for (int e = tid; e < WW * WW; e += blockDim.x) {
        int r = e / WW, c = e - (e / WW) * WW;
        win[e] = load_sample(ref, ix - BEFORE + c, iy - BEFORE + r);
    }
    __syncthreads();
    if (!hfrac && !vfrac) {
        for (int e = tid; e < T * T; e += blockDim.x) {
            int y = e / T, x = e % T;
            if (x < tw && y < th) store_sample(dst, x0 + x, y0 + y, win[(y + BEFORE) * WW + x + BEFORE]);
        }
        return;
    }
    if (!vfrac || !hfrac) {
        // one-dimensional: (sum + 32) >> 6 clipped (inter.rs:939-1002)
        for (int e = tid; e < T * T; e += blockDim.x) {
            int y = e / T, x = e % T;
            if (x < tw && y < th) {
                int sum = 0;
                if (hfrac) {
#pragma unroll
                    for (int k = 0; k < TAPS; k++) sum += cx[k] * win[(y + BEFORE) * WW + x + k];
                } else {
#pragma unroll
                    for (int k = 0; k < TAPS; k++) sum += cy[k] * win[(y + k) * WW + x + BEFORE];
                }
                store_sample(dst, x0 + x, y0 + y, clampi((sum + 32) >> 6, 0, maxv));
            }
        }
        return;
    }
    // two-dimensional (inter.rs:1004-1050): horizontal pass over th + HALO rows
    const int shift1 = 6 - head_room;
    const int off1 = -8192 * (1 << shift1);
    for (int e = tid; e < WW * T; e += blockDim.x) {
        int r = e / T, x = e % T;
        int sum = 0;
#pragma unroll
        for (int k = 0; k < TAPS; k++) sum += cx[k] * win[r * WW + x + k];
        tmp[e] = (sum + off1) >> shift1;
    }
    __syncthreads();
    const int shift2 = 6 + head_room;
    const int off2 = (1 << (shift2 - 1)) + (8192 << 6);
    for (int e = tid; e < T * T; e += blockDim.x) {
        int y = e / T, x = e % T;
        if (x < tw && y < th) {
            int sum = 0;
#pragma unroll
            for (int k = 0; k < TAPS; k++) sum += cy[k] * tmp[(y + k) * T + x];
            store_sample(dst, x0 + x, y0 + y, clampi((sum + off2) >> shift2, 0, maxv));
        }
    }
}

__global__ __launch_bounds__(256) void mc_kernel(const rmh_mc_frame* __restrict__ frames,
                                                 const rmh_pu_desc* __restrict__ pus,
                                                 const McItem* __restrict__ items) {
    __shared__ int win[23 * 23];
    __shared__ int tmp[23 * 16];
    const McItem it = items[blockIdx.x];
    const rmh_pu_desc pu = pus[it.pu];
    const rmh_mc_frame& f = frames[pu.frame];
    const int w = 1 << (pu.log2_w & 15), h = 1 << (pu.log2_h & 15);
    const bool alt = (pu.log2_w & RMH_PU_ALT_HPEL) != 0;
    const int tx0 = it.tx * 16, ty0 = it.ty * 16;
    const int tw = min(16, w - tx0), th = min(16, h - ty0);
    const int x0 = pu.x + tx0, y0 = pu.y + ty0;

    // luma: 1/16 sample
    {
        const int fx = pu.mv_x & 15, fy = pu.mv_y & 15;
        const int ix = x0 + (pu.mv_x >> 4), iy = y0 + (pu.mv_y >> 4);
        const int8_t* cx = (alt && fx == 8) ? kLumaAltHpel : kLuma8[fx];
        const int8_t* cy = (alt && fy == 8) ? kLumaAltHpel : kLuma8[fy];
        const int vw = min(tw, (int)f.dst[0].width - x0), vh = min(th, (int)f.dst[0].height - y0);
        interp_tile<8, 16>(f.ref[0], f.dst[0], x0, y0, vw, vh, ix, iy, cx, cy, fx != 0, fy != 0, win, tmp);
    }
    // chroma 4:2:0: the same MV at 1/32 sample
    const int fx = pu.mv_x & 31, fy = pu.mv_y & 31;
    const int cx0 = x0 / 2, cy0 = y0 / 2;
    const int ix = cx0 + (pu.mv_x >> 5), iy = cy0 + (pu.mv_y >> 5);
    for (int c = 1; c < 3; c++) {
        __syncthreads();
        const int vw = min(tw / 2, (int)f.dst[c].width - cx0), vh = min(th / 2, (int)f.dst[c].height - cy0);
        interp_tile<4, 8>(f.ref[c], f.dst[c], cx0, cy0, vw, vh, ix, iy, kChroma4[fx], kChroma4[fy], fx != 0, fy != 0,
                          win, tmp);
    }
}

}  // namespace

hipError_t launch_mc(const rmh_mc_frame* frames, const rmh_pu_desc* pus, const McItem* items, uint32_t n_items,
                     hipStream_t stream) {
    if (n_items == 0) return hipSuccess;
    hipLaunchKernelGGL(mc_kernel, dim3(n_items), dim3(256), 0, stream, frames, pus, items);
    return hipGetLastError();
}

}  // namespace rmh

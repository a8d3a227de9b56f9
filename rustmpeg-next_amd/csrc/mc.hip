// mc.hip - motion compensation for gfx950: 8-tap luma and 4-tap 4:2:0
// chroma fractional-sample interpolation, uni-prediction (mc_kernel) and
// bi-prediction with CU-level weights (mc_bi_kernel).
//
// Semantics follow interpolate_luma_block / interpolate_chroma_block_420
// (crates/rustmpeg-video/src/vvc/inter.rs:886-1182) at bit depth bd with
// head_room = 14 - bd (the spec-consistent second-stage shift 6 + head_room;
// see the note on inter.rs:2046 in the design notes), the uni-prediction path
// of predict_inter_pu (inter.rs:1734-1848) and write_prediction
// (decode_general.rs:2641-2667).
//
// Mapping: one wave per 16x16 luma tile of a PU and the matching 8x8 tiles of
// both chroma planes, four waves per workgroup. The three reference windows
// (tile + 7 / + 3 samples of halo, edge-replicated) are staged in LDS as
// int16 before any filtering; the horizontal pass writes 14-bit intermediates
// to LDS; the vertical pass writes the prediction.
#include <hip/hip_runtime.h>

#include "gmem.h"
#include "kernels.h"
#include "xcd.h"

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

// Sample access with the sample size BPS (1: 8-bit planes, 2: up to 12-bit)
// fixed at compile time; a launch covers items of one sample size.
template <int BPS>
__device__ __forceinline__ int load_sample(const rmh_plane& p, int x, int y) {
    x = clampi(x, 0, (int)p.width - 1);
    y = clampi(y, 0, (int)p.height - 1);
    const uint32_t o = (uint32_t)y * p.stride + (uint32_t)x;
    return BPS == 1 ? (int)gsrc<uint8_t>(p.data)[o] : (int)gsrc<uint16_t>(p.data)[o];
}

template <int BPS>
__device__ __forceinline__ void store_sample(const rmh_plane& p, int x, int y, int v) {
    const uint32_t o = (uint32_t)y * p.stride + (uint32_t)x;
    if (BPS == 1)
        gdst<uint8_t>(p.data)[o] = (uint8_t)v;
    else
        gdst<uint16_t>(p.data)[o] = (uint16_t)v;
}

__device__ __forceinline__ int dot2(uint32_t a, uint32_t b, int acc) {
    typedef short short2v __attribute__((ext_vector_type(2)));
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(short2v, a), __builtin_bit_cast(short2v, b), acc, false);
}

// (hi:lo) >> 16: the int16 pair starting one sample after `lo`'s first.
__device__ __forceinline__ uint32_t mid(uint32_t hi, uint32_t lo) { return __builtin_amdgcn_alignbyte(hi, lo, 2); }

__device__ __forceinline__ uint32_t pack2(int a, int b) { return (uint32_t)(uint16_t)a | ((uint32_t)(uint16_t)b << 16); }

// Geometry of one wave's LDS (int16): reference windows row-major with an even
// pitch, so sample pairs (x, x + 1) at even x are aligned dwords, and the
// horizontal-pass intermediates column-major, so vertical pairs are too.
constexpr int kLR = 23, kLP = 24;  // luma window rows, pitch (16 + 7 samples, + 1 pad)
constexpr int kCR = 11, kCP = 12;  // chroma window rows, pitch (8 + 3, + 1 pad)
struct WaveLds {
    int16_t luma[kLR * kLP];
    int16_t ltmp[16 * kLP];        // [x][r]
    int16_t chroma[2][kCR * kCP];
    int16_t ctmp[2][8 * kCP];      // [x][r]
};

// One plane's tile of an item; every field is the same for all lanes of the
// wave. Taps are kept as int16 pairs (c0, c1), (c2, c3), ...
template <int TAPS>
struct Tile {
    int x0, y0;    // first output sample in dst
    int tw, th;    // output size (clipped to the picture)
    int ix, iy;    // integer reference position of the first output sample
    uint32_t cx[TAPS / 2];
    uint32_t cy[TAPS / 2];
    int sx[TAPS];  // horizontal taps as ints (bi-prediction samples)
    int sy[TAPS];  // vertical taps as ints (vertical-only tiles)
    bool hf, vf;   // fractional horizontally / vertically
};

// Phase 1: window rows iy - BEFORE .. (ROWS) and columns ix - BEFORE ..
// (PITCH, the last one only as padding), clamped to the picture. Lane (c, r0)
// holds column c of rows r0, r0 + 64 / LANES_X, ... fetch issues the loads
// into registers; put writes them to LDS, so a wave can have the loads of all
// its windows in flight before the first LDS write waits for them.
template <int ROWS, int LANES_X>
struct Window {
    static constexpr int RSTEP = 64 / LANES_X, N = (ROWS + RSTEP - 1) / RSTEP;
    int v[N];
};

template <int TAPS, int ROWS, int PITCH, int LANES_X, int BPS>
__device__ __forceinline__ void fetch(const rmh_plane& ref, const Tile<TAPS>& t, Window<ROWS, LANES_X>& w, int lane) {
    using W = Window<ROWS, LANES_X>;
    constexpr int BEFORE = TAPS / 2 - 1;
    const int ox = t.ix - BEFORE, oy = t.iy - BEFORE;
    const int c = lane & (LANES_X - 1), r0 = lane / LANES_X;
    if (c >= PITCH) return;
    const bool inside = ox >= 0 && oy >= 0 && ox + PITCH <= (int)ref.width && oy + ROWS <= (int)ref.height;
    if (inside) {
        const uint32_t base = (uint32_t)(oy + r0) * ref.stride + (uint32_t)(ox + c);
        const uint32_t step = W::RSTEP * ref.stride;
#pragma unroll
        for (int k = 0; k < W::N; k++)
            if (r0 + W::RSTEP * k < ROWS)
                w.v[k] = BPS == 1 ? (int)gsrc<uint8_t>(ref.data)[base + k * step]
                                  : (int)gsrc<uint16_t>(ref.data)[base + k * step];
    } else {
#pragma unroll
        for (int k = 0; k < W::N; k++)
            if (r0 + W::RSTEP * k < ROWS) w.v[k] = load_sample<BPS>(ref, ox + c, oy + r0 + W::RSTEP * k);
    }
}

template <int ROWS, int PITCH, int LANES_X>
__device__ __forceinline__ void put(const Window<ROWS, LANES_X>& w, int16_t* win, int lane) {
    using W = Window<ROWS, LANES_X>;
    const int c = lane & (LANES_X - 1), r0 = lane / LANES_X;
    if (c >= PITCH) return;
#pragma unroll
    for (int k = 0; k < W::N; k++)
        if (r0 + W::RSTEP * k < ROWS) win[(r0 + W::RSTEP * k) * PITCH + c] = (int16_t)w.v[k];
}

// Horizontal FIR of output pair (x, x + 1), x even, on a row of int16 samples:
// TAPS / 2 + 1 aligned dword reads, pairs at odd offsets by alignbyte.
template <int TAPS>
__device__ __forceinline__ void hpair(const int16_t* row, const uint32_t (&c)[TAPS / 2], int& s0, int& s1) {
    const uint32_t* d = reinterpret_cast<const uint32_t*>(row);
    uint32_t v[TAPS / 2 + 1];
#pragma unroll
    for (int k = 0; k <= TAPS / 2; k++) v[k] = d[k];
    s0 = 0;
    s1 = 0;
#pragma unroll
    for (int k = 0; k < TAPS / 2; k++) {
        s0 = dot2(v[k], c[k], s0);
        s1 = dot2(mid(v[k + 1], v[k]), c[k], s1);
    }
}

// Phase 2 (two-dimensional tiles only): horizontal pass over th + TAPS - 1
// rows into column-major intermediates, (sum - (8192 << shift1)) >> shift1
// (interpolate_luma_block 2-D branch, vvc/inter.rs:1004-1030), for NPL planes
// of one tile geometry (the two chroma planes share their tile). Lane item e:
// plane e / (RMAX * NPAIR), then pair xp = e % NPAIR, row e / NPAIR; a fixed
// number of items per lane.
template <int TAPS, int PITCH, int NPAIR, int RMAX, int NPL, int WSTRIDE, int TSTRIDE>
__device__ __forceinline__ void hpass(const Tile<TAPS>& t, const int16_t* win, int16_t* tmp, int shift1, int lane) {
    if (!(t.hf && t.vf)) return;
    constexpr int PER = RMAX * NPAIR, N = (NPL * PER + 63) / 64;
    const int rows = t.th + TAPS - 1;
    const int off = 8192 << shift1;
#pragma unroll
    for (int k = 0; k < N; k++) {
        const int e0 = lane + 64 * k;
        const int pl = NPL == 1 ? 0 : e0 / PER, e = e0 - pl * PER;
        const int xp = e % NPAIR, r = e / NPAIR, x = 2 * xp;
        if (e0 < NPL * PER && r < rows && x < t.tw) {
            int s0, s1;
            hpair<TAPS>(win + pl * WSTRIDE + r * PITCH + x, t.cx, s0, s1);
            int16_t* o = tmp + pl * TSTRIDE + x * PITCH + r;
            o[0] = (int16_t)((s0 - off) >> shift1);
            o[PITCH] = (int16_t)((s1 - off) >> shift1);
        }
    }
}

// Phase 3 (uni-prediction): the four output samples (x, y), (x + 1, y),
// (x, y + 1), (x + 1, y + 1), x and y even, of one plane's tile, clipped to
// [0, maxv]: a copy, a one-dimensional filter (sum + 32) >> 6
// (vvc/inter.rs:939-1002), or the vertical pass over the intermediates
// (sum + 2^(shift2-1) + (8192 << 6)) >> shift2, shift2 = 6 + head_room.
template <int TAPS, int PITCH>
__device__ __forceinline__ void uni_quad(const Tile<TAPS>& t, const int16_t* win, const int16_t* tmp, int x, int y,
                                         int head_room, int maxv, int (&v)[4]) {
    constexpr int BEFORE = TAPS / 2 - 1;
    if (t.hf && t.vf) {
        const int shift2 = 6 + head_room;
        const int off2 = (1 << (shift2 - 1)) + (8192 << 6);
        int s[4];
        hpair<TAPS>(tmp + x * PITCH + y, t.cy, s[0], s[2]);
        hpair<TAPS>(tmp + (x + 1) * PITCH + y, t.cy, s[1], s[3]);
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = clampi((s[q] + off2) >> shift2, 0, maxv);
    } else if (t.hf) {
        int s[4];
        hpair<TAPS>(win + (y + BEFORE) * PITCH + x, t.cx, s[0], s[1]);
        hpair<TAPS>(win + (y + 1 + BEFORE) * PITCH + x, t.cx, s[2], s[3]);
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = clampi((s[q] + 32) >> 6, 0, maxv);
    } else if (t.vf) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int sum = 0;
#pragma unroll
            for (int k = 0; k < TAPS; k++) sum += t.sy[k] * win[(y + (q >> 1) + k) * PITCH + x + (q & 1) + BEFORE];
            v[q] = clampi((sum + 32) >> 6, 0, maxv);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = win[(y + (q >> 1) + BEFORE) * PITCH + x + (q & 1) + BEFORE];
    }
}

// Stores the 2 x 2 samples v at (x, y) of dst, those at x + 1 / y + 1 only
// when present; pair: sample pairs (x, x + 1) are aligned for one 2-sample store.
template <int BPS>
__device__ __forceinline__ void store_quad(const rmh_plane& dst, int x, int y, const int (&v)[4], bool x1, bool y1,
                                           bool pair) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
        if (h == 1 && !y1) break;
        const uint32_t o = (uint32_t)(y + h) * dst.stride + (uint32_t)x;
        if (pair && x1) {
            if (BPS == 1)
                *(RMH_GLOBAL uint16_t*)(gdst<uint8_t>(dst.data) + o) = (uint16_t)(v[2 * h] | (v[2 * h + 1] << 8));
            else
                *(RMH_GLOBAL uint32_t*)(gdst<uint16_t>(dst.data) + o) = pack2(v[2 * h], v[2 * h + 1]);
        } else {
            store_sample<BPS>(dst, x, y + h, v[2 * h]);
            if (x1) store_sample<BPS>(dst, x + 1, y + h, v[2 * h + 1]);
        }
    }
}

// Whether sample pairs at even tile columns are aligned in dst.
template <int BPS>
__device__ __forceinline__ bool pairs_aligned(const rmh_plane& dst, int x0) {
    return (((uintptr_t)dst.data & (2 * BPS - 1)) | (dst.stride & 1) | (x0 & 1)) == 0;
}

template <int TAPS>
__device__ __forceinline__ void set_taps(Tile<TAPS>& t, const int8_t* cx, const int8_t* cy) {
#pragma unroll
    for (int k = 0; k < TAPS / 2; k++) {
        t.cx[k] = pack2(cx[2 * k], cx[2 * k + 1]);
        t.cy[k] = pack2(cy[2 * k], cy[2 * k + 1]);
    }
#pragma unroll
    for (int k = 0; k < TAPS; k++) {
        t.sx[k] = cx[k];
        t.sy[k] = cy[k];
    }
}

// The 14-bit bi-prediction intermediates of the 2 x 2 output samples
// (x, y), (x + 1, y), (x, y + 1), (x + 1, y + 1), x and y even, of a tile
// (for_bi_pred branches of interpolate_luma_block_16 / interpolate_chroma_
// block_format_16, vvc/inter.rs:2046-2133, 2582-2669): a copy gives
// (s << head_room) - 8192, a one-dimensional filter (sum >> shift1) - 8192,
// a two-dimensional one the vertical sum over the horizontal pass's
// intermediates (which already hold (sum >> shift1) - 8192), >> 6. Pairs
// along the filter direction come from hpair's aligned dword reads.
template <int TAPS, int PITCH>
__device__ __forceinline__ void bi_quad(const Tile<TAPS>& t, const int16_t* win, const int16_t* tmp, int x, int y,
                                        int head_room, int shift1, int (&v)[4]) {
    constexpr int BEFORE = TAPS / 2 - 1;
    if (t.hf && t.vf) {
        int s0, s1, s2, s3;
        hpair<TAPS>(tmp + x * PITCH + y, t.cy, s0, s2);
        hpair<TAPS>(tmp + (x + 1) * PITCH + y, t.cy, s1, s3);
        v[0] = s0 >> 6;
        v[1] = s1 >> 6;
        v[2] = s2 >> 6;
        v[3] = s3 >> 6;
    } else if (t.hf) {
        int s0, s1, s2, s3;
        hpair<TAPS>(win + (y + BEFORE) * PITCH + x, t.cx, s0, s1);
        hpair<TAPS>(win + (y + 1 + BEFORE) * PITCH + x, t.cx, s2, s3);
        v[0] = (s0 >> shift1) - 8192;
        v[1] = (s1 >> shift1) - 8192;
        v[2] = (s2 >> shift1) - 8192;
        v[3] = (s3 >> shift1) - 8192;
    } else if (t.vf) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int s = 0;
#pragma unroll
            for (int k = 0; k < TAPS; k++) s += t.sy[k] * win[(y + (q >> 1) + k) * PITCH + x + (q & 1) + BEFORE];
            v[q] = (s >> shift1) - 8192;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++)
            v[q] = ((int)win[(y + (q >> 1) + BEFORE) * PITCH + x + (q & 1) + BEFORE] << head_room) - 8192;
    }
}

// Four work items per 256-lane workgroup, one per wave: the 16x16 luma tile
// of a PU (smaller for 8-wide or 8-tall PUs) and its two chroma tiles. The
// loads of all three reference windows are issued before the first LDS write.
// Outputs go in 2 x 2 quads: the 64 quads of the luma tile one per lane, the
// 2 x 16 of the chroma tiles on lanes 0 .. 31. Item fields, frame planes and
// filter taps are wave-uniform and read with scalar loads.
// NW: waves (work items) per workgroup, chosen per plan by the autotuner.
template <int BPS, int NW>
__global__ __launch_bounds__(64 * NW) void mc_kernel(const rmh_mc_frame* __restrict__ frames,
                                                     const McItem* __restrict__ items, uint32_t n_items) {
    __shared__ __align__(16) WaveLds lds[NW];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t item = xcd_local_block(blockIdx.x, gridDim.x) * NW + wave;
    const bool active = item < n_items;
    WaveLds& L = lds[wave];
    Tile<8> tl{};
    Tile<4> tc{};
    const rmh_mc_frame* f = frames;
    int bd = 8;
    if (active) {
        const McItem it = items[item];
        f = &frames[it.frame];
        const bool alt = it.flags & 1;
        // luma: 1/16 sample
        const int fx = it.mv_x & 15, fy = it.mv_y & 15;
        tl.x0 = it.x0;
        tl.y0 = it.y0;
        tl.tw = it.tw;
        tl.th = it.th;
        tl.ix = tl.x0 + (it.mv_x >> 4);
        tl.iy = tl.y0 + (it.mv_y >> 4);
        set_taps<8>(tl, (alt && fx == 8) ? kLumaAltHpel : kLuma8[fx], (alt && fy == 8) ? kLumaAltHpel : kLuma8[fy]);
        tl.hf = fx != 0;
        tl.vf = fy != 0;
        // chroma 4:2:0: the same MV at 1/32 sample
        const int cfx = it.mv_x & 31, cfy = it.mv_y & 31;
        tc.x0 = it.x0 / 2;
        tc.y0 = it.y0 / 2;
        tc.tw = it.cw;
        tc.th = it.ch;
        tc.ix = tc.x0 + (it.mv_x >> 5);
        tc.iy = tc.y0 + (it.mv_y >> 5);
        set_taps<4>(tc, kChroma4[cfx], kChroma4[cfy]);
        tc.hf = cfx != 0;
        tc.vf = cfy != 0;
        bd = f->ref[0].bit_depth;
        const bool hl = tl.tw > 0 && tl.th > 0, hc = tc.tw > 0 && tc.th > 0;
        Window<kLR, 32> wl;
        Window<kCR, 16> wc0, wc1;
        if (hl) fetch<8, kLR, kLP, 32, BPS>(f->ref[0], tl, wl, lane);
        if (hc) {
            fetch<4, kCR, kCP, 16, BPS>(f->ref[1], tc, wc0, lane);
            fetch<4, kCR, kCP, 16, BPS>(f->ref[2], tc, wc1, lane);
        }
        if (hl) put<kLR, kLP, 32>(wl, L.luma, lane);
        if (hc) {
            put<kCR, kCP, 16>(wc0, L.chroma[0], lane);
            put<kCR, kCP, 16>(wc1, L.chroma[1], lane);
        }
    }
    __syncthreads();
    const int head_room = 14 - bd;
    const int shift1 = 6 - head_room;
    if (active && tl.tw > 0 && tl.th > 0) hpass<8, kLP, 8, kLR, 1, 0, 0>(tl, L.luma, L.ltmp, shift1, lane);
    if (active && tc.tw > 0 && tc.th > 0)
        hpass<4, kCP, 4, kCR, 2, kCR * kCP, 8 * kCP>(tc, L.chroma[0], L.ctmp[0], shift1, lane);
    __syncthreads();
    if (!active) return;
    const int maxv = (1 << bd) - 1;
    if (tl.tw > 0 && tl.th > 0) {
        const int x = 2 * (lane & 7), y = 2 * (lane >> 3);
        if (x < tl.tw && y < tl.th) {
            int v[4];
            uni_quad<8, kLP>(tl, L.luma, L.ltmp, x, y, head_room, maxv, v);
            store_quad<BPS>(f->dst[0], tl.x0 + x, tl.y0 + y, v, x + 1 < tl.tw, y + 1 < tl.th,
                            pairs_aligned<BPS>(f->dst[0], tl.x0));
        }
    }
    if (tc.tw > 0 && tc.th > 0 && lane < 32) {
        const int pl = lane >> 4, x = 2 * (lane & 3), y = 2 * ((lane >> 2) & 3);
        if (x < tc.tw && y < tc.th) {
            int v[4];
            uni_quad<4, kCP>(tc, L.chroma[pl], L.ctmp[pl], x, y, head_room, maxv, v);
            store_quad<BPS>(f->dst[1 + pl], tc.x0 + x, tc.y0 + y, v, x + 1 < tc.tw, y + 1 < tc.th,
                            pairs_aligned<BPS>(f->dst[1 + pl], tc.x0));
        }
    }
}

// Luma and chroma tiles of one list of a bi item.
__device__ __forceinline__ void bi_tiles(const McBiItem& it, int mv_x, int mv_y, Tile<8>& tl, Tile<4>& tc) {
    const bool alt = it.flags & 1;
    const int fx = mv_x & 15, fy = mv_y & 15;
    tl.x0 = it.x0;
    tl.y0 = it.y0;
    tl.tw = it.tw;
    tl.th = it.th;
    tl.ix = tl.x0 + (mv_x >> 4);
    tl.iy = tl.y0 + (mv_y >> 4);
    set_taps<8>(tl, (alt && fx == 8) ? kLumaAltHpel : kLuma8[fx], (alt && fy == 8) ? kLumaAltHpel : kLuma8[fy]);
    tl.hf = fx != 0;
    tl.vf = fy != 0;
    const int cfx = mv_x & 31, cfy = mv_y & 31;
    tc.x0 = it.x0 / 2;
    tc.y0 = it.y0 / 2;
    tc.tw = it.cw;
    tc.th = it.ch;
    tc.ix = tc.x0 + (mv_x >> 5);
    tc.iy = tc.y0 + (mv_y >> 5);
    set_taps<4>(tc, kChroma4[cfx], kChroma4[cfy]);
    tc.hf = cfx != 0;
    tc.vf = cfy != 0;
}

// Weighted average of the two lists' intermediates at the 2 x 2 output
// samples (x, y) .. (x + 1, y + 1) of one plane's tile
// (weighted_average_bi_16, vvc/inter.rs:2676-2696).
template <int TAPS, int PITCH, int BPS>
__device__ __forceinline__ void bi_output_quad(const rmh_plane& dst, const Tile<TAPS>& t0, const Tile<TAPS>& t1,
                                               const int16_t* win0, const int16_t* tmp0, const int16_t* win1,
                                               const int16_t* tmp1, int x, int y, int head_room, int w0, int w1,
                                               int maxv) {
    const int shift1 = 6 - head_room, shift = head_room + 3;
    const int off = (1 << (shift - 1)) + (8192 << 3);
    int p0[4], p1[4], v[4];
    bi_quad<TAPS, PITCH>(t0, win0, tmp0, x, y, head_room, shift1, p0);
    bi_quad<TAPS, PITCH>(t1, win1, tmp1, x, y, head_room, shift1, p1);
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = clampi((p0[q] * w0 + p1[q] * w1 + off) >> shift, 0, maxv);
    store_quad<BPS>(dst, t0.x0 + x, t0.y0 + y, v, x + 1 < t0.tw, y + 1 < t0.th, pairs_aligned<BPS>(dst, t0.x0));
}

// Bi-prediction: as mc_kernel, with the windows of both lists fetched before
// any LDS write (two LDS slices per wave) and the weighted average fused into
// the output quads.
template <int BPS, int NW>
__global__ __launch_bounds__(64 * NW) void mc_bi_kernel(const rmh_mc_bi_frame* __restrict__ frames,
                                                        const McBiItem* __restrict__ items, uint32_t n_items) {
    __shared__ __align__(16) WaveLds lds[NW][2];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t item = xcd_local_block(blockIdx.x, gridDim.x) * NW + wave;
    const bool active = item < n_items;
    WaveLds& L0 = lds[wave][0];
    WaveLds& L1 = lds[wave][1];
    Tile<8> tl0{}, tl1{};
    Tile<4> tc0{}, tc1{};
    const rmh_mc_bi_frame* f = frames;
    int bd = 8, w1 = 4;
    if (active) {
        const McBiItem it = items[item];
        f = &frames[it.frame];
        w1 = it.w1;
        bi_tiles(it, it.mv0_x, it.mv0_y, tl0, tc0);
        bi_tiles(it, it.mv1_x, it.mv1_y, tl1, tc1);
        bd = f->ref0[0].bit_depth;
        const bool hl = tl0.tw > 0 && tl0.th > 0, hc = tc0.tw > 0 && tc0.th > 0;
        Window<kLR, 32> wl0, wl1;
        Window<kCR, 16> wc[4];
        if (hl) {
            fetch<8, kLR, kLP, 32, BPS>(f->ref0[0], tl0, wl0, lane);
            fetch<8, kLR, kLP, 32, BPS>(f->ref1[0], tl1, wl1, lane);
        }
        if (hc) {
            fetch<4, kCR, kCP, 16, BPS>(f->ref0[1], tc0, wc[0], lane);
            fetch<4, kCR, kCP, 16, BPS>(f->ref0[2], tc0, wc[1], lane);
            fetch<4, kCR, kCP, 16, BPS>(f->ref1[1], tc1, wc[2], lane);
            fetch<4, kCR, kCP, 16, BPS>(f->ref1[2], tc1, wc[3], lane);
        }
        if (hl) {
            put<kLR, kLP, 32>(wl0, L0.luma, lane);
            put<kLR, kLP, 32>(wl1, L1.luma, lane);
        }
        if (hc) {
            put<kCR, kCP, 16>(wc[0], L0.chroma[0], lane);
            put<kCR, kCP, 16>(wc[1], L0.chroma[1], lane);
            put<kCR, kCP, 16>(wc[2], L1.chroma[0], lane);
            put<kCR, kCP, 16>(wc[3], L1.chroma[1], lane);
        }
    }
    __syncthreads();
    const int head_room = 14 - bd;
    const int shift1 = 6 - head_room;
    if (active && tl0.tw > 0 && tl0.th > 0) {
        hpass<8, kLP, 8, kLR, 1, 0, 0>(tl0, L0.luma, L0.ltmp, shift1, lane);
        hpass<8, kLP, 8, kLR, 1, 0, 0>(tl1, L1.luma, L1.ltmp, shift1, lane);
    }
    if (active && tc0.tw > 0 && tc0.th > 0) {
        hpass<4, kCP, 4, kCR, 2, kCR * kCP, 8 * kCP>(tc0, L0.chroma[0], L0.ctmp[0], shift1, lane);
        hpass<4, kCP, 4, kCR, 2, kCR * kCP, 8 * kCP>(tc1, L1.chroma[0], L1.ctmp[0], shift1, lane);
    }
    __syncthreads();
    if (!active) return;
    const int maxv = (1 << bd) - 1, w0 = 8 - w1;
    if (tl0.tw > 0 && tl0.th > 0) {
        const int x = 2 * (lane & 7), y = 2 * (lane >> 3);
        if (x < tl0.tw && y < tl0.th)
            bi_output_quad<8, kLP, BPS>(f->dst[0], tl0, tl1, L0.luma, L0.ltmp, L1.luma, L1.ltmp, x, y, head_room, w0,
                                        w1, maxv);
    }
    if (tc0.tw > 0 && tc0.th > 0 && lane < 32) {
        const int pl = lane >> 4, x = 2 * (lane & 3), y = 2 * ((lane >> 2) & 3);
        if (x < tc0.tw && y < tc0.th)
            bi_output_quad<4, kCP, BPS>(f->dst[1 + pl], tc0, tc1, L0.chroma[pl], L0.ctmp[pl], L1.chroma[pl],
                                        L1.ctmp[pl], x, y, head_room, w0, w1, maxv);
    }
}

}  // namespace

namespace {

template <int NW>
void launch_mc_bi_nw(const rmh_mc_bi_frame* frames, const McBiItem* items, uint32_t n_items, uint32_t n_items8,
                     hipStream_t stream) {
    if (n_items8 > 0)
        hipLaunchKernelGGL((mc_bi_kernel<1, NW>), dim3((n_items8 + NW - 1) / NW), dim3(64 * NW), 0, stream, frames,
                           items, n_items8);
    if (n_items > n_items8)
        hipLaunchKernelGGL((mc_bi_kernel<2, NW>), dim3((n_items - n_items8 + NW - 1) / NW), dim3(64 * NW), 0, stream,
                           frames, items + n_items8, n_items - n_items8);
}

template <int NW>
void launch_mc_nw(const rmh_mc_frame* frames, const McItem* items, uint32_t n_items, uint32_t n_items8,
                  hipStream_t stream) {
    if (n_items8 > 0)
        hipLaunchKernelGGL((mc_kernel<1, NW>), dim3((n_items8 + NW - 1) / NW), dim3(64 * NW), 0, stream, frames, items,
                           n_items8);
    if (n_items > n_items8)
        hipLaunchKernelGGL((mc_kernel<2, NW>), dim3((n_items - n_items8 + NW - 1) / NW), dim3(64 * NW), 0, stream,
                           frames, items + n_items8, n_items - n_items8);
}

}  // namespace

hipError_t launch_mc_bi(const rmh_mc_bi_frame* frames, const McBiItem* items, uint32_t n_items, uint32_t n_items8,
                        hipStream_t stream, int waves) {
    if (waves == 1)
        launch_mc_bi_nw<1>(frames, items, n_items, n_items8, stream);
    else if (waves == 2)
        launch_mc_bi_nw<2>(frames, items, n_items, n_items8, stream);
    else
        launch_mc_bi_nw<4>(frames, items, n_items, n_items8, stream);
    return hipGetLastError();
}

hipError_t launch_mc(const rmh_mc_frame* frames, const McItem* items, uint32_t n_items, uint32_t n_items8,
                     hipStream_t stream, int waves) {
    if (waves == 1)
        launch_mc_nw<1>(frames, items, n_items, n_items8, stream);
    else if (waves == 2)
        launch_mc_nw<2>(frames, items, n_items, n_items8, stream);
    else
        launch_mc_nw<4>(frames, items, n_items, n_items8, stream);
    return hipGetLastError();
}

}  // namespace rmh

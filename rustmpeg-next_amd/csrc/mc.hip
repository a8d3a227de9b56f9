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

// Phase 1: stage window rows iy - BEFORE .. (ROWS) and columns ix - BEFORE ..
// (PITCH, the last one only as padding), clamped to the picture, into LDS.
// Lane (c, r0) loads column c of rows r0, r0 + 64 / LANES_X, ...
template <int TAPS, int ROWS, int PITCH, int LANES_X, int BPS>
__device__ __forceinline__ void stage(const rmh_plane& ref, const Tile<TAPS>& t, int16_t* win, int lane) {
    constexpr int BEFORE = TAPS / 2 - 1, RSTEP = 64 / LANES_X;
    const int ox = t.ix - BEFORE, oy = t.iy - BEFORE;
    const int c = lane & (LANES_X - 1), r0 = lane / LANES_X;
    if (c >= PITCH) return;
    const bool inside = ox >= 0 && oy >= 0 && ox + PITCH <= (int)ref.width && oy + ROWS <= (int)ref.height;
    if (inside) {
        const uint32_t base = (uint32_t)(oy + r0) * ref.stride + (uint32_t)(ox + c);
        const uint32_t step = RSTEP * ref.stride;
        if (BPS == 1) {
            const RMH_GLOBAL uint8_t* src = gsrc<uint8_t>(ref.data) + base;
#pragma unroll 4
            for (int r = r0, o = 0; r < ROWS; r += RSTEP, o += step) win[r * PITCH + c] = (int16_t)src[o];
        } else {
            const RMH_GLOBAL uint16_t* src = gsrc<uint16_t>(ref.data) + base;
#pragma unroll 4
            for (int r = r0, o = 0; r < ROWS; r += RSTEP, o += step) win[r * PITCH + c] = (int16_t)src[o];
        }
    } else {
#pragma unroll 4
        for (int r = r0; r < ROWS; r += RSTEP) win[r * PITCH + c] = (int16_t)load_sample<BPS>(ref, ox + c, oy + r);
    }
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
// (interpolate_luma_block 2-D branch, vvc/inter.rs:1004-1030). Lane item e:
// pair xp = e % NPAIR, row e / NPAIR.
template <int TAPS, int PITCH, int NPAIR>
__device__ __forceinline__ void hpass(const Tile<TAPS>& t, const int16_t* win, int16_t* tmp, int shift1, int lane) {
    if (!(t.hf && t.vf)) return;
    const int rows = t.th + TAPS - 1;
    const int off = 8192 << shift1;
    for (int e = lane; e < rows * NPAIR; e += 64) {
        const int xp = e % NPAIR, r = e / NPAIR, x = 2 * xp;
        if (x >= t.tw) continue;
        int s0, s1;
        hpair<TAPS>(win + r * PITCH + x, t.cx, s0, s1);
        tmp[x * PITCH + r] = (int16_t)((s0 - off) >> shift1);
        tmp[(x + 1) * PITCH + r] = (int16_t)((s1 - off) >> shift1);
    }
}

// Phase 3: the output samples of one plane's tile (copy, one-dimensional or
// vertical pass), T = full tile width.
template <int TAPS, int PITCH, int T, int BPS>
__device__ __forceinline__ void output(const rmh_plane& dst, const Tile<TAPS>& t, const int16_t* win, const int16_t* tmp,
                                       int head_room, int maxv, int lane) {
    constexpr int BEFORE = TAPS / 2 - 1;
    if (t.hf && t.vf) {
        // vertical pass over the intermediates, output pair (y, y + 1):
        // (sum + 2^(shift2-1) + (8192 << 6)) >> shift2, shift2 = 6 + head_room
        const int shift2 = 6 + head_room;
        const int off2 = (1 << (shift2 - 1)) + (8192 << 6);
        for (int e = lane; e < T * (T / 2); e += 64) {
            const int x = e % T, y = 2 * (e / T);
            if (x >= t.tw || y >= t.th) continue;
            int s0, s1;
            hpair<TAPS>(tmp + x * PITCH + y, t.cy, s0, s1);
            store_sample<BPS>(dst, t.x0 + x, t.y0 + y, clampi((s0 + off2) >> shift2, 0, maxv));
            if (y + 1 < t.th) store_sample<BPS>(dst, t.x0 + x, t.y0 + y + 1, clampi((s1 + off2) >> shift2, 0, maxv));
        }
    } else if (t.hf) {
        // horizontal only: (sum + 32) >> 6 clipped (vvc/inter.rs:939-970), output pair (x, x + 1)
        for (int e = lane; e < T * (T / 2); e += 64) {
            const int xp = e % (T / 2), y = e / (T / 2), x = 2 * xp;
            if (x >= t.tw || y >= t.th) continue;
            int s0, s1;
            hpair<TAPS>(win + (y + BEFORE) * PITCH + x, t.cx, s0, s1);
            store_sample<BPS>(dst, t.x0 + x, t.y0 + y, clampi((s0 + 32) >> 6, 0, maxv));
            if (x + 1 < t.tw) store_sample<BPS>(dst, t.x0 + x + 1, t.y0 + y, clampi((s1 + 32) >> 6, 0, maxv));
        }
    } else if (t.vf) {
        // vertical only (vvc/inter.rs:971-1002)
        for (int e = lane; e < T * T; e += 64) {
            const int x = e % T, y = e / T;
            if (x >= t.tw || y >= t.th) continue;
            int sum = 0;
#pragma unroll
            for (int k = 0; k < TAPS; k++) sum += t.sy[k] * win[(y + k) * PITCH + x + BEFORE];
            store_sample<BPS>(dst, t.x0 + x, t.y0 + y, clampi((sum + 32) >> 6, 0, maxv));
        }
    } else {
        for (int e = lane; e < T * T; e += 64) {
            const int x = e % T, y = e / T;
            if (x >= t.tw || y >= t.th) continue;
            store_sample<BPS>(dst, t.x0 + x, t.y0 + y, win[(y + BEFORE) * PITCH + x + BEFORE]);
        }
    }
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
// of a PU (smaller for 8-wide or 8-tall PUs) and its two chroma tiles. All
// three reference windows are requested before any filtering, so a wave
// waits for memory once per item. Item fields, frame planes and filter taps
// are wave-uniform and read with scalar loads.
template <int BPS>
__global__ __launch_bounds__(256) void mc_kernel(const rmh_mc_frame* __restrict__ frames,
                                                 const McItem* __restrict__ items, uint32_t n_items) {
    __shared__ __align__(16) WaveLds lds[4];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t item = xcd_local_block(blockIdx.x, gridDim.x) * 4 + wave;
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
        if (tl.tw > 0 && tl.th > 0) stage<8, kLR, kLP, 32, BPS>(f->ref[0], tl, L.luma, lane);
        if (tc.tw > 0 && tc.th > 0) {
            stage<4, kCR, kCP, 16, BPS>(f->ref[1], tc, L.chroma[0], lane);
            stage<4, kCR, kCP, 16, BPS>(f->ref[2], tc, L.chroma[1], lane);
        }
    }
    __syncthreads();
    const int head_room = 14 - bd;
    const int shift1 = 6 - head_room;
    if (active && tl.tw > 0 && tl.th > 0) hpass<8, kLP, 8>(tl, L.luma, L.ltmp, shift1, lane);
    if (active && tc.tw > 0 && tc.th > 0) {
        hpass<4, kCP, 4>(tc, L.chroma[0], L.ctmp[0], shift1, lane);
        hpass<4, kCP, 4>(tc, L.chroma[1], L.ctmp[1], shift1, lane);
    }
    __syncthreads();
    if (!active) return;
    const int maxv = (1 << bd) - 1;
    if (tl.tw > 0 && tl.th > 0) output<8, kLP, 16, BPS>(f->dst[0], tl, L.luma, L.ltmp, head_room, maxv, lane);
    if (tc.tw > 0 && tc.th > 0) {
        output<4, kCP, 8, BPS>(f->dst[1], tc, L.chroma[0], L.ctmp[0], head_room, maxv, lane);
        output<4, kCP, 8, BPS>(f->dst[2], tc, L.chroma[1], L.ctmp[1], head_room, maxv, lane);
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

// Weighted average of the two lists' intermediates over one plane's tile
// (weighted_average_bi_16, vvc/inter.rs:2676-2696), T = full tile width; a
// lane takes 2 x 2 output samples.
template <int TAPS, int PITCH, int T, int BPS>
__device__ __forceinline__ void bi_output(const rmh_plane& dst, const Tile<TAPS>& t0, const Tile<TAPS>& t1,
                                          const int16_t* win0, const int16_t* tmp0, const int16_t* win1,
                                          const int16_t* tmp1, int head_room, int w0, int w1, int maxv, int lane) {
    const int shift1 = 6 - head_room, shift = head_room + 3;
    const int off = (1 << (shift - 1)) + (8192 << 3);
    for (int e = lane; e < (T / 2) * (T / 2); e += 64) {
        const int x = 2 * (e % (T / 2)), y = 2 * (e / (T / 2));
        if (x >= t0.tw || y >= t0.th) continue;
        int p0[4], p1[4];
        bi_quad<TAPS, PITCH>(t0, win0, tmp0, x, y, head_room, shift1, p0);
        bi_quad<TAPS, PITCH>(t1, win1, tmp1, x, y, head_room, shift1, p1);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int xq = x + (q & 1), yq = y + (q >> 1);
            if (xq < t0.tw && yq < t0.th)
                store_sample<BPS>(dst, t0.x0 + xq, t0.y0 + yq, clampi((p0[q] * w0 + p1[q] * w1 + off) >> shift, 0, maxv));
        }
    }
}

// Bi-prediction: as mc_kernel, with the windows of both lists staged before
// any filtering (two LDS slices per wave) and the weighted average fused into
// the output pass.
template <int BPS>
__global__ __launch_bounds__(256) void mc_bi_kernel(const rmh_mc_bi_frame* __restrict__ frames,
                                                    const McBiItem* __restrict__ items, uint32_t n_items) {
    __shared__ __align__(16) WaveLds lds[4][2];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t item = xcd_local_block(blockIdx.x, gridDim.x) * 4 + wave;
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
        if (tl0.tw > 0 && tl0.th > 0) {
            stage<8, kLR, kLP, 32, BPS>(f->ref0[0], tl0, L0.luma, lane);
            stage<8, kLR, kLP, 32, BPS>(f->ref1[0], tl1, L1.luma, lane);
        }
        if (tc0.tw > 0 && tc0.th > 0) {
            stage<4, kCR, kCP, 16, BPS>(f->ref0[1], tc0, L0.chroma[0], lane);
            stage<4, kCR, kCP, 16, BPS>(f->ref0[2], tc0, L0.chroma[1], lane);
            stage<4, kCR, kCP, 16, BPS>(f->ref1[1], tc1, L1.chroma[0], lane);
            stage<4, kCR, kCP, 16, BPS>(f->ref1[2], tc1, L1.chroma[1], lane);
        }
    }
    __syncthreads();
    const int head_room = 14 - bd;
    const int shift1 = 6 - head_room;
    if (active && tl0.tw > 0 && tl0.th > 0) {
        hpass<8, kLP, 8>(tl0, L0.luma, L0.ltmp, shift1, lane);
        hpass<8, kLP, 8>(tl1, L1.luma, L1.ltmp, shift1, lane);
    }
    if (active && tc0.tw > 0 && tc0.th > 0) {
        hpass<4, kCP, 4>(tc0, L0.chroma[0], L0.ctmp[0], shift1, lane);
        hpass<4, kCP, 4>(tc0, L0.chroma[1], L0.ctmp[1], shift1, lane);
        hpass<4, kCP, 4>(tc1, L1.chroma[0], L1.ctmp[0], shift1, lane);
        hpass<4, kCP, 4>(tc1, L1.chroma[1], L1.ctmp[1], shift1, lane);
    }
    __syncthreads();
    if (!active) return;
    const int maxv = (1 << bd) - 1, w0 = 8 - w1;
    if (tl0.tw > 0 && tl0.th > 0)
        bi_output<8, kLP, 16, BPS>(f->dst[0], tl0, tl1, L0.luma, L0.ltmp, L1.luma, L1.ltmp, head_room, w0, w1, maxv, lane);
    if (tc0.tw > 0 && tc0.th > 0) {
        for (int c = 0; c < 2; c++)
            bi_output<4, kCP, 8, BPS>(f->dst[1 + c], tc0, tc1, L0.chroma[c], L0.ctmp[c], L1.chroma[c], L1.ctmp[c], head_room,
                                 w0, w1, maxv, lane);
    }
}

}  // namespace

hipError_t launch_mc_bi(const rmh_mc_bi_frame* frames, const McBiItem* items, uint32_t n_items, uint32_t n_items8,
                        hipStream_t stream) {
    if (n_items8 > 0)
        hipLaunchKernelGGL(mc_bi_kernel<1>, dim3((n_items8 + 3) / 4), dim3(256), 0, stream, frames, items, n_items8);
    if (n_items > n_items8)
        hipLaunchKernelGGL(mc_bi_kernel<2>, dim3((n_items - n_items8 + 3) / 4), dim3(256), 0, stream, frames,
                           items + n_items8, n_items - n_items8);
    return hipGetLastError();
}

hipError_t launch_mc(const rmh_mc_frame* frames, const McItem* items, uint32_t n_items, uint32_t n_items8,
                     hipStream_t stream) {
    if (n_items8 > 0)
        hipLaunchKernelGGL(mc_kernel<1>, dim3((n_items8 + 3) / 4), dim3(256), 0, stream, frames, items, n_items8);
    if (n_items > n_items8)
        hipLaunchKernelGGL(mc_kernel<2>, dim3((n_items - n_items8 + 3) / 4), dim3(256), 0, stream, frames,
                           items + n_items8, n_items - n_items8);
    return hipGetLastError();
}

}  // namespace rmh

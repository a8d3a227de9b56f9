// sao.hip - sample adaptive offset over every plane of a batch of pictures.
//
// apply_sao (crates/rustmpeg-video/src/vvc/decode_general.rs:5189-5222) runs
// apply_sao_region_bounded_16 (vvc/filter.rs:1954-2043) on every CTB of the
// visible picture, reading neighbours from an unmodified copy. Here the copy
// is the source plane itself and the output is a separate plane: every
// sample of dst is written once (offset applied where the CTB's SAO type and
// the visible-boundary rule say so, copied otherwise), so a plane costs one
// read and one write of HBM; the +-1 neighbour rows and columns come from L2.
//
// One workgroup covers one 256-lane segment of kSaoRows rows of one plane
// (found through a per-workgroup job table); a lane covers 8 bytes of each
// row (4 samples of 16-bit storage, 8 of 8-bit). Rows, CTB row and neighbour
// rows are uniform over the workgroup.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gmem.h"
#include "kernels.h"

namespace rmh {
namespace {

__device__ __forceinline__ int sgn(int v) { return (v > 0) - (v < 0); }

// Samples x0 - 1 .. x0 + K of one row into w[0 .. K + 1]: the lane's own
// K samples with one access, the two neighbours from the adjacent lanes
// (which hold x0 - K .. x0 - 1 and x0 + K .. x0 + 2K - 1), read from memory
// only at the wave's two ends. Samples outside the plane read as 0 (the edge
// rule never uses them). Every lane of the wave must call this together.
template <typename T, int K>
__device__ __forceinline__ void load_window(const RMH_GLOBAL T* row, int x0, int width, bool vec, bool live,
                                            int (&w)[K + 2]) {
    if (live && vec && x0 + K <= width) {
        T v[K];
        *reinterpret_cast<u32x2*>(v) = *reinterpret_cast<const RMH_GLOBAL u32x2*>(row + x0);
#pragma unroll
        for (int k = 0; k < K; k++) w[k + 1] = v[k];
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) w[k + 1] = live && x0 + k < width ? (int)row[x0 + k] : 0;
    }
    const int lane = threadIdx.x & 63;
    const int left = __shfl_up(w[K], 1, 64), right = __shfl_down(w[1], 1, 64);
    w[0] = lane > 0 ? left : (live && x0 > 0 ? (int)row[x0 - 1] : 0);
    w[K + 1] = lane < 63 ? right : (live && x0 + K < width ? (int)row[x0 + K] : 0);
    if (x0 + K >= width) w[K + 1] = 0;
}

// Workgroup b (of the blocks from block0) runs kSaoRows rows of one 256-lane
// segment of the plane of job block_job[b]; every row of the group lies in
// one CTB row (CTBs are at least 8 rows of a component, kSaoRows divides 8).
template <typename T>
__global__ __launch_bounds__(256) void sao_kernel(const SaoJob* __restrict__ jobs, const uint32_t* __restrict__ block_job,
                                                  uint32_t block0) {
    constexpr int K = 8 / (int)sizeof(T);
    constexpr int R = kSaoRows;
    const uint32_t b = blockIdx.x + block0;
    const SaoJob& J = jobs[block_job[b]];
    const uint32_t rel = b - J.first_block;
    const int yb = (int)(rel / J.blocks_per_row);
    const int x0 = (int)((rel - (uint32_t)yb * J.blocks_per_row) * 256u + threadIdx.x) * K;
    const int width = (int)J.width, height = (int)J.height;
    const bool live = x0 < width;  // lanes past the row only take part in the exchanges
    const RMH_GLOBAL T* src = gsrc<T>(J.src);
    RMH_GLOBAL T* dst = gdst<T>(J.dst);
    const size_t pitch = J.stride;
    const bool vec = J.vec != 0;
    const int y0 = yb * R;
    const int rows = min(R, height - y0);
    const int vis_w = (int)J.vis_w, vis_h = (int)J.vis_h;

    // the lane's CTB parameters (rmh_sao_param as 5 dwords: type |
    // band_or_class << 8, offsets[4])
    int type = RMH_SAO_OFF, boc = 0, o0 = 0, o1 = 0, o2 = 0, o3 = 0;
    if (live && y0 < vis_h && x0 < vis_w) {
        const int cl = (int)J.ctb_log2;
        const RMH_GLOBAL int32_t* pw5 =
            gsrc<int32_t>(J.params) + (((size_t)(y0 >> cl) * J.ctb_cols + (size_t)(x0 >> cl)) * 3 + J.comp) * 5;
        const int32_t head = pw5[0];
        type = head & 255;
        boc = (head >> 8) & 255;
        o0 = pw5[1];
        o1 = pw5[2];
        o2 = pw5[3];
        o3 = pw5[4];
    }
    const int cls = boc & 3;
    const bool edge = type == RMH_SAO_EDGE;
    const int dx = cls == 1 ? 0 : (cls == 3 ? -1 : 1), dy = (edge && cls != 0) ? 1 : 0;
    const int maxv = (1 << J.bd) - 1;
    const int bdc = J.bd < 8 ? 8 : (J.bd > 12 ? 12 : (int)J.bd);

    // rows y0 - 1 .. y0 + R: the outer two only where a lane of the wave has
    // a vertical or diagonal edge class
    int win[R + 2][K + 2];
    const bool outer = __builtin_amdgcn_ballot_w64(live && dy != 0) != 0;
#pragma unroll
    for (int r = 0; r < R + 2; r++) {
        const int y = y0 - 1 + r;
        const bool need = (r >= 1 && r <= R) || outer;
        if (need && y >= 0 && y < height) {
            load_window<T, K>(src + (size_t)y * pitch, x0, width, vec, live, win[r]);
        } else {
#pragma unroll
            for (int k = 0; k < K + 2; k++) win[r][k] = 0;
        }
    }
    if (!live) return;

    const int pwid = min(vis_w, width), ph = min(vis_h, height);
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r >= rows) break;
        const int y = y0 + r;
        const int(&w0)[K + 2] = win[r + 1];
        int out[K];
#pragma unroll
        for (int k = 0; k < K; k++) out[k] = w0[k + 1];
        if (y < vis_h) {
            if (type == RMH_SAO_BAND) {
                const int shift = bdc - 5;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int sv = w0[k + 1];
                    const int band = ((sv >> shift) - boc) & 31;
                    if (x0 + k < vis_w && band < 4) {
                        const int off = band == 0 ? o0 : band == 1 ? o1 : band == 2 ? o2 : o3;
                        out[k] = min(max(sv + off, 0), maxv);
                    }
                }
            } else if (edge && y - dy >= 0 && y + dy < ph) {
                // rows y - dy and y + dy inside the visible picture; the
                // neighbours are selected among constant-index registers
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int x = x0 + k;
                    const bool ok = x < vis_w && x - dx >= 0 && x - dx < pwid && x + dx >= 0 && x + dx < pwid;
                    const int c = w0[k + 1];
                    // a = (x - dx, y - dy), b = (x + dx, y + dy)
                    const int am = dy ? win[r][k] : win[r + 1][k], a0 = dy ? win[r][k + 1] : win[r + 1][k + 1],
                              ap = dy ? win[r][k + 2] : win[r + 1][k + 2];
                    const int bm = dy ? win[r + 2][k] : win[r + 1][k], b0 = dy ? win[r + 2][k + 1] : win[r + 1][k + 1],
                              bp = dy ? win[r + 2][k + 2] : win[r + 1][k + 2];
                    const int ea = dx == 1 ? am : (dx == 0 ? a0 : ap), eb = dx == 1 ? bp : (dx == 0 ? b0 : bm);
                    const int e = sgn(c - ea) + sgn(c - eb);  // edge_idx - 2
                    const int off = e == -2 ? o0 : e == -1 ? o1 : e == 1 ? o2 : e == 2 ? o3 : 0;
                    if (ok) out[k] = min(max(c + off, 0), maxv);
                }
            }
        }
        RMH_GLOBAL T* drow = dst + (size_t)y * pitch;
        if (vec && x0 + K <= width) {
            T v[K];
#pragma unroll
            for (int k = 0; k < K; k++) v[k] = (T)out[k];
            *reinterpret_cast<RMH_GLOBAL u32x2*>(drow + x0) = *reinterpret_cast<const u32x2*>(v);
        } else {
#pragma unroll
            for (int k = 0; k < K; k++)
                if (x0 + k < width) drow[x0 + k] = (T)out[k];
        }
    }
}

// ---------------------------------------------------------------------------
// CTB-uniform form, for planes whose CTBs are at least 16 samples on a side:
// a wave covers an LA*4 x (64 / LA) block of one CTB (LA = min(16, CTB / 4)
// lanes across, 4 samples each), so the SAO type, class and offsets are
// uniform over the wave (scalar loads and branches) and every edge class runs
// its own straight-line code. Neighbour samples come from the lanes beside,
// above and below by shuffles, or from memory at the block's edges. Offsets
// that fit int8 are looked up with one byte permute.
// ---------------------------------------------------------------------------

// The lane's 4 samples of row y (zeros outside the plane) into v[0..3].
template <typename T>
__device__ __forceinline__ void sao_load4(const RMH_GLOBAL T* plane, size_t pitch, int x0, int y, int width, int height,
                                          bool vec, int (&v)[4]) {
    if (y < 0 || y >= height || x0 >= width) {
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = 0;
        return;
    }
    const RMH_GLOBAL T* row = plane + (size_t)y * pitch;
    if (vec && x0 + 4 <= width) {
        if constexpr (sizeof(T) == 2) {
            const u32x2 d = *reinterpret_cast<const RMH_GLOBAL u32x2*>(row + x0);
            v[0] = (int)(d.x & 0xffff);
            v[1] = (int)(d.x >> 16);
            v[2] = (int)(d.y & 0xffff);
            v[3] = (int)(d.y >> 16);
        } else {
            const uint32_t d = *reinterpret_cast<const RMH_GLOBAL uint32_t*>(row + x0);
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = (int)((d >> (8 * k)) & 255);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = x0 + k < width ? (int)row[x0 + k] : 0;
    }
}

// w[0..5] = samples x0 - 1 .. x0 + 4 of row y: the lane's own four, the two
// beside it from the neighbouring lanes of the block row (lx - 1, lx + 1),
// or from memory at the block's left and right edges. Called by every lane
// of the wave together.
template <typename T, int LA>
__device__ __forceinline__ void sao_window(const RMH_GLOBAL T* plane, size_t pitch, int x0, int y, int width,
                                           int height, bool vec, int lx, int (&w)[6]) {
    int v[4];
    sao_load4<T>(plane, pitch, x0, y, width, height, vec, v);
#pragma unroll
    for (int k = 0; k < 4; k++) w[k + 1] = v[k];
    const int left = __shfl_up(v[3], 1, 64), right = __shfl_down(v[0], 1, 64);
    const bool row_ok = y >= 0 && y < height;
    const RMH_GLOBAL T* row = plane + (size_t)(row_ok ? y : 0) * pitch;
    w[0] = lx > 0 ? left : (row_ok && x0 > 0 && x0 - 1 < width ? (int)row[x0 - 1] : 0);
    w[5] = lx < LA - 1 ? right : (row_ok && x0 + 4 < width ? (int)row[x0 + 4] : 0);
    if (x0 + 4 >= width) w[5] = 0;
}

__device__ __forceinline__ int sext8(uint32_t v) { return (int)(int8_t)(uint8_t)(v & 255); }

// One edge class over the lane's 4 samples; w0 is the window of row y, the
// windows of rows y - 1 and y + 1 are loaded here for the classes that need
// them (every lane of the wave takes the same class).
template <int DX, int DY, typename T, int LA>
__device__ __forceinline__ void sao_edge(const RMH_GLOBAL T* src, size_t pitch, int x0, int y, int width, int height,
                                         bool vec, int lx, const int (&w0)[6], bool interior, int vis_w, int pwid,
                                         int ph, bool small, uint32_t lut_lo, uint32_t lut_hi, const int (&o)[4],
                                         int maxv, int (&out)[4]) {
    int wa[6], wb[6];
    if constexpr (DY == 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) wa[k] = wb[k] = w0[k];
    } else {
        sao_window<T, LA>(src, pitch, x0, y - DY, width, height, vec, lx, wa);
        sao_window<T, LA>(src, pitch, x0, y + DY, width, height, vec, lx, wb);
    }
    const bool rows_ok = y - DY >= 0 && y + DY < ph;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int c = w0[k + 1], ea = wa[k + 1 - DX], eb = wb[k + 1 + DX];
        const int e = min(max(c - ea, -1), 1) + min(max(c - eb, -1), 1);  // edge_idx - 2
        int off;
        if (small) {
            off = sext8(__builtin_amdgcn_perm(lut_hi, lut_lo, (uint32_t)(e + 2) | 0x0c0c0c00u));
        } else {
            off = e == -2 ? o[0] : e == -1 ? o[1] : e == 1 ? o[2] : e == 2 ? o[3] : 0;
        }
        const int x = x0 + k;
        const bool ok = interior || (rows_ok && x < vis_w && x - DX >= 0 && x - DX < pwid && x + DX >= 0 && x + DX < pwid);
        if (ok) out[k] = min(max(c + off, 0), maxv);
    }
}

template <typename T, int LA>
__device__ __forceinline__ void sao_ctb_block(const SaoJob& J, uint32_t tile) {
    constexpr int TH = 64 / LA, TW = 4 * LA;
    const int lane = threadIdx.x & 63, lx = lane % LA, ly = lane / LA;
    // tile = ty * tiles_x + tx without an integer division (the scalar unit has
    // none): a double-precision estimate, exact after one correction
    const uint32_t tiles_x = J.tiles_x;
    uint32_t tyu = (uint32_t)((double)tile * J.inv_tiles_x);
    int r = (int)(tile - tyu * tiles_x);
    if (r < 0) {
        tyu--;
        r += (int)tiles_x;
    } else if (r >= (int)tiles_x) {
        tyu++;
        r -= (int)tiles_x;
    }
    const int tx = __builtin_amdgcn_readfirstlane(r), ty = __builtin_amdgcn_readfirstlane((int)tyu);
    const int X0 = tx * TW, Y0 = ty * TH;
    const int x0 = X0 + 4 * lx, y = Y0 + ly;
    const int width = (int)J.width, height = (int)J.height;
    const RMH_GLOBAL T* src = gsrc<T>(J.src);
    RMH_GLOBAL T* dst = gdst<T>(J.dst);
    const size_t pitch = J.stride;
    const bool vec = J.vec != 0;
    const int vis_w = (int)J.vis_w, vis_h = (int)J.vis_h;

    int w0[6];
    sao_window<T, LA>(src, pitch, x0, y, width, height, vec, lx, w0);
    int out[4];
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = w0[k + 1];

    if (X0 < vis_w && Y0 < vis_h) {
        // the block's CTB parameters, uniform over the wave
        const int cl = (int)J.ctb_log2;
        const RMH_GLOBAL int32_t* pw5 =
            gsrc<int32_t>(J.params) + (((size_t)(Y0 >> cl) * J.ctb_cols + (size_t)(X0 >> cl)) * 3 + J.comp) * 5;
        const int head = __builtin_amdgcn_readfirstlane(pw5[0]);
        const int type = head & 255, boc = (head >> 8) & 255;
        int o[4];
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = __builtin_amdgcn_readfirstlane(pw5[1 + k]);
        const bool small = o[0] >= -128 && o[0] <= 127 && o[1] >= -128 && o[1] <= 127 && o[2] >= -128 &&
                           o[2] <= 127 && o[3] >= -128 && o[3] <= 127;
        const int maxv = (1 << J.bd) - 1;
        const bool in_vis = X0 + TW <= vis_w && Y0 + TH <= vis_h;
        if (type == RMH_SAO_BAND) {
            const int bd = J.bd < 8 ? 8 : (J.bd > 12 ? 12 : (int)J.bd);
            const int shift = bd - 5;
            // bytes 0-3: offsets of the four bands; byte 4 (lut_hi): 0
            const uint32_t lut_lo = (uint32_t)(o[0] & 255) | (uint32_t)(o[1] & 255) << 8 | (uint32_t)(o[2] & 255) << 16 |
                                    (uint32_t)(o[3] & 255) << 24;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int sv = w0[k + 1];
                const int band = ((sv >> shift) - boc) & 31;
                const int sel = min(band, 4);
                int off;
                if (small)
                    off = sext8(__builtin_amdgcn_perm(0u, lut_lo, (uint32_t)sel | 0x0c0c0c00u));
                else
                    off = sel == 0 ? o[0] : sel == 1 ? o[1] : sel == 2 ? o[2] : sel == 3 ? o[3] : 0;
                if ((in_vis || (x0 + k < vis_w && y < vis_h)) && band < 4) out[k] = min(max(sv + off, 0), maxv);
            }
        } else if (type == RMH_SAO_EDGE) {
            const int cls = boc & 3;
            const int pwid = min(vis_w, width), ph = min(vis_h, height);
            // edge categories 1..4 (edge_idx 0, 1, 3, 4) as bytes 0, 1, 3, 4; byte 2: 0
            const uint32_t lut_lo = (uint32_t)(o[0] & 255) | (uint32_t)(o[1] & 255) << 8 | (uint32_t)(o[2] & 255) << 24;
            const uint32_t lut_hi = (uint32_t)(o[3] & 255);
            // every sample of the block and both neighbours inside the visible picture
            const int dxa = cls == 1 ? 0 : 1, dya = cls == 0 ? 0 : 1;
            const bool interior = in_vis && X0 - dxa >= 0 && X0 + TW + dxa <= pwid && Y0 - dya >= 0 &&
                                  Y0 + TH + dya <= ph;
            const bool live_y = y < vis_h;
            if (cls == 0)
                sao_edge<1, 0, T, LA>(src, pitch, x0, y, width, height, vec, lx, w0, interior, live_y ? vis_w : 0,
                                      pwid, ph, small, lut_lo, lut_hi, o, maxv, out);
            else if (cls == 1)
                sao_edge<0, 1, T, LA>(src, pitch, x0, y, width, height, vec, lx, w0, interior, live_y ? vis_w : 0,
                                      pwid, ph, small, lut_lo, lut_hi, o, maxv, out);
            else if (cls == 2)
                sao_edge<1, 1, T, LA>(src, pitch, x0, y, width, height, vec, lx, w0, interior, live_y ? vis_w : 0,
                                      pwid, ph, small, lut_lo, lut_hi, o, maxv, out);
            else
                sao_edge<-1, 1, T, LA>(src, pitch, x0, y, width, height, vec, lx, w0, interior, live_y ? vis_w : 0,
                                       pwid, ph, small, lut_lo, lut_hi, o, maxv, out);
        }
    }

    if (y >= height || x0 >= width) return;
    RMH_GLOBAL T* drow = dst + (size_t)y * pitch;
    if (vec && x0 + 4 <= width) {
        if constexpr (sizeof(T) == 2) {
            *reinterpret_cast<RMH_GLOBAL u32x2*>(drow + x0) =
                u32x2{(uint32_t)out[0] | (uint32_t)out[1] << 16, (uint32_t)out[2] | (uint32_t)out[3] << 16};
        } else {
            *reinterpret_cast<RMH_GLOBAL uint32_t*>(drow + x0) =
                (uint32_t)out[0] | (uint32_t)out[1] << 8 | (uint32_t)out[2] << 16 | (uint32_t)out[3] << 24;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (x0 + k < width) drow[x0 + k] = (T)out[k];
    }
}

// Workgroup b: wave w runs blocks kSaoBlocksPerWave * (4 (b - first_block)
// + w) + i of the job block_job[b].
template <typename T>
__global__ __launch_bounds__(256) void sao_ctb_kernel(const SaoJob* __restrict__ jobs,
                                                      const uint32_t* __restrict__ block_job, uint32_t block0) {
    const uint32_t b = blockIdx.x + block0;
    const SaoJob& J = jobs[block_job[b]];
    const uint32_t first =
        ((b - J.first_block) * 4 + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) * kSaoBlocksPerWave;
#pragma unroll
    for (int i = 0; i < kSaoBlocksPerWave; i++) {
        const uint32_t tile = first + i;
        if (tile >= J.tiles) return;  // the whole wave
        switch (J.lanes_across) {
            case 4: sao_ctb_block<T, 4>(J, tile); break;
            case 8: sao_ctb_block<T, 8>(J, tile); break;
            default: sao_ctb_block<T, 16>(J, tile); break;
        }
    }
}

}  // namespace

hipError_t launch_sao(const SaoJob* jobs, const uint32_t* block_job, const uint32_t (&range)[5], hipStream_t stream) {
    // block ranges: [0] row form 8-bit, [1] CTB form 8-bit, [2] row form 16-bit, [3] CTB form 16-bit
    if (range[1] > range[0])
        hipLaunchKernelGGL(sao_kernel<uint8_t>, dim3(range[1] - range[0]), dim3(256), 0, stream, jobs, block_job, range[0]);
    if (range[2] > range[1])
        hipLaunchKernelGGL(sao_ctb_kernel<uint8_t>, dim3(range[2] - range[1]), dim3(256), 0, stream, jobs, block_job,
                           range[1]);
    if (range[3] > range[2])
        hipLaunchKernelGGL(sao_kernel<uint16_t>, dim3(range[3] - range[2]), dim3(256), 0, stream, jobs, block_job,
                           range[2]);
    if (range[4] > range[3])
        hipLaunchKernelGGL(sao_ctb_kernel<uint16_t>, dim3(range[4] - range[3]), dim3(256), 0, stream, jobs, block_job,
                           range[3]);
    return hipGetLastError();
}

}  // namespace rmh

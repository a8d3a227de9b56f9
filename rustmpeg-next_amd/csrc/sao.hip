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
        const int cs = (int)J.ctb;
        const RMH_GLOBAL int32_t* pw5 =
            gsrc<int32_t>(J.params) + (((size_t)(y0 / cs) * J.ctb_cols + (size_t)(x0 / cs)) * 3 + J.comp) * 5;
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

}  // namespace

hipError_t launch_sao(const SaoJob* jobs, const uint32_t* block_job, uint32_t blocks8, uint32_t blocks,
                      hipStream_t stream) {
    if (blocks8 > 0)
        hipLaunchKernelGGL(sao_kernel<uint8_t>, dim3(blocks8), dim3(256), 0, stream, jobs, block_job, 0u);
    if (blocks > blocks8)
        hipLaunchKernelGGL(sao_kernel<uint16_t>, dim3(blocks - blocks8), dim3(256), 0, stream, jobs, block_job, blocks8);
    return hipGetLastError();
}

}  // namespace rmh

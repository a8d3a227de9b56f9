// deblock.hip - luma and chroma deblocking of whole pictures for gfx950.
//
// Semantics follow deblock_luma_edge_bd / deblock_chroma_edge
// (crates/rustmpeg-video/src/vvc/filter.rs:184-1057) driven in the order of
// apply_deblock (vvc/decode_general.rs:4977-5187): every vertical edge of a
// picture, then every horizontal edge.
//
// Mapping: one lane per 4-sample edge segment (luma) or two-line decision
// group (chroma); one launch for all vertical edges of all pictures, a second
// for all horizontal edges. Inside one direction the reference's sequential
// order is equivalent to edge-parallel execution: with the VVC filter lengths
// a segment writes at most maxFilterLength samples per side and reads at most
// one more, so no segment reads what another segment of the same direction
// writes. Each lane therefore writes back only the samples its filter may
// modify.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace rmh {
namespace {

__constant__ int16_t kTc[66] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,   0,   0,   0,   0,   0,   0,   0,
                                0,  3,  4,  4,  4,  4,  5,  5,  5,  5,   7,   7,   8,   9,   10,  10,  11,
                                13, 14, 15, 17, 19, 21, 24, 25, 29, 33,  36,  41,  45,  51,  57,  64,  71,
                                80, 89, 100, 112, 125, 141, 157, 177, 198, 222, 250, 280, 314, 352, 395};
__constant__ int16_t kBeta[64] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,
                                  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 22, 24,
                                  26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56,
                                  58, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78, 80, 82, 84, 86, 88};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }

__device__ __forceinline__ int ld(const rmh_plane& p, int x, int y) {
    size_t o = (size_t)y * p.stride + x;
    return p.bytes_per_sample == 1 ? (int)((const uint8_t*)p.data)[o] : (int)((const uint16_t*)p.data)[o];
}
__device__ __forceinline__ void st(const rmh_plane& p, int x, int y, int v) {
    size_t o = (size_t)y * p.stride + x;
    if (p.bytes_per_sample == 1)
        ((uint8_t*)p.data)[o] = (uint8_t)v;
    else
        ((uint16_t*)p.data)[o] = (uint16_t)v;
}

__device__ __forceinline__ void tc_beta(int qp, int bs, int beta_off, int tc_off, int bd, int& tc, int& beta) {
    int it = clampi(qp + 2 * (bs - 1) + 2 * tc_off, 0, 65);
    int ib = clampi(qp + 2 * beta_off, 0, 63);
    tc = bd < 10 ? (kTc[it] + (1 << (9 - bd))) >> (10 - bd) : kTc[it] << (bd - 10);
    beta = kBeta[ib] << (bd - 8);
}

// luma_sample_decision (vvc/filter.rs:276-304)
__device__ __forceinline__ bool dsam(int p0, int p3, int q0, int q3, int dpq, int sp, int sq, int spq, bool pl, bool ql,
                                     int beta, int tc) {
    if (pl) sp = (sp + iabs(p3 - p0) + 1) >> 1;
    if (ql) sq = (sq + iabs(q3 - q0) + 1) >> 1;
    int t1 = (pl || ql) ? (3 * beta) >> 5 : beta >> 3;
    int t2 = (pl || ql) ? beta >> 4 : beta >> 2;
    return dpq < t2 && sp + sq < t1 && spq < ((5 * tc + 1) >> 1);
}

#define PP(l, i) (l)[7 - (i)]
#define QQ(l, j) (l)[8 + (j)]

// Filters one luma segment held as L[4][16] = p7..p0 q0..q7 in registers.
// Returns the number of P and Q samples that may have changed (0 = none).
__device__ __forceinline__ void luma_segment(int (*L)[16], int qp, int bs, int beta_off, int tc_off, int bd, int len_p, int len_q,
                             bool at_ctb, int& np_out, int& nq_out) {
    np_out = nq_out = 0;
    const int maxv = (1 << bd) - 1;
    int tc, beta;
    tc_beta(qp, bs, beta_off, tc_off, bd, tc, beta);
    if (beta == 0 || tc == 0) return;
    const int* l0 = L[0];
    const int* l3 = L[3];
    int dp0 = iabs(PP(l0, 2) - 2 * PP(l0, 1) + PP(l0, 0)), dp3 = iabs(PP(l3, 2) - 2 * PP(l3, 1) + PP(l3, 0));
    int dq0 = iabs(QQ(l0, 2) - 2 * QQ(l0, 1) + QQ(l0, 0)), dq3 = iabs(QQ(l3, 2) - 2 * QQ(l3, 1) + QQ(l3, 0));
    int dpq0 = dp0 + dq0, dpq3 = dp3 + dq3;
    int sp0 = iabs(PP(l0, 3) - PP(l0, 0)), sp3 = iabs(PP(l3, 3) - PP(l3, 0));
    int sq0 = iabs(QQ(l0, 0) - QQ(l0, 3)), sq3 = iabs(QQ(l3, 0) - QQ(l3, 3));
    int spq0 = iabs(PP(l0, 0) - QQ(l0, 0)), spq3 = iabs(PP(l3, 0) - QQ(l3, 0));
    int lp = clampi(len_p, 1, 7), lq = clampi(len_q, 1, 7);
    bool pl = lp > 3 && !at_ctb, ql = lq > 3;
    int de = 0;
    bool dep = false, deq = false;
    if (pl || ql) {
        int dp0l = pl ? (dp0 + iabs(PP(l0, 5) - 2 * PP(l0, 4) + PP(l0, 3)) + 1) >> 1 : dp0;
        int dp3l = pl ? (dp3 + iabs(PP(l3, 5) - 2 * PP(l3, 4) + PP(l3, 3)) + 1) >> 1 : dp3;
        int dq0l = ql ? (dq0 + iabs(QQ(l0, 5) - 2 * QQ(l0, 4) + QQ(l0, 3)) + 1) >> 1 : dq0;
        int dq3l = ql ? (dq3 + iabs(QQ(l3, 5) - 2 * QQ(l3, 4) + QQ(l3, 3)) + 1) >> 1 : dq3;
        if (!pl) lp = 3;
        if (!ql) lq = 3;
        int sp0l = lp == 7 ? sp0 + iabs(PP(l0, 7) - PP(l0, 6) - PP(l0, 5) + PP(l0, 4)) : sp0;
        int sp3l = lp == 7 ? sp3 + iabs(PP(l3, 7) - PP(l3, 6) - PP(l3, 5) + PP(l3, 4)) : sp3;
        int sq0l = lq == 7 ? sq0 + iabs(QQ(l0, 4) - QQ(l0, 5) - QQ(l0, 6) + QQ(l0, 7)) : sq0;
        int sq3l = lq == 7 ? sq3 + iabs(QQ(l3, 4) - QQ(l3, 5) - QQ(l3, 6) + QQ(l3, 7)) : sq3;
        int d0l = dp0l + dq0l, d3l = dp3l + dq3l;
        if (d0l + d3l < beta) {
            bool s0 = dsam(pl ? PP(l0, lp) : 0, pl ? PP(l0, 3) : 0, ql ? QQ(l0, lq) : 0, ql ? QQ(l0, 3) : 0, 2 * d0l,
                           sp0l, sq0l, spq0, pl, ql, beta, tc);
            bool s3 = dsam(pl ? PP(l3, lp) : 0, pl ? PP(l3, 3) : 0, ql ? QQ(l3, lq) : 0, ql ? QQ(l3, 3) : 0, 2 * d3l,
                           sp3l, sq3l, spq3, pl, ql, beta, tc);
            if (s0 && s3) de = 3;
        }
    }
    if (de != 3) {
        if (dpq0 + dpq3 >= beta) return;
        bool s0 = false, s3 = false;
        if (lp > 2 && lq > 2) {
            s0 = dsam(0, 0, 0, 0, 2 * dpq0, sp0, sq0, spq0, false, false, beta, tc);
            s3 = dsam(0, 0, 0, 0, 2 * dpq3, sp3, sq3, spq3, false, false, beta, tc);
        }
        de = 1;
        if (s0 && s3) {
            de = 2;
            lp = 3;
            lq = 3;
        }
        int thr = (beta + (beta >> 1)) >> 3;
        dep = lp > 1 && lq > 1 && (dp0 + dp3) < thr;
        deq = lp > 1 && lq > 1 && (dq0 + dq3) < thr;
    }
#pragma unroll
    for (int li = 0; li < 4; li++) {
        int* line = L[li];
        int m[16];
#pragma unroll
        for (int k = 0; k < 16; k++) m[k] = line[k];
        if (de == 3) {
            // long_tap_luma_line (vvc/filter.rs:477-582)
            const int np = lp, nq = lq;
            int rm;
            if (np == 5 && nq == 5)
                rm = (2 * (PP(m, 0) + QQ(m, 0) + PP(m, 1) + QQ(m, 1) + PP(m, 2) + QQ(m, 2)) + PP(m, 3) + QQ(m, 3) +
                      PP(m, 4) + QQ(m, 4) + 8) >> 4;
            else if ((np == 7 && nq == 5) || (np == 5 && nq == 7))
                rm = (2 * (PP(m, 0) + QQ(m, 0) + PP(m, 1) + QQ(m, 1)) + PP(m, 2) + QQ(m, 2) + PP(m, 3) + QQ(m, 3) +
                      PP(m, 4) + QQ(m, 4) + PP(m, 5) + QQ(m, 5) + 8) >> 4;
            else if ((np == 5 && nq == 3) || (np == 3 && nq == 5))
                rm = (PP(m, 0) + QQ(m, 0) + PP(m, 1) + QQ(m, 1) + PP(m, 2) + QQ(m, 2) + PP(m, 3) + QQ(m, 3) + 4) >> 3;
            else if (np == 7 && nq == 3)
                rm = (2 * (PP(m, 0) + QQ(m, 0)) + QQ(m, 0) + 2 * (QQ(m, 1) + QQ(m, 2)) + PP(m, 1) + QQ(m, 1) +
                      PP(m, 2) + PP(m, 3) + PP(m, 4) + PP(m, 5) + PP(m, 6) + 8) >> 4;
            else if (np == 3 && nq == 7)
                rm = (2 * (PP(m, 0) + QQ(m, 0)) + PP(m, 0) + 2 * (PP(m, 1) + PP(m, 2)) + PP(m, 1) + QQ(m, 1) +
                      QQ(m, 2) + QQ(m, 3) + QQ(m, 4) + QQ(m, 5) + QQ(m, 6) + 8) >> 4;
            else
                rm = (2 * (PP(m, 0) + QQ(m, 0)) + PP(m, 1) + QQ(m, 1) + PP(m, 2) + QQ(m, 2) + PP(m, 3) + QQ(m, 3) +
                      PP(m, 4) + QQ(m, 4) + PP(m, 5) + QQ(m, 5) + PP(m, 6) + QQ(m, 6) + 8) >> 4;
            int refp = (PP(m, np) + PP(m, np - 1) + 1) >> 1;
            int refq = (QQ(m, nq) + QQ(m, nq - 1) + 1) >> 1;
#pragma unroll
            for (int pos = 0; pos < 7; pos++) {
                if (pos < np) {
                    int f = np == 7 ? 59 - 9 * pos : np == 5 ? (pos == 0 ? 58 : pos == 1 ? 45 : pos == 2 ? 32 : pos == 3 ? 19 : 6)
                                                             : (pos == 0 ? 53 : pos == 1 ? 32 : 11);
                    int t = np == 7 ? (pos < 5 ? 6 - pos : 1) : np == 5 ? 6 - pos : 6 - 2 * pos;
                    int o = PP(m, pos), cv = (tc * t) >> 1;
                    int v = (rm * f + refp * (64 - f) + 32) >> 6;
                    line[7 - pos] = clampi(clampi(v, o - cv, o + cv), 0, maxv);
                }
                if (pos < nq) {
                    int f = nq == 7 ? 59 - 9 * pos : nq == 5 ? (pos == 0 ? 58 : pos == 1 ? 45 : pos == 2 ? 32 : pos == 3 ? 19 : 6)
                                                             : (pos == 0 ? 53 : pos == 1 ? 32 : 11);
                    int t = nq == 7 ? (pos < 5 ? 6 - pos : 1) : nq == 5 ? 6 - pos : 6 - 2 * pos;
                    int o = QQ(m, pos), cv = (tc * t) >> 1;
                    int v = (rm * f + refq * (64 - f) + 32) >> 6;
                    line[8 + pos] = clampi(clampi(v, o - cv, o + cv), 0, maxv);
                }
            }
        } else if (de == 2) {
            // strong short filter (vvc/filter.rs:590-611)
            line[7] = clampi((PP(m, 2) + 2 * PP(m, 1) + 2 * PP(m, 0) + 2 * QQ(m, 0) + QQ(m, 1) + 4) >> 3,
                             PP(m, 0) - 3 * tc, PP(m, 0) + 3 * tc);
            line[6] = clampi((PP(m, 2) + PP(m, 1) + PP(m, 0) + QQ(m, 0) + 2) >> 2, PP(m, 1) - 2 * tc, PP(m, 1) + 2 * tc);
            line[5] = clampi((2 * PP(m, 3) + 3 * PP(m, 2) + PP(m, 1) + PP(m, 0) + QQ(m, 0) + 4) >> 3, PP(m, 2) - tc,
                             PP(m, 2) + tc);
            line[8] = clampi((PP(m, 1) + 2 * PP(m, 0) + 2 * QQ(m, 0) + 2 * QQ(m, 1) + QQ(m, 2) + 4) >> 3,
                             QQ(m, 0) - 3 * tc, QQ(m, 0) + 3 * tc);
            line[9] = clampi((PP(m, 0) + QQ(m, 0) + QQ(m, 1) + QQ(m, 2) + 2) >> 2, QQ(m, 1) - 2 * tc, QQ(m, 1) + 2 * tc);
            line[10] = clampi((PP(m, 0) + QQ(m, 0) + QQ(m, 1) + 3 * QQ(m, 2) + 2 * QQ(m, 3) + 4) >> 3, QQ(m, 2) - tc,
                              QQ(m, 2) + tc);
        } else {
            // weak filter (vvc/filter.rs:613-630)
            int delta = (9 * (QQ(m, 0) - PP(m, 0)) - 3 * (QQ(m, 1) - PP(m, 1)) + 8) >> 4;
            if (iabs(delta) < tc * 10) {
                delta = clampi(delta, -tc, tc);
                line[7] = clampi(PP(m, 0) + delta, 0, maxv);
                line[8] = clampi(QQ(m, 0) - delta, 0, maxv);
                int tc2 = tc >> 1;
                if (dep) line[6] = clampi(PP(m, 1) + clampi((((PP(m, 2) + PP(m, 0) + 1) >> 1) - PP(m, 1) + delta) >> 1, -tc2, tc2), 0, maxv);
                if (deq) line[9] = clampi(QQ(m, 1) + clampi((((QQ(m, 2) + QQ(m, 0) + 1) >> 1) - QQ(m, 1) - delta) >> 1, -tc2, tc2), 0, maxv);
            }
        }
    }
    np_out = de == 3 ? lp : de == 2 ? 3 : 2;
    nq_out = de == 3 ? lq : de == 2 ? 3 : 2;
}

// deblock_chroma_edge with two lines (vvc/filter.rs:935-1057). Returns the
// number of samples per side that may have changed.
__device__ __forceinline__ int chroma_segment(int (*L)[8], int qp, int bs, int beta_off, int tc_off, int len_p, int len_q, int bd) {
    len_p = len_p < 3 ? len_p : 3;
    len_q = len_q < 3 ? len_q : 3;
    if (len_p == 1 && len_q == 1 && bs != 2) return 0;
    const int maxv = (1 << bd) - 1;
    int tc, beta;
    tc_beta(qp, bs, beta_off, tc_off, bd, tc, beta);
    if (len_q == 3) {
        int a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            a[k] = L[0][k];
            b[k] = L[1][k];
        }
        if (len_p == 1) {
            a[0] = a[1] = a[2];
            b[0] = b[1] = b[2];
        }
        int dp0 = iabs(a[1] - 2 * a[2] + a[3]), dp1 = iabs(b[1] - 2 * b[2] + b[3]);
        int dq0 = iabs(a[6] - 2 * a[5] + a[4]), dq1 = iabs(b[6] - 2 * b[5] + b[4]);
        int dpq0 = dp0 + dq0, dpq1 = dp1 + dq1;
        auto cdec = [&](int p0, int p3, int q0, int q3, int dpq) {
            return dpq < (beta >> 2) && iabs(p3 - p0) + iabs(q0 - q3) < (beta >> 3) && iabs(p0 - q0) < ((5 * tc + 1) >> 1);
        };
        bool strong = (dpq0 + dpq1) < beta && cdec(a[3], a[0], a[4], a[7], 2 * dpq0) && cdec(b[3], b[0], b[4], b[7], 2 * dpq1);
        if (!strong) len_p = len_q = 1;
    }
#pragma unroll
    for (int li = 0; li < 2; li++) {
        int* l = L[li];
        int m[8];
#pragma unroll
        for (int k = 0; k < 8; k++) m[k] = l[k];
        if (len_p == 3 && len_q == 3) {
            l[1] = clampi(clampi((3 * m[0] + 2 * m[1] + m[2] + m[3] + m[4] + 4) >> 3, m[1] - tc, m[1] + tc), 0, maxv);
            l[2] = clampi(clampi((2 * m[0] + m[1] + 2 * m[2] + m[3] + m[4] + m[5] + 4) >> 3, m[2] - tc, m[2] + tc), 0, maxv);
            l[3] = clampi(clampi((m[0] + m[1] + m[2] + 2 * m[3] + m[4] + m[5] + m[6] + 4) >> 3, m[3] - tc, m[3] + tc), 0, maxv);
            l[4] = clampi(clampi((m[1] + m[2] + m[3] + 2 * m[4] + m[5] + m[6] + m[7] + 4) >> 3, m[4] - tc, m[4] + tc), 0, maxv);
            l[5] = clampi(clampi((m[2] + m[3] + m[4] + 2 * m[5] + m[6] + 2 * m[7] + 4) >> 3, m[5] - tc, m[5] + tc), 0, maxv);
            l[6] = clampi(clampi((m[3] + m[4] + m[5] + 2 * m[6] + 3 * m[7] + 4) >> 3, m[6] - tc, m[6] + tc), 0, maxv);
        } else if (len_p == 1 && len_q == 3) {
            l[3] = clampi(clampi((3 * m[2] + 2 * m[3] + m[4] + m[5] + m[6] + 4) >> 3, m[3] - tc, m[3] + tc), 0, maxv);
            l[4] = clampi(clampi((2 * m[2] + m[3] + 2 * m[4] + m[5] + m[6] + m[7] + 4) >> 3, m[4] - tc, m[4] + tc), 0, maxv);
            l[5] = clampi(clampi((m[2] + m[3] + m[4] + 2 * m[5] + m[6] + 2 * m[7] + 4) >> 3, m[5] - tc, m[5] + tc), 0, maxv);
            l[6] = clampi(clampi((m[3] + m[4] + m[5] + 2 * m[6] + 3 * m[7] + 4) >> 3, m[6] - tc, m[6] + tc), 0, maxv);
        } else {
            int delta = clampi((4 * (m[4] - m[3]) + m[2] - m[5] + 4) >> 3, -tc, tc);
            l[3] = clampi(m[3] + delta, 0, maxv);
            l[4] = clampi(m[4] - delta, 0, maxv);
        }
    }
    return 3;
}

// One lane per edge segment of every plane of every frame; `dir` 0 = vertical
// edges, 1 = horizontal edges.
__global__ __launch_bounds__(256) void deblock_kernel(const rmh_dbk_frame* __restrict__ frames, int dir) {
    const rmh_dbk_frame& f = frames[blockIdx.y];
    const int plane = blockIdx.z;
    const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= f.grid_w * f.grid_h) return;
    const int i = cell % f.grid_w, j = cell / f.grid_w;
    const rmh_plane& P = f.planes[plane];
    const int w = P.width, h = P.height, bd = P.bit_depth;
    const uint8_t e = f.edges[(size_t)cell * 4 + (plane == 0 ? 0 : 2) + dir];
    const int bs = e & RMH_EDGE_BS_MASK;
    if (!bs) return;
    const int len_p = 2 * ((e >> RMH_EDGE_LENP_SHIFT) & 3) + 1, len_q = 2 * ((e >> RMH_EDGE_LENQ_SHIFT) & 3) + 1;
    const int qp = f.qp[plane], boff = f.beta_offset_div2[plane], toff = f.tc_offset_div2[plane];
    if (plane == 0) {
        const int x = 4 * i, y = 4 * j;
        if ((dir == 0 && (i == 0 || x >= w || y >= h)) || (dir == 1 && (j == 0 || y >= h || x >= w))) return;
        int L[4][16];
#pragma unroll
        for (int li = 0; li < 4; li++)
#pragma unroll
            for (int k = 0; k < 16; k++)
                L[li][k] = dir == 0 ? ld(P, clampi(x - 8 + k, 0, w - 1), min(y + li, h - 1))
                                    : ld(P, min(x + li, w - 1), clampi(y - 8 + k, 0, h - 1));
        int np, nq;
        luma_segment(L, qp, bs, boff, toff, bd, len_p, len_q, dir == 1 && (e & RMH_EDGE_CTB_ROW), np, nq);
        if (np == 0) return;
#pragma unroll
        for (int li = 0; li < 4; li++)
#pragma unroll
            for (int k = 1; k < 15; k++) {
                if (k < 8 - np || k >= 8 + nq) continue;
                if (dir == 0)
                    st(P, clampi(x - 8 + k, 0, w - 1), min(y + li, h - 1), L[li][k]);
                else
                    st(P, min(x + li, w - 1), clampi(y - 8 + k, 0, h - 1), L[li][k]);
            }
    } else {
        const int x = 2 * i, y = 2 * j;
        if ((dir == 0 && (i == 0 || x >= w || y + 2 > h)) || (dir == 1 && (j == 0 || y >= h || x + 2 > w))) return;
        int L[2][8];
#pragma unroll
        for (int li = 0; li < 2; li++)
#pragma unroll
            for (int k = 0; k < 8; k++)
                L[li][k] = dir == 0 ? ld(P, clampi(x - 4 + k, 0, w - 1), min(y + li, h - 1))
                                    : ld(P, min(x + li, w - 1), clampi(y - 4 + k, 0, h - 1));
        if (!chroma_segment(L, qp, bs, boff, toff, len_p, len_q, bd)) return;
#pragma unroll
        for (int li = 0; li < 2; li++)
#pragma unroll
            for (int k = 1; k < 7; k++) {
                if (dir == 0)
                    st(P, clampi(x - 4 + k, 0, w - 1), min(y + li, h - 1), L[li][k]);
                else
                    st(P, min(x + li, w - 1), clampi(y - 4 + k, 0, h - 1), L[li][k]);
            }
    }
}

}  // namespace

hipError_t launch_deblock(const rmh_dbk_frame* frames, uint32_t n_frames, uint32_t max_cells, hipStream_t stream) {
    if (n_frames == 0 || max_cells == 0) return hipSuccess;
    dim3 grid((max_cells + 255) / 256, n_frames, 3);
    hipLaunchKernelGGL(deblock_kernel, grid, dim3(256), 0, stream, frames, 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(deblock_kernel, grid, dim3(256), 0, stream, frames, 1);
    return hipGetLastError();
}

}  // namespace rmh

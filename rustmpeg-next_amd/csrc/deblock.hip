// deblock.hip - luma and chroma deblocking of whole pictures for gfx950.
//
// Semantics follow deblock_luma_edge_bd / deblock_chroma_edge
// (crates/rustmpeg-video/src/vvc/filter.rs:184-1057) driven in the order of
// apply_deblock (vvc/decode_general.rs:4977-5187): every vertical edge of a
// picture, then every horizontal edge.
//
// Mapping: a copy of every plane into a shadow, then one workgroup per 64x32
// luma tile (with its 32x16 chroma tiles) that filters both edge directions
// in LDS from the shadow and writes the tile back once (deblock_tile_kernel
// below states why this equals the reference order).
#include <hip/hip_runtime.h>

#include "gmem.h"
#include "kernels.h"
#include "xcd.h"

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

// One of three values by a filter length n in {3, 5, 7}, with masks rather
// than a select of array elements (which the compiler turns into an indexed
// load and moves the line to scratch memory).
__device__ __forceinline__ int by_len(int n, int v3, int v5, int v7) {
    return (-(int)(n == 3) & v3) | (-(int)(n == 5) & v5) | (-(int)(n == 7) & v7);
}
// p_n / q_n and p_(n-1) / q_(n-1) of a line for n in {3, 5, 7}
__device__ __forceinline__ int p_at(const int* l, int n) { return by_len(n, PP(l, 3), PP(l, 5), PP(l, 7)); }
__device__ __forceinline__ int q_at(const int* l, int n) { return by_len(n, QQ(l, 3), QQ(l, 5), QQ(l, 7)); }
__device__ __forceinline__ int p_before(const int* l, int n) { return by_len(n, PP(l, 2), PP(l, 4), PP(l, 6)); }
__device__ __forceinline__ int q_before(const int* l, int n) { return by_len(n, QQ(l, 2), QQ(l, 4), QQ(l, 6)); }

// Decision of one 4-line luma segment (deblock_luma_edge_bd, vvc/filter.rs:638-679,
// decide_luma_edge, vvc/filter.rs:311-468) from its lines 0 and 3.
struct LumaDecision {
    int de;        // 1 weak, 2 strong short, 3 long
    int lp, lq;    // filter lengths in use
    bool dep, deq; // weak filter also changes p1 / q1
    int tc;
};

__device__ __forceinline__ bool luma_decide(const int* l0, const int* l3, int qp, int bs, int beta_off, int tc_off, int bd,
                                            int len_p, int len_q, bool at_ctb, LumaDecision& dd) {
    int tc, beta;
    tc_beta(qp, bs, beta_off, tc_off, bd, tc, beta);
    if (beta == 0 || tc == 0) return false;
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
            bool s0 = dsam(pl ? p_at(l0, lp) : 0, pl ? PP(l0, 3) : 0, ql ? q_at(l0, lq) : 0, ql ? QQ(l0, 3) : 0, 2 * d0l,
                           sp0l, sq0l, spq0, pl, ql, beta, tc);
            bool s3 = dsam(pl ? p_at(l3, lp) : 0, pl ? PP(l3, 3) : 0, ql ? q_at(l3, lq) : 0, ql ? QQ(l3, 3) : 0, 2 * d3l,
                           sp3l, sq3l, spq3, pl, ql, beta, tc);
            if (s0 && s3) de = 3;
        }
    }
    if (de != 3) {
        if (dpq0 + dpq3 >= beta) return false;
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
    dd.de = de;
    dd.lp = lp;
    dd.lq = lq;
    dd.dep = dep;
    dd.deq = deq;
    dd.tc = tc;
    return true;
}

// Filters one line (p7..p0 q0..q7) of a segment with the segment's decision
// (long_tap_luma_line / short_filter_luma_line, vvc/filter.rs:477-631).
__device__ __forceinline__ void luma_filter_line(int* line, const LumaDecision& dd, int maxv) {
    const int de = dd.de, tc = dd.tc;
    const int lp = dd.lp, lq = dd.lq;
    const bool dep = dd.dep, deq = dd.deq;
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
        int refp = (p_at(m, np) + p_before(m, np) + 1) >> 1;
        int refq = (q_at(m, nq) + q_before(m, nq) + 1) >> 1;
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

// Samples per side that a decision may change.
__device__ __forceinline__ int luma_np(const LumaDecision& d) { return d.de == 3 ? d.lp : d.de == 2 ? 3 : 2; }
__device__ __forceinline__ int luma_nq(const LumaDecision& d) { return d.de == 3 ? d.lq : d.de == 2 ? 3 : 2; }

// deblock_chroma_edge (vvc/filter.rs:935-1057) for one line m (p3..q3) of a
// two-line segment whose other line is o; the decision depends on both lines
// symmetrically, so both lanes of a segment reach the same one. Returns
// false when the segment is not filtered.
__device__ __forceinline__ bool chroma_line(int* m, const int* o, int qp, int bs, int beta_off, int tc_off, int len_p,
                                            int len_q, int bd) {
    len_p = len_p < 3 ? len_p : 3;
    len_q = len_q < 3 ? len_q : 3;
    if (len_p == 1 && len_q == 1 && bs != 2) return false;
    const int maxv = (1 << bd) - 1;
    int tc, beta;
    tc_beta(qp, bs, beta_off, tc_off, bd, tc, beta);
    if (len_q == 3) {
        int a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            a[k] = m[k];
            b[k] = o[k];
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
    int v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = m[k];
    if (len_p == 3 && len_q == 3) {
        m[1] = clampi(clampi((3 * v[0] + 2 * v[1] + v[2] + v[3] + v[4] + 4) >> 3, v[1] - tc, v[1] + tc), 0, maxv);
        m[2] = clampi(clampi((2 * v[0] + v[1] + 2 * v[2] + v[3] + v[4] + v[5] + 4) >> 3, v[2] - tc, v[2] + tc), 0, maxv);
        m[3] = clampi(clampi((v[0] + v[1] + v[2] + 2 * v[3] + v[4] + v[5] + v[6] + 4) >> 3, v[3] - tc, v[3] + tc), 0, maxv);
        m[4] = clampi(clampi((v[1] + v[2] + v[3] + 2 * v[4] + v[5] + v[6] + v[7] + 4) >> 3, v[4] - tc, v[4] + tc), 0, maxv);
        m[5] = clampi(clampi((v[2] + v[3] + v[4] + 2 * v[5] + v[6] + 2 * v[7] + 4) >> 3, v[5] - tc, v[5] + tc), 0, maxv);
        m[6] = clampi(clampi((v[3] + v[4] + v[5] + 2 * v[6] + 3 * v[7] + 4) >> 3, v[6] - tc, v[6] + tc), 0, maxv);
    } else if (len_p == 1 && len_q == 3) {
        m[3] = clampi(clampi((3 * v[2] + 2 * v[3] + v[4] + v[5] + v[6] + 4) >> 3, v[3] - tc, v[3] + tc), 0, maxv);
        m[4] = clampi(clampi((2 * v[2] + v[3] + 2 * v[4] + v[5] + v[6] + v[7] + 4) >> 3, v[4] - tc, v[4] + tc), 0, maxv);
        m[5] = clampi(clampi((v[2] + v[3] + v[4] + 2 * v[5] + v[6] + 2 * v[7] + 4) >> 3, v[5] - tc, v[5] + tc), 0, maxv);
        m[6] = clampi(clampi((v[3] + v[4] + v[5] + 2 * v[6] + 3 * v[7] + 4) >> 3, v[6] - tc, v[6] + tc), 0, maxv);
    } else {
        int delta = clampi((4 * (v[4] - v[3]) + v[2] - v[5] + 4) >> 3, -tc, tc);
        m[3] = clampi(v[3] + delta, 0, maxv);
        m[4] = clampi(v[4] - delta, 0, maxv);
    }
    return true;
}

// ---------------------------------------------------------------------------
// Copy of every plane into its shadow (the tiled kernel reads the shadow and
// writes the picture, so no workgroup reads samples another one has written).
// Rows are copied in the widest unit that the two layouts allow.
__global__ __launch_bounds__(256) void shadow_copy_kernel(const rmh_dbk_frame* __restrict__ frames,
                                                          const rmh_plane* __restrict__ shadows) {
    const rmh_plane& src = frames[blockIdx.y].planes[blockIdx.z];
    const rmh_plane& dst = shadows[blockIdx.y * 3 + blockIdx.z];
    const uint32_t row_bytes = src.width * src.bytes_per_sample;
    const uint32_t sp = src.stride * src.bytes_per_sample, dp = dst.stride * dst.bytes_per_sample;
    const uintptr_t al = (uintptr_t)src.data | (uintptr_t)dst.data | sp | dp | row_bytes;
    for (uint32_t y = blockIdx.x * 4 + (threadIdx.x >> 6); y < src.height; y += gridDim.x * 4) {
        const RMH_GLOBAL uint8_t* s = gsrc<uint8_t>(src.data) + (size_t)y * sp;
        RMH_GLOBAL uint8_t* d = gdst<uint8_t>(dst.data) + (size_t)y * dp;
        const uint32_t lane = threadIdx.x & 63;
        if ((al & 15) == 0) {
            for (uint32_t o = lane * 16; o < row_bytes; o += 1024)
                *(RMH_GLOBAL u32x4*)(d + o) = *(const RMH_GLOBAL u32x4*)(s + o);
        } else if ((al & 7) == 0) {
            for (uint32_t o = lane * 8; o < row_bytes; o += 512)
                *(RMH_GLOBAL u32x2*)(d + o) = *(const RMH_GLOBAL u32x2*)(s + o);
        } else if ((al & 3) == 0) {
            for (uint32_t o = lane * 4; o < row_bytes; o += 256)
                *(RMH_GLOBAL uint32_t*)(d + o) = *(const RMH_GLOBAL uint32_t*)(s + o);
        } else {
            for (uint32_t o = lane; o < row_bytes; o += 64) d[o] = s[o];
        }
    }
}

// ---------------------------------------------------------------------------
// Tiled deblocking: one workgroup per 64x32 luma tile of a frame and the
// matching 32x16 tiles of both chroma planes.
//
//   1. the tile plus an 8-sample luma (4-sample chroma) halo, clamped at the
//      picture edges, is read from the shadow into LDS;
//   2. every vertical edge segment that can change a sample the tile or its
//      horizontal pass reads is filtered in LDS (rows of the halo included);
//   3. rows outside the picture are re-filled from the nearest picture row,
//      as the reference's clamped gather reads them;
//   4. every horizontal edge segment that can change a tile sample is
//      filtered in LDS;
//   5. the tile is written to the picture, each sample once.
// Segments of one direction touch disjoint samples for VVC filter lengths
// (length 1 on edges off the 8-sample grid, at most 3 (chroma) / 7 (luma) on
// it), so filtering them in parallel equals the reference's sequential order,
// and a tile's halo segments compute exactly what the neighbouring tile
// computes for the same samples.
// ---------------------------------------------------------------------------

constexpr int kTW = 64, kTH = 32;                      // luma tile
constexpr int kLWW = kTW + 16, kLWH = kTH + 16;        // luma window (8 halo each side)
constexpr int kCTW = kTW / 2, kCTH = kTH / 2;          // chroma tile
constexpr int kCWW = kCTW + 8, kCWH = kCTH + 8;        // chroma window (4 halo each side)

// Sample access with the sample size BPS (1: 8-bit, 2: up to 12-bit) fixed
// at compile time; a launch covers the tiles of frames of one sample size.
template <int BPS>
__device__ __forceinline__ int ld32(const rmh_plane& p, int x, int y) {
    const uint32_t o = (uint32_t)y * p.stride + (uint32_t)x;
    return BPS == 1 ? (int)gsrc<uint8_t>(p.data)[o] : (int)gsrc<uint16_t>(p.data)[o];
}
template <int BPS>
__device__ __forceinline__ void st32(const rmh_plane& p, int x, int y, int v) {
    const uint32_t o = (uint32_t)y * p.stride + (uint32_t)x;
    if (BPS == 1)
        gdst<uint8_t>(p.data)[o] = (uint8_t)v;
    else
        gdst<uint16_t>(p.data)[o] = (uint16_t)v;
}

// Chunked window reads and tile writes. A chunk is CH samples of one row
// (CH * BPS bytes), aligned whenever the plane's base address and row pitch
// are: a chunk inside the picture's columns moves as one vector access, a
// chunk crossing a picture column edge sample by sample (reads clamped as the
// reference's gather clamps them). Samples sit in LDS as int16, two per dword.
template <int CH>
struct Packed {
    uint32_t d[CH / 2];
};

__device__ __forceinline__ uint32_t widen_lo(uint32_t v) { return (v & 0xffu) | ((v & 0xff00u) << 8); }
__device__ __forceinline__ uint32_t widen_hi(uint32_t v) { return ((v >> 16) & 0xffu) | ((v >> 8) & 0xff0000u); }
__device__ __forceinline__ uint32_t narrow(uint32_t a, uint32_t b) {
    return (a & 0xffu) | ((a >> 8) & 0xff00u) | ((b & 0xffu) << 16) | ((b << 8) & 0xff000000u);
}

template <int CH, int BPS>
__device__ __forceinline__ bool chunks_aligned(const rmh_plane& p) {
    return (((uintptr_t)p.data | (uintptr_t)p.stride * BPS) & (CH * BPS - 1)) == 0;
}

// Samples x .. x + CH - 1 of row y (already clamped into the picture).
template <int CH, int BPS>
__device__ __forceinline__ Packed<CH> read_chunk(const rmh_plane& p, int x, int y, bool vec) {
    Packed<CH> o;
    const int w = p.width;
    if (vec && x >= 0 && x + CH <= w) {
        const RMH_GLOBAL uint8_t* src = gsrc<uint8_t>(p.data) + ((size_t)y * p.stride + x) * BPS;
        if constexpr (BPS == 2 && CH == 8) {
            const u32x4 v = *(const RMH_GLOBAL u32x4*)src;
            o.d[0] = v.x; o.d[1] = v.y; o.d[2] = v.z; o.d[3] = v.w;
        } else if constexpr (BPS == 2) {
            const u32x2 v = *(const RMH_GLOBAL u32x2*)src;
            o.d[0] = v.x; o.d[1] = v.y;
        } else if constexpr (CH == 8) {
            const u32x2 v = *(const RMH_GLOBAL u32x2*)src;
            o.d[0] = widen_lo(v.x); o.d[1] = widen_hi(v.x); o.d[2] = widen_lo(v.y); o.d[3] = widen_hi(v.y);
        } else {
            const uint32_t v = *(const RMH_GLOBAL uint32_t*)src;
            o.d[0] = widen_lo(v); o.d[1] = widen_hi(v);
        }
    } else {
#pragma unroll
        for (int j = 0; j < CH / 2; j++)
            o.d[j] = (uint32_t)ld32<BPS>(p, clampi(x + 2 * j, 0, w - 1), y) |
                     ((uint32_t)ld32<BPS>(p, clampi(x + 2 * j + 1, 0, w - 1), y) << 16);
    }
    return o;
}

template <int CH>
__device__ __forceinline__ void put_chunk(int16_t* dst, const Packed<CH>& v) {
    if constexpr (CH == 8)
        *reinterpret_cast<u32x4*>(dst) = u32x4{v.d[0], v.d[1], v.d[2], v.d[3]};
    else
        *reinterpret_cast<u32x2*>(dst) = u32x2{v.d[0], v.d[1]};
}

// Samples x .. x + 7 of row y from LDS (src 8-byte aligned) to the picture;
// samples at or beyond column w are not written.
template <int BPS>
__device__ __forceinline__ void write_chunk8(const rmh_plane& p, int x, int y, const int16_t* src, bool vec) {
    const u32x2 a = reinterpret_cast<const u32x2*>(src)[0], b = reinterpret_cast<const u32x2*>(src)[1];
    if (vec && x + 8 <= (int)p.width) {
        RMH_GLOBAL uint8_t* dst = gdst<uint8_t>(p.data) + ((size_t)y * p.stride + x) * BPS;
        if (BPS == 2)
            *(RMH_GLOBAL u32x4*)dst = u32x4{a.x, a.y, b.x, b.y};
        else
            *(RMH_GLOBAL u32x2*)dst = u32x2{narrow(a.x, a.y), narrow(b.x, b.y)};
    } else {
        const uint32_t d[4] = {a.x, a.y, b.x, b.y};
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (x + j < (int)p.width) st32<BPS>(p, x + j, y, (int)((d[j >> 1] >> (16 * (j & 1))) & 0xffffu));
    }
}

// Rows of the window outside the picture take the nearest picture row.
template <int WW, int WH>
__device__ __forceinline__ void refill_rows(int wy0, int h, int16_t* win) {
    for (int e = threadIdx.x; e < WW * WH; e += 256) {
        const int r = e / WW, c = e - r * WW, y = wy0 + r;
        if (y < 0 || y >= h) win[e] = win[(clampi(y, 0, h - 1) - wy0) * WW + c];
    }
}

// Geometry of one tile's segment slots. Vertical pass: luma x = X0 .. X0 + 64
// step 4 over window rows in 4-row segments; chroma x = CX0 .. CX0 + 32 step
// 2 over window rows in 2-row segments (the edge map's chroma cells are 2x2
// chroma samples). Horizontal pass: luma y = Y0 .. Y0 + 32 step 4 over the
// tile's columns in 4-column segments; chroma y = CY0 .. CY0 + 16 step 2 over
// 2-column segments.
constexpr int kLVX = kTW / 4 + 1, kLVY = kLWH / 4;   // 17 x 12
constexpr int kCVX = kCTW / 2 + 1, kCVY = kCWH / 2;  // 17 x 12
constexpr int kLHX = kTW / 4, kLHY = kTH / 4 + 1;    // 16 x 9
constexpr int kCHX = kCTW / 2, kCHY = kCTH / 2 + 1;  // 16 x 9
constexpr int kNV = kLVX * kLVY + 2 * kCVX * kCVY;
constexpr int kNH = kLHX * kLHY + 2 * kCHX * kCHY;
constexpr int kMaxQ = kNV > kNH ? kNV : kNH;

// The edge-map cells a tile needs: luma cells x / 4 = X0 / 4 .. + 16 and
// y / 4 = Y0 / 4 - 2 .. + 11 (chroma cells are the same cells: chroma
// sample 2i is luma sample 4i).
constexpr int kEX = kLVX, kEY = kLVY;  // 17 x 12

struct TileCtx {
    const rmh_dbk_frame* f;
    const uint32_t* cells;  // LDS, kEY x kEX, the 4 edge bytes of a cell
    int W, H, CW, CH, X0, Y0, CX0, CY0, bd;
};

__device__ __forceinline__ uint8_t cell_byte(const TileCtx& t, int cx, int cy, int b) {
    return (uint8_t)(t.cells[(cy - (t.Y0 / 4 - 2)) * kEX + (cx - t.X0 / 4)] >> (8 * b));
}

// Slot -> (plane, edge position). Returns the edge byte, 0 when the slot has
// no edge to filter (outside the picture or bS 0).
__device__ __forceinline__ uint8_t slot_edge(const TileCtx& t, int dir, int s, int& plane, int& x, int& y) {
    const int nl = dir == 0 ? kLVX * kLVY : kLHX * kLHY;
    const int cx = dir == 0 ? kCVX : kCHX, cy = dir == 0 ? kCVY : kCHY;
    if (s < nl) {
        plane = 0;
        if (dir == 0) {
            x = t.X0 + 4 * (s % kLVX);
            y = t.Y0 - 8 + 4 * (s / kLVX);
            if (x == 0 || x >= t.W || y < 0 || y >= t.H) return 0;
        } else {
            x = t.X0 + 4 * (s % kLHX);
            y = t.Y0 + 4 * (s / kLHX);
            if (y == 0 || y >= t.H || x >= t.W) return 0;
        }
        const uint8_t e = cell_byte(t, x >> 2, y >> 2, dir);
        return (e & RMH_EDGE_BS_MASK) ? e : 0;
    }
    const int u = s - nl, c = u / (cx * cy), v = u % (cx * cy);
    plane = 1 + c;
    if (dir == 0) {
        x = t.CX0 + 2 * (v % kCVX);
        y = t.CY0 - 4 + 2 * (v / kCVX);
        if (x == 0 || x >= t.CW || y < 0 || y + 2 > t.CH) return 0;
    } else {
        x = t.CX0 + 2 * (v % kCHX);
        y = t.CY0 + 2 * (v / kCHX);
        if (y == 0 || y >= t.CH || x + 2 > t.CW) return 0;
    }
    const uint8_t e = cell_byte(t, x >> 1, y >> 1, 2 + dir);
    return (e & RMH_EDGE_BS_MASK) ? e : 0;
}

// Lane exchange inside a pair (lanes 2m, 2m + 1) of a wave; both lanes must
// be active.
__device__ __forceinline__ int pair_other(int v) { return __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false); }

// Slot -> (plane, x, y) without the edge-map read.
__device__ __forceinline__ void slot_pos(const TileCtx& t, int dir, int s, int& plane, int& x, int& y) {
    const int nl = dir == 0 ? kLVX * kLVY : kLHX * kLHY;
    const int cx = dir == 0 ? kCVX : kCHX, cy = dir == 0 ? kCVY : kCHY;
    if (s < nl) {
        plane = 0;
        x = dir == 0 ? t.X0 + 4 * (s % kLVX) : t.X0 + 4 * (s % kLHX);
        y = dir == 0 ? t.Y0 - 8 + 4 * (s / kLVX) : t.Y0 + 4 * (s / kLHX);
        return;
    }
    const int u = s - nl, c = u / (cx * cy), v = u % (cx * cy);
    plane = 1 + c;
    x = dir == 0 ? t.CX0 + 2 * (v % kCVX) : t.CX0 + 2 * (v % kCHX);
    y = dir == 0 ? t.CY0 - 4 + 2 * (v / kCVX) : t.CY0 + 2 * (v / kCHX);
}

// Line li (0..3) of one luma segment of the LDS window: dir 0 (vertical edge
// at x) row y + li, element k (p7 .. q7) at base[k], base 8-byte aligned;
// dir 1 (horizontal edge at y) column x + li clamped into the picture, as the
// reference's gather clamps it, element k at base[k * kLWW].
__device__ __forceinline__ int16_t* luma_line_base(const TileCtx& t, int dir, int x, int y, int li, int16_t* lw) {
    return dir == 0 ? lw + (y + li - (t.Y0 - 8)) * kLWW + (x - t.X0)
                    : lw + (y - t.Y0) * kLWW + 8 + (min(x + li, t.W - 1) - t.X0);
}

__device__ __forceinline__ void read_luma_line(int dir, const int16_t* base, int (&line)[16]) {
    if (dir == 0) {
        const uint2* b8 = reinterpret_cast<const uint2*>(base);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint2 v = b8[k];
            line[4 * k] = (int)(v.x & 0xffff);
            line[4 * k + 1] = (int)(v.x >> 16);
            line[4 * k + 2] = (int)(v.y & 0xffff);
            line[4 * k + 3] = (int)(v.y >> 16);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 16; k++) line[k] = base[k * kLWW];
    }
}

// A luma segment's decision in one dword: slot (bits 0-7), lp (8-10), lq
// (11-13), dep (14), deq (15), tc (16-27). The filter form (weak, strong
// short, long) is the queue the segment sits in.
__device__ __forceinline__ uint32_t pack_decision(int slot, const LumaDecision& d) {
    return (uint32_t)slot | ((uint32_t)d.lp << 8) | ((uint32_t)d.lq << 11) | ((uint32_t)d.dep << 14) |
           ((uint32_t)d.deq << 15) | ((uint32_t)d.tc << 16);
}
__device__ __forceinline__ LumaDecision unpack_decision(uint32_t v, int de) {
    LumaDecision d;
    d.de = de;
    d.lp = (v >> 8) & 7;
    d.lq = (v >> 11) & 7;
    d.dep = (v >> 14) & 1;
    d.deq = (v >> 15) & 1;
    d.tc = (v >> 16) & 0xfff;
    return d;
}

constexpr int kLQ = kLVX * kLVY;      // luma segments per pass, at most
constexpr int kCQ = 2 * kCVX * kCVY;  // chroma segments per pass, at most

// Decision of one luma segment (one lane per segment) from its lines 0 and
// 3; a filtered segment goes to the queue of its filter form, qd[de - 1].
__device__ __forceinline__ void luma_segment_decide(const TileCtx& t, int dir, uint32_t q, const int16_t* lw,
                                                    uint32_t (*qd)[kLQ], int* counts) {
    const rmh_dbk_frame& f = *t.f;
    int plane, x, y;
    const int slot = (int)(q & 0xffff);
    slot_pos(t, dir, slot, plane, x, y);
    const uint8_t e = (uint8_t)(q >> 16);
    const int bs = e & RMH_EDGE_BS_MASK;
    const int len_p = 2 * ((e >> RMH_EDGE_LENP_SHIFT) & 3) + 1, len_q = 2 * ((e >> RMH_EDGE_LENQ_SHIFT) & 3) + 1;
    int l0[16], l3[16];
    read_luma_line(dir, luma_line_base(t, dir, x, y, 0, const_cast<int16_t*>(lw)), l0);
    read_luma_line(dir, luma_line_base(t, dir, x, y, 3, const_cast<int16_t*>(lw)), l3);
    LumaDecision d;
    if (!luma_decide(l0, l3, f.qp[0], bs, f.beta_offset_div2[0], f.tc_offset_div2[0], t.bd, len_p, len_q,
                     dir == 1 && (e & RMH_EDGE_CTB_ROW) != 0, d))
        return;
    const uint32_t v = pack_decision(slot, d);
    // one queue address per atomic, so that each is one atomic per wave
    if (d.de == 1) qd[0][atomicAdd(&counts[2], 1)] = v;
    if (d.de == 2) qd[1][atomicAdd(&counts[3], 1)] = v;
    if (d.de == 3) qd[2][atomicAdd(&counts[4], 1)] = v;
}

// Line li (0..3) of a luma segment with its decision: filtered and written
// back. A line whose column repeats the previous one (clamped at the right
// picture edge) gives the same result and is not written twice.
__device__ __forceinline__ void luma_segment_filter(const TileCtx& t, int dir, uint32_t v, int de, int li, int16_t* lw) {
    int plane, x, y;
    slot_pos(t, dir, (int)(v & 0xff), plane, x, y);
    if (dir == 1 && li > 0 && x + li > t.W - 1) return;
    const LumaDecision d = unpack_decision(v, de);
    int16_t* base = luma_line_base(t, dir, x, y, li, lw);
    const int ks = dir == 0 ? 1 : kLWW;
    int line[16];
    read_luma_line(dir, base, line);
    const int np = luma_np(d), nq = luma_nq(d), maxv = (1 << t.bd) - 1;
    luma_filter_line(line, d, maxv);
#pragma unroll
    for (int k = 1; k < 15; k++)
        if (k >= 8 - np && k < 8 + nq) base[k * ks] = (int16_t)line[k];
}

// Line li (0, 1) of one chroma segment; the two lanes of a pair hold its two lines.
__device__ __forceinline__ void chroma_line_lds(const TileCtx& t, int dir, int plane, int x, int y, uint8_t e, int li,
                                                int16_t* cw) {
    const rmh_dbk_frame& f = *t.f;
    const int bs = e & RMH_EDGE_BS_MASK;
    const int len_p = 2 * ((e >> RMH_EDGE_LENP_SHIFT) & 3) + 1, len_q = 2 * ((e >> RMH_EDGE_LENQ_SHIFT) & 3) + 1;
    int16_t* base;
    int ks;
    if (dir == 0) {
        base = cw + (y + li - (t.CY0 - 4)) * kCWW + (x - t.CX0);  // column x - 4
        ks = 1;
    } else {
        base = cw + (y - t.CY0) * kCWW + (x + li - t.CX0 + 4);    // row y - 4
        ks = kCWW;
    }
    int m[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; k++) m[k] = base[k * ks];
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = pair_other(m[k]);
    if (!chroma_line(m, o, f.qp[plane], bs, f.beta_offset_div2[plane], f.tc_offset_div2[plane], len_p, len_q, t.bd))
        return;
#pragma unroll
    for (int k = 1; k < 7; k++) base[k * ks] = (int16_t)m[k];
}

// Queue entries: slot | edge byte << 16. Luma runs in two steps: decisions
// with one lane per segment (into one queue per filter form), then the
// filters with one lane per line, form by form, so that the lanes of a wave
// run the same filter. Chroma segments (two lanes each) are filtered during
// the luma decisions.
__device__ __forceinline__ void tile_pass(const TileCtx& t, int dir, int16_t* lw, int16_t (*cw)[kCWH * kCWW],
                                          uint32_t* ql, uint32_t* qc, uint32_t (*qd)[kLQ], int* counts) {
    constexpr int NS = (kMaxQ + 255) / 256;
    const int nslots = dir == 0 ? kNV : kNH;
    const int nl = dir == 0 ? kLVX * kLVY : kLHX * kLHY;
    if (threadIdx.x < 5) counts[threadIdx.x] = 0;
    __syncthreads();
    uint8_t eb[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
        const int s = threadIdx.x + 256 * i;
        int plane, x, y;
        eb[i] = s < nslots ? slot_edge(t, dir, s, plane, x, y) : 0;
    }
#pragma unroll
    for (int i = 0; i < NS; i++) {
        const int s = threadIdx.x + 256 * i;
        if (eb[i]) {
            const uint32_t q = (uint32_t)s | ((uint32_t)eb[i] << 16);
            if (s < nl)
                ql[atomicAdd(&counts[0], 1)] = q;
            else
                qc[atomicAdd(&counts[1], 1)] = q;
        }
    }
    __syncthreads();
    const int n_l = counts[0], n_c = counts[1];
    for (int i = threadIdx.x; i < n_l; i += 256) luma_segment_decide(t, dir, ql[i], lw, qd, counts);
    // pairs of lanes per chroma segment (256 is even, so a pair is active as a whole)
    for (int i = threadIdx.x; i < 2 * n_c; i += 256) {
        int plane, x, y;
        const uint32_t q = qc[i >> 1];
        slot_pos(t, dir, (int)(q & 0xffff), plane, x, y);
        chroma_line_lds(t, dir, plane, x, y, (uint8_t)(q >> 16), i & 1, cw[plane - 1]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int n = counts[2 + k];
        for (int i = threadIdx.x; i < 4 * n; i += 256) luma_segment_filter(t, dir, qd[k][i >> 2], k + 1, i & 3, lw);
    }
    __syncthreads();
}

// One workgroup per tile. Tiles are numbered frame by frame (in the order
// frame_of[] gives, 8-bit frames first) in raster order; tile_start[k] is the
// first tile of the k-th frame of that order. A launch covers tiles
// first_tile .. + gridDim.x, all on frames of sample size BPS, handed to the
// XCDs in contiguous bands so that tiles sharing halo rows and columns share
// an L2.
template <int BPS>
__global__ __launch_bounds__(256) void deblock_tile_kernel(const rmh_dbk_frame* __restrict__ frames,
                                                           const rmh_plane* __restrict__ shadows,
                                                           const uint32_t* __restrict__ tile_start,
                                                           const uint32_t* __restrict__ frame_of, uint32_t n_frames,
                                                           uint32_t first_tile) {
    __shared__ __align__(16) int16_t lw[kLWH * kLWW];
    __shared__ __align__(16) int16_t cwin[2][kCWH * kCWW];
    __shared__ uint32_t ql[kLQ], qc[kCQ];
    __shared__ uint32_t qd[3][kLQ];
    __shared__ uint32_t cells[kEY * kEX];
    __shared__ int counts[5];
    const uint32_t tile = first_tile + xcd_local_block(blockIdx.x, gridDim.x);
    uint32_t lo = 0, hi = n_frames;  // tile_start[lo] <= tile < tile_start[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tile_start[mid] <= tile) lo = mid; else hi = mid;
    }
    const uint32_t fi = frame_of[lo];
    TileCtx t;
    t.f = &frames[fi];
    t.cells = cells;
    const rmh_plane& PY = t.f->planes[0];
    t.W = PY.width;
    t.H = PY.height;
    const int tiles_x = (t.W + kTW - 1) / kTW, k = (int)(tile - tile_start[lo]);
    t.X0 = (k % tiles_x) * kTW;
    t.Y0 = (k / tiles_x) * kTH;
    t.CX0 = t.X0 / 2;
    t.CY0 = t.Y0 / 2;
    t.CW = t.f->planes[1].width;
    t.CH = t.f->planes[1].height;
    t.bd = PY.bit_depth;
    const rmh_plane* sh = shadows + fi * 3;

    // 1. edge cells and windows
    if (threadIdx.x < kEY * kEX) {
        const rmh_dbk_frame& f = *t.f;
        const int cx = t.X0 / 4 + threadIdx.x % kEX, cy = t.Y0 / 4 - 2 + threadIdx.x / kEX;
        uint32_t v = 0;
        if (cx < (int)f.grid_w && cy >= 0 && cy < (int)f.grid_h) {
            const RMH_GLOBAL uint8_t* e = gsrc<uint8_t>(f.edges) + ((size_t)cy * f.grid_w + cx) * 4;
            if (((uintptr_t)f.edges & 3) == 0)
                v = *(const RMH_GLOBAL uint32_t*)e;
            else
                v = (uint32_t)e[0] | ((uint32_t)e[1] << 8) | ((uint32_t)e[2] << 16) | ((uint32_t)e[3] << 24);
        }
        cells[threadIdx.x] = v;
    }
    {
        // luma window in 8-sample chunks (10 per row), both chroma windows in
        // 4-sample chunks (10 per row); all loads are issued before the LDS writes
        constexpr int NL = kLWH * kLWW / 8, NC = kCWH * kCWW / 4;  // 480, 240
        const bool vl = chunks_aligned<8, BPS>(sh[0]);
        const bool vc = chunks_aligned<4, BPS>(sh[1]) && chunks_aligned<4, BPS>(sh[2]);
        Packed<8> a[2];
        Packed<4> b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int e = threadIdx.x + 256 * i;
            if (e < NL) {
                const int r = e / 10, c = e - 10 * r;
                a[i] = read_chunk<8, BPS>(sh[0], t.X0 - 8 + 8 * c, clampi(t.Y0 - 8 + r, 0, t.H - 1), vl);
            }
            if (e < 2 * NC) {
                const int pl = e >= NC, q = e - NC * pl, r = q / 10, c = q - 10 * r;
                b[i] = read_chunk<4, BPS>(sh[1 + pl], t.CX0 - 4 + 4 * c, clampi(t.CY0 - 4 + r, 0, t.CH - 1), vc);
            }
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int e = threadIdx.x + 256 * i;
            if (e < NL) put_chunk<8>(lw + 8 * e, a[i]);
            if (e < 2 * NC) put_chunk<4>(&cwin[0][0] + 4 * e, b[i]);  // the two planes: one array of 2 x 240 chunks
        }
    }
    __syncthreads();
    // 2. vertical edges
    tile_pass(t, 0, lw, cwin, ql, qc, qd, counts);
    // 3. rows outside the picture read as the nearest (filtered) picture row
    if (t.Y0 - 8 < 0 || t.Y0 + kTH + 8 > t.H) refill_rows<kLWW, kLWH>(t.Y0 - 8, t.H, lw);
    if (t.CY0 - 4 < 0 || t.CY0 + kCTH + 4 > t.CH) {
        refill_rows<kCWW, kCWH>(t.CY0 - 4, t.CH, cwin[0]);
        refill_rows<kCWW, kCWH>(t.CY0 - 4, t.CH, cwin[1]);
    }
    __syncthreads();
    // 4. horizontal edges
    tile_pass(t, 1, lw, cwin, ql, qc, qd, counts);
    // 5. write the tile back (picture samples only)
    // in 8-sample chunks: 256 luma chunks (8 per row), then 2 x 64 chroma chunks (4 per row)
    {
        constexpr int NL = kTW * kTH / 8, NC = kCTW * kCTH / 8;
        for (int e = threadIdx.x; e < NL + 2 * NC; e += 256) {
            if (e < NL) {
                const int r = e >> 3, c = e & 7, y = t.Y0 + r;
                if (y < t.H)
                    write_chunk8<BPS>(PY, t.X0 + 8 * c, y, lw + (r + 8) * kLWW + 8 * c + 8, chunks_aligned<8, BPS>(PY));
            } else {
                const int pl = (e - NL) >= NC, q = e - NL - NC * pl, r = q >> 2, c = q & 3, y = t.CY0 + r;
                const rmh_plane& P = t.f->planes[1 + pl];
                if (y < t.CH)
                    write_chunk8<BPS>(P, t.CX0 + 8 * c, y, cwin[pl] + (r + 4) * kCWW + 8 * c + 4,
                                      chunks_aligned<8, BPS>(P));
            }
        }
    }
}

}  // namespace

namespace {

// luma_lens (decode_general.rs:4916-4923): a side of at most 4 samples forces
// both lengths to 1, else 7 for a side >= 32 and 3 below. Returns the two
// length codes (length = 2 c + 1) as lenP | lenQ << 2.
__device__ __forceinline__ int luma_len_codes(int p, int q) {
    if (p <= 4 || q <= 4) return 0;
    return (p >= 32 ? 3 : 1) | ((q >= 32 ? 3 : 1) << 2);
}

// chroma_lens (decode_general.rs:4929-4935), as codes lenP | lenQ << 2.
__device__ __forceinline__ int chroma_len_codes(int p, int q, bool ctb) {
    if (p >= 8 && q >= 8) return ctb ? (0 | (1 << 2)) : (1 | (1 << 2));
    return 0;
}

// deblock_bs (decode_general.rs:4939-4975) over derive_boundary_strength
// (vvc/filter.rs:202-226) for a luma TU edge between cells p and q.
__device__ __forceinline__ int edge_bs(const rmh_edge_cell& p, const rmh_edge_cell& q) {
    if ((p.flags | q.flags) & RMH_CELL_INTRA) return 2;
    if ((p.flags | q.flags) & RMH_CELL_CBF) return 1;
    if (p.ref_poc != q.ref_poc || abs(p.mv_x - q.mv_x) >= 8 || abs(p.mv_y - q.mv_y) >= 8) return 1;
    return 0;
}

__global__ __launch_bounds__(256) void edge_map_kernel(const rmh_edge_frame* __restrict__ frames,
                                                       const uint32_t* __restrict__ cell_start, uint32_t n_frames,
                                                       uint32_t total) {
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= total) return;
    // the frame of cell g: binary search of the running sum
    uint32_t lo = 0, hi = n_frames;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cell_start[mid] <= g) lo = mid; else hi = mid;
    }
    const rmh_edge_frame f = frames[lo];
    const uint32_t k = g - cell_start[lo];
    const int i = (int)(k % f.grid_w), j = (int)(k / f.grid_w);
    const int gw = (int)f.grid_w;
    const rmh_edge_cell c = f.cells[k];
    const int x = 4 * i, y = 4 * j;
    const int cw = (int)((f.vis_w + 1) / 2), ch = (int)((f.vis_h + 1) / 2);  // visible chroma size
    uint32_t out = 0;
    // luma vertical edge at x, rows y .. y + 3 (x = 4, 8, ... < vis_w)
    if (i > 0 && x < (int)f.vis_w && y < (int)f.vis_h) {
        const rmh_edge_cell p = f.cells[k - 1];
        if (p.tu_y != 0xffffffffu && c.tu_y != 0xffffffffu && p.tu_y != c.tu_y)
            out |= (uint32_t)(edge_bs(p, c) | (luma_len_codes(p.tu_wy, c.tu_wy) << 2));
    }
    // luma horizontal edge at y, columns x .. x + 3
    if (j > 0 && y < (int)f.vis_h && x < (int)f.vis_w) {
        const rmh_edge_cell p = f.cells[k - gw];
        if (p.tu_y != 0xffffffffu && c.tu_y != 0xffffffffu && p.tu_y != c.tu_y)
            out |= (uint32_t)(edge_bs(p, c) | (luma_len_codes(p.tu_hy, c.tu_hy) << 2) |
                              ((y % (int)f.ctb_size) == 0 ? RMH_EDGE_CTB_ROW : 0)) << 8;
    }
    // chroma vertical edge at chroma x = 2 i (on the 8-sample chroma grid),
    // chroma rows 2 j, 2 j + 1 (while 2 j + 2 <= visible chroma height)
    const int xc = 2 * i, yc = 2 * j;
    if (i > 0 && (2 * xc) % 16 == 0 && xc < cw && yc + 2 <= ch) {
        const rmh_edge_cell p = f.cells[k - 1];
        if (p.tu_c != 0xffffffffu && c.tu_c != 0xffffffffu && p.tu_c != c.tu_c)
            out |= (uint32_t)(2 | (chroma_len_codes(p.tu_wc, c.tu_wc, false) << 2)) << 16;
    }
    // chroma horizontal edge at chroma y = 2 j, chroma columns 2 i, 2 i + 1
    if (j > 0 && (2 * yc) % 16 == 0 && yc < ch && xc + 2 <= cw) {
        const rmh_edge_cell p = f.cells[k - gw];
        if (p.tu_c != 0xffffffffu && c.tu_c != 0xffffffffu && p.tu_c != c.tu_c)
            out |= (uint32_t)(2 | (chroma_len_codes(p.tu_hc, c.tu_hc, (yc % ((int)f.ctb_size / 2)) == 0) << 2)) << 24;
    }
    gdst<uint32_t>(f.edges)[k] = out;
}

}  // namespace

hipError_t launch_edge_map(const rmh_edge_frame* frames, const uint32_t* cell_start, uint32_t n_frames,
                           uint32_t total_cells, hipStream_t stream) {
    if (total_cells == 0) return hipSuccess;
    hipLaunchKernelGGL(edge_map_kernel, dim3((total_cells + 255) / 256), dim3(256), 0, stream, frames, cell_start,
                       n_frames, total_cells);
    return hipGetLastError();
}

uint32_t deblock_tiles(uint32_t width, uint32_t height) {
    return ((width + kTW - 1) / kTW) * ((height + kTH - 1) / kTH);
}

hipError_t launch_deblock(const rmh_dbk_frame* frames, const rmh_plane* shadows, const uint32_t* tile_start,
                          const uint32_t* frame_of, uint32_t n_frames, uint32_t total_tiles, uint32_t tiles8,
                          uint32_t max_rows, bool copy_shadows, hipStream_t stream) {
    if (n_frames == 0 || total_tiles == 0) return hipSuccess;
    if (copy_shadows) {
        hipLaunchKernelGGL(shadow_copy_kernel, dim3((max_rows + 3) / 4, n_frames, 3), dim3(256), 0, stream, frames,
                           shadows);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (tiles8 > 0)
        hipLaunchKernelGGL(deblock_tile_kernel<1>, dim3(tiles8), dim3(256), 0, stream, frames, shadows, tile_start,
                           frame_of, n_frames, 0u);
    if (total_tiles > tiles8)
        hipLaunchKernelGGL(deblock_tile_kernel<2>, dim3(total_tiles - tiles8), dim3(256), 0, stream, frames, shadows,
                           tile_start, frame_of, n_frames, tiles8);
    return hipGetLastError();
}

}  // namespace rmh

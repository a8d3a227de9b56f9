/*
 * orc_deblock.c - deblocking stage of the CPU oracle (test infrastructure only).
 *
 * Restates the luma and chroma edge filters of
 * crates/rustmpeg-video/src/vvc/filter.rs:184-1057 and the whole-picture
 * drivers: apply_deblock (vvc/decode_general.rs:4977-5187), with the per-edge
 * bS and filter lengths supplied by an edge map, and deblock_luma_plane_16 /
 * deblock_chroma_plane_16 (vvc/filter.rs:2344-2527). Each driver keeps the
 * reference's order and its gather / filter / write-all-samples-back steps.
 */
#include <stdlib.h>
#include <string.h>

#include "vvc_oracle.h"

/* DBF_TC_TABLE / DBF_BETA_TABLE (vvc/filter.rs:184-198, H.266 Table 43). */
static const int TC_TABLE[66] = {0,  0,  0,  0,  0,  0,  0,  0,  0,   0,   0,   0,   0,   0,   0,   0,   0,
                                 0,  3,  4,  4,  4,  4,  5,  5,  5,   5,   7,   7,   8,   9,   10,  10,  11,
                                 13, 14, 15, 17, 19, 21, 24, 25, 29,  33,  36,  41,  45,  51,  57,  64,  71,
                                 80, 89, 100, 112, 125, 141, 157, 177, 198, 222, 250, 280, 314, 352, 395};
static const int BETA_TABLE[64] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,
                                   6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 22, 24,
                                   26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56,
                                   58, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78, 80, 82, 84, 86, 88};

static inline int iabs(int v) { return v < 0 ? -v : v; }
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

/* Line layout [p7..p0, q0..q7] (vvc/filter.rs:263-273). */
#define P(l, i) ((l)[7 - (i)])
#define Q(l, j) ((l)[8 + (j)])

void orc_tc_beta(int qp, int bs, int beta_off, int tc_off, int bd, int* tc, int* beta) {
    if (bd < 8) bd = 8;
    if (bd > 12) bd = 12;
    int idx_tc = clampi(qp + 2 * (bs - 1) + 2 * tc_off, 0, 65);
    int idx_b = clampi(qp + 2 * beta_off, 0, 63);
    *tc = bd < 10 ? (TC_TABLE[idx_tc] + (1 << (9 - bd))) >> (10 - bd) : TC_TABLE[idx_tc] << (bd - 10);
    *beta = BETA_TABLE[idx_b] << (bd - 8);
}

/* luma_sample_decision (vvc/filter.rs:276-304). */
static int luma_sample_decision(int p0, int p3, int q0, int q3, int dpq, int sp, int sq, int spq,
                                int side_p_large, int side_q_large, int beta, int tc) {
    if (side_p_large) sp = (sp + iabs(p3 - p0) + 1) >> 1;
    if (side_q_large) sq = (sq + iabs(q3 - q0) + 1) >> 1;
    int thr1, thr2;
    if (side_p_large || side_q_large) {
        thr1 = (3 * beta) >> 5;
        thr2 = beta >> 4;
    } else {
        thr1 = beta >> 3;
        thr2 = beta >> 2;
    }
    return dpq < thr2 && sp + sq < thr1 && spq < ((5 * tc + 1) >> 1);
}

typedef struct {
    int de, dep, deq, len_p, len_q, tc;
} luma_decision;

/* decide_luma_edge (vvc/filter.rs:311-468); returns 0 for "no filtering". */
static int decide_luma_edge(const int32_t* lines, int beta, int tc, int len_p_in, int len_q_in, int at_ctb,
                            luma_decision* d) {
    const int32_t* l0 = lines;
    const int32_t* l3 = lines + 3 * 16;
    int dp0 = iabs(P(l0, 2) - 2 * P(l0, 1) + P(l0, 0));
    int dp3 = iabs(P(l3, 2) - 2 * P(l3, 1) + P(l3, 0));
    int dq0 = iabs(Q(l0, 2) - 2 * Q(l0, 1) + Q(l0, 0));
    int dq3 = iabs(Q(l3, 2) - 2 * Q(l3, 1) + Q(l3, 0));
    int dpq0 = dp0 + dq0, dpq3 = dp3 + dq3;
    int sp0 = iabs(P(l0, 3) - P(l0, 0)), sp3 = iabs(P(l3, 3) - P(l3, 0));
    int sq0 = iabs(Q(l0, 0) - Q(l0, 3)), sq3 = iabs(Q(l3, 0) - Q(l3, 3));
    int spq0 = iabs(P(l0, 0) - Q(l0, 0)), spq3 = iabs(P(l3, 0) - Q(l3, 0));
    int len_p = clampi(len_p_in, 1, 7), len_q = clampi(len_q_in, 1, 7);
    int side_p_large = len_p > 3 && !at_ctb;
    int side_q_large = len_q > 3;

    if (side_p_large || side_q_large) {
        int dp0_l = side_p_large ? (dp0 + iabs(P(l0, 5) - 2 * P(l0, 4) + P(l0, 3)) + 1) >> 1 : dp0;
        int dp3_l = side_p_large ? (dp3 + iabs(P(l3, 5) - 2 * P(l3, 4) + P(l3, 3)) + 1) >> 1 : dp3;
        int dq0_l = side_q_large ? (dq0 + iabs(Q(l0, 5) - 2 * Q(l0, 4) + Q(l0, 3)) + 1) >> 1 : dq0;
        int dq3_l = side_q_large ? (dq3 + iabs(Q(l3, 5) - 2 * Q(l3, 4) + Q(l3, 3)) + 1) >> 1 : dq3;
        if (!side_p_large) len_p = 3;
        if (!side_q_large) len_q = 3;
        int sp0_l = len_p == 7 ? sp0 + iabs(P(l0, 7) - P(l0, 6) - P(l0, 5) + P(l0, 4)) : sp0;
        int sp3_l = len_p == 7 ? sp3 + iabs(P(l3, 7) - P(l3, 6) - P(l3, 5) + P(l3, 4)) : sp3;
        int sq0_l = len_q == 7 ? sq0 + iabs(Q(l0, 4) - Q(l0, 5) - Q(l0, 6) + Q(l0, 7)) : sq0;
        int sq3_l = len_q == 7 ? sq3 + iabs(Q(l3, 4) - Q(l3, 5) - Q(l3, 6) + Q(l3, 7)) : sq3;
        int d0_l = dp0_l + dq0_l, d3_l = dp3_l + dq3_l;
        if (d0_l + d3_l < beta) {
            int s0 = luma_sample_decision(side_p_large ? P(l0, len_p) : 0, side_p_large ? P(l0, 3) : 0,
                                          side_q_large ? Q(l0, len_q) : 0, side_q_large ? Q(l0, 3) : 0,
                                          2 * d0_l, sp0_l, sq0_l, spq0, side_p_large, side_q_large, beta, tc);
            int s3 = luma_sample_decision(side_p_large ? P(l3, len_p) : 0, side_p_large ? P(l3, 3) : 0,
                                          side_q_large ? Q(l3, len_q) : 0, side_q_large ? Q(l3, 3) : 0,
                                          2 * d3_l, sp3_l, sq3_l, spq3, side_p_large, side_q_large, beta, tc);
            if (s0 && s3) {
                d->de = 3;
                d->dep = d->deq = 1;
                d->len_p = len_p;
                d->len_q = len_q;
                d->tc = tc;
                return 1;
            }
        }
    }
    int dp = dp0 + dp3, dq = dq0 + dq3;
    if (dpq0 + dpq3 >= beta) return 0;
    int s0 = 0, s3 = 0;
    if (len_p > 2 && len_q > 2) {
        s0 = luma_sample_decision(0, 0, 0, 0, 2 * dpq0, sp0, sq0, spq0, 0, 0, beta, tc);
        s3 = luma_sample_decision(0, 0, 0, 0, 2 * dpq3, sp3, sq3, spq3, 0, 0, beta, tc);
    }
    int de = 1;
    if (s0 && s3) {
        de = 2;
        len_p = 3;
        len_q = 3;
    }
    int side_thr = (beta + (beta >> 1)) >> 3;
    int dep = len_p > 1 && len_q > 1 && dp < side_thr;
    int deq = len_p > 1 && len_q > 1 && dq < side_thr;
    if (de == 1) {
        len_p = 1 + dep;
        len_q = 1 + deq;
    }
    d->de = de;
    d->dep = dep;
    d->deq = deq;
    d->len_p = len_p;
    d->len_q = len_q;
    d->tc = tc;
    return 1;
}

/* long_tap_luma_line (vvc/filter.rs:477-582), VTM/VVdeC refMiddle for 7/3 and 3/7. */
static void long_tap_line(int32_t* line, int len_p, int len_q, int tc, int max_val) {
    int32_t m[16];
    memcpy(m, line, sizeof(m));
    int np = len_p < 7 ? len_p : 7, nq = len_q < 7 ? len_q : 7;
#define p(i) m[7 - (i)]
#define q(j) m[8 + (j)]
    int ref_middle;
    if (np == 5 && nq == 5)
        ref_middle = (2 * (p(0) + q(0) + p(1) + q(1) + p(2) + q(2)) + p(3) + q(3) + p(4) + q(4) + 8) >> 4;
    else if ((np == 7 && nq == 5) || (np == 5 && nq == 7))
        ref_middle = (2 * (p(0) + q(0) + p(1) + q(1)) + p(2) + q(2) + p(3) + q(3) + p(4) + q(4) + p(5) +
                      q(5) + 8) >> 4;
    else if ((np == 5 && nq == 3) || (np == 3 && nq == 5))
        ref_middle = (p(0) + q(0) + p(1) + q(1) + p(2) + q(2) + p(3) + q(3) + 4) >> 3;
    else if (np == 7 && nq == 3)
        ref_middle = (2 * (p(0) + q(0)) + q(0) + 2 * (q(1) + q(2)) + p(1) + q(1) + p(2) + p(3) + p(4) +
                      p(5) + p(6) + 8) >> 4;
    else if (np == 3 && nq == 7)
        ref_middle = (2 * (p(0) + q(0)) + p(0) + 2 * (p(1) + p(2)) + p(1) + q(1) + q(2) + q(3) + q(4) +
                      q(5) + q(6) + 8) >> 4;
    else
        ref_middle = (2 * (p(0) + q(0)) + p(1) + q(1) + p(2) + q(2) + p(3) + q(3) + p(4) + q(4) + p(5) +
                      q(5) + p(6) + q(6) + 8) >> 4;
    int ref_p = (p(np) + p(np - 1) + 1) >> 1;
    int ref_q = (q(nq) + q(nq - 1) + 1) >> 1;
    static const int C7[7] = {59, 50, 41, 32, 23, 14, 5}, T7[7] = {6, 5, 4, 3, 2, 1, 1};
    static const int C5[5] = {58, 45, 32, 19, 6}, T5[5] = {6, 5, 4, 3, 2};
    static const int C3[3] = {53, 32, 11}, T3[3] = {6, 4, 2};
    const int* cp = np == 7 ? C7 : np == 5 ? C5 : C3;
    const int* tp = np == 7 ? T7 : np == 5 ? T5 : T3;
    const int* cq = nq == 7 ? C7 : nq == 5 ? C5 : C3;
    const int* tq = nq == 7 ? T7 : nq == 5 ? T5 : T3;
    int lp = np == 7 ? 7 : np == 5 ? 5 : 3, lq = nq == 7 ? 7 : nq == 5 ? 5 : 3;
    for (int pos = 0; pos < lp; pos++) {
        int orig = p(pos), cval = (tc * tp[pos]) >> 1;
        int interp = (ref_middle * cp[pos] + ref_p * (64 - cp[pos]) + 32) >> 6;
        line[7 - pos] = clampi(clampi(interp, orig - cval, orig + cval), 0, max_val);
    }
    for (int pos = 0; pos < lq; pos++) {
        int orig = q(pos), cval = (tc * tq[pos]) >> 1;
        int interp = (ref_middle * cq[pos] + ref_q * (64 - cq[pos]) + 32) >> 6;
        line[8 + pos] = clampi(clampi(interp, orig - cval, orig + cval), 0, max_val);
    }
#undef p
#undef q
}

/* short_filter_luma_line (vvc/filter.rs:586-631). */
static void short_line(int32_t* line, const luma_decision* e, int max_val) {
    int32_t m[16];
    memcpy(m, line, sizeof(m));
    int tc = e->tc;
    if (e->de == 2) {
        line[7] = clampi((P(m, 2) + 2 * P(m, 1) + 2 * P(m, 0) + 2 * Q(m, 0) + Q(m, 1) + 4) >> 3, P(m, 0) - 3 * tc,
                         P(m, 0) + 3 * tc);
        line[6] = clampi((P(m, 2) + P(m, 1) + P(m, 0) + Q(m, 0) + 2) >> 2, P(m, 1) - 2 * tc, P(m, 1) + 2 * tc);
        line[5] = clampi((2 * P(m, 3) + 3 * P(m, 2) + P(m, 1) + P(m, 0) + Q(m, 0) + 4) >> 3, P(m, 2) - tc,
                         P(m, 2) + tc);
        line[8] = clampi((P(m, 1) + 2 * P(m, 0) + 2 * Q(m, 0) + 2 * Q(m, 1) + Q(m, 2) + 4) >> 3, Q(m, 0) - 3 * tc,
                         Q(m, 0) + 3 * tc);
        line[9] = clampi((P(m, 0) + Q(m, 0) + Q(m, 1) + Q(m, 2) + 2) >> 2, Q(m, 1) - 2 * tc, Q(m, 1) + 2 * tc);
        line[10] = clampi((P(m, 0) + Q(m, 0) + Q(m, 1) + 3 * Q(m, 2) + 2 * Q(m, 3) + 4) >> 3, Q(m, 2) - tc,
                          Q(m, 2) + tc);
        return;
    }
    int delta = (9 * (Q(m, 0) - P(m, 0)) - 3 * (Q(m, 1) - P(m, 1)) + 8) >> 4;
    if (iabs(delta) >= tc * 10) return;
    delta = clampi(delta, -tc, tc);
    line[7] = clampi(P(m, 0) + delta, 0, max_val);
    line[8] = clampi(Q(m, 0) - delta, 0, max_val);
    int tc2 = tc >> 1;
    if (e->dep) {
        int dp = clampi(((((P(m, 2) + P(m, 0) + 1) >> 1) - P(m, 1) + delta) >> 1), -tc2, tc2);
        line[6] = clampi(P(m, 1) + dp, 0, max_val);
    }
    if (e->deq) {
        int dq = clampi(((((Q(m, 2) + Q(m, 0) + 1) >> 1) - Q(m, 1) - delta) >> 1), -tc2, tc2);
        line[9] = clampi(Q(m, 1) + dq, 0, max_val);
    }
}

/* The per-line luma filters and the sample decisions, exported so that tests
 * can compare them one by one with the reference's expressions. */
void orc_dbk_long_line(int32_t* line, int len_p, int len_q, int tc, int max_val) {
    long_tap_line(line, len_p, len_q, tc, max_val);
}

void orc_dbk_short_line(int32_t* line, int de, int dep, int deq, int tc, int max_val) {
    luma_decision d = {de, dep, deq, 0, 0, tc};
    short_line(line, &d, max_val);
}

int orc_dbk_luma_sample_decision(int p0, int p3, int q0, int q3, int dpq, int sp, int sq, int spq, int side_p_large,
                                 int side_q_large, int beta, int tc) {
    return luma_sample_decision(p0, p3, q0, q3, dpq, sp, sq, spq, side_p_large, side_q_large, beta, tc);
}

/* deblock_luma_edge_bd (vvc/filter.rs:638-679). */
void orc_deblock_luma_edge(int32_t* lines, int qp_p, int qp_q, int bs, int beta_off, int tc_off, int bd,
                           int len_p, int len_q, int at_ctb) {
    if (bs == 0) return;
    if (bd < 8) bd = 8;
    if (bd > 12) bd = 12;
    int max_val = (1 << bd) - 1;
    int qp = (qp_p + qp_q + 1) >> 1, tc, beta;
    orc_tc_beta(qp, bs, beta_off, tc_off, bd, &tc, &beta);
    if (beta == 0 || tc == 0) return;
    luma_decision d;
    if (!decide_luma_edge(lines, beta, tc, len_p, len_q, at_ctb, &d)) return;
    for (int l = 0; l < 4; l++) {
        if (d.de == 3)
            long_tap_line(lines + 16 * l, d.len_p, d.len_q, d.tc, max_val);
        else
            short_line(lines + 16 * l, &d, max_val);
    }
}

/* One chroma line [p3..p0, q0..q3] by the filter form the lengths select:
 * chroma_strong_line (3, 3), chroma_weak_line_p1q3 (1, 3), chroma_weak_line
 * otherwise (vvc/filter.rs:986-1057). */
void orc_dbk_chroma_line(int32_t* line, int len_p, int len_q, int tc, int max_val) {
    int32_t m[8];
    memcpy(m, line, sizeof(m));
#define CL(o, v) clampi(clampi((v), (o)-tc, (o) + tc), 0, max_val)
    if (len_p == 3 && len_q == 3) {
        line[1] = CL(m[1], (3 * m[0] + 2 * m[1] + m[2] + m[3] + m[4] + 4) >> 3);
        line[2] = CL(m[2], (2 * m[0] + m[1] + 2 * m[2] + m[3] + m[4] + m[5] + 4) >> 3);
        line[3] = CL(m[3], (m[0] + m[1] + m[2] + 2 * m[3] + m[4] + m[5] + m[6] + 4) >> 3);
        line[4] = CL(m[4], (m[1] + m[2] + m[3] + 2 * m[4] + m[5] + m[6] + m[7] + 4) >> 3);
        line[5] = CL(m[5], (m[2] + m[3] + m[4] + 2 * m[5] + m[6] + 2 * m[7] + 4) >> 3);
        line[6] = CL(m[6], (m[3] + m[4] + m[5] + 2 * m[6] + 3 * m[7] + 4) >> 3);
    } else if (len_p == 1 && len_q == 3) {
        line[3] = CL(m[3], (3 * m[2] + 2 * m[3] + m[4] + m[5] + m[6] + 4) >> 3);
        line[4] = CL(m[4], (2 * m[2] + m[3] + 2 * m[4] + m[5] + m[6] + m[7] + 4) >> 3);
        line[5] = CL(m[5], (m[2] + m[3] + m[4] + 2 * m[5] + m[6] + 2 * m[7] + 4) >> 3);
        line[6] = CL(m[6], (m[3] + m[4] + m[5] + 2 * m[6] + 3 * m[7] + 4) >> 3);
    } else {
        int delta = clampi((4 * (m[4] - m[3]) + m[2] - m[5] + 4) >> 3, -tc, tc);
        line[3] = clampi(m[3] + delta, 0, max_val);
        line[4] = clampi(m[4] - delta, 0, max_val);
    }
#undef CL
}

/* chroma_sample_decision (vvc/filter.rs:907-919). */
static int chroma_sample_decision(int p0, int p3, int q0, int q3, int dpq, int beta, int tc) {
    return dpq < (beta >> 2) && iabs(p3 - p0) + iabs(q0 - q3) < (beta >> 3) && iabs(p0 - q0) < ((5 * tc + 1) >> 1);
}

int orc_dbk_chroma_sample_decision(int p0, int p3, int q0, int q3, int dpq, int beta, int tc) {
    return chroma_sample_decision(p0, p3, q0, q3, dpq, beta, tc);
}

/* deblock_chroma_edge (vvc/filter.rs:935-1057); lines of [p3..p0, q0..q3]. */
void orc_deblock_chroma_edge(int32_t* lines, int n, int qp, int bs, int beta_off, int tc_off, int len_p, int len_q,
                             int bd) {
    if (bs == 0 || (n != 2 && n != 4)) return;
    len_p = len_p < 3 ? len_p : 3;
    len_q = len_q < 3 ? len_q : 3;
    if (len_p == 1 && len_q == 1 && bs != 2) return;
    if (bd < 8) bd = 8;
    if (bd > 12) bd = 12;
    int max_val = (1 << bd) - 1, tc, beta;
    orc_tc_beta(qp, bs, beta_off, tc_off, bd, &tc, &beta);
    if (len_q == 3) {
        int32_t a[8], b[8];
        memcpy(a, lines, sizeof(a));
        memcpy(b, lines + 8 * (n - 1), sizeof(b));
        if (len_p == 1) {
            a[0] = a[1] = a[2];
            b[0] = b[1] = b[2];
        }
        int dp0 = iabs(a[1] - 2 * a[2] + a[3]), dp1 = iabs(b[1] - 2 * b[2] + b[3]);
        int dq0 = iabs(a[6] - 2 * a[5] + a[4]), dq1 = iabs(b[6] - 2 * b[5] + b[4]);
        int dpq0 = dp0 + dq0, dpq1 = dp1 + dq1;
        int strong = (dpq0 + dpq1) < beta && chroma_sample_decision(a[3], a[0], a[4], a[7], 2 * dpq0, beta, tc) &&
                     chroma_sample_decision(b[3], b[0], b[4], b[7], 2 * dpq1, beta, tc);
        if (!strong) {
            len_p = 1;
            len_q = 1;
        }
    }
    for (int l = 0; l < n; l++) orc_dbk_chroma_line(lines + 8 * l, len_p, len_q, tc, max_val);
}

static inline int pget(const rmh_plane* p, int x, int y) {
    if (p->bytes_per_sample == 1) return ((const uint8_t*)p->data)[(size_t)y * p->stride + x];
    return ((const uint16_t*)p->data)[(size_t)y * p->stride + x];
}

static inline void pset(const rmh_plane* p, int x, int y, int v) {
    if (p->bytes_per_sample == 1)
        ((uint8_t*)p->data)[(size_t)y * p->stride + x] = (uint8_t)v;
    else
        ((uint16_t*)p->data)[(size_t)y * p->stride + x] = (uint16_t)v;
}

static inline int lenc(int code) { return 2 * code + 1; }

/* apply_deblock (decode_general.rs:4977-5187) with bS and lengths from the edge map. */
static void deblock_frame(const rmh_dbk_frame* f) {
    const rmh_plane* Y = &f->planes[0];
    int w = (int)Y->width, h = (int)Y->height, bd = Y->bit_depth, maxv = (1 << bd) - 1;
    uint32_t gw = f->grid_w;
    int32_t lines[64];
    /* Luma vertical edges: x ascending, then y. */
    for (int x = 4; x < w; x += 4)
        for (int y = 0; y < h; y += 4) {
            uint8_t e = f->edges[((size_t)(y / 4) * gw + (size_t)(x / 4)) * 4 + 0];
            int bs = e & RMH_EDGE_BS_MASK;
            if (!bs) continue;
            for (int i = 0; i < 4; i++) {
                int yy = y + i < h - 1 ? y + i : h - 1;
                for (int k = 0; k < 16; k++) lines[i * 16 + k] = pget(Y, clampi(x - 8 + k, 0, w - 1), yy);
            }
            orc_deblock_luma_edge(lines, f->qp[0], f->qp[0], bs, f->beta_offset_div2[0], f->tc_offset_div2[0], bd,
                                  lenc((e >> RMH_EDGE_LENP_SHIFT) & 3), lenc((e >> RMH_EDGE_LENQ_SHIFT) & 3), 0);
            for (int i = 0; i < 4; i++) {
                int yy = y + i < h - 1 ? y + i : h - 1;
                for (int k = 0; k < 16; k++) pset(Y, clampi(x - 8 + k, 0, w - 1), yy, clampi(lines[i * 16 + k], 0, maxv));
            }
        }
    /* Luma horizontal edges: y ascending, then x. */
    for (int y = 4; y < h; y += 4)
        for (int x = 0; x < w; x += 4) {
            uint8_t e = f->edges[((size_t)(y / 4) * gw + (size_t)(x / 4)) * 4 + 1];
            int bs = e & RMH_EDGE_BS_MASK;
            if (!bs) continue;
            for (int i = 0; i < 4; i++) {
                int xx = x + i < w - 1 ? x + i : w - 1;
                for (int k = 0; k < 16; k++) lines[i * 16 + k] = pget(Y, xx, clampi(y - 8 + k, 0, h - 1));
            }
            orc_deblock_luma_edge(lines, f->qp[0], f->qp[0], bs, f->beta_offset_div2[0], f->tc_offset_div2[0], bd,
                                  lenc((e >> RMH_EDGE_LENP_SHIFT) & 3), lenc((e >> RMH_EDGE_LENQ_SHIFT) & 3),
                                  (e & RMH_EDGE_CTB_ROW) != 0);
            for (int i = 0; i < 4; i++) {
                int xx = x + i < w - 1 ? x + i : w - 1;
                for (int k = 0; k < 16; k++) pset(Y, xx, clampi(y - 8 + k, 0, h - 1), clampi(lines[i * 16 + k], 0, maxv));
            }
        }
    /* Chroma: Cb then Cr; vertical then horizontal; two-line decision groups. */
    for (int c = 1; c < 3; c++) {
        const rmh_plane* C = &f->planes[c];
        int cw = (int)C->width, ch = (int)C->height, cbd = C->bit_depth, cmax = (1 << cbd) - 1;
        for (int x = 2; x < cw; x += 2)
            for (int y = 0; y + 2 <= ch; y += 2) {
                uint8_t e = f->edges[((size_t)(y / 2) * gw + (size_t)(x / 2)) * 4 + 2];
                int bs = e & RMH_EDGE_BS_MASK;
                if (!bs) continue;
                for (int i = 0; i < 2; i++) {
                    int yy = y + i < ch - 1 ? y + i : ch - 1;
                    for (int k = 0; k < 8; k++) lines[i * 8 + k] = pget(C, clampi(x - 4 + k, 0, cw - 1), yy);
                }
                orc_deblock_chroma_edge(lines, 2, f->qp[c], bs, f->beta_offset_div2[c], f->tc_offset_div2[c],
                                        lenc((e >> RMH_EDGE_LENP_SHIFT) & 3), lenc((e >> RMH_EDGE_LENQ_SHIFT) & 3), cbd);
                for (int i = 0; i < 2; i++) {
                    int yy = y + i < ch - 1 ? y + i : ch - 1;
                    for (int k = 0; k < 8; k++) pset(C, clampi(x - 4 + k, 0, cw - 1), yy, clampi(lines[i * 8 + k], 0, cmax));
                }
            }
        for (int y = 2; y < ch; y += 2)
            for (int x = 0; x + 2 <= cw; x += 2) {
                uint8_t e = f->edges[((size_t)(y / 2) * gw + (size_t)(x / 2)) * 4 + 3];
                int bs = e & RMH_EDGE_BS_MASK;
                if (!bs) continue;
                for (int i = 0; i < 2; i++) {
                    int xx = x + i < cw - 1 ? x + i : cw - 1;
                    for (int k = 0; k < 8; k++) lines[i * 8 + k] = pget(C, xx, clampi(y - 4 + k, 0, ch - 1));
                }
                orc_deblock_chroma_edge(lines, 2, f->qp[c], bs, f->beta_offset_div2[c], f->tc_offset_div2[c],
                                        lenc((e >> RMH_EDGE_LENP_SHIFT) & 3), lenc((e >> RMH_EDGE_LENQ_SHIFT) & 3), cbd);
                for (int i = 0; i < 2; i++) {
                    int xx = x + i < cw - 1 ? x + i : cw - 1;
                    for (int k = 0; k < 8; k++) pset(C, xx, clampi(y - 4 + k, 0, ch - 1), clampi(lines[i * 8 + k], 0, cmax));
                }
            }
    }
}

extern int orc_threads;

size_t orc_deblock_frames(const rmh_dbk_frame* frames, size_t n_frames) {
    /* frames are independent; within a frame the reference's order holds */
#pragma omp parallel for num_threads(orc_threads) schedule(dynamic, 1) if (orc_threads > 1)
    for (size_t i = 0; i < n_frames; i++) deblock_frame(&frames[i]);
    return 0;
}

/* luma_filter_length (vvc/filter.rs:235-243). */
static int luma_filter_length(int size) { return size <= 4 ? 1 : size >= 32 ? 7 : 3; }

/* deblock_luma_plane_16 (vvc/filter.rs:2344-2431). */
void orc_deblock_luma_plane(uint16_t* plane, int w, int h, int bd, int cu_step, int ctb, int qp, int bs,
                            int beta_off, int tc_off) {
    int step = cu_step > 8 ? cu_step : 8;
    ctb = ctb > step ? ctb : step;
    int max_len = luma_filter_length(step), maxv = (1 << bd) - 1;
    int32_t lines[64];
    for (int x = step; x + 8 <= w && x >= 8; x += step)
        for (int y = 0; y + 4 <= h; y += 4) {
            for (int i = 0; i < 4; i++)
                for (int k = 0; k < 16; k++) lines[i * 16 + k] = plane[(size_t)(y + i) * w + x - 8 + k];
            orc_deblock_luma_edge(lines, qp, qp, bs, beta_off, tc_off, bd, max_len, max_len, 0);
            for (int i = 0; i < 4; i++)
                for (int k = 0; k < 16; k++) plane[(size_t)(y + i) * w + x - 8 + k] = (uint16_t)clampi(lines[i * 16 + k], 0, maxv);
        }
    for (int y = step; y + 8 <= h && y >= 8; y += step) {
        int at_ctb = y % ctb == 0;
        for (int bx = 0; bx + 4 <= w; bx += 4) {
            for (int i = 0; i < 4; i++)
                for (int k = 0; k < 16; k++) lines[i * 16 + k] = plane[(size_t)(y - 8 + k) * w + bx + i];
            orc_deblock_luma_edge(lines, qp, qp, bs, beta_off, tc_off, bd, max_len, max_len, at_ctb);
            for (int i = 0; i < 4; i++)
                for (int k = 0; k < 16; k++) plane[(size_t)(y - 8 + k) * w + bx + i] = (uint16_t)clampi(lines[i * 16 + k], 0, maxv);
        }
    }
}

/* deblock_chroma_plane_16 (vvc/filter.rs:2433-2527) for 4:2:0. */
void orc_deblock_chroma_plane(uint16_t* plane, int w, int h, int bd, int cu_step, int ctb_c, int qp, int bs,
                              int beta_off, int tc_off) {
    if (bs == 0 || w == 0 || h == 0) return;
    int maxv = (1 << bd) - 1;
    int step = cu_step > 8 ? cu_step : 8;
    int ctb = ctb_c > step ? ctb_c : step;
    int len = step >= 8 ? 3 : 1;
    int32_t lines[16];
    for (int x = step; x >= 4 && x + 4 <= w; x += step)
        for (int y = 0; y + 2 <= h; y += 2) {
            for (int i = 0; i < 2; i++)
                for (int k = 0; k < 8; k++) lines[i * 8 + k] = plane[(size_t)(y + i) * w + x - 4 + k];
            orc_deblock_chroma_edge(lines, 2, qp, bs, beta_off, tc_off, len, len, bd);
            for (int i = 0; i < 2; i++)
                for (int k = 0; k < 8; k++) plane[(size_t)(y + i) * w + x - 4 + k] = (uint16_t)clampi(lines[i * 8 + k], 0, maxv);
        }
    for (int y = step; y >= 4 && y + 4 <= h; y += step) {
        int lp = y % ctb == 0 ? 1 : len;
        for (int bx = 0; bx + 2 <= w; bx += 2) {
            for (int i = 0; i < 2; i++)
                for (int k = 0; k < 8; k++) lines[i * 8 + k] = plane[(size_t)(y - 4 + k) * w + bx + i];
            orc_deblock_chroma_edge(lines, 2, qp, bs, beta_off, tc_off, lp, len, bd);
            for (int i = 0; i < 2; i++)
                for (int k = 0; k < 8; k++) plane[(size_t)(y - 4 + k) * w + bx + i] = (uint16_t)clampi(lines[i * 8 + k], 0, maxv);
        }
    }
}

/* ---- edge map: the edges apply_deblock filters, with their bS and lengths ---- */

/* luma_lens (vvc/decode_general.rs:4916-4923). */
static void luma_lens(int size_p, int size_q, int* lp, int* lq) {
    if (size_p <= 4 || size_q <= 4) {
        *lp = *lq = 1;
        return;
    }
    *lp = size_p >= 32 ? 7 : 3;
    *lq = size_q >= 32 ? 7 : 3;
}

/* chroma_lens (vvc/decode_general.rs:4929-4935). */
static void chroma_lens(int size_p, int size_q, int ctb_boundary, int* lp, int* lq) {
    if (size_p >= 8 && size_q >= 8) {
        *lp = ctb_boundary ? 1 : 3;
        *lq = 3;
    } else {
        *lp = *lq = 1;
    }
}

/* derive_boundary_strength (vvc/filter.rs:202-226). */
static int derive_boundary_strength(int p_intra, int q_intra, int p_cbf, int q_cbf, int is_tu_edge, int mvp_x,
                                    int mvp_y, int mvq_x, int mvq_y, int ref_p, int ref_q) {
    if (p_intra || q_intra) return 2;
    if (is_tu_edge && (p_cbf || q_cbf)) return 1;
    if (ref_p != ref_q || iabs(mvp_x - mvq_x) >= 8 || iabs(mvp_y - mvq_y) >= 8) return 1;
    return 0;
}

/* deblock_bs (vvc/decode_general.rs:4939-4975) for the cells either side of
 * a luma TU edge (the cells carry the MV and POC that `field` selects). */
static int deblock_bs(const rmh_edge_cell* p, const rmh_edge_cell* q) {
    return derive_boundary_strength(p->flags & RMH_CELL_INTRA, q->flags & RMH_CELL_INTRA, p->flags & RMH_CELL_CBF,
                                    q->flags & RMH_CELL_CBF, 1, p->mv_x, p->mv_y, q->mv_x, q->mv_y, p->ref_poc,
                                    q->ref_poc);
}

static uint8_t edge_byte(int bs, int lp, int lq, int ctb) {
    return (uint8_t)((bs & 3) | (((lp - 1) / 2) << 2) | (((lq - 1) / 2) << 4) | (ctb ? RMH_EDGE_CTB_ROW : 0));
}

/* The edge loops of apply_deblock (vvc/decode_general.rs:4977-5187): luma
 * edges on the 4-sample grid where the TU ids differ; chroma edges on the
 * chroma 2-sample grid, kept only on the 8-sample chroma grid, where the
 * chroma TU ids differ, always with bs = 2. Every byte of the map is written
 * (0 where no edge is filtered). */
size_t orc_edge_map(const rmh_edge_frame* frames, size_t n_frames) {
    for (size_t fi = 0; fi < n_frames; fi++) {
        const rmh_edge_frame* f = &frames[fi];
        const int gw = (int)f->grid_w, w = (int)f->vis_w, h = (int)f->vis_h;
        const rmh_edge_cell* cell = f->cells;
        uint8_t* e = f->edges;
        memset(e, 0, (size_t)gw * f->grid_h * 4);
        int lp, lq;
        for (int x = 4; x < w; x += 4)
            for (int y = 0; y < h; y += 4) {
                const rmh_edge_cell *l = &cell[(y / 4) * gw + x / 4 - 1], *r = &cell[(y / 4) * gw + x / 4];
                if (l->tu_y != UINT32_MAX && r->tu_y != UINT32_MAX && l->tu_y != r->tu_y) {
                    luma_lens(l->tu_wy, r->tu_wy, &lp, &lq);
                    e[((y / 4) * gw + x / 4) * 4 + 0] = edge_byte(deblock_bs(l, r), lp, lq, 0);
                }
            }
        for (int y = 4; y < h; y += 4)
            for (int x = 0; x < w; x += 4) {
                const rmh_edge_cell *t = &cell[(y / 4 - 1) * gw + x / 4], *b = &cell[(y / 4) * gw + x / 4];
                if (t->tu_y != UINT32_MAX && b->tu_y != UINT32_MAX && t->tu_y != b->tu_y) {
                    luma_lens(t->tu_hy, b->tu_hy, &lp, &lq);
                    e[((y / 4) * gw + x / 4) * 4 + 1] = edge_byte(deblock_bs(t, b), lp, lq, y % (int)f->ctb_size == 0);
                }
            }
        const int cw = (w + 1) / 2, ch = (h + 1) / 2;
        for (int x = 2; x < cw; x += 2) {
            if ((2 * x) % 16 != 0) continue;
            for (int y = 0; y + 2 <= ch; y += 2) {
                const rmh_edge_cell *l = &cell[(y / 2) * gw + x / 2 - 1], *r = &cell[(y / 2) * gw + x / 2];
                if (l->tu_c != UINT32_MAX && r->tu_c != UINT32_MAX && l->tu_c != r->tu_c) {
                    chroma_lens(l->tu_wc, r->tu_wc, 0, &lp, &lq);
                    e[((y / 2) * gw + x / 2) * 4 + 2] = edge_byte(2, lp, lq, 0);
                }
            }
        }
        for (int y = 2; y < ch; y += 2) {
            if ((2 * y) % 16 != 0) continue;
            for (int x = 0; x + 2 <= cw; x += 2) {
                const rmh_edge_cell *t = &cell[(y / 2 - 1) * gw + x / 2], *b = &cell[(y / 2) * gw + x / 2];
                if (t->tu_c != UINT32_MAX && b->tu_c != UINT32_MAX && t->tu_c != b->tu_c) {
                    chroma_lens(t->tu_hc, b->tu_hc, y % ((int)f->ctb_size / 2) == 0, &lp, &lq);
                    e[((y / 2) * gw + x / 2) * 4 + 3] = edge_byte(2, lp, lq, 0);
                }
            }
        }
    }
    return 0;
}

/*
 * orc_mc.c - inter prediction stage of the CPU oracle (test infrastructure only).
 *
 * Restates interpolate_luma_block / interpolate_chroma_block_420
 * (crates/rustmpeg-video/src/vvc/inter.rs:886-1182), generalised from 8 bits to
 * bit depth bd with head_room = IF_INTERNAL_PREC - bd (for bd = 8 the function is
 * the reference's 8-bit one term for term), the uni-prediction path and the
 * bi-prediction path of predict_inter_pu (inter.rs:1734-1848) with its
 * 8-bit refinements DMVR (refine_dmvr_subblock, inter.rs:1341-1459) and BDOF
 * (apply_bdof_luma, inter.rs:1204-1339), weighted_average_bi / weighted_average_bi_16
 * (inter.rs:1189-1201, 2676-2696) and write_prediction
 * (decode_general.rs:2641-2667).
 *
 * The reference's 16-bit functions (inter.rs:2010-2134, 2543-2671) use
 * shift2 = 12 - bd for the uni-prediction two-dimensional case, which
 * saturates every non-zero sample; orc_interp_luma_16_as_written restates that
 * function as written so a test can show the defect and show that every other
 * branch agrees with the generalised function.
 */
#include <stdlib.h>
#include <string.h>

#include "vvc_oracle.h"

#define IF_INTERNAL_PREC 14
#define IF_FILTER_PREC 6
#define IF_INTERNAL_OFFS 8192

/* LUMA_FILTER_8TAP (inter.rs:57-74, H.266 Table 33). */
static const int LUMA8[16][8] = {
    {0, 0, 0, 64, 0, 0, 0, 0},        {0, 1, -3, 63, 4, -2, 1, 0},     {-1, 2, -5, 62, 8, -3, 1, 0},
    {-1, 3, -8, 60, 13, -4, 1, 0},    {-1, 4, -10, 58, 17, -5, 1, 0},  {-1, 4, -11, 52, 26, -8, 3, -1},
    {-1, 3, -9, 47, 31, -10, 4, -1},  {-1, 4, -11, 45, 34, -10, 4, -1}, {-1, 4, -11, 40, 40, -11, 4, -1},
    {-1, 4, -10, 34, 45, -11, 4, -1}, {-1, 4, -10, 31, 47, -9, 3, -1}, {-1, 3, -8, 26, 52, -11, 4, -1},
    {0, 1, -5, 17, 58, -10, 4, -1},   {0, 1, -4, 13, 60, -8, 3, -1},   {0, 1, -3, 8, 62, -5, 2, -1},
    {0, 1, -2, 4, 63, -3, 1, 0}};
/* LUMA_ALT_HPEL_FILTER_8TAP (inter.rs:80). */
static const int LUMA_ALT_HPEL[8] = {0, 3, 9, 20, 20, 9, 3, 0};
/* CHROMA_FILTER_4TAP (vvc/predict.rs:86-119, H.266 Table 34). */
static const int CHROMA4[32][4] = {
    {0, 64, 0, 0},    {-1, 63, 2, 0},   {-2, 62, 4, 0},   {-2, 60, 7, -1},  {-2, 58, 10, -2},
    {-3, 57, 12, -2}, {-4, 56, 14, -2}, {-4, 55, 15, -2}, {-4, 54, 16, -2}, {-5, 53, 18, -2},
    {-6, 52, 20, -2}, {-6, 49, 24, -3}, {-6, 46, 28, -4}, {-5, 44, 29, -4}, {-4, 42, 30, -4},
    {-4, 39, 33, -4}, {-4, 36, 36, -4}, {-4, 33, 39, -4}, {-4, 30, 42, -4}, {-4, 29, 44, -5},
    {-4, 28, 46, -6}, {-3, 24, 49, -6}, {-2, 20, 52, -6}, {-2, 18, 53, -5}, {-2, 16, 54, -4},
    {-2, 15, 55, -4}, {-2, 14, 56, -4}, {-2, 12, 57, -3}, {-2, 10, 58, -2}, {-1, 7, 60, -2},
    {0, 4, 62, -2},   {0, 2, 63, -1}};

/* sample_clamped (inter.rs:886-890). */
static inline int32_t smp(const uint16_t* p, int stride, int pw, int ph, int x, int y) {
    if (x < 0) x = 0;
    if (x > pw - 1) x = pw - 1;
    if (y < 0) y = 0;
    if (y > ph - 1) y = ph - 1;
    return p[(size_t)y * (size_t)stride + (size_t)x];
}

static inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : v > hi ? hi : v; }

/* Arithmetic right shift of a signed MV (Rust `>>` on i32). */
static inline int asr(int v, int s) { return v >= 0 ? v >> s : -((-v + (1 << s) - 1) >> s); }

int orc_interp_luma(const uint16_t* plane, int stride, int pw, int ph, int bd, int bx, int by, int w,
                    int h, int mv_x, int mv_y, int bi, int alt_hpel, int affine, int32_t* out) {
    if (affine) return -1; /* the affine filter path is not part of this oracle */
    int int_x = bx + asr(mv_x, 4), int_y = by + asr(mv_y, 4);
    int fx = mv_x & 15, fy = mv_y & 15;
    int head_room = IF_INTERNAL_PREC - bd;
    int32_t maxv = (1 << bd) - 1;
    const int* cx = (alt_hpel && fx == 8) ? LUMA_ALT_HPEL : LUMA8[fx];
    const int* cy = (alt_hpel && fy == 8) ? LUMA_ALT_HPEL : LUMA8[fy];

    if (fx == 0 && fy == 0) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int32_t s = smp(plane, stride, pw, ph, int_x + x, int_y + y);
                out[y * w + x] = bi ? s * (1 << head_room) - IF_INTERNAL_OFFS : s;
            }
        return 0;
    }
    if (fy == 0 || fx == 0) {
        int horizontal = fy == 0;
        const int* c = horizontal ? cx : cy;
        int shift = bi ? IF_FILTER_PREC - head_room : IF_FILTER_PREC;
        int32_t offset = bi ? -IF_INTERNAL_OFFS * (1 << shift) : 1 << (shift - 1);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int32_t sum = 0;
                for (int k = 0; k < 8; k++) {
                    int tx = int_x + x + (horizontal ? k - 3 : 0);
                    int ty = int_y + y + (horizontal ? 0 : k - 3);
                    sum += smp(plane, stride, pw, ph, tx, ty) * c[k];
                }
                int32_t val = (sum + offset) >> shift;
                out[y * w + x] = bi ? val : clampi(val, 0, maxv);
            }
        return 0;
    }
    int ext_h = h + 7;
    int32_t* tmp = malloc(sizeof(int32_t) * (size_t)(w * ext_h));
    int shift1 = IF_FILTER_PREC - head_room;
    int32_t offset1 = -IF_INTERNAL_OFFS * (1 << shift1);
    for (int y = 0; y < ext_h; y++)
        for (int x = 0; x < w; x++) {
            int32_t sum = 0;
            for (int k = 0; k < 8; k++)
                sum += smp(plane, stride, pw, ph, int_x + x + k - 3, int_y - 3 + y) * cx[k];
            tmp[y * w + x] = (sum + offset1) >> shift1;
        }
    int shift2 = bi ? IF_FILTER_PREC : IF_FILTER_PREC + head_room;
    int32_t offset2 = bi ? 0 : (1 << (shift2 - 1)) + (IF_INTERNAL_OFFS << IF_FILTER_PREC);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int32_t sum = 0;
            for (int k = 0; k < 8; k++) sum += tmp[(y + k) * w + x] * cy[k];
            int32_t val = (sum + offset2) >> shift2;
            out[y * w + x] = bi ? val : clampi(val, 0, maxv);
        }
    free(tmp);
    return 0;
}

int orc_interp_chroma(const uint16_t* plane, int stride, int pw, int ph, int bd, int cxp, int cyp,
                      int cw, int ch, int mv_x, int mv_y, int bi, int32_t* out) {
    int int_x = cxp + asr(mv_x, 5), int_y = cyp + asr(mv_y, 5);
    int fx = mv_x & 31, fy = mv_y & 31;
    int head_room = IF_INTERNAL_PREC - bd;
    int32_t maxv = (1 << bd) - 1;
    if (fx == 0 && fy == 0) {
        for (int y = 0; y < ch; y++)
            for (int x = 0; x < cw; x++) {
                int32_t s = smp(plane, stride, pw, ph, int_x + x, int_y + y);
                out[y * cw + x] = bi ? s * (1 << head_room) - IF_INTERNAL_OFFS : s;
            }
        return 0;
    }
    if (fy == 0 || fx == 0) {
        int horizontal = fy == 0;
        const int* c = CHROMA4[horizontal ? fx : fy];
        int shift = bi ? IF_FILTER_PREC - head_room : IF_FILTER_PREC;
        int32_t offset = bi ? -IF_INTERNAL_OFFS * (1 << shift) : 1 << (shift - 1);
        for (int y = 0; y < ch; y++)
            for (int x = 0; x < cw; x++) {
                int32_t sum = 0;
                for (int k = 0; k < 4; k++) {
                    int tx = int_x + x + (horizontal ? k - 1 : 0);
                    int ty = int_y + y + (horizontal ? 0 : k - 1);
                    sum += smp(plane, stride, pw, ph, tx, ty) * c[k];
                }
                int32_t val = (sum + offset) >> shift;
                out[y * cw + x] = bi ? val : clampi(val, 0, maxv);
            }
        return 0;
    }
    const int* cxf = CHROMA4[fx];
    const int* cyf = CHROMA4[fy];
    int ext_h = ch + 3;
    int32_t* tmp = malloc(sizeof(int32_t) * (size_t)(cw * ext_h));
    int shift1 = IF_FILTER_PREC - head_room;
    int32_t offset1 = -IF_INTERNAL_OFFS * (1 << shift1);
    for (int y = 0; y < ext_h; y++)
        for (int x = 0; x < cw; x++) {
            int32_t sum = 0;
            for (int k = 0; k < 4; k++)
                sum += smp(plane, stride, pw, ph, int_x + x - 1 + k, int_y - 1 + y) * cxf[k];
            tmp[y * cw + x] = (sum + offset1) >> shift1;
        }
    int shift2 = bi ? IF_FILTER_PREC : IF_FILTER_PREC + head_room;
    int32_t offset2 = bi ? 0 : (1 << (shift2 - 1)) + (IF_INTERNAL_OFFS << IF_FILTER_PREC);
    for (int y = 0; y < ch; y++)
        for (int x = 0; x < cw; x++) {
            int32_t sum = 0;
            for (int k = 0; k < 4; k++) sum += tmp[(y + k) * cw + x] * cyf[k];
            int32_t val = (sum + offset2) >> shift2;
            out[y * cw + x] = bi ? val : clampi(val, 0, maxv);
        }
    free(tmp);
    return 0;
}

/* interpolate_luma_block_16 as written (inter.rs:2010-2134), regular filter only. */
int orc_interp_luma_16_as_written(const uint16_t* plane, int stride, int pw, int ph, int bd, int bx,
                                  int by, int w, int h, int mv_x, int mv_y, int bi, int32_t* out) {
    int dx = asr(mv_x, 4), dy = asr(mv_y, 4);
    int fx = mv_x & 15, fy = mv_y & 15;
    if (bd < 8) bd = 8;
    if (bd > 12) bd = 12;
    int32_t maxv = (1 << bd) - 1;
    int head_room = 14 - bd < 2 ? 2 : 14 - bd;
    int shift1 = 6 - head_room;
    int shift2 = bi ? 6 : 12 - bd;
    /* At bd = 12 the as-written shift2 is 0 and the reference's
     * `1i32 << (shift2 - 1)` shifts by -1: a debug build panics there and a
     * release build masks the amount (1 << 31). The rounding term has no
     * defined reference value then; 0 keeps the C shift defined. Only the
     * uni-prediction 2-D branch reads add2. */
    int32_t add2 = (bi || shift2 <= 0) ? 0 : 1 << (shift2 - 1);
    int32_t boff = 1 << (bd + head_room - 1);
    if (fx == 0 && fy == 0) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int32_t s = smp(plane, stride, pw, ph, bx + x + dx, by + y + dy);
                out[y * w + x] = bi ? s * (1 << head_room) - boff : s;
            }
        return 0;
    }
    if (fy == 0 || fx == 0) {
        int horizontal = fy == 0;
        const int* f = LUMA8[horizontal ? fx : fy];
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int32_t sum = 0;
                for (int k = 0; k < 8; k++) {
                    int tx = bx + x + dx + (horizontal ? k - 3 : 0);
                    int ty = by + y + dy + (horizontal ? 0 : k - 3);
                    sum += f[k] * smp(plane, stride, pw, ph, tx, ty);
                }
                out[y * w + x] = bi ? (sum >> shift1) - boff : clampi((sum + 32) >> 6, 0, maxv);
            }
        return 0;
    }
    int th = h + 7;
    int32_t* tmp = malloc(sizeof(int32_t) * (size_t)(w * th));
    for (int ty = 0; ty < th; ty++)
        for (int x = 0; x < w; x++) {
            int32_t sum = 0;
            for (int k = 0; k < 8; k++)
                sum += LUMA8[fx][k] * smp(plane, stride, pw, ph, bx + x + dx + k - 3, by + ty - 3 + dy);
            tmp[ty * w + x] = sum >> shift1;
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int32_t sum = 0;
            for (int k = 0; k < 8; k++) sum += LUMA8[fy][k] * tmp[(y + k) * w + x];
            out[y * w + x] = bi ? (sum >> shift2) - boff : clampi((sum + add2) >> shift2, 0, maxv);
        }
    free(tmp);
    return 0;
}

/* BCW_WEIGHTS (inter.rs:128). */
static const int BCW_WEIGHTS[5] = {-2, 3, 4, 5, 10};
#define BCW_DEFAULT 2

/* right_shift_msb (inter.rs:1204-1211): numer >> floor(log2(denom)), 0 for denom <= 0. */
static int right_shift_msb(int numer, int denom) {
    if (denom <= 0) return 0;
    int l = 31 - __builtin_clz((unsigned)denom);
    return numer >> l;
}

static int sgn_mul(int s, int v) { return s < 0 ? -v : s == 0 ? 0 : v; }

/* apply_bdof_luma (inter.rs:1218-1339), 8-bit: pred0 / pred1 are the 14-bit
 * intermediates of a w x h block (w, h multiples of 4). The block is padded by
 * one sample on each side by replication, the gradients are taken on the
 * inner samples and replicated to the border, and each 4x4 unit's motion
 * offset comes from the sums over its 6x6 window. */
int orc_bdof_luma(const int32_t* pred0, const int32_t* pred1, int w, int h, int32_t* dst) {
    if (w < 4 || h < 4 || (w & 3) || (h & 3)) return -1;
    const int ew = w + 2, eh = h + 2;
    int32_t* buf = malloc(sizeof(int32_t) * (size_t)(ew * eh) * 6);
    int32_t* p[2] = {buf, buf + ew * eh};
    int32_t* gx[2] = {buf + 2 * ew * eh, buf + 3 * ew * eh};
    int32_t* gy[2] = {buf + 4 * ew * eh, buf + 5 * ew * eh};
    for (int l = 0; l < 2; l++) {
        const int32_t* src = l ? pred1 : pred0;
        int32_t* q = p[l];
        for (int y = 0; y < h; y++) {
            q[(y + 1) * ew] = src[y * w];
            for (int x = 0; x < w; x++) q[(y + 1) * ew + x + 1] = src[y * w + x];
            q[(y + 1) * ew + w + 1] = src[y * w + w - 1];
        }
        for (int x = 0; x < ew; x++) {
            q[x] = q[ew + x];
            q[(h + 1) * ew + x] = q[h * ew + x];
        }
        int32_t *a = gx[l], *b = gy[l];
        for (int y = 1; y <= h; y++) {
            for (int x = 1; x <= w; x++) {
                const int i = y * ew + x;
                a[i] = (q[i + 1] >> 6) - (q[i - 1] >> 6);
                b[i] = (q[i + ew] >> 6) - (q[i - ew] >> 6);
            }
            a[y * ew] = a[y * ew + 1];
            a[y * ew + w + 1] = a[y * ew + w];
            b[y * ew] = b[y * ew + 1];
            b[y * ew + w + 1] = b[y * ew + w];
        }
        for (int x = 0; x < ew; x++) {
            a[x] = a[ew + x];
            a[(h + 1) * ew + x] = a[h * ew + x];
            b[x] = b[ew + x];
            b[(h + 1) * ew + x] = b[h * ew + x];
        }
    }
    const int shift = IF_INTERNAL_PREC - 8 + 1;
    const int offset = (1 << (shift - 1)) + 2 * IF_INTERNAL_OFFS;
    const int limit = 15;
    for (int yu = 0; yu < h / 4; yu++)
        for (int xu = 0; xu < w / 4; xu++) {
            int sgx = 0, sgy = 0, sdx = 0, sdy = 0, sxy = 0;
            for (int dy = 0; dy < 6; dy++)
                for (int dx = 0; dx < 6; dx++) {
                    const int i = (yu * 4 + dy) * ew + xu * 4 + dx;
                    const int tgx = (gx[0][i] + gx[1][i]) >> 1;
                    const int tgy = (gy[0][i] + gy[1][i]) >> 1;
                    const int tdi = (p[1][i] >> 4) - (p[0][i] >> 4);
                    sgx += tgx < 0 ? -tgx : tgx;
                    sgy += tgy < 0 ? -tgy : tgy;
                    sdx += sgn_mul(tgx, tdi);
                    sdy += sgn_mul(tgy, tdi);
                    sxy += sgn_mul(tgy, tgx);
                }
            int vx = sgx == 0 ? 0 : clampi(right_shift_msb(4 * sdx, sgx), -limit, limit);
            int t = (sxy * vx) >> 1;
            int vy = sgy == 0 ? 0 : clampi(right_shift_msb(4 * sdy - t, sgy), -limit, limit);
            for (int dy = 0; dy < 4; dy++)
                for (int dx = 0; dx < 4; dx++) {
                    const int y = yu * 4 + dy, x = xu * 4 + dx, i = (y + 1) * ew + x + 1;
                    const int b = vx * (gx[0][i] - gx[1][i]) + vy * (gy[0][i] - gy[1][i]);
                    dst[y * w + x] = clampi((p[0][i] + p[1][i] + b + offset) >> shift, 0, 255);
                }
        }
    free(buf);
    return 0;
}

/* dmvr_subpel_1d (inter.rs:1341-1363): the parabolic sub-sample offset in
 * 1/16 sample from the costs at -1, 0, +1 (s1, s0, s2). */
static int dmvr_subpel_1d(int64_t s0, int64_t s1, int64_t s2) {
    const int64_t den = s1 + s2 - 2 * s0;
    if (den == 0) return 0;
    if (s1 != s0 && s2 != s0) {
        const int64_t num = (s1 - s2) * 8; /* ((s1 - s2) << 4) >> 1, exact */
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

/* bilinear_pred of refine_dmvr_subblock: BILINEAR_FILTER_PREC4 taps (16 - f, f)
 * (inter.rs:106-123) at the MV's fraction, offset (ox, oy) integer samples. */
static void dmvr_bilinear(const uint16_t* pl, int stride, int pw, int ph, int bx, int by, int dx, int dy, int mvx,
                          int mvy, int ox, int oy, int32_t* out) {
    const int ix = bx + asr(mvx, 4) + ox, iy = by + asr(mvy, 4) + oy;
    const int fx = mvx & 15, fy = mvy & 15;
    for (int y = 0; y < dy; y++)
        for (int x = 0; x < dx; x++) {
            const int px = ix + x, py = iy + y;
            const int h0 = (smp(pl, stride, pw, ph, px, py) * (16 - fx) + smp(pl, stride, pw, ph, px + 1, py) * fx + 8) >> 4;
            const int h1 =
                (smp(pl, stride, pw, ph, px, py + 1) * (16 - fx) + smp(pl, stride, pw, ph, px + 1, py + 1) * fx + 8) >> 4;
            out[y * dx + x] = (h0 * (16 - fy) + h1 * fy + 8) >> 4;
        }
}

/* refine_dmvr_subblock (inter.rs:1371-1459) over a dx x dy block at (bx, by):
 * bilateral SAD on every second row over the 5x5 integer offsets (L0 + o,
 * L1 - o), the centre cost reduced by a quarter, early exit when it is below
 * dx * dy, then the parabolic sub-sample step around a best offset inside the
 * 3x3 core. mv = {mv0_x, mv0_y, mv1_x, mv1_y}, refined in place (clipped to
 * the 18-bit storage range). */
void orc_dmvr_refine(const uint16_t* r0, const uint16_t* r1, int stride, int pw, int ph, int bx, int by, int dx,
                     int dy, int32_t* mv) {
    int32_t* a = malloc(sizeof(int32_t) * (size_t)(dx * dy) * 2);
    int32_t* b = a + dx * dy;
    int64_t sads[5][5];
    for (int oy = -2; oy <= 2; oy++)
        for (int ox = -2; ox <= 2; ox++) {
            dmvr_bilinear(r0, stride, pw, ph, bx, by, dx, dy, mv[0], mv[1], ox, oy, a);
            dmvr_bilinear(r1, stride, pw, ph, bx, by, dx, dy, mv[2], mv[3], -ox, -oy, b);
            int64_t c = 0;
            for (int y = 0; y < dy; y += 2)
                for (int x = 0; x < dx; x++) {
                    const int d = a[y * dx + x] - b[y * dx + x];
                    c += d < 0 ? -d : d;
                }
            sads[oy + 2][ox + 2] = c;
        }
    free(a);
    int64_t min_cost = sads[2][2];
    min_cost -= min_cost >> 2;
    sads[2][2] = min_cost;
    if (min_cost < (int64_t)dx * dy) return;
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
    int ddx = bix * 16, ddy = biy * 16;
    if (bix > -2 && bix < 2 && biy > -2 && biy < 2) {
        const int ux = bix + 2, uy = biy + 2;
        ddx += dmvr_subpel_1d(sads[uy][ux], sads[uy][ux - 1], sads[uy][ux + 1]);
        ddy += dmvr_subpel_1d(sads[uy][ux], sads[uy - 1][ux], sads[uy + 1][ux]);
    }
    mv[0] = clampi(mv[0] + ddx, -(1 << 17), (1 << 17) - 1);
    mv[1] = clampi(mv[1] + ddy, -(1 << 17), (1 << 17) - 1);
    mv[2] = clampi(mv[2] - ddx, -(1 << 17), (1 << 17) - 1);
    mv[3] = clampi(mv[3] - ddy, -(1 << 17), (1 << 17) - 1);
}

void orc_weighted_average_bi(const int32_t* p0, const int32_t* p1, size_t n, int bcw_idx, int bd,
                             int32_t* out) {
    int w1 = BCW_WEIGHTS[bcw_idx > 4 ? 4 : bcw_idx];
    int w0 = 8 - w1;
    int head_room = IF_INTERNAL_PREC - bd;
    int shift_num = head_room + 3;
    int32_t offset = (1 << (shift_num - 1)) + (IF_INTERNAL_OFFS << 3);
    int32_t maxv = (1 << bd) - 1;
    for (size_t i = 0; i < n; i++) out[i] = clampi((p0[i] * w0 + p1[i] * w1 + offset) >> shift_num, 0, maxv);
}

static int plane_get(const rmh_plane* p, int x, int y) {
    if (p->bytes_per_sample == 1) return ((const uint8_t*)p->data)[(size_t)y * p->stride + x];
    return ((const uint16_t*)p->data)[(size_t)y * p->stride + x];
}

static void plane_set(const rmh_plane* p, int x, int y, int v) {
    if (p->bytes_per_sample == 1)
        ((uint8_t*)p->data)[(size_t)y * p->stride + x] = (uint8_t)v;
    else
        ((uint16_t*)p->data)[(size_t)y * p->stride + x] = (uint16_t)v;
}

/* A uint16 copy of a plane so the interpolators above read one sample type. */
static uint16_t* widen(const rmh_plane* p) {
    uint16_t* o = malloc(sizeof(uint16_t) * (size_t)p->width * p->height);
    for (uint32_t y = 0; y < p->height; y++)
        for (uint32_t x = 0; x < p->width; x++) o[(size_t)y * p->width + x] = (uint16_t)plane_get(p, (int)x, (int)y);
    return o;
}

extern int orc_threads;

size_t orc_mc_batch(const rmh_mc_frame* frames, size_t n_frames, const rmh_pu_desc* pus, size_t n_pu) {
    uint16_t** wide = calloc(n_frames * 3, sizeof(uint16_t*));
    /* PUs up to the first one with a bad frame index are predicted */
    size_t bad = 0, n_ok = n_pu;
    for (size_t i = 0; i < n_pu; i++) {
        if (pus[i].frame >= n_frames) {
            bad = i + 1;
            n_ok = i;
            break;
        }
        for (int c = 0; c < 3; c++)
            if (!wide[pus[i].frame * 3 + c]) wide[pus[i].frame * 3 + c] = widen(&frames[pus[i].frame].ref[c]);
    }
#pragma omp parallel for num_threads(orc_threads) schedule(dynamic, 16) if (orc_threads > 1)
    for (size_t i = 0; i < n_ok; i++) {
        const rmh_pu_desc* pu = &pus[i];
        const rmh_mc_frame* f = &frames[pu->frame];
        int w = 1 << (pu->log2_w & 15), h = 1 << (pu->log2_h & 15);
        int alt = (pu->log2_w & RMH_PU_ALT_HPEL) != 0;
        int bd = f->ref[0].bit_depth;
        int32_t* buf = malloc(sizeof(int32_t) * (size_t)(w * h));
        const rmh_plane* r = &f->ref[0];
        orc_interp_luma(wide[pu->frame * 3], (int)r->width, (int)r->width, (int)r->height, bd, pu->x, pu->y, w, h,
                        pu->mv_x, pu->mv_y, 0, alt, 0, buf);
        int maxv = (1 << bd) - 1;
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++)
                if (pu->x + x < (int)f->dst[0].width && pu->y + y < (int)f->dst[0].height)
                    plane_set(&f->dst[0], pu->x + x, pu->y + y, clampi(buf[y * w + x], 0, maxv));
        int cw = w / 2, ch = h / 2;
        for (int c = 1; c < 3; c++) {
            const rmh_plane* rc = &f->ref[c];
            orc_interp_chroma(wide[pu->frame * 3 + c], (int)rc->width, (int)rc->width, (int)rc->height,
                              rc->bit_depth, pu->x / 2, pu->y / 2, cw, ch, pu->mv_x, pu->mv_y, 0, buf);
            int cmax = (1 << rc->bit_depth) - 1;
            for (int y = 0; y < ch; y++)
                for (int x = 0; x < cw; x++)
                    if (pu->x / 2 + x < (int)f->dst[c].width && pu->y / 2 + y < (int)f->dst[c].height)
                        plane_set(&f->dst[c], pu->x / 2 + x, pu->y / 2 + y, clampi(buf[y * cw + x], 0, cmax));
        }
        free(buf);
    }
    for (size_t i = 0; i < n_frames * 3; i++) free(wide[i]);
    free(wide);
    return bad;
}

/* Writes the w x h block `buf` (already clipped sample values) at (x0, y0) of
 * plane p, skipping samples outside it (write_prediction). */
static void put_block(const rmh_plane* p, int x0, int y0, int w, int h, const int32_t* buf) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            if (x0 + x < (int)p->width && y0 + y < (int)p->height) plane_set(p, x0 + x, y0 + y, buf[y * w + x]);
}

size_t orc_mc_bi_batch(const rmh_mc_bi_frame* frames, size_t n_frames, const rmh_bi_pu_desc* pus, size_t n_pu) {
    uint16_t** wide = calloc(n_frames * 6, sizeof(uint16_t*));
    size_t bad = 0;
    for (size_t i = 0; i < n_pu && !bad; i++) {
        const rmh_bi_pu_desc* pu = &pus[i];
        if (pu->frame >= n_frames) {
            bad = i + 1;
            break;
        }
        const rmh_mc_bi_frame* f = &frames[pu->frame];
        uint16_t** wf = wide + (size_t)pu->frame * 6;
        for (int c = 0; c < 3; c++) {
            if (!wf[c]) wf[c] = widen(&f->ref0[c]);
            if (!wf[3 + c]) wf[3 + c] = widen(&f->ref1[c]);
        }
        int w = 1 << (pu->log2_w & 15), h = 1 << (pu->log2_h & 15);
        int alt = (pu->log2_w & RMH_PU_ALT_HPEL) != 0;
        int bd = f->ref0[0].bit_depth;
        int32_t* p0 = malloc(sizeof(int32_t) * (size_t)(w * h));
        int32_t* p1 = malloc(sizeof(int32_t) * (size_t)(w * h));
        int32_t* out = malloc(sizeof(int32_t) * (size_t)(w * h));
        const rmh_plane* r = &f->ref0[0];
        /* the refinements of predict_inter_pu's bi branch (inter.rs:1763-1800):
         * DMVR with default BCW weights on blocks of at least 8x8 and 128
         * samples, then BDOF with default weights; the caller's flags stand for
         * the enables and the POC conditions */
        int32_t mv[4] = {pu->mv0_x, pu->mv0_y, pu->mv1_x, pu->mv1_y};
        if ((pu->refine & RMH_BI_DMVR) && pu->bcw_idx == BCW_DEFAULT && w >= 8 && h >= 8 && w * h >= 128)
            orc_dmvr_refine(wf[0], wf[3], (int)r->width, (int)r->width, (int)r->height, pu->x, pu->y, w, h, mv);
        orc_interp_luma(wf[0], (int)r->width, (int)r->width, (int)r->height, bd, pu->x, pu->y, w, h, mv[0], mv[1], 1,
                        alt, 0, p0);
        orc_interp_luma(wf[3], (int)r->width, (int)r->width, (int)r->height, bd, pu->x, pu->y, w, h, mv[2], mv[3], 1,
                        alt, 0, p1);
        if ((pu->refine & RMH_BI_BDOF) && pu->bcw_idx == BCW_DEFAULT)
            orc_bdof_luma(p0, p1, w, h, out);
        else
            orc_weighted_average_bi(p0, p1, (size_t)(w * h), pu->bcw_idx, bd, out);
        put_block(&f->dst[0], pu->x, pu->y, w, h, out);
        int cw = w / 2, ch = h / 2;
        for (int c = 1; c < 3; c++) {
            const rmh_plane* rc = &f->ref0[c];
            orc_interp_chroma(wf[c], (int)rc->width, (int)rc->width, (int)rc->height, bd, pu->x / 2, pu->y / 2, cw,
                              ch, mv[0], mv[1], 1, p0);
            orc_interp_chroma(wf[3 + c], (int)rc->width, (int)rc->width, (int)rc->height, bd, pu->x / 2, pu->y / 2,
                              cw, ch, mv[2], mv[3], 1, p1);
            orc_weighted_average_bi(p0, p1, (size_t)(cw * ch), pu->bcw_idx, bd, out);
            put_block(&f->dst[c], pu->x / 2, pu->y / 2, cw, ch, out);
        }
        free(p0);
        free(p1);
        free(out);
    }
    for (size_t i = 0; i < n_frames * 6; i++) free(wide[i]);
    free(wide);
    return bad;
}

/* ------------------------------------------------------------------ GPM */

/* GEO_DIS and GEO_ANGLE_TO_MASK (inter.rs:131-140; VTM Rom.cpp:708-711). */
static const int GEO_DIS[32] = {8, 8, 8, 8, 4, 4, 2, 1, 0, -1, -2, -4, -4, -8, -8, -8,
                                -8, -8, -8, -8, -4, -4, -2, -1, 0, 1, 2, 4, 4, 8, 8, 8};
static const int GEO_ANGLE_TO_MASK[32] = {0, -1, 1, 2, 3, 4, -1, -1, 5, -1, -1, 4, 3, 2, 1, -1,
                                          0, -1, 1, 2, 3, 4, -1, -1, 5, -1, -1, 4, 3, 2, 1, -1};

/* derive_gpm_params_table (inter.rs:1641-1671): the (angle, distance) pair of
 * each of the 64 split directions. */
void orc_gpm_params(uint8_t* angle, uint8_t* distance) {
    int mode = 0;
    for (int a = 0; a < 32; a++)
        for (int d = 0; d < 4; d++) {
            const int mask = GEO_ANGLE_TO_MASK[a];
            if ((d == 0 && a >= 16) || ((d == 2 || d == 0) && (mask == 0 || mask == 5)) || mask == -1) continue;
            if (mode < 64) {
                angle[mode] = (uint8_t)a;
                distance[mode] = (uint8_t)d;
                mode++;
            }
        }
}

/* predict_gpm_block (inter.rs:1673-1730): the w x h blend of two 8-bit
 * predictions with the split direction's mask, positions scaled by
 * (1 << scale_x, 1 << scale_y) onto the luma grid. */
int orc_gpm_blend(const int32_t* p0, const int32_t* p1, int w, int h, int split_dir, int scale_x, int scale_y,
                  int32_t* out) {
    uint8_t ang[64], dis[64];
    orc_gpm_params(ang, dis);
    const int angle = ang[split_dir & 63], distance = dis[split_dir & 63];
    const int lw = w << scale_x, lh = h << scale_y;
    int ox = -(lw >> 1), oy = -(lh >> 1);
    if (distance > 0) {
        const int am = angle % 16;
        if (am == 8 || (am != 0 && lh >= lw)) {
            const int d = (distance * lh) >> 3;
            oy += angle < 16 ? d : -d;
        } else {
            const int d = (distance * lw) >> 3;
            ox += angle < 16 ? d : -d;
        }
    }
    const int dx = angle, dy = (dx + 8) % 32;
    for (int y = 0; y < h; y++) {
        const int ly = y << scale_y;
        const int look_y = (((ly + oy) * 2) + 1) * GEO_DIS[dy];
        for (int x = 0; x < w; x++) {
            const int lx = x << scale_x;
            const int widx = (((lx + ox) * 2) + 1) * GEO_DIS[dx] + look_y;
            const int w0 = clampi((32 + widx + 4) >> 3, 0, 8);
            out[y * w + x] = clampi((w0 * p0[y * w + x] + (8 - w0) * p1[y * w + x] + 4) >> 3, 0, 255);
        }
    }
    return 0;
}

/* The GPM blend of every PU of a batch: luma blocks at (x, y), chroma at
 * (x / 2, y / 2) with half the size, each read from the two prediction planes
 * and written to dst (clipped to the plane), as the general decoder's GPM CU
 * does (decode_general.rs:2716-2741). Returns 0, or 1 + the index of the first
 * PU with a bad frame index. */
size_t orc_gpm_batch(const rmh_gpm_frame* frames, size_t n_frames, const rmh_gpm_pu* pus, size_t n_pu, int luma_scale,
                     int chroma_scale) {
    for (size_t i = 0; i < n_pu; i++) {
        const rmh_gpm_pu* pu = &pus[i];
        if (pu->frame >= n_frames) return i + 1;
        const rmh_gpm_frame* f = &frames[pu->frame];
        const int w = 1 << pu->log2_w, h = 1 << pu->log2_h;
        int32_t* a = malloc(sizeof(int32_t) * (size_t)(w * h) * 3);
        int32_t *b = a + w * h, *o = b + w * h;
        for (int c = 0; c < 3; c++) {
            const int sx = c ? 1 : 0;
            const int bw = w >> sx, bh = h >> sx, x0 = pu->x >> sx, y0 = pu->y >> sx;
            const rmh_plane* d = &f->dst[c];
            for (int y = 0; y < bh; y++)
                for (int x = 0; x < bw; x++) {
                    const int px = x0 + x < (int)d->width ? x0 + x : (int)d->width - 1;
                    const int py = y0 + y < (int)d->height ? y0 + y : (int)d->height - 1;
                    a[y * bw + x] = plane_get(&f->pred0[c], px, py);
                    b[y * bw + x] = plane_get(&f->pred1[c], px, py);
                }
            orc_gpm_blend(a, b, bw, bh, pu->split_dir, c ? chroma_scale : luma_scale, c ? chroma_scale : luma_scale, o);
            put_block(d, x0, y0, bw, bh, o);
        }
        free(a);
    }
    return 0;
}


/*
 * orc_residual.c - residual stage of the CPU oracle (test infrastructure only).
 *
 * Restates crates/rustmpeg-video/src/vvc/dct.rs (quantisers, transforms),
 * vvc/transform.rs (transform-skip), vvc/lfnst.rs (LFNST) and the TU driver
 * inverse_residual / residual_of (vvc/decode_general.rs:4100-4236).
 */
#include <stdlib.h>
#include <string.h>

#include "vvc_oracle.h"

static const int INV_QUANT_SCALES[6] = {40, 45, 51, 57, 64, 72};         /* vvc/dct.rs:37 */
static const int INV_QUANT_SCALES_SQRT2[6] = {57, 64, 72, 80, 90, 102};  /* vvc/dct.rs:40 */
static const int QUANT_SCALES[6] = {26214, 23302, 20560, 18396, 16384, 14564};
static const int QUANT_SCALES_SQRT2[6] = {18396, 16384, 14564, 13107, 11651, 10280};

#define MAX_LOG2_TR_DYNAMIC_RANGE 15
#define TRANSFORM_MATRIX_SHIFT 6

static int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

static int is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

static int32_t clip16(int64_t v) {  /* clip_transform, vvc/dct.rs:291-293 */
    return (int32_t)(v < -32768 ? -32768 : v > 32767 ? 32767 : v);
}

/* i64 << s with the wrapping of Rust's `<<` on i64. */
static int64_t shl64(int64_t v, int s) { return (int64_t)((uint64_t)v << s); }

static int clamp_bd(int bd) { return bd < 8 ? 8 : bd > 12 ? 12 : bd; }

/* is_supported_rect (vvc/dct.rs:265-271). */
static int supported_rect(int w, int h) {
    return w >= 1 && w <= 64 && h >= 1 && h <= 64 && w * h >= 2 && is_pow2(w) && is_pow2(h);
}

/* transform_shift_rect_bd with the sqrt(2) adjustment (vvc/dct.rs:858-860, 1167-1169). */
static int tr_shift_bd(int w, int h, int bd, int* sqrt2) {
    int s = ilog2(w) + ilog2(h);
    *sqrt2 = s & 1;
    return MAX_LOG2_TR_DYNAMIC_RANGE - bd - (s >> 1) - (*sqrt2 ? 1 : 0);
}

/* dequantize_rect_bd (vvc/dct.rs:1151-1187). */
int orc_dequant_rect(int w, int h, const int32_t* levels, int qp, int bd, int32_t* out) {
    if (!supported_rect(w, h)) return -1;
    if (qp < 4) qp = 4;
    int qp_per = qp / 6, qp_rem = qp % 6, sqrt2;
    int tr_shift = tr_shift_bd(w, h, bd, &sqrt2);
    int right_shift = 6 - (tr_shift + qp_per);
    int64_t scale = sqrt2 ? INV_QUANT_SCALES_SQRT2[qp_rem] : INV_QUANT_SCALES[qp_rem];
    for (int i = 0; i < w * h; i++) {
        int64_t value = (int64_t)levels[i] * scale;
        int64_t scaled = right_shift > 0 ? (value + ((int64_t)1 << (right_shift - 1))) >> right_shift
                                         : shl64(value, -right_shift);
        out[i] = clip16(scaled);
    }
    return 0;
}

/* dequantize_dep_quant_bd (vvc/dct.rs:1191-1248) over CoeffScan raster order. */
int orc_dequant_dep_quant(int w, int h, const int32_t* levels, int qp, int bd, int32_t* out) {
    if (!supported_rect(w, h)) return -1;
    int32_t scan[4096];
    int n = orc_coeff_scan(w, h, scan);
    memset(out, 0, sizeof(int32_t) * (size_t)(w * h));
    int last = -1;
    for (int s = n - 1; s >= 0; s--)
        if (scan[s] < w * h && levels[scan[s]] != 0) {
            last = s;
            break;
        }
    if (last < 0) return 0;
    if (qp < 4) qp = 4;
    int qp_dq = qp + 1;
    int qp_per = qp_dq / 6, qp_rem = qp_dq % 6, sqrt2;
    int tr_shift = tr_shift_bd(w, h, bd, &sqrt2);
    int shift = 7 - qp_per - tr_shift;
    int64_t inv_q_scale = sqrt2 ? INV_QUANT_SCALES_SQRT2[qp_rem] : INV_QUANT_SCALES[qp_rem];
    if (shift < 0) inv_q_scale = shl64(inv_q_scale, -shift);
    int64_t add = shift > 0 ? ((int64_t)1 << (shift - 1)) : 0;
    int rshift = shift > 0 ? shift : 0;
    int state = 0;
    for (int s = last; s >= 0; s--) {
        int rpos = scan[s];
        int32_t level = levels[rpos];
        if (level != 0) {
            int64_t q_idx = 2 * (int64_t)level + (level > 0 ? -(state >> 1) : (state >> 1));
            out[rpos] = clip16((q_idx * inv_q_scale + add) >> rshift);
        }
        state = (32040 >> ((state << 2) + ((level & 1) << 1))) & 3;
    }
    return 0;
}

/* dequantize_ts (vvc/transform.rs:229-246): the i64 result is narrowed to i32
 * before the clamp, as the reference's `as i32` does. */
int32_t orc_dequant_ts(int32_t level, int qp) {
    if (qp < 4) qp = 4;
    int qp_per = qp / 6, qp_rem = qp % 6;
    int right_shift = 6 - qp_per;
    int64_t value = (int64_t)level * INV_QUANT_SCALES[qp_rem];
    int64_t scaled = right_shift > 0 ? (value + ((int64_t)1 << (right_shift - 1))) >> right_shift
                                     : shl64(value, -right_shift);
    int32_t narrow = (int32_t)(uint32_t)(uint64_t)scaled;
    return narrow < -32768 ? -32768 : narrow > 32767 ? 32767 : narrow;
}

/* quantize_ts (vvc/transform.rs:214-226). */
int32_t orc_quant_ts(int32_t residual, int qp) {
    if (qp < 4) qp = 4;
    int qp_per = qp / 6, qp_rem = qp % 6;
    int i_qbits = 14 + qp_per;
    int64_t add = (int64_t)171 << (i_qbits - 9);
    int sign = residual > 0 ? 1 : residual < 0 ? -1 : 0;
    int64_t a = residual < 0 ? -(int64_t)residual : residual;
    int32_t level = (int32_t)((a * QUANT_SCALES[qp_rem] + add) >> i_qbits);
    int64_t v = (int64_t)sign * level;
    return (int32_t)(v < -32768 ? -32768 : v > 32767 ? 32767 : v);
}

/* quantize_rect_bd (vvc/dct.rs:1113-1147). */
int orc_quant_rect(int w, int h, const int32_t* coeffs, int qp, int bd, int is_intra, int32_t* out) {
    if (!supported_rect(w, h)) return -1;
    if (qp < 4) qp = 4;
    int qp_per = qp / 6, qp_rem = qp % 6, sqrt2;
    int tr_shift = tr_shift_bd(w, h, bd, &sqrt2);
    int i_qbits = 14 + qp_per + tr_shift;
    int64_t offset_num = is_intra ? 171 : 85;
    int64_t add = offset_num << (i_qbits - 9 > 0 ? i_qbits - 9 : 0);
    int64_t scale = sqrt2 ? QUANT_SCALES_SQRT2[qp_rem] : QUANT_SCALES[qp_rem];
    int sh = i_qbits > 0 ? i_qbits : 0;
    for (int i = 0; i < w * h; i++) {
        int64_t c = coeffs[i];
        int64_t level = ((c < 0 ? -c : c) * scale + add) >> sh;
        out[i] = clip16(c < 0 ? -level : level);
    }
    return 0;
}

/* quantize_dep_quant_bd (vvc/dct.rs:1252-1301). */
int orc_quant_dep_quant(int w, int h, const int32_t* coeffs, int qp, int bd, int32_t* out) {
    if (!supported_rect(w, h)) return -1;
    int32_t scan[4096];
    int n = orc_coeff_scan(w, h, scan);
    if (qp < 4) qp = 4;
    int qp_dq = qp + 1;
    int qp_per = qp_dq / 6, qp_rem = qp_dq % 6, sqrt2;
    int tr_shift = tr_shift_bd(w, h, bd, &sqrt2);
    int q_shift = 13 + qp_per + tr_shift;
    int64_t q_scale = sqrt2 ? QUANT_SCALES_SQRT2[qp_rem] : QUANT_SCALES[qp_rem];
    int64_t half = (int64_t)1 << (q_shift - 1 > 0 ? q_shift - 1 : 0);
    int qs = q_shift > 0 ? q_shift : 0;
    memset(out, 0, sizeof(int32_t) * (size_t)(w * h));
    int state = 0;
    for (int s = n - 1; s >= 0; s--) {
        int rpos = scan[s];
        if (rpos >= w * h) continue;
        int64_t c = coeffs[rpos];
        int64_t q_idx_est = ((c < 0 ? -c : c) * q_scale + half) >> qs;
        int64_t offset = state >> 1;
        int32_t abs_level = q_idx_est <= offset ? 0 : (int32_t)((q_idx_est + offset + 1) >> 1);
        int64_t level = c < 0 ? -(int64_t)abs_level : abs_level;
        if (level < -32768) level = -32768;
        if (level > 32767) level = 32767;
        out[rpos] = (int32_t)level;
        state = (32040 >> ((state << 2) + (((int32_t)level & 1) << 1))) & 3;
    }
    return 0;
}

/* quantize_dct / dequantize_dct for square DCT-2 blocks at 8 bits (vvc/dct.rs:563-613). */
static int square_supported(int n) { return n >= 4 && n <= 32 && is_pow2(n); }

int orc_quant_dct(int n, const int32_t* coeffs, int qp, int32_t* out) {
    if (!square_supported(n)) return -1;
    if (qp < 4) qp = 4;
    int qp_per = qp / 6, qp_rem = qp % 6;
    int i_qbits = 14 + qp_per + (MAX_LOG2_TR_DYNAMIC_RANGE - 8 - ilog2(n));
    int64_t add = (int64_t)171 << (i_qbits - 9);
    for (int i = 0; i < n * n; i++) {
        int64_t c = coeffs[i];
        int64_t level = ((c < 0 ? -c : c) * QUANT_SCALES[qp_rem] + add) >> i_qbits;
        out[i] = clip16(c < 0 ? -level : level);
    }
    return 0;
}

int orc_dequant_dct(int n, const int32_t* levels, int qp, int32_t* out) {
    if (!square_supported(n)) return -1;
    if (qp < 4) qp = 4;
    int qp_per = qp / 6, qp_rem = qp % 6;
    int right_shift = 6 - ((MAX_LOG2_TR_DYNAMIC_RANGE - 8 - ilog2(n)) + qp_per);
    int tib = MAX_LOG2_TR_DYNAMIC_RANGE + 1;
    if (32 + right_shift - 15 < tib) tib = 32 + right_shift - 15;
    int64_t in_min = -((int64_t)1 << (tib - 1)), in_max = ((int64_t)1 << (tib - 1)) - 1;
    for (int i = 0; i < n * n; i++) {
        int64_t l = levels[i];
        if (l < in_min) l = in_min;
        if (l > in_max) l = in_max;
        int64_t value = l * INV_QUANT_SCALES[qp_rem];
        int64_t scaled = right_shift > 0 ? (value + ((int64_t)1 << (right_shift - 1))) >> right_shift
                                         : shl64(value, -right_shift);
        out[i] = clip16(scaled);
    }
    return 0;
}

/* Default high-frequency zero-out (vvc/dct.rs:1063-1072). */
static int default_skip(int type, int n) {
    if (type != RMH_TR_DCT2 && n == 32) return 16;
    return n > 32 ? n - 32 : 0;
}

/* inverse_transform_rect_bd (vvc/dct.rs:996-1109). */
int orc_inverse_transform(int w, int h, const int32_t* c, int th, int tv, int esw, int esh, int bd,
                          int32_t* out) {
    bd = clamp_bd(bd);
    int32_t mv[4096], mh[4096];
    if (w == 1 && h > 1) {
        if (orc_transform_matrix(tv, h, mv)) return -1;
        int skip_h = default_skip(tv, h);
        if (esh > skip_h) skip_h = esh;
        if (skip_h >= h) return -1;
        int shift = TRANSFORM_MATRIX_SHIFT + MAX_LOG2_TR_DYNAMIC_RANGE - bd;
        int64_t add = (int64_t)1 << (shift - 1);
        int cut = h - skip_h;
        for (int j = 0; j < h; j++) {
            int64_t sum = 0;
            for (int k = 0; k < cut; k++) sum += (int64_t)mv[k * h + j] * c[k];
            out[j] = clip16((sum + add) >> shift);
        }
        return 0;
    }
    if (h == 1 && w > 1) {
        if (orc_transform_matrix(th, w, mh)) return -1;
        int skip_w = default_skip(th, w);
        if (esw > skip_w) skip_w = esw;
        if (skip_w >= w) return -1;
        int shift = TRANSFORM_MATRIX_SHIFT + MAX_LOG2_TR_DYNAMIC_RANGE - bd;
        int64_t add = (int64_t)1 << (shift - 1);
        int cut = w - skip_w;
        for (int x = 0; x < w; x++) {
            int64_t sum = 0;
            for (int k = 0; k < cut; k++) sum += (int64_t)mh[k * w + x] * c[k];
            out[x] = clip16((sum + add) >> shift);
        }
        return 0;
    }
    if (orc_transform_matrix(th, w, mh) || orc_transform_matrix(tv, h, mv)) return -1;
    int skip_w = default_skip(th, w), skip_h = default_skip(tv, h);
    if (esw > skip_w) skip_w = esw;
    if (esh > skip_h) skip_h = esh;
    if (skip_w >= w || skip_h >= h) return -1;
    int shift1 = TRANSFORM_MATRIX_SHIFT + 1;
    int shift2 = (TRANSFORM_MATRIX_SHIFT + MAX_LOG2_TR_DYNAMIC_RANGE - 1) - bd;
    int64_t add1 = (int64_t)1 << (shift1 - 1), add2 = (int64_t)1 << (shift2 - 1);
    int cw = w - skip_w, ch = h - skip_h;
    int32_t* tmp = calloc((size_t)(w * h), sizeof(int32_t));
    for (int i = 0; i < cw; i++)
        for (int j = 0; j < h; j++) {
            int64_t sum = 0;
            for (int k = 0; k < ch; k++) sum += (int64_t)mv[k * h + j] * c[k * w + i];
            tmp[i * h + j] = clip16((sum + add1) >> shift1);
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int64_t sum = 0;
            for (int k = 0; k < cw; k++) sum += (int64_t)mh[k * w + x] * tmp[k * h + y];
            out[y * w + x] = clip16((sum + add2) >> shift2);
        }
    free(tmp);
    return 0;
}

/* forward_transform_rect_bd (vvc/dct.rs:865-991). */
int orc_forward_transform(int w, int h, const int32_t* r, int th, int tv, int esw, int esh, int bd,
                          int32_t* out) {
    bd = clamp_bd(bd);
    int32_t mv[4096], mh[4096];
    if (w == 1 && h > 1) {
        if (orc_transform_matrix(tv, h, mv)) return -1;
        int skip_h = default_skip(tv, h);
        if (esh > skip_h) skip_h = esh;
        if (skip_h >= h) return -1;
        int shift = (ilog2(h) + bd + TRANSFORM_MATRIX_SHIFT) - MAX_LOG2_TR_DYNAMIC_RANGE;
        int64_t add = shift > 0 ? ((int64_t)1 << (shift - 1)) : 0;
        memset(out, 0, sizeof(int32_t) * (size_t)h);
        for (int v = 0; v < h - skip_h; v++) {
            int64_t sum = 0;
            for (int y = 0; y < h; y++) sum += (int64_t)mv[v * h + y] * r[y];
            out[v] = clip16(shift >= 0 ? (sum + add) >> shift : shl64(sum, -shift));
        }
        return 0;
    }
    if (h == 1 && w > 1) {
        if (orc_transform_matrix(th, w, mh)) return -1;
        int skip_w = default_skip(th, w);
        if (esw > skip_w) skip_w = esw;
        if (skip_w >= w) return -1;
        int shift = (ilog2(w) + bd + TRANSFORM_MATRIX_SHIFT) - MAX_LOG2_TR_DYNAMIC_RANGE;
        int64_t add = shift > 0 ? ((int64_t)1 << (shift - 1)) : 0;
        memset(out, 0, sizeof(int32_t) * (size_t)w);
        for (int k = 0; k < w - skip_w; k++) {
            int64_t sum = 0;
            for (int n = 0; n < w; n++) sum += (int64_t)mh[k * w + n] * r[n];
            out[k] = clip16(shift >= 0 ? (sum + add) >> shift : shl64(sum, -shift));
        }
        return 0;
    }
    if (orc_transform_matrix(th, w, mh) || orc_transform_matrix(tv, h, mv)) return -1;
    int skip_w = default_skip(th, w), skip_h = default_skip(tv, h);
    if (esw > skip_w) skip_w = esw;
    if (esh > skip_h) skip_h = esh;
    if (skip_w >= w || skip_h >= h) return -1;
    int shift1 = (ilog2(w) + bd + TRANSFORM_MATRIX_SHIFT) - MAX_LOG2_TR_DYNAMIC_RANGE;
    int shift2 = ilog2(h) + TRANSFORM_MATRIX_SHIFT;
    int64_t add1 = shift1 > 0 ? ((int64_t)1 << (shift1 - 1)) : 0;
    int64_t add2 = (int64_t)1 << (shift2 - 1);
    int cw = w - skip_w, ch = h - skip_h;
    int32_t* tmp = calloc((size_t)(w * h), sizeof(int32_t));
    for (int y = 0; y < h; y++)
        for (int k = 0; k < cw; k++) {
            int64_t sum = 0;
            for (int n = 0; n < w; n++) sum += (int64_t)mh[k * w + n] * r[y * w + n];
            tmp[k * h + y] = (int32_t)((sum + add1) >> shift1);
        }
    memset(out, 0, sizeof(int32_t) * (size_t)(w * h));
    for (int v = 0; v < ch; v++)
        for (int k = 0; k < cw; k++) {
            int64_t sum = 0;
            for (int y = 0; y < h; y++) sum += (int64_t)mv[v * h + y] * tmp[k * h + y];
            out[v * w + k] = clip16((sum + add2) >> shift2);
        }
    free(tmp);
    return 0;
}

/* forward_transform (vvc/dct.rs:296-342): square 4..32, 8-bit, no final clip. */
int orc_forward_transform_square(int n, const int32_t* r, int th, int tv, int skip_w, int skip_h,
                                 int32_t* out) {
    int32_t mh[1024], mv[1024];
    if (!square_supported(n)) return -1;
    if (orc_transform_matrix(th, n, mh) || orc_transform_matrix(tv, n, mv)) return -1;
    if (skip_w >= n || skip_h >= n) return -1;
    int lg = ilog2(n);
    int shift1 = (lg + 8 + TRANSFORM_MATRIX_SHIFT) - MAX_LOG2_TR_DYNAMIC_RANGE;
    int shift2 = lg + TRANSFORM_MATRIX_SHIFT;
    int64_t add1 = (int64_t)1 << (shift1 - 1), add2 = (int64_t)1 << (shift2 - 1);
    int cw = n - skip_w, ch = n - skip_h;
    int32_t tmp[1024];
    memset(tmp, 0, sizeof(tmp));
    for (int y = 0; y < n; y++)
        for (int k = 0; k < cw; k++) {
            int64_t sum = 0;
            for (int j = 0; j < n; j++) sum += (int64_t)mh[k * n + j] * r[y * n + j];
            tmp[k * n + y] = (int32_t)((sum + add1) >> shift1);
        }
    memset(out, 0, sizeof(int32_t) * (size_t)(n * n));
    for (int v = 0; v < ch; v++)
        for (int k = 0; k < cw; k++) {
            int64_t sum = 0;
            for (int y = 0; y < n; y++) sum += (int64_t)mv[v * n + y] * tmp[k * n + y];
            out[v * n + k] = (int32_t)((sum + add2) >> shift2);
        }
    return 0;
}

/* parameters_rect (vvc/lfnst.rs:111-129). */
static int lfnst_params(int w, int h, int idx, int* sb, int* tr_size, int* zero_out) {
    if (idx < 1 || idx > 2) return -1;
    if (w < 4 || h < 4) return -1;
    *sb = (w >= 8 && h >= 8) ? 8 : 4;
    int small = (w == 4 && h == 4) || (w == 8 && h == 8);
    *tr_size = *sb == 4 ? 16 : 48;
    *zero_out = small ? 8 : 16;
    return 0;
}

/* g_coefTopLeftDiagScan8x8 (vvc/lfnst.rs:16-81) as raster positions in an 8-wide block. */
static void top_left_scan(int* xs, int* ys) {
    int32_t r[64];
    orc_coeff_scan(8, 8, r);
    for (int i = 0; i < 64; i++) {
        xs[i] = r[i] % 8;
        ys[i] = r[i] / 8;
    }
}

/* inverse_lfnst_rect (vvc/lfnst.rs:219-297). */
int orc_inverse_lfnst(int w, int h, int32_t* coeffs, int idx, int mode) {
    int sb, tr_size, zero_out;
    if (lfnst_params(w, h, idx, &sb, &tr_size, &zero_out)) return -1;
    int set = orc_lfnst_set(mode);
    if (set < 0) return -1;
    int tp = orc_lfnst_transposes(mode);
    int xs[64], ys[64];
    top_left_scan(xs, ys);
    int32_t src[16], out[48];
    for (int i = 0; i < 16; i++) src[i] = coeffs[ys[i] * w + xs[i]];
    memset(out, 0, sizeof(out));
    for (int j = 0; j < tr_size; j++) {
        int64_t sum = 0;
        for (int i = 0; i < zero_out; i++)
            sum += (int64_t)src[i] * orc_lfnst_coeff(set, idx - 1, sb == 8, i, j);
        out[j] = clip16((sum + 64) >> 7);
    }
    if (tp) {
        if (sb == 4) {
            for (int y = 0; y < 4; y++)
                for (int x = 0; x < 4; x++) coeffs[y * w + x] = out[y + 4 * x];
        } else {
            for (int y = 0; y < 8; y++) {
                for (int x = 0; x < 4; x++) coeffs[y * w + x] = out[y + 8 * x];
                if (y < 4)
                    for (int x = 4; x < 8; x++) coeffs[y * w + x] = out[y + 32 + 4 * (x - 4)];
            }
        }
    } else {
        int pos = 0;
        for (int y = 0; y < sb; y++) {
            int stride = y < 4 ? sb : 4;
            for (int x = 0; x < stride; x++) coeffs[y * w + x] = out[pos++];
        }
    }
    return 0;
}

/* forward_lfnst_rect (vvc/lfnst.rs:132-205). */
int orc_forward_lfnst(int w, int h, int32_t* coeffs, int idx, int mode) {
    int sb, tr_size, zero_out;
    if (lfnst_params(w, h, idx, &sb, &tr_size, &zero_out)) return -1;
    int set = orc_lfnst_set(mode);
    if (set < 0) return -1;
    int tp = orc_lfnst_transposes(mode);
    int32_t src[48];
    memset(src, 0, sizeof(src));
    if (tp) {
        if (sb == 4) {
            for (int y = 0; y < 4; y++)
                for (int x = 0; x < 4; x++) src[y + 4 * x] = coeffs[y * w + x];
        } else {
            for (int y = 0; y < 8; y++) {
                for (int x = 0; x < 4; x++) src[y + 8 * x] = coeffs[y * w + x];
                if (y < 4)
                    for (int x = 4; x < 8; x++) src[y + 32 + 4 * (x - 4)] = coeffs[y * w + x];
            }
        }
    } else {
        int pos = 0;
        for (int y = 0; y < sb; y++) {
            int stride = y < 4 ? sb : 4;
            for (int x = 0; x < stride; x++) src[pos++] = coeffs[y * w + x];
        }
    }
    int32_t out[48];
    memset(out, 0, sizeof(out));
    for (int j = 0; j < zero_out; j++) {
        int64_t sum = 0;
        for (int i = 0; i < tr_size; i++)
            sum += (int64_t)src[i] * orc_lfnst_coeff(set, idx - 1, sb == 8, j, i);
        out[j] = (int32_t)((sum + 64) >> 7);
    }
    int xs[64], ys[64];
    top_left_scan(xs, ys);
    for (int k = 0; k < tr_size; k++) coeffs[ys[k] * w + xs[k]] = out[k];
    return 0;
}

/* lfnst_skips (decode_general.rs:4233-4236). */
static void lfnst_skips(int w, int h, int* sw, int* sh) {
    int nz = (w == 4 || h == 4) ? 4 : 8;
    *sw = w - (nz < w ? nz : w);
    *sh = h - (nz < h ? nz : h);
}

/* residual_of + inverse_residual (decode_general.rs:4100-4199) for one TU. */
int orc_residual_tu(const rmh_tu_desc* t, const int16_t* coeff, int16_t* residual) {
    if (t->log2_w > 6 || t->log2_h > 6) return -1;
    int w = 1 << t->log2_w, h = 1 << t->log2_h;
    if (w * h < 2 || t->bit_depth < 8 || t->bit_depth > 12) return -1;
    if (t->tr_hor > 2 || t->tr_ver > 2 || t->lfnst_idx > 2 || t->lfnst_mode > 94) return -1;
    if (t->flags & ~(RMH_TU_TRANSFORM_SKIP | RMH_TU_DEP_QUANT)) return -1;
    int n = w * h;
    int32_t* lv = malloc(sizeof(int32_t) * (size_t)n);
    int32_t* dq = malloc(sizeof(int32_t) * (size_t)n);
    int32_t* out = malloc(sizeof(int32_t) * (size_t)n);
    int rc = 0;
    for (int i = 0; i < n; i++) lv[i] = coeff[i];
    if (t->flags & RMH_TU_TRANSFORM_SKIP) {
        for (int i = 0; i < n; i++) out[i] = orc_dequant_ts(lv[i], t->qp);
    } else {
        rc = (t->flags & RMH_TU_DEP_QUANT) ? orc_dequant_dep_quant(w, h, lv, t->qp, t->bit_depth, dq)
                                           : orc_dequant_rect(w, h, lv, t->qp, t->bit_depth, dq);
        int sw = t->skip_w, sh = t->skip_h;
        if (!rc && t->lfnst_idx) {
            rc = orc_inverse_lfnst(w, h, dq, t->lfnst_idx, t->lfnst_mode);
            int lw, lh;
            lfnst_skips(w, h, &lw, &lh);
            if (lw > sw) sw = lw;
            if (lh > sh) sh = lh;
        }
        if (!rc) rc = orc_inverse_transform(w, h, dq, t->tr_hor, t->tr_ver, sw, sh, t->bit_depth, out);
    }
    if (!rc)
        for (int i = 0; i < n; i++) residual[i] = (int16_t)out[i];
    free(lv);
    free(dq);
    free(out);
    return rc;
}

/* Threads of the batch drivers (orc_residual_batch, orc_recon_batch,
 * orc_mc_batch, orc_deblock_frames): 1 (the default) runs them serially in
 * the reference's order; more runs independent TUs, PUs or frames on an
 * OpenMP team (the CPU baseline over all host cores). Results do not
 * depend on it: no two TUs, PUs or frames of a valid batch share an output. */
int orc_threads = 1;

void orc_set_threads(int n) { orc_threads = n < 1 ? 1 : n; }

size_t orc_residual_batch(const rmh_tu_desc* tus, size_t n_tu, const int16_t* coeff, size_t n_coeff,
                          int16_t* residual) {
    for (size_t i = 0; i < n_tu; i++) {
        const rmh_tu_desc* t = &tus[i];
        if (t->log2_w > 6 || t->log2_h > 6) return i + 1;
        size_t n = (size_t)1 << (t->log2_w + t->log2_h);
        if ((size_t)t->offset + n > n_coeff) return i + 1;
    }
    size_t bad = 0;
#pragma omp parallel for num_threads(orc_threads) schedule(dynamic, 64) if (orc_threads > 1) reduction(max : bad)
    for (size_t i = 0; i < n_tu; i++)
        if (orc_residual_tu(&tus[i], coeff + tus[i].offset, residual + tus[i].offset)) bad = bad > i + 1 ? bad : i + 1;
    return bad;
}

/* quantize_ts_with_mode (vvc/transform.rs:415-427): transform-skip
 * quantisation with the intra (171) or inter (85) rounding offset. */
int32_t orc_quant_ts_mode(int32_t residual, int qp, int is_intra) {
    if (qp < 4) qp = 4;
    int qp_per = qp / 6, qp_rem = qp % 6;
    int i_qbits = 14 + qp_per;
    int64_t add = (int64_t)(is_intra ? 171 : 85) << (i_qbits - 9);
    int sign = residual > 0 ? 1 : residual < 0 ? -1 : 0;
    int64_t a = residual < 0 ? -(int64_t)residual : residual;
    int32_t level = (int32_t)((a * QUANT_SCALES[qp_rem] + add) >> i_qbits);
    int64_t v = (int64_t)sign * level;
    return (int32_t)(v < -32768 ? -32768 : v > 32767 ? 32767 : v);
}

/* The encoder's forward path for one TU, in the order of plan_primary
 * (vvc/slice.rs:718-761) and the 8..12-bit encoder (slice.rs:4279-4296):
 * forward_transform_rect_bd with the TU's zero-out (widened to the LFNST
 * region when lfnst_idx, as the decoder's inverse_residual does), then
 * forward_lfnst_rect, then quantize_rect_bd (intra rounding unless
 * RMH_TU_QUANT_INTER) or quantize_dep_quant_bd; transform-skip TUs through
 * quantize_ts_with_mode per sample. */
int orc_forward_tu(const rmh_tu_desc* t, const int16_t* residual, int16_t* levels) {
    if (t->log2_w > 6 || t->log2_h > 6) return -1;
    int w = 1 << t->log2_w, h = 1 << t->log2_h;
    if (w * h < 2 || t->bit_depth < 8 || t->bit_depth > 12) return -1;
    if (t->tr_hor > 2 || t->tr_ver > 2 || t->lfnst_idx > 2 || t->lfnst_mode > 94) return -1;
    if (t->flags & ~(RMH_TU_TRANSFORM_SKIP | RMH_TU_DEP_QUANT | RMH_TU_QUANT_INTER)) return -1;
    const int intra = !(t->flags & RMH_TU_QUANT_INTER);
    int n = w * h;
    int32_t* r = malloc(sizeof(int32_t) * (size_t)n);
    int32_t* c = malloc(sizeof(int32_t) * (size_t)n);
    int32_t* lv = malloc(sizeof(int32_t) * (size_t)n);
    int rc = 0;
    for (int i = 0; i < n; i++) r[i] = residual[i];
    if (t->flags & RMH_TU_TRANSFORM_SKIP) {
        for (int i = 0; i < n; i++) lv[i] = orc_quant_ts_mode(r[i], t->qp, intra);
    } else {
        int sw = t->skip_w, sh = t->skip_h;
        if (t->lfnst_idx) {
            int lw, lh;
            lfnst_skips(w, h, &lw, &lh);
            if (lw > sw) sw = lw;
            if (lh > sh) sh = lh;
        }
        rc = orc_forward_transform(w, h, r, t->tr_hor, t->tr_ver, sw, sh, t->bit_depth, c);
        if (!rc && t->lfnst_idx) rc = orc_forward_lfnst(w, h, c, t->lfnst_idx, t->lfnst_mode);
        if (!rc)
            rc = (t->flags & RMH_TU_DEP_QUANT) ? orc_quant_dep_quant(w, h, c, t->qp, t->bit_depth, lv)
                                               : orc_quant_rect(w, h, c, t->qp, t->bit_depth, intra, lv);
    }
    if (!rc)
        for (int i = 0; i < n; i++) levels[i] = (int16_t)lv[i];
    free(r);
    free(c);
    free(lv);
    return rc;
}

size_t orc_forward_batch(const rmh_tu_desc* tus, size_t n_tu, const int16_t* residual, size_t n_coeff,
                         int16_t* levels) {
    for (size_t i = 0; i < n_tu; i++) {
        const rmh_tu_desc* t = &tus[i];
        if (t->log2_w > 6 || t->log2_h > 6) return i + 1;
        size_t n = (size_t)1 << (t->log2_w + t->log2_h);
        if ((size_t)t->offset + n > n_coeff) return i + 1;
    }
    size_t bad = 0;
#pragma omp parallel for num_threads(orc_threads) schedule(dynamic, 64) if (orc_threads > 1) reduction(max : bad)
    for (size_t i = 0; i < n_tu; i++)
        if (orc_forward_tu(&tus[i], residual + tus[i].offset, levels + tus[i].offset)) bad = bad > i + 1 ? bad : i + 1;
    return bad;
}

/* Residual added to the prediction held in the picture and clipped to the
 * sample range: the add of reconstruct_tu (vvc/decode_general.rs:4086-4092,
 * `(pred + residual).clamp(0, max_val)`), which reconstruct_tus applies to the
 * inter prediction written by write_prediction (decode_general.rs:2641-2667,
 * 2866-2870). Planes are host memory; TUs in descriptor order. */
/* pred + res clipped to the sample range, over a w x h block at (x0, y0). */
static void add_block(const rmh_plane* p, int x0, int y0, int w, int h, const int32_t* res) {
    const int maxv = (1 << p->bit_depth) - 1;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const size_t at = (size_t)(y0 + y) * p->stride + x0 + x;
            int v = res[y * w + x];
            if (p->bytes_per_sample == 1) {
                uint8_t* s = (uint8_t*)p->data + at;
                v += *s;
                *s = (uint8_t)(v < 0 ? 0 : v > maxv ? maxv : v);
            } else {
                uint16_t* s = (uint16_t*)p->data + at;
                v += *s;
                *s = (uint16_t)(v < 0 ? 0 : v > maxv ? maxv : v);
            }
        }
}

size_t orc_recon_batch(const rmh_tu_desc* tus, size_t n_tu, const int16_t* coeff, size_t n_coeff,
                       const rmh_tu_place* place, const rmh_recon_frame* frames, size_t n_frames) {
    for (size_t i = 0; i < n_tu; i++) {
        const rmh_tu_desc* t = &tus[i];
        if (t->log2_w > 6 || t->log2_h > 6) return i + 1;
        const int w = 1 << t->log2_w, h = 1 << t->log2_h;
        if ((size_t)t->offset + (size_t)w * h > n_coeff) return i + 1;
        if (place[i].frame >= n_frames || place[i].plane > 2) return i + 1;
        const rmh_plane* p = &frames[place[i].frame].planes[place[i].plane];
        if ((uint32_t)place[i].x + w > p->width || (uint32_t)place[i].y + h > p->height) return i + 1;
        if (t->joint & 3) {
            if (place[i].plane != ((t->joint & 3) == 1 ? 2 : 1)) return i + 1;
            const rmh_plane* q = &frames[place[i].frame].planes[3 - place[i].plane];
            if (q->width != p->width || q->height != p->height) return i + 1;
        }
    }
    size_t bad = 0;
#pragma omp parallel for num_threads(orc_threads) schedule(dynamic, 64) if (orc_threads > 1) reduction(max : bad)
    for (size_t i = 0; i < n_tu; i++) {
        int16_t res[64 * 64];
        const rmh_tu_desc* t = &tus[i];
        const int w = 1 << t->log2_w, h = 1 << t->log2_h;
        const rmh_plane* p = &frames[place[i].frame].planes[place[i].plane];
        if (orc_residual_tu(t, coeff + t->offset, res)) {
            bad = bad > i + 1 ? bad : i + 1;
            continue;
        }
        int32_t r32[64 * 64];
        for (int k = 0; k < w * h; k++) r32[k] = res[k];
        add_block(p, place[i].x, place[i].y, w, h, r32);
        if (t->joint & 3) {
            /* reconstruct_joint_cbcr (vvc/decode_general.rs:3908-3981): with
             * mode = +-3 / +-1 / +-2 for TuCResMode 1 / 2 / 3 (the sign from
             * ph_joint_cbcr_sign_flag), the other component is res_cb = res_cr >> 1,
             * res_cr = res_cb >> 1 or res_cr = res_cb, each negated first
             * (-res >> 1) under the sign, then added to its prediction and
             * clamped like the coded one. */
            const int mode = t->joint & 3, neg = (t->joint >> 2) & 1;
            for (int k = 0; k < w * h; k++) {
                const int32_t r = neg ? -r32[k] : r32[k]; /* i32 in the reference: -(-32768) stays 32768 */
                r32[k] = mode == 3 ? r : r >> 1;
            }
            add_block(&frames[place[i].frame].planes[3 - place[i].plane], place[i].x, place[i].y, w, h, r32);
        }
    }
    return bad;
}

/*
 * orc_intra.c - the intra-profile subset codec of the CPU oracle (test
 * infrastructure only): encoder decisions and reconstruction of
 * crates/rustmpeg-video/src/vvc/slice.rs:325-455 (encode_picture),
 * slice.rs:640-1097 (plan_primary, encode_chroma, encode_luma, candidate
 * checks, reconstruct_block) and the PLANAR / DC predictors of
 * vvc/predict.rs:1084-1300, 1688-1760, 2619-2730.
 *
 * Entropy coding is not restated: the encoder hands its decisions (intra
 * mode, transform choice, quantised levels) straight to the decoder as the TU
 * batch the residual stage consumes. CTU = CU = TU = 32 luma / 16 chroma,
 * 8-bit 4:2:0, one slice.
 */
#include <stdlib.h>
#include <string.h>

#include "vvc_oracle.h"

typedef struct {
    int32_t above[128];
    int32_t left[128];
    int32_t top_left;
} refs_t;

static int ilog2i(int v) {
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

/* gather_references (vvc/predict.rs:1084-1139) */
static void gather(const uint8_t* p, int pw, int ph, int bx, int by, int w, int h, refs_t* r) {
    int above_len = by == 0 ? 0 : (pw - bx < 2 * w ? pw - bx : 2 * w);
    int left_len = bx == 0 ? 0 : (ph - by < h ? ph - by : h);
    memset(r, 0, sizeof *r);
    for (int i = 0; i < above_len; i++) r->above[i] = p[(by - 1) * pw + bx + i];
    for (int i = 0; i < left_len; i++) r->left[i] = p[(by + i) * pw + bx - 1];
    if (above_len > 0 && left_len > 0)
        r->top_left = p[(by - 1) * pw + bx - 1];
    else if (above_len > 0)
        r->top_left = r->above[0];
    else if (left_len > 0)
        r->top_left = r->left[0];
    else
        r->top_left = 128;
    if (above_len == 0) {
        int v = left_len > 0 ? r->left[0] : 128;
        for (int i = 0; i < 2 * w; i++) r->above[i] = v;
    } else {
        for (int i = above_len; i < 2 * w; i++) r->above[i] = r->above[above_len - 1];
    }
    if (left_len == 0) {
        int v = above_len > 0 ? r->above[0] : 128;
        for (int i = 0; i < 2 * h; i++) r->left[i] = v;
    } else {
        for (int i = left_len; i < 2 * h; i++) r->left[i] = r->left[left_len - 1];
    }
}

/* RefSamples::filtered (vvc/predict.rs:1233-1273) */
static void filter_refs(const refs_t* s, int w, int h, refs_t* o) {
    int w2 = 2 * w, h2 = 2 * h;
    *o = *s;
    o->top_left = (2 * s->top_left + s->above[0] + s->left[0] + 2) >> 2;
    for (int j = 0; j < w2 - 1; j++) {
        int prev = j == 0 ? s->top_left : s->above[j - 1];
        o->above[j] = (prev + 2 * s->above[j] + s->above[j + 1] + 2) >> 2;
    }
    o->above[w2 - 1] = s->above[w2 - 1];
    for (int j = 0; j < h2 - 1; j++) {
        int prev = j == 0 ? s->top_left : s->left[j - 1];
        o->left[j] = (prev + 2 * s->left[j] + s->left[j + 1] + 2) >> 2;
    }
    o->left[h2 - 1] = s->left[h2 - 1];
}

/* predict_block for PLANAR (mode 0) / DC (mode 1): reference filter for luma
 * planar with w*h > 32, then the predictor, then PDPC (vvc/predict.rs:2619-2646). */
void orc_intra_predict(const uint8_t* plane, int pw, int ph, int bx, int by, int w, int h, int is_luma, int mode,
                       int32_t* pred) {
    refs_t raw, filt;
    gather(plane, pw, ph, bx, by, w, h, &raw);
    const refs_t* r = &raw;
    if (mode == 0 && is_luma && w * h > 32) {
        filter_refs(&raw, w, h, &filt);
        r = &filt;
    }
    int lw = ilog2i(w), lh = ilog2i(h);
    if (mode == 0) {
        /* predict_planar (vvc/predict.rs:1688-1719) */
        int offset = 1 << (lw + lh), fs = 1 + lw + lh;
        int32_t top[65], left[65], bottom[64], right[64];
        for (int i = 0; i <= w; i++) top[i] = r->above[i];
        for (int i = 0; i <= h; i++) left[i] = r->left[i];
        int tr = top[w], bl = left[h];
        for (int k = 0; k < w; k++) bottom[k] = bl - top[k];
        for (int k = 0; k < h; k++) right[k] = tr - left[k];
        for (int i = 0; i <= w; i++) top[i] <<= lh;
        for (int i = 0; i <= h; i++) left[i] <<= lw;
        for (int y = 0; y < h; y++) {
            int hor = left[y];
            for (int x = 0; x < w; x++) {
                hor += right[y];
                top[x] += bottom[x];
                pred[y * w + x] = ((hor << lh) + (top[x] << lw) + offset) >> fs;
            }
        }
    } else {
        /* predict_dc (vvc/predict.rs:1722-1736) */
        int denom = w == h ? w << 1 : (w > h ? w : h);
        int ds = ilog2i(denom), doff = denom >> 1, sum = 0;
        if (w >= h)
            for (int i = 0; i < w; i++) sum += r->above[i];
        if (w <= h)
            for (int i = 0; i < h; i++) sum += r->left[i];
        int dc = (sum + doff) >> ds;
        for (int i = 0; i < w * h; i++) pred[i] = dc;
    }
    /* apply_pdpc (vvc/predict.rs:1739-1754) */
    int scale = (lw + lh >= 2 ? lw + lh - 2 : 0) >> 2;
    for (int y = 0; y < h; y++) {
        int sy = (y * 2) >> scale;
        int wt = 32 >> (sy < 31 ? sy : 31);
        int left = r->left[y];
        for (int x = 0; x < w; x++) {
            int sx = (x * 2) >> scale;
            int wl = 32 >> (sx < 31 ? sx : 31);
            int val = pred[y * w + x];
            pred[y * w + x] = val + ((wl * (left - val) + wt * (r->above[x] - val) + 32) >> 6);
        }
    }
}

/* choose_intra_mode (vvc/predict.rs:2706-2730): DC only when its SAD is strictly lower. */
static int choose_mode(const uint8_t* rec, const uint8_t* src, int pw, int ph, int bx, int by, int n) {
    int32_t pred[1024];
    uint64_t sad[2];
    for (int m = 0; m < 2; m++) {
        orc_intra_predict(rec, pw, ph, bx, by, n, n, 1, m, pred);
        uint64_t s = 0;
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                int d = src[(by + y) * pw + bx + x] - pred[y * n + x];
                s += (uint64_t)(d < 0 ? -d : d);
            }
        sad[m] = s;
    }
    return sad[1] < sad[0] ? 1 : 0;
}

int orc_intra_choose_mode(const uint8_t* rec, const uint8_t* src, int pw, int ph, int bx, int by, int n) {
    return choose_mode(rec, src, pw, ph, bx, by, n);
}

typedef struct {
    int32_t levels[1024];
    uint8_t samples[1024];
    int cbf, ts, mts_idx, lfnst_idx;
} plan_t;

static const int MTS_H[5] = {RMH_TR_DCT2, RMH_TR_DST7, RMH_TR_DCT8, RMH_TR_DST7, RMH_TR_DCT8};
static const int MTS_V[5] = {RMH_TR_DCT2, RMH_TR_DST7, RMH_TR_DST7, RMH_TR_DCT8, RMH_TR_DCT8};

static void skips_for(int n, int mts_idx, int lfnst_idx, int* sw, int* sh) {
    if (lfnst_idx) {
        int nz = n == 4 ? 4 : 8; /* lfnst_skips, slice.rs:674-682 (square blocks) */
        *sw = *sh = n - (nz < n ? nz : n);
    } else {
        *sw = (n == 32 && MTS_H[mts_idx] != RMH_TR_DCT2) ? 16 : 0;
        *sh = (n == 32 && MTS_V[mts_idx] != RMH_TR_DCT2) ? 16 : 0;
    }
}

/* plan_primary (slice.rs:716-766) */
static int plan_primary(const int32_t* pred, const int32_t* res, int n, int qp, int mts_idx, int lfnst_idx, int mode,
                        plan_t* pl) {
    int sw, sh;
    int32_t coeffs[1024], dq[1024], rec[1024];
    skips_for(n, mts_idx, lfnst_idx, &sw, &sh);
    if (orc_forward_transform(n, n, res, MTS_H[mts_idx], MTS_V[mts_idx], sw, sh, 8, coeffs)) return -1;
    if (lfnst_idx && orc_forward_lfnst(n, n, coeffs, lfnst_idx, mode)) return -1;
    orc_quant_rect(n, n, coeffs, qp, 8, 1, pl->levels);
    pl->cbf = 0;
    for (int i = 0; i < n * n; i++) pl->cbf |= pl->levels[i] != 0;
    orc_dequant_rect(n, n, pl->levels, qp, 8, dq);
    if (lfnst_idx && orc_inverse_lfnst(n, n, dq, lfnst_idx, mode)) return -1;
    if (orc_inverse_transform(n, n, dq, MTS_H[mts_idx], MTS_V[mts_idx], sw, sh, 8, rec)) return -1;
    for (int i = 0; i < n * n; i++) {
        int v = pred[i] + rec[i];
        pl->samples[i] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
    pl->ts = 0;
    pl->mts_idx = mts_idx;
    pl->lfnst_idx = lfnst_idx;
    return 0;
}

/* plan_transform_skip (slice.rs:693-713) */
static void plan_ts(const int32_t* pred, const int32_t* res, int n, int qp, plan_t* pl) {
    pl->cbf = 0;
    for (int i = 0; i < n * n; i++) {
        int32_t lv = orc_quant_ts(res[i], qp);
        pl->levels[i] = lv;
        pl->cbf |= lv != 0;
        int v = pred[i] + orc_dequant_ts(lv, qp);
        pl->samples[i] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
    pl->ts = 1;
    pl->mts_idx = pl->lfnst_idx = 0;
}

/* scan_pos_last (slice.rs:894-898); -1 when the block is all zero */
static int scan_pos_last(int n, const int32_t* levels) {
    int32_t scan[1024];
    int len = orc_coeff_scan(n, n, scan);
    for (int s = len - 1; s >= 0; s--)
        if (levels[scan[s]] != 0) return s;
    return -1;
}

static int max_lfnst_pos(int n) { return (n == 4 || n == 8) ? 7 : 15; }

/* violates_mts_constraint (slice.rs:909-919) */
static int violates_mts(int n, const int32_t* levels) {
    for (int i = 0; i < n * n; i++)
        if (levels[i] && ((i % n) / 4 > 3 || (i / n) / 4 > 3)) return 1;
    return 0;
}

static uint64_t block_sse(const uint8_t* src, int pw, int bx, int by, int n, const uint8_t* s) {
    uint64_t t = 0;
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
            int64_t d = (int64_t)src[(by + y) * pw + bx + x] - s[y * n + x];
            t += (uint64_t)(d * d);
        }
    return t;
}

static void residual_of(const uint8_t* src, int pw, int bx, int by, int n, const int32_t* pred, int32_t* res) {
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) res[y * n + x] = src[(by + y) * pw + bx + x] - pred[y * n + x];
}

static void apply_plan(uint8_t* rec, int pw, int bx, int by, int n, const plan_t* pl) {
    for (int y = 0; y < n; y++) memcpy(rec + (by + y) * pw + bx, pl->samples + y * n, (size_t)n);
}

/* CuConstraints::update violates_lfnst for chroma (slice.rs:933-951) */
static int chroma_violates_lfnst(int n, const plan_t* pl) {
    if (!pl->cbf || pl->ts) return 0;
    int last = scan_pos_last(n, pl->levels);
    return last >= 0 && n >= 4 && last > max_lfnst_pos(n);
}

static void emit_tu(rmh_tu_desc* t, int16_t* coeff, uint32_t* off, int n, int qp, const plan_t* pl, int mode) {
    memset(t, 0, sizeof *t);
    t->offset = *off;
    t->log2_w = t->log2_h = (uint8_t)ilog2i(n);
    t->qp = (uint8_t)qp;
    t->bit_depth = 8;
    t->tr_hor = (uint8_t)MTS_H[pl->mts_idx];
    t->tr_ver = (uint8_t)MTS_V[pl->mts_idx];
    t->lfnst_idx = (uint8_t)pl->lfnst_idx;
    t->lfnst_mode = (uint8_t)mode; /* PLANAR = 0, DC = 1 (lfnst_mode, slice.rs:685-691) */
    t->flags = pl->ts ? RMH_TU_TRANSFORM_SKIP : 0;
    for (int i = 0; i < n * n; i++) coeff[*off + i] = (int16_t)pl->levels[i];
    *off += (uint32_t)(n * n);
}

/* Encoder decisions of encode_picture (slice.rs:325-455) for one 8-bit 4:2:0
 * frame (w, h multiples of 32). Outputs, per CTU in raster order, the mode and
 * three TUs (Y 32x32, Cb 16x16, Cr 16x16; a TU without coded coefficients is
 * still listed, all levels zero), the coefficients, and the encoder's own
 * reconstruction. Returns the number of CTUs, or -1. */
int orc_intra_encode(const uint8_t* sy, const uint8_t* su, const uint8_t* sv, int w, int h, int qp,
                     uint8_t* ry, uint8_t* ru, uint8_t* rv, uint8_t* modes, rmh_tu_desc* tus, int16_t* coeff) {
    if (w % 32 || h % 32 || w <= 0 || h <= 0) return -1;
    int luma_qp = qp < 4 ? 4 : qp;
    int cq = qp < 0 ? 0 : qp > 63 ? 63 : qp;
    cq = cq <= 26 ? cq : cq - 1; /* derived_chroma_qp, vvc/hls.rs:778-781 */
    int chroma_qp = cq < 4 ? 4 : cq;
    int luma_ts = luma_qp == 4, chroma_ts = chroma_qp == 4;
    int cw = w / 2;
    memcpy(ry, sy, (size_t)w * h);
    memcpy(ru, su, (size_t)cw * (h / 2));
    memcpy(rv, sv, (size_t)cw * (h / 2));
    uint32_t off = 0;
    int ctus = 0;
    int32_t pred[1024], res[1024];
    plan_t* cand = malloc(sizeof(plan_t) * 2);
    plan_t* best = malloc(sizeof(plan_t));
    plan_t* chroma = malloc(sizeof(plan_t) * 2);
    for (int cy0 = 0; cy0 < h; cy0 += 32)
        for (int cx0 = 0; cx0 < w; cx0 += 32) {
            int mode = choose_mode(ry, sy, w, h, cx0, cy0, 32);
            modes[ctus] = (uint8_t)mode;
            /* encode_chroma (slice.rs:771-807) */
            int bx = cx0 / 2, by = cy0 / 2;
            for (int c = 0; c < 2; c++) {
                const uint8_t* s = c ? sv : su;
                uint8_t* r = c ? rv : ru;
                orc_intra_predict(r, cw, h / 2, bx, by, 16, 16, 0, mode, pred);
                residual_of(s, cw, bx, by, 16, pred, res);
                if (chroma_ts)
                    plan_ts(pred, res, 16, chroma_qp, &chroma[c]);
                else
                    plan_primary(pred, res, 16, chroma_qp, 0, 0, mode, &chroma[c]);
            }
            int chroma_can_lfnst = !(chroma_violates_lfnst(16, &chroma[0]) || chroma_violates_lfnst(16, &chroma[1]));
            /* encode_luma (slice.rs:810-874) */
            orc_intra_predict(ry, w, h, cx0, cy0, 32, 32, 1, mode, pred);
            residual_of(sy, w, cx0, cy0, 32, pred, res);
            if (luma_ts) {
                plan_ts(pred, res, 32, luma_qp, best);
            } else {
                plan_primary(pred, res, 32, luma_qp, 0, 0, mode, best);
                uint64_t best_sse = block_sse(sy, w, cx0, cy0, 32, best->samples);
                for (int m = 1; m <= 4; m++) {
                    plan_primary(pred, res, 32, luma_qp, m, 0, mode, &cand[0]);
                    int last = scan_pos_last(32, cand[0].levels);
                    if (!cand[0].cbf || last < 1 || violates_mts(32, cand[0].levels)) continue;
                    uint64_t sse = block_sse(sy, w, cx0, cy0, 32, cand[0].samples);
                    if (sse < best_sse) {
                        best_sse = sse;
                        *best = cand[0];
                    }
                }
                if (chroma_can_lfnst)
                    for (int l = 1; l <= 2; l++) {
                        plan_primary(pred, res, 32, luma_qp, 0, l, mode, &cand[0]);
                        int last = scan_pos_last(32, cand[0].levels);
                        if (!cand[0].cbf || last < 1 || last > max_lfnst_pos(32)) continue;
                        uint64_t sse = block_sse(sy, w, cx0, cy0, 32, cand[0].samples);
                        if (sse < best_sse) {
                            best_sse = sse;
                            *best = cand[0];
                        }
                    }
            }
            apply_plan(ry, w, cx0, cy0, 32, best);
            apply_plan(ru, cw, bx, by, 16, &chroma[0]);
            apply_plan(rv, cw, bx, by, 16, &chroma[1]);
            emit_tu(&tus[3 * ctus + 0], coeff, &off, 32, luma_qp, best, mode);
            emit_tu(&tus[3 * ctus + 1], coeff, &off, 16, chroma_qp, &chroma[0], mode);
            emit_tu(&tus[3 * ctus + 2], coeff, &off, 16, chroma_qp, &chroma[1], mode);
            ctus++;
        }
    free(cand);
    free(best);
    free(chroma);
    return ctus;
}

/* Decoder reconstruction (decode_picture -> reconstruct_block, slice.rs:458-632,
 * 1050-1097) on the CPU, CTUs in raster order, from the TU batch. */
int orc_intra_decode(int w, int h, const uint8_t* modes, const rmh_tu_desc* tus, const int16_t* coeff, uint8_t* ry,
                     uint8_t* ru, uint8_t* rv) {
    int32_t pred[1024];
    int16_t res[1024];
    int cw = w / 2, ctus = 0;
    for (int cy0 = 0; cy0 < h; cy0 += 32)
        for (int cx0 = 0; cx0 < w; cx0 += 32) {
            for (int c = 0; c < 3; c++) {
                const rmh_tu_desc* t = &tus[3 * ctus + c];
                int n = c ? 16 : 32, pw = c ? cw : w, ph = c ? h / 2 : h;
                int bx = c ? cx0 / 2 : cx0, by = c ? cy0 / 2 : cy0;
                uint8_t* p = c == 0 ? ry : c == 1 ? ru : rv;
                orc_intra_predict(p, pw, ph, bx, by, n, n, c == 0, modes[ctus], pred);
                if (orc_residual_tu(t, coeff + t->offset, res)) return -1;
                for (int y = 0; y < n; y++)
                    for (int x = 0; x < n; x++) {
                        int v = pred[y * n + x] + res[y * n + x];
                        p[(by + y) * pw + bx + x] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
                    }
            }
            ctus++;
        }
    return ctus;
}

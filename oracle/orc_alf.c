/*
 * orc_alf.c - CPU oracle for luma mapping (LMCS) and the adaptive loop
 * filters (ALF, CCALF).
 *
 * TEST INFRASTRUCTURE ONLY (see vvc_oracle.h). Scalar restatement of
 *   LmcsReshaper16::from_aps        crates/rustmpeg-video/src/vvc/filter.rs:2628-2705
 *   apply_{forward,inverse}_plane_16                                filter.rs:2707-2722
 *   chroma_scale_for_luma, scale_chroma_residual                    filter.rs:2724-2750
 *   derive_alf_class_4x4_vb_samples                                 filter.rs:1580-1691
 *   apply_alf                       crates/rustmpeg-video/src/vvc/decode_general.rs:5225-5333
 *   ccalf_block                                                     decode_general.rs:5335-5409
 *   alf_luma_block                                                  decode_general.rs:5411-5566
 *   alf_chroma_block                                                decode_general.rs:5568-5660
 * with the tables ALF_FIXED_FILTER_COEFF, ALF_CLASS_TO_FILTER_MAPPING
 * (filter.rs:1374, 1443; alf_data.h), ALF_TRANSPOSE_PERM (filter.rs:2036-2041)
 * and the clip tables ALF_CLIP_VALS_{8,10,12}BIT (filter.rs:1509, 1872, 1876).
 */
#include <stdlib.h>
#include <string.h>

#include "alf_data.h"
#include "vvc_oracle.h"

#define LMCS_FP_PREC 11 /* filter.rs:22 */

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
static int mini(int a, int b) { return a < b ? a : b; }
static int maxi(int a, int b) { return a > b ? a : b; }
static int absi(int v) { return v < 0 ? -v : v; }

static int px(const rmh_plane* p, int x, int y) {
    const uint8_t* row = (const uint8_t*)p->data + (size_t)y * p->stride * p->bytes_per_sample;
    return p->bytes_per_sample == 1 ? row[x] : ((const uint16_t*)row)[x];
}

static void put(const rmh_plane* p, int x, int y, int v) {
    uint8_t* row = (uint8_t*)p->data + (size_t)y * p->stride * p->bytes_per_sample;
    if (p->bytes_per_sample == 1)
        row[x] = (uint8_t)v;
    else
        ((uint16_t*)row)[x] = (uint16_t)v;
}

/* ---- LMCS ---- */

/* LmcsReshaper16::from_aps (filter.rs:2628-2705). */
int orc_lmcs_tables(const rmh_lmcs_aps* aps, int bit_depth, uint16_t* fwd_lut, uint16_t* inv_lut,
                    int32_t* chroma_adj_lut, int32_t* reshape_pivot_out) {
    const int bd = clampi(bit_depth, 8, 12);
    const int max_val = (1 << bd) - 1;
    const int log2_cw = bd - 4;
    const int init_cw = 1 << log2_cw;
    const int min_bin = mini(aps->min_bin_idx, 15);
    const int max_bin = maxi(mini(aps->max_bin_idx, 15), min_bin);
    int bin_cw[16] = {0};
    for (int i = min_bin; i <= max_bin; i++) bin_cw[i] = maxi(init_cw + aps->bin_cw_delta[i], 0);
    int input_pivot[17] = {0}, reshape_pivot[17] = {0};
    int fwd_scale[16], inv_scale[16], cadj[16];
    for (int i = 0; i < 16; i++) {
        reshape_pivot[i + 1] = reshape_pivot[i] + bin_cw[i];
        input_pivot[i + 1] = input_pivot[i] + init_cw;
        fwd_scale[i] = ((bin_cw[i] << LMCS_FP_PREC) + (1 << (log2_cw - 1))) >> log2_cw;
        if (bin_cw[i] == 0) {
            inv_scale[i] = 0;
            cadj[i] = 1 << LMCS_FP_PREC;
        } else {
            inv_scale[i] = (init_cw << LMCS_FP_PREC) / bin_cw[i];
            const int denom = maxi(bin_cw[i] + aps->chroma_res_scaling_offset, 1);
            cadj[i] = (init_cw << LMCS_FP_PREC) / denom;
        }
    }
    for (int s = 0; s <= max_val; s++) {
        const int iy = mini(s >> log2_cw, 15);
        const int fv = reshape_pivot[iy] +
                       ((fwd_scale[iy] * (s - input_pivot[iy]) + (1 << (LMCS_FP_PREC - 1))) >> LMCS_FP_PREC);
        if (fwd_lut) fwd_lut[s] = (uint16_t)clampi(fv, 0, max_val);
        /* get_pwl_idx_inv: the first bin from min_bin whose upper pivot exceeds s */
        int idx = min_bin;
        while (idx <= max_bin) {
            if (s < reshape_pivot[idx + 1]) break;
            idx++;
        }
        idx = mini(idx, 15);
        /* i32 arithmetic wrapping as in a release build of the reference */
        const int32_t t = (int32_t)((uint32_t)inv_scale[idx] * (uint32_t)(s - reshape_pivot[idx]) +
                                    (uint32_t)(1 << (LMCS_FP_PREC - 1)));
        const int iv = input_pivot[idx] + (t >> LMCS_FP_PREC);
        if (inv_lut) inv_lut[s] = (uint16_t)clampi(iv, 0, max_val);
    }
    if (chroma_adj_lut) memcpy(chroma_adj_lut, cadj, sizeof(cadj));
    if (reshape_pivot_out) memcpy(reshape_pivot_out, reshape_pivot, sizeof(reshape_pivot));
    return 0;
}

/* apply_forward_plane_16 / apply_inverse_plane_16 (filter.rs:2707-2722): every
 * sample of the plane through the LUT, indices above the last entry clamped. */
int orc_lmcs_map(const rmh_plane* src, const rmh_plane* dst, const uint16_t* lut, int bit_depth) {
    const int cap = (1 << clampi(bit_depth, 8, 12)) - 1;
    for (uint32_t y = 0; y < src->height; y++)
        for (uint32_t x = 0; x < src->width; x++) put(dst, (int)x, (int)y, lut[mini(px(src, (int)x, (int)y), cap)]);
    return 0;
}

/* chroma_scale_for_luma + scale_chroma_residual (filter.rs:2724-2750). */
int orc_lmcs_scale_chroma(const rmh_lmcs_aps* aps, int bit_depth, int avg_luma, int32_t* residual, int n) {
    int32_t cadj[16], pivot[17];
    orc_lmcs_tables(aps, bit_depth, NULL, NULL, cadj, pivot);
    const int min_bin = mini(aps->min_bin_idx, 15);
    const int max_bin = maxi(mini(aps->max_bin_idx, 15), min_bin);
    int idx = min_bin;
    while (idx <= max_bin) {
        if (avg_luma < pivot[idx + 1]) break;
        idx++;
    }
    const int scale = cadj[mini(idx, 15)];
    const int max_val = (1 << clampi(bit_depth, 8, 12)) - 1;
    for (int i = 0; i < n; i++) {
        const int v = clampi(residual[i], -max_val - 1, max_val);
        const int sign = v < 0 ? -1 : 1;
        residual[i] = sign * ((absi(v) * scale + (1 << (LMCS_FP_PREC - 1))) >> LMCS_FP_PREC);
    }
    return scale;
}

/* ---- ALF ---- */

static const int ALF_TRANSPOSE_PERM[4][12] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11},
    {9, 4, 10, 8, 1, 5, 11, 7, 3, 0, 2, 6},
    {0, 3, 2, 1, 8, 7, 6, 5, 4, 9, 10, 11},
    {9, 8, 10, 4, 3, 7, 11, 5, 1, 0, 2, 6},
};

static const int* clip_table(int bd) {
    static const int t8[4] = {256, 32, 8, 2}, t10[4] = {1024, 128, 32, 8}, t12[4] = {4096, 512, 128, 32};
    return bd >= 12 ? t12 : bd >= 10 ? t10 : t8;
}

/* Clamped read of the visible region lw x lh. */
static int get_c(const rmh_plane* p, int x, int y, int lw, int lh) {
    return px(p, clampi(x, 0, lw - 1), clampi(y, 0, lh - 1));
}

/* derive_alf_class_4x4_vb_samples (filter.rs:1580-1691) over a plane read
 * through get_c with the visible extent bw x bh. */
int orc_alf_classify(const rmh_plane* plane, int bx, int by, int bw, int bh, int vb_h, int vb_pos, int* transpose) {
    static const int TH[16] = {0, 1, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 4};
    static const int TRANSPOSE_TABLE[8] = {0, 1, 0, 2, 2, 3, 1, 3};
    const int lw = mini(bw, (int)plane->width), lh = mini(bh, (int)plane->height);
    const int bd = clampi(plane->bit_depth, 8, 12);
#define P(xx, yy) get_c(plane, (xx), (yy), lw, lh)
    int sum_v = 0, sum_h = 0, sum_d0 = 0, sum_d1 = 0;
    int cls_row_vb = 0;
    if (vb_h > 0 && by % vb_h == vb_pos - 4) /* vb_pos.wrapping_sub(4): vb_pos >= 4 here */
        cls_row_vb = 1;
    else if (vb_h > 0 && by % vb_h == vb_pos)
        cls_row_vb = 2;
    for (int dy = -2; dy < 6; dy += 2) {
        if ((cls_row_vb == 1 && dy == 4) || (cls_row_vb == 2 && dy == -2)) continue;
        for (int dx = -2; dx < 6; dx += 2) {
            const int x = bx + dx, y = by + dy;
            int down = y - 1, up = y + 1, up2 = y + 2;
            if (vb_h > 0 && y > 0) {
                const int yy = y % vb_h;
                if (vb_pos >= 2 && yy == vb_pos - 2) {
                    up2 = y + 1;
                } else if (yy == vb_pos) {
                    down = y;
                }
            }
            const int y0 = P(x, y) << 1;
            const int yup1 = P(x + 1, y + 1) << 1;
            sum_v += absi(y0 - P(x, down) - P(x, up)) + absi(yup1 - P(x + 1, y) - P(x + 1, up2));
            sum_h += absi(y0 - P(x + 1, y) - P(x - 1, y)) + absi(yup1 - P(x + 2, y + 1) - P(x, y + 1));
            sum_d0 += absi(y0 - P(x - 1, down) - P(x + 1, up)) + absi(yup1 - P(x, y) - P(x + 2, up2));
            sum_d1 += absi(y0 - P(x - 1, up) - P(x + 1, down)) + absi(yup1 - P(x, up2) - P(x + 2, y));
        }
    }
#undef P
    const int temp_act = sum_v + sum_h;
    int activity_scale = 64;
    if (vb_h > 0 && vb_pos >= 4) {
        const int yy = by % vb_h;
        if (yy == vb_pos - 4 || yy == vb_pos) activity_scale = 96;
    }
    const int activity = clampi((temp_act * activity_scale) >> (bd + 4), 0, 15);
    int class_idx = TH[activity];
    int hv1, hv0, dir_hv, d1, d0, dir_d;
    if (sum_v > sum_h) {
        hv1 = sum_v; hv0 = sum_h; dir_hv = 1;
    } else {
        hv1 = sum_h; hv0 = sum_v; dir_hv = 3;
    }
    if (sum_d0 > sum_d1) {
        d1 = sum_d0; d0 = sum_d1; dir_d = 0;
    } else {
        d1 = sum_d1; d0 = sum_d0; dir_d = 2;
    }
    int hvd1, hvd0, main_dir, sec_dir;
    if ((int64_t)d1 * hv0 > (int64_t)hv1 * d0) {
        hvd1 = d1; hvd0 = d0; main_dir = dir_d; sec_dir = dir_hv;
    } else {
        hvd1 = hv1; hvd0 = hv0; main_dir = dir_hv; sec_dir = dir_d;
    }
    int strength = 0;
    if (hvd1 > 2 * hvd0) strength = 1;
    if (hvd1 * 2 > 9 * hvd0) strength = 2;
    if (strength > 0) class_idx += (((main_dir & 1) << 1) + strength) * 5;
    *transpose = TRANSPOSE_TABLE[main_dir * 2 + (sec_dir >> 1)];
    return class_idx;
}

/* The coefficients and clip values alf_luma_block applies to a block of
 * class class_idx and transposition `transpose` (decode_general.rs:5454-5483):
 * mode 1..16 fixed set mode - 1, 17 the APS. */
void orc_alf_luma_filter(int mode, int class_idx, int transpose, const rmh_alf_aps* aps, int bit_depth, int32_t* fc,
                         int32_t* fclip) {
    const int* ct = clip_table(clampi(bit_depth, 8, 12));
    int coeffs[12], clips[12] = {0};
    if (mode != RMH_ALF_LUMA_APS) {
        const int fi = ORC_ALF_CLASS_MAP[mini(mode - 1, 15)][mini(class_idx, 24)];
        for (int k = 0; k < 12; k++) coeffs[k] = ORC_ALF_FIXED[fi][k];
    } else {
        const int fi = mini(aps->class_to_filter[mini(class_idx, 24)], maxi(aps->n_luma - 1, 0));
        for (int k = 0; k < 12; k++) {
            coeffs[k] = aps->luma_coeff[fi][k];
            clips[k] = aps->luma_clip[fi][k];
        }
    }
    const int* perm = ALF_TRANSPOSE_PERM[transpose & 3];
    for (int k = 0; k < 12; k++) {
        fc[k] = coeffs[perm[k]];
        fclip[k] = ct[clips[perm[k]] & 3];
    }
}

/* alf_luma_block (decode_general.rs:5411-5566). mode: 1..16 fixed set, 17 APS. */
static void alf_luma_block(const rmh_plane* src, const rmh_plane* dst, int x0, int y0, int w, int h, int vis_w,
                           int vis_h, int vb_h, int vb_pos, int mode, const rmh_alf_aps* aps) {
    /* tap pairs (column offset, row slot) with row slots 0 = current, 1/+1, 2/-1,
     * 3/+2, 4/-2, 5/+3, 6/-3 (the TERMS table of alf_luma_block) */
    static const int TERMS[12][4] = {{0, 5, 0, 6},  {1, 3, -1, 4}, {0, 3, 0, 4},  {-1, 3, 1, 4},
                                     {2, 1, -2, 2}, {1, 1, -1, 2}, {0, 1, 0, 2},  {-1, 1, 1, 2},
                                     {-2, 1, 2, 2}, {3, 0, -3, 0}, {2, 0, -2, 0}, {1, 0, -1, 0}};
    const int bd = clampi(src->bit_depth, 8, 12);
    const int max_val = (1 << dst->bit_depth) - 1;
    const int bw = mini(vis_w, (int)src->width), bh = mini(vis_h, (int)src->height);
    for (int by = y0; by < y0 + h; by += 4) {
        for (int bx = x0; bx < x0 + w; bx += 4) {
            int transpose = 0;
            const int class_idx = orc_alf_classify(src, bx, by, bw, bh, vb_h, vb_pos, &transpose);
            int32_t fc[12], fclip[12];
            orc_alf_luma_filter(mode, class_idx, transpose, aps, bd, fc, fclip);
            const int xe = mini(bx + 4, x0 + w), ye = mini(by + 4, y0 + h);
            for (int y = by; y < ye; y++) {
                for (int x = bx; x < xe; x++) {
                    const int cur = get_c(src, x, y, bw, bh);
                    const int yv = vb_h > 0 ? y % vb_h : -1;
                    int rows[7] = {y, y + 1, y - 1, y + 2, y - 2, y + 3, y - 3};
                    if (vb_h > 0 && yv < vb_pos && yv + 4 >= vb_pos) {
                        if (yv == vb_pos - 1) {
                            rows[1] = rows[0];
                            rows[2] = rows[0];
                        }
                        if (yv + 2 >= vb_pos) {
                            rows[3] = rows[1];
                            rows[4] = rows[2];
                        }
                        if (yv + 3 >= vb_pos) {
                            rows[5] = rows[3];
                            rows[6] = rows[4];
                        }
                    } else if (vb_h > 0 && yv >= vb_pos && yv <= vb_pos + 3) {
                        if (yv == vb_pos) {
                            rows[2] = rows[0];
                            rows[1] = rows[0];
                        }
                        if (yv <= vb_pos + 1) {
                            rows[4] = rows[2];
                            rows[3] = rows[1];
                        }
                        if (yv <= vb_pos + 2) {
                            rows[6] = rows[4];
                            rows[5] = rows[3];
                        }
                    }
                    const int near = vb_h > 0 && ((yv < vb_pos && yv + 1 >= vb_pos) || yv == vb_pos);
                    int sum = 0;
                    for (int k = 0; k < 12; k++) {
                        const int a = get_c(src, x + TERMS[k][0], rows[TERMS[k][1]], bw, bh);
                        const int b = get_c(src, x + TERMS[k][2], rows[TERMS[k][3]], bw, bh);
                        sum += fc[k] * (clampi(a - cur, -fclip[k], fclip[k]) + clampi(b - cur, -fclip[k], fclip[k]));
                    }
                    const int shift = near ? 10 : 7;
                    put(dst, x, y, clampi(cur + ((sum + (1 << (shift - 1))) >> shift), 0, max_val));
                }
            }
        }
    }
}

/* alf_chroma_block (decode_general.rs:5568-5660). */
static void alf_chroma_block(const rmh_plane* src, const rmh_plane* dst, const rmh_alf_aps* aps, int alt, int x0,
                             int y0, int w, int h, int vis_w, int vis_h, int vb_h, int vb_pos) {
    static const int TERMS[6][4] = {{0, 3, 0, 4}, {1, 1, -1, 2}, {0, 1, 0, 2},
                                    {-1, 1, 1, 2}, {2, 0, -2, 0}, {1, 0, -1, 0}};
    if (aps->n_chroma == 0) return;
    const int idx = mini(alt, aps->n_chroma - 1);
    const int bd = clampi(src->bit_depth, 8, 12);
    const int max_val = (1 << dst->bit_depth) - 1;
    const int* table = clip_table(bd);
    const int bw = mini(vis_w, (int)src->width), bh = mini(vis_h, (int)src->height);
    for (int y = y0; y < mini(y0 + h, (int)dst->height); y++) {
        for (int x = x0; x < mini(x0 + w, (int)dst->width); x++) {
            const int cur = get_c(src, x, y, bw, bh);
            const int yv = vb_h > 0 ? y % vb_h : -1;
            int rows[5] = {y, y + 1, y - 1, y + 2, y - 2};
            if (vb_h > 0 && yv < vb_pos && yv + 2 >= vb_pos) {
                if (yv + 1 == vb_pos) {
                    rows[1] = rows[0];
                    rows[2] = rows[0];
                }
                if (yv + 2 >= vb_pos) {
                    rows[3] = rows[1];
                    rows[4] = rows[2];
                }
            } else if (vb_h > 0 && yv >= vb_pos && yv <= vb_pos + 1) {
                if (yv == vb_pos) {
                    rows[2] = rows[0];
                    rows[1] = rows[0];
                }
                if (yv <= vb_pos + 1) {
                    rows[4] = rows[2];
                    rows[3] = rows[1];
                }
            }
            const int near = vb_h > 0 && ((yv < vb_pos && yv + 1 >= vb_pos) || yv == vb_pos);
            int sum = 0;
            for (int k = 0; k < 6; k++) {
                const int a = get_c(src, x + TERMS[k][0], rows[TERMS[k][1]], bw, bh);
                const int b = get_c(src, x + TERMS[k][2], rows[TERMS[k][3]], bw, bh);
                const int clk = table[aps->chroma_clip[idx][k] & 3];
                sum += aps->chroma_coeff[idx][k] * (clampi(a - cur, -clk, clk) + clampi(b - cur, -clk, clk));
            }
            const int shift = near ? 10 : 7;
            put(dst, x, y, clampi(cur + ((sum + (1 << (shift - 1))) >> shift), 0, max_val));
        }
    }
}

/* ccalf_block (decode_general.rs:5335-5409): adds the CCALF term from the
 * pre-ALF luma to the chroma plane in place. */
static void ccalf_block(const rmh_plane* src_y, const rmh_plane* chroma, const int16_t* coeffs, int x0, int y0, int w,
                        int h, int cw, int ch, int vis_lw, int vis_lh, int ctu_l) {
    const int bd = clampi(src_y->bit_depth, 8, 12);
    const int max_val = (1 << chroma->bit_depth) - 1;
    const int offset = 1 << (bd - 1);
    const int lw = mini(vis_lw, (int)src_y->width), lh = mini(vis_lh, (int)src_y->height);
    const int vb_pos = maxi(ctu_l - 4, 0);
#define Y(xx, yy) get_c(src_y, (xx), (yy), lw, lh)
    for (int y = y0; y < mini(y0 + h, ch); y++) {
        const int ly = y * 2;
        const int pos = ly & (ctu_l - 1);
        int above, below, below2;
        if (pos == maxi(vb_pos - 2, 0) || pos == vb_pos + 1) {
            above = ly - 1; below = ly + 1; below2 = ly + 1;
        } else if (pos == maxi(vb_pos - 1, 0) || pos == vb_pos) {
            above = ly; below = ly; below2 = ly;
        } else {
            above = ly - 1; below = ly + 1; below2 = ly + 2;
        }
        for (int x = x0; x < mini(x0 + w, cw); x++) {
            const int lx = x * 2;
            const int cur = Y(lx, ly);
            int sum = coeffs[0] * (Y(lx, above) - cur);
            sum += coeffs[1] * (Y(lx - 1, ly) - cur);
            sum += coeffs[2] * (Y(lx + 1, ly) - cur);
            sum += coeffs[3] * (Y(lx - 1, below) - cur);
            sum += coeffs[4] * (Y(lx, below) - cur);
            sum += coeffs[5] * (Y(lx + 1, below) - cur);
            sum += coeffs[6] * (Y(lx, below2) - cur);
            const int v = clampi((sum + 64) >> 7, -offset, offset - 1);
            put(chroma, x, y, clampi(px(chroma, x, y) + v, 0, max_val));
        }
    }
#undef Y
}

static void copy_plane(const rmh_plane* src, const rmh_plane* dst) {
    for (uint32_t y = 0; y < src->height; y++)
        for (uint32_t x = 0; x < src->width; x++) put(dst, (int)x, (int)y, px(src, (int)x, (int)y));
}

/* apply_alf (decode_general.rs:5225-5333): dst starts as a copy of src (the
 * decoder filters g.y from its clone), then every CTB in raster order. */
int orc_alf_frame(const rmh_alf_frame* f) {
    for (int c = 0; c < 3; c++) copy_plane(&f->src[c], &f->dst[c]);
    const int ctu = (int)f->ctb_size, vis_w = (int)f->vis_w, vis_h = (int)f->vis_h;
    const int cvis_w = (vis_w + 1) / 2, cvis_h = (vis_h + 1) / 2;
    const int rows = (vis_h + ctu - 1) / ctu, cols = (vis_w + ctu - 1) / ctu;
#pragma omp parallel for schedule(dynamic)
    for (int gy = 0; gy < rows; gy++) {
        for (int gx = 0; gx < cols; gx++) {
            const rmh_alf_ctb* s = &f->ctbs[(size_t)gy * f->ctb_cols + gx];
            const int x0 = gx * ctu, y0 = gy * ctu;
            const int w = mini(ctu, vis_w - x0), h = mini(ctu, vis_h - y0);
            if (s->luma >= 1 && s->luma <= RMH_ALF_LUMA_APS && (s->luma != RMH_ALF_LUMA_APS || f->luma_aps))
                alf_luma_block(&f->src[0], &f->dst[0], x0, y0, w, h, vis_w, vis_h, ctu, maxi(ctu - 4, 0), s->luma,
                               f->luma_aps);
            for (int comp = 1; comp < 3; comp++) {
                const int alt = comp == 1 ? s->cb : s->cr;
                if (alt && f->chroma_aps)
                    alf_chroma_block(&f->src[comp], &f->dst[comp], f->chroma_aps, alt - 1, x0 / 2, y0 / 2,
                                     maxi(w / 2, 1), maxi(h / 2, 1), cvis_w, cvis_h, ctu / 2, maxi(ctu / 2 - 2, 0));
            }
            for (int comp = 1; comp < 3; comp++) {
                const int idc = comp == 1 ? s->cc_cb : s->cc_cr;
                const rmh_alf_aps* aps = comp == 1 ? f->cc_cb_aps : f->cc_cr_aps;
                if (idc == 0 || !aps) continue;
                const int n = comp == 1 ? aps->n_cc_cb : aps->n_cc_cr;
                if (idc > n) continue;
                const int16_t* coeffs = comp == 1 ? aps->cc_cb[idc - 1] : aps->cc_cr[idc - 1];
                ccalf_block(&f->src[0], &f->dst[comp], coeffs, x0 / 2, y0 / 2, maxi(w / 2, 1), maxi(h / 2, 1),
                            cvis_w, cvis_h, vis_w, vis_h, ctu);
            }
        }
    }
    return 0;
}

/*
 * orc_sao.c - CPU oracle for sample adaptive offset (SAO).
 *
 * TEST INFRASTRUCTURE ONLY (see vvc_oracle.h). Scalar restatement of
 *   apply_sao_region_bounded     crates/rustmpeg-video/src/vvc/filter.rs:1276-1345 (8-bit)
 *   apply_sao_region_bounded_16  crates/rustmpeg-video/src/vvc/filter.rs:1954-2043
 *   apply_sao                    crates/rustmpeg-video/src/vvc/decode_general.rs:5189-5222
 */
#include <string.h>

#include "vvc_oracle.h"

static int px(const rmh_plane* p, int x, int y) {
    const uint8_t* row = (const uint8_t*)p->data + (size_t)y * p->stride * p->bytes_per_sample;
    return p->bytes_per_sample == 1 ? row[x] : ((const uint16_t*)row)[x];
}

static void put(rmh_plane* p, int x, int y, int v) {
    uint8_t* row = (uint8_t*)p->data + (size_t)y * p->stride * p->bytes_per_sample;
    if (p->bytes_per_sample == 1)
        row[x] = (uint8_t)v;
    else
        ((uint16_t*)row)[x] = (uint16_t)v;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
static int sgn(int v) { return (v > 0) - (v < 0); }
static int mini(int a, int b) { return a < b ? a : b; }

/* apply_sao_region_bounded_16 (filter.rs:1954-2043). src and dst have the
 * same size; only samples of the region that the SAO type changes are
 * written. */
int orc_sao_region_bounded16(const rmh_plane* src, rmh_plane* dst, int bx, int by, int w, int h, int bw, int bh,
                             const rmh_sao_param* sao) {
    const int bd = clampi(src->bit_depth, 8, 12);
    const int max_val = (1 << src->bit_depth) - 1;
    const int band_shift = bd - 5;
    const int w_lim = mini(bx + w, (int)src->width), h_lim = mini(by + h, (int)src->height);
    if (sao->type == RMH_SAO_OFF) return 0;
    if (sao->type == RMH_SAO_BAND) {
        for (int y = by; y < h_lim; y++)
            for (int x = bx; x < w_lim; x++) {
                const int s = px(src, x, y);
                const int band = (int)((uint8_t)((uint8_t)(s >> band_shift) - sao->band_or_class) & 31);
                if (band < 4) put(dst, x, y, clampi(s + sao->offsets[band], 0, max_val));
            }
        return 0;
    }
    if (sao->type != RMH_SAO_EDGE) return -1;
    int dx, dy;
    switch (sao->band_or_class & 3) {
        case 0: dx = 1; dy = 0; break;
        case 1: dx = 0; dy = 1; break;
        case 2: dx = 1; dy = 1; break;
        default: dx = -1; dy = 1; break;
    }
    const int pw = mini(bw, (int)src->width), ph = mini(bh, (int)src->height);
    for (int y = by; y < h_lim; y++)
        for (int x = bx; x < w_lim; x++) {
            const int ax = x - dx, ay = y - dy, nx = x + dx, ny = y + dy;
            if (ax < 0 || ay < 0 || ax >= pw || ay >= ph || nx < 0 || ny < 0 || nx >= pw || ny >= ph) continue;
            const int c = px(src, x, y), a = px(src, ax, ay), b = px(src, nx, ny);
            const int edge_idx = 2 + sgn(c - a) + sgn(c - b);
            const int off = edge_idx == 0   ? sao->offsets[0]
                            : edge_idx == 1 ? sao->offsets[1]
                            : edge_idx == 3 ? sao->offsets[2]
                            : edge_idx == 4 ? sao->offsets[3]
                                            : 0;
            put(dst, x, y, clampi(c + off, 0, max_val));
        }
    return 0;
}

/* apply_sao_region_bounded (filter.rs:1276-1345), the 8-bit form: a 32-entry
 * band table and a write of every region sample for band offset, the 45
 * degree class as (dx, dy) = (-1, 1), results clamped to 0..255. */
int orc_sao_region_bounded8(const rmh_plane* src, rmh_plane* dst, int bx, int by, int w, int h, int bw, int bh,
                            const rmh_sao_param* sao) {
    if (src->bytes_per_sample != 1 || dst->bytes_per_sample != 1) return -1;
    const int w_lim = mini(bx + w, (int)src->width), h_lim = mini(by + h, (int)src->height);
    if (sao->type == RMH_SAO_OFF) return 0;
    if (sao->type == RMH_SAO_BAND) {
        int table[32] = {0};
        for (int i = 0; i < 4; i++) table[(sao->band_or_class + i) & 31] = sao->offsets[i];
        for (int y = by; y < h_lim; y++)
            for (int x = bx; x < w_lim; x++) {
                const int s = px(src, x, y);
                put(dst, x, y, clampi(s + table[(s >> 3) & 31], 0, 255));
            }
        return 0;
    }
    if (sao->type != RMH_SAO_EDGE) return -1;
    const int lut[5] = {sao->offsets[0], sao->offsets[1], 0, sao->offsets[2], sao->offsets[3]};
    int dx, dy;
    switch (sao->band_or_class & 3) {
        case 0: dx = 1; dy = 0; break;
        case 1: dx = 0; dy = 1; break;
        case 2: dx = 1; dy = 1; break;
        default: dx = -1; dy = 1; break;
    }
    const int pw = mini(bw, (int)src->width), ph = mini(bh, (int)src->height);
    for (int y = by; y < h_lim; y++)
        for (int x = bx; x < w_lim; x++) {
            const int ax = x - dx, ay = y - dy, nx = x + dx, ny = y + dy;
            if (ax < 0 || ay < 0 || ax >= pw || ay >= ph || nx < 0 || ny < 0 || nx >= pw || ny >= ph) continue;
            const int c = px(src, x, y);
            const int e = 2 + sgn(c - px(src, ax, ay)) + sgn(c - px(src, nx, ny));
            put(dst, x, y, clampi(c + lut[e], 0, 255));
        }
    return 0;
}

/* apply_sao (decode_general.rs:5189-5222) for one picture: dst starts as a
 * copy of src (the decoder filters its own picture, reading a clone), then
 * every CTB of the visible region runs apply_sao_region_bounded_16 per
 * component. */
int orc_sao_frame(const rmh_sao_frame* f) {
    for (int c = 0; c < 3; c++) {
        const rmh_plane* s = &f->src[c];
        rmh_plane* d = (rmh_plane*)&f->dst[c];
        for (uint32_t y = 0; y < s->height; y++)
            memcpy((uint8_t*)d->data + (size_t)y * d->stride * d->bytes_per_sample,
                   (const uint8_t*)s->data + (size_t)y * s->stride * s->bytes_per_sample,
                   (size_t)s->width * s->bytes_per_sample);
    }
    const int ctu = (int)f->ctb_size;
    const int vis_w = (int)f->vis_w, vis_h = (int)f->vis_h;
    const int cvis_w = (vis_w + 1) / 2, cvis_h = (vis_h + 1) / 2;
    for (int gy = 0; gy < (vis_h + ctu - 1) / ctu; gy++)
        for (int gx = 0; gx < (vis_w + ctu - 1) / ctu; gx++) {
            const rmh_sao_param* sao = f->params + ((size_t)gy * f->ctb_cols + gx) * 3;
            const int x0 = gx * ctu, y0 = gy * ctu;
            const int w = mini(ctu, vis_w - x0), h = mini(ctu, vis_h - y0);
            if (orc_sao_region_bounded16(&f->src[0], (rmh_plane*)&f->dst[0], x0, y0, w, h, vis_w, vis_h, &sao[0]))
                return -1;
            const int cx0 = x0 / 2, cy0 = y0 / 2;
            const int cw = mini(ctu / 2, cvis_w - cx0 > 0 ? cvis_w - cx0 : 0);
            const int ch = mini(ctu / 2, cvis_h - cy0 > 0 ? cvis_h - cy0 : 0);
            if (cw > 0 && ch > 0)
                for (int c = 1; c < 3; c++)
                    if (orc_sao_region_bounded16(&f->src[c], (rmh_plane*)&f->dst[c], cx0, cy0, cw, ch, cvis_w, cvis_h,
                                                 &sao[c]))
                        return -1;
        }
    return 0;
}

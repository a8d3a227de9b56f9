/* orc_cabac.c - oracle (test infrastructure only): the CABAC DECODING side,
 * used to read back what the library's encoder writes.
 *
 * The library codes with the reference's encoder side (BinEncoder,
 * encode_residual, encode_coding_tree_split); this file restates the
 * reference's decoder side, so an encode -> decode round trip compares two
 * independent halves of the reference:
 *   orc_ctx_*             CtxModel::init / update / get_lps   vvc/cabac.rs:41-99
 *   BinDecoder            BinDecoder                          vvc/cabac.rs:761-1007
 *   decode_last           decode_last_sig_coeff               vvc/residual.rs:359-433
 *   orc_decode_residual   decode_residual_with_options        vvc/residual.rs:537-607
 *   decode_subblock       decode_subblock                     vvc/residual.rs:745-887
 *   can_split_tree        can_split_coding_tree               vvc/transform.rs:537-622
 *   split_ctx             derive_split_contexts               vvc/transform.rs:624-715
 *   decode_split          decode_coding_tree_split            vvc/slice.rs:3895-3946
 *   decode_luma_mode      decode_full_luma_intra_mode's MPM branch (vvc/slice.rs:1690-1720)
 * orc_decode_luma_stream reads the luma-only substream of
 * rmh_rd_encode_luma_stream and reconstructs it with the dequantisation and
 * inverse transform restated in orc_residual.c and the RD search's PLANAR/DC
 * predictor (restated below from its description: nearest-available
 * substitution, no filtering, no PDPC; not a reference function).
 *
 * The coefficient scan is orc_coeff_scan (orc_tables.c). */
#include <stdlib.h>
#include <string.h>

#include "cabac_data.h"
#include "vvc_oracle.h"

/* ---- context models ---- */

typedef struct {
    uint16_t s0, s1;
    uint8_t r0, r1;
} orc_model;

static const uint8_t RENORM[32] = {6, 5, 4, 4, 3, 3, 3, 3, 2, 2, 2, 2, 2, 2, 2, 2,
                                   1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};

static orc_model orc_ctx_init(int qp, int init_id, int ws) {
    int slope = (init_id >> 3) - 4;
    int offset = (init_id & 7) * 18 + 1;
    int st = ((slope * (qp - 16)) >> 1) + offset;
    if (st < 1) st = 1;
    if (st > 127) st = 127;
    uint32_t p1 = (uint32_t)st << 8;
    orc_model m;
    m.r0 = (uint8_t)(2 + ((ws >> 2) & 3));
    m.r1 = (uint8_t)(3 + m.r0 + (ws & 3));
    m.s0 = (uint16_t)(p1 & 0x7FE0);
    m.s1 = (uint16_t)(p1 & 0x7FFE);
    return m;
}

static int orc_ctx_state(const orc_model* m) { return (m->s0 + m->s1) >> 8; }

static uint32_t orc_ctx_lps(const orc_model* m, uint32_t range) {
    int q = orc_ctx_state(m);
    if (q & 0x80) q ^= 0xFF;
    return ((((uint32_t)q >> 2) * (range >> 5)) >> 1) + 4;
}

static void orc_ctx_update(orc_model* m, int bin) {
    uint32_t s0 = m->s0, s1 = m->s1;
    uint32_t n0 = s0 - ((s0 >> m->r0) & 0x7FE0);
    uint32_t n1 = s1 - ((s1 >> m->r1) & 0x7FFE);
    if (bin) {
        n0 += (0x7FFFu >> m->r0) & 0x7FE0;
        n1 += (0x7FFFu >> m->r1) & 0x7FFE;
    }
    m->s0 = (uint16_t)(n0 & 0xFFFF);
    m->s1 = (uint16_t)(n1 & 0xFFFF);
}

static void orc_ctx_store(orc_model* models, int slice, int qp) {
    if (qp < 0) qp = 0;
    if (qp > 63) qp = 63;
    for (int i = 0; i < ORC_NUM_CONTEXTS; i++)
        models[i] = orc_ctx_init(qp, ORC_CTX_INIT[slice][i][0], ORC_CTX_INIT[slice][i][1]);
}

/* ---- BinDecoder ---- */

typedef struct {
    const uint8_t* data;
    size_t n, pos;
    uint32_t range, value;
    int bits_needed;
    uint8_t last_byte;
} BinDecoder;

static void dec_init(BinDecoder* d, const uint8_t* data, size_t n) {
    uint32_t first = n > 0 ? data[0] : 0, second = n > 1 ? data[1] : 0;
    d->data = data;
    d->n = n;
    d->pos = n < 2 ? n : 2;
    d->range = 510;
    d->value = (first << 8) + second;
    d->bits_needed = -8;
    d->last_byte = (uint8_t)second;
}

static uint32_t dec_read_byte(BinDecoder* d) {
    if (d->pos < d->n) {
        uint8_t b = d->data[d->pos++];
        d->last_byte = b;
        return b;
    }
    return 0;
}

static int dec_bin(BinDecoder* d, orc_model* m) {
    int mps = orc_ctx_state(m) >= 128;
    uint32_t lps = orc_ctx_lps(m, d->range);
    d->range -= lps;
    uint32_t scaled = d->range << 7;
    int bin;
    if (d->value < scaled) {
        bin = mps;
        if (d->range < 256) {
            d->range <<= 1;
            d->value <<= 1;
            if (++d->bits_needed >= 0) {
                d->value += dec_read_byte(d) << d->bits_needed;
                d->bits_needed -= 8;
            }
        }
    } else {
        bin = !mps;
        int nb = RENORM[lps >> 3];
        d->value -= scaled;
        d->value <<= nb;
        d->range = lps << nb;
        d->bits_needed += nb;
        if (d->bits_needed >= 0) {
            d->value += dec_read_byte(d) << d->bits_needed;
            d->bits_needed -= 8;
        }
    }
    orc_ctx_update(m, bin);
    return bin;
}

static int dec_bin_ep(BinDecoder* d) {
    d->value <<= 1;
    if (++d->bits_needed >= 0) {
        d->value += dec_read_byte(d);
        d->bits_needed = -8;
    }
    uint32_t scaled = d->range << 7;
    if (d->value >= scaled) {
        d->value -= scaled;
        return 1;
    }
    return 0;
}

static uint32_t dec_bins_ep(BinDecoder* d, uint32_t n) {
    uint32_t v = 0;
    for (uint32_t i = 0; i < n; i++) v = (v << 1) | (uint32_t)dec_bin_ep(d);
    return v;
}

static uint32_t dec_unary_max_ep(BinDecoder* d, uint32_t max_symbol) {
    uint32_t s = 0;
    while (s < max_symbol && dec_bin_ep(d)) s++;
    return s;
}

static int dec_bin_trm(BinDecoder* d) {
    d->range -= 2;
    uint32_t scaled = d->range << 7;
    if (d->value >= scaled) return 1;
    if (d->range < 256) {
        d->range <<= 1;
        d->value <<= 1;
        if (++d->bits_needed == 0) {
            d->value += dec_read_byte(d);
            d->bits_needed = -8;
        }
    }
    return 0;
}

/* 0 when the stream ends in the stop bit and alignment zeros. */
static int dec_finish(const BinDecoder* d) {
    int shift = 8 + d->bits_needed;
    if (shift >= 0 && (((uint32_t)d->last_byte << shift) & 0xFF) != 0x80) return -1;
    return 0;
}

static uint32_t dec_rem_abs_ep(BinDecoder* d, uint32_t k, uint32_t cutoff, uint32_t max_log2) {
    uint32_t max_prefix = 32 - max_log2;
    uint32_t prefix = 0, cw;
    do {
        prefix++;
        cw = (uint32_t)dec_bin_ep(d);
    } while (cw != 0 && prefix < max_prefix);
    prefix -= 1 - cw;
    uint32_t length = k, offset;
    if (prefix < cutoff) {
        offset = prefix << k;
    } else {
        offset = ((1u << (prefix - cutoff)) + cutoff - 1) << k;
        length += prefix == max_prefix ? max_log2 - k : prefix - cutoff;
    }
    return offset + dec_bins_ep(d, length);
}

/* ---- residual_coding (decoder side) ---- */

enum {
    C_SPLIT_FLAG = 0, C_SPLIT_QT_FLAG = 9, C_SPLIT_HV_FLAG = 15, C_SPLIT_12_FLAG = 20,
    C_INTRA_LUMA_MPM_FLAG = 40, C_INTRA_LUMA_PLANAR_FLAG = 41, C_QT_CBF_Y = 77,
    C_SIG_COEFF_GROUP_LUMA = 86, C_SIG_COEFF_GROUP_CHROMA = 88,
    C_SIG_FLAG_LUMA = 90, C_SIG_FLAG_CHROMA = 102, C_SIG_FLAG_LUMA_1 = 110, C_SIG_FLAG_CHROMA_1 = 122,
    C_SIG_FLAG_LUMA_2 = 130, C_SIG_FLAG_CHROMA_2 = 142, C_PAR_FLAG_LUMA = 150, C_PAR_FLAG_CHROMA = 171,
    C_GTX2_LUMA = 182, C_GTX2_CHROMA = 203, C_GTX1_LUMA = 214, C_GTX1_CHROMA = 235,
    C_LAST_X_LUMA = 246, C_LAST_X_CHROMA = 266, C_LAST_Y_LUMA = 269, C_LAST_Y_CHROMA = 289
};

/* The context-set offsets above, for the CPU test that compares them with
 * the reference's ctx module (vvc/cabac.rs:307-398). */
int orc_cabac_ctx_offset(int i) {
    static const int v[] = {C_SPLIT_FLAG, C_SPLIT_QT_FLAG, C_SPLIT_HV_FLAG, C_SPLIT_12_FLAG,
                            C_INTRA_LUMA_MPM_FLAG, C_INTRA_LUMA_PLANAR_FLAG, C_QT_CBF_Y,
                            C_SIG_COEFF_GROUP_LUMA, C_SIG_COEFF_GROUP_CHROMA, C_SIG_FLAG_LUMA,
                            C_SIG_FLAG_CHROMA, C_SIG_FLAG_LUMA_1, C_SIG_FLAG_CHROMA_1, C_SIG_FLAG_LUMA_2,
                            C_SIG_FLAG_CHROMA_2, C_PAR_FLAG_LUMA, C_PAR_FLAG_CHROMA, C_GTX2_LUMA,
                            C_GTX2_CHROMA, C_GTX1_LUMA, C_GTX1_CHROMA, C_LAST_X_LUMA, C_LAST_X_CHROMA,
                            C_LAST_Y_LUMA, C_LAST_Y_CHROMA};
    return i >= 0 && i < (int)(sizeof v / sizeof v[0]) ? v[i] : -1;
}

static const int PREFIX_CTX[8] = {0, 0, 0, 3, 6, 10, 15, 21};
static const int MIN_IN_GROUP[14] = {0, 1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96};
static const int GROUP_IDX[64] = {0, 1, 2, 3, 4, 4, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7, 8, 8, 8, 8, 8, 8,
                                  8, 8, 9, 9, 9, 9, 9, 9, 9, 9, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10,
                                  10, 10, 10, 10, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11};
static const int GO_RICE_PARS[32] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 2, 2,
                                     2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3};

typedef struct {
    int w, h, log2_cg, wg, hg, n;
    int32_t raster[1024];
    int cg_raster[1024];
} Scan;

static void make_scan(Scan* s, int w, int h) {
    int a, b;
    orc_cg_log2(w, h, &a, &b);
    s->w = w;
    s->h = h;
    s->log2_cg = a + b;
    s->wg = (w < 32 ? w : 32) >> a;
    s->hg = (h < 32 ? h : 32) >> b;
    s->n = orc_coeff_scan(w, h, s->raster);
    for (int g = 0; g < (s->n >> s->log2_cg); g++) {
        int r = s->raster[g << s->log2_cg];
        s->cg_raster[g] = ((r / w) >> b) * s->wg + ((r % w) >> a);
    }
}

typedef struct {
    int diag, sum1;
} Tmpl;

static int sig_ctx_id_abs(const Scan* s, const int32_t* c, int pos, int luma, int state, Tmpl* t) {
    int r = s->raster[pos], x = r % s->w, y = r / s->w, w = s->w, h = s->h, base = r;
    int sum_abs = 0, num_pos = 0;
#define UPD(i)                                  \
    do {                                        \
        int a_ = abs(c[(i)]);                   \
        int m_ = 4 + (a_ & 1);                  \
        sum_abs += a_ < m_ ? a_ : m_;           \
        num_pos += a_ != 0;                     \
    } while (0)
    if (x + 1 < w) {
        UPD(base + 1);
        if (x + 2 < w) UPD(base + 2);
        if (y + 1 < h) UPD(base + w + 1);
    }
    if (y + 1 < h) {
        UPD(base + w);
        if (y + 2 < h) UPD(base + 2 * w);
    }
#undef UPD
    int diag = x + y;
    int ofs = (sum_abs + 1) >> 1;
    if (ofs > 3) ofs = 3;
    ofs += (diag < 2) * 4;
    if (luma) ofs += (diag < 5) * 4;
    t->diag = diag;
    t->sum1 = sum_abs - num_pos;
    int set = state - 1 < 0 ? 0 : state - 1;
    if (set > 2) set = 2;
    static const int L[3] = {C_SIG_FLAG_LUMA, C_SIG_FLAG_LUMA_1, C_SIG_FLAG_LUMA_2};
    static const int C[3] = {C_SIG_FLAG_CHROMA, C_SIG_FLAG_CHROMA_1, C_SIG_FLAG_CHROMA_2};
    return (luma ? L[set] : C[set]) + ofs;
}

static int ctx_offset_abs(const Tmpl* t, int luma) {
    if (t->diag < 0) return 0;
    int o = (t->sum1 < 4 ? t->sum1 : 4) + 1;
    if (t->diag == 0)
        o += luma ? 15 : 5;
    else if (luma)
        o += t->diag < 3 ? 10 : t->diag < 10 ? 5 : 0;
    return o;
}

static uint32_t derive_rice(const Scan* s, const int32_t* c, int pos, int base_level) {
    int r = s->raster[pos], x = r % s->w, y = r / s->w, w = s->w, h = s->h;
    int sum = 0;
    if (x + 1 < w) {
        sum += abs(c[r + 1]);
        if (x + 2 < w) sum += abs(c[r + 2]);
        if (y + 1 < h) sum += abs(c[r + w + 1]);
    }
    if (y + 1 < h) {
        sum += abs(c[r + w]);
        if (y + 2 < h) sum += abs(c[r + 2 * w]);
    }
    sum -= 5 * base_level;
    if (sum < 0) sum = 0;
    if (sum > 31) sum = 31;
    return (uint32_t)GO_RICE_PARS[sum];
}

static int next_state(int table, int state, int level) { return (table >> ((state << 2) + ((level & 1) << 1))) & 3; }

static int decode_last(BinDecoder* d, orc_model* m, const Scan* s, int luma) {
    int w = s->w, h = s->h;
    int maxx = GROUP_IDX[(w < 32 ? w : 32) - 1], maxy = GROUP_IDX[(h < 32 ? h : 32) - 1];
    int bx = luma ? C_LAST_X_LUMA : C_LAST_X_CHROMA, by = luma ? C_LAST_Y_LUMA : C_LAST_Y_CHROMA;
    int lw = 0, lh = 0;
    while ((1 << lw) < w) lw++;
    while ((1 << lh) < h) lh++;
    int ox, sx, oy, sy;
    if (luma) {
        ox = PREFIX_CTX[lw];
        sx = (lw + 1) >> 2;
        oy = PREFIX_CTX[lh];
        sy = (lh + 1) >> 2;
    } else {
        ox = oy = 0;
        sx = (w >> 3) < 2 ? (w >> 3) : 2;
        sy = (h >> 3) < 2 ? (h >> 3) : 2;
    }
    int px = 0, py = 0;
    while (px < maxx && dec_bin(d, &m[bx + ox + (px >> sx)])) px++;
    while (py < maxy && dec_bin(d, &m[by + oy + (py >> sy)])) py++;
    if (px > 3) {
        int cnt = ((px - 2) >> 1) - 1, v = 0;
        for (int i = cnt; i >= 0; i--) v += dec_bin_ep(d) << i;
        px = MIN_IN_GROUP[px] + v;
    }
    if (py > 3) {
        int cnt = ((py - 2) >> 1) - 1, v = 0;
        for (int i = cnt; i >= 0; i--) v += dec_bin_ep(d) << i;
        py = MIN_IN_GROUP[py] + v;
    }
    if (px >= w || py >= h) return -1;
    for (int p = 0; p < s->n; p++)
        if (s->raster[p] == py * w + px) return p;
    return -1;
}

static int decode_subblock(BinDecoder* d, orc_model* m, const Scan* s, int32_t* c, int min_sub, int max_sub,
                           int last, int luma, int table, int* state, int sign_hiding, Tmpl* t, int* remaining) {
    int is_last = (last >> s->log2_cg) == (min_sub >> s->log2_cg);
    int first_sig = is_last ? last : max_sub;
    int next = first_sig;
    int sub_set = min_sub >> s->log2_cg;
    int infer = next != last ? (sub_set != 0 ? min_sub : -1) : next;
    int first_nz = next, last_nz = -1, nnz = 0;
    int sig_blk[16];
    int gt1b = luma ? C_GTX1_LUMA : C_GTX1_CHROMA, parb = luma ? C_PAR_FLAG_LUMA : C_PAR_FLAG_CHROMA;
    int gt2b = luma ? C_GTX2_LUMA : C_GTX2_CHROMA;

    while (next >= min_sub && *remaining >= 4) {
        int blk = s->raster[next];
        int sig = nnz == 0 && next == infer;
        if (!sig) {
            sig = dec_bin(d, &m[sig_ctx_id_abs(s, c, next, luma, *state, t)]);
            (*remaining)--;
        } else if (next != last) {
            sig_ctx_id_abs(s, c, next, luma, *state, t);
        }
        if (sig) {
            int o = ctx_offset_abs(t, luma);
            sig_blk[nnz++] = blk;
            first_nz = next;
            if (next > last_nz) last_nz = next;
            int gt1 = dec_bin(d, &m[gt1b + o]);
            (*remaining)--;
            int par = 0, gt2 = 0;
            if (gt1) {
                par = dec_bin(d, &m[parb + o]);
                gt2 = dec_bin(d, &m[gt2b + o]);
                *remaining -= 2;
            }
            c[blk] += 1 + par + gt1 + (gt2 << 1);
        }
        *state = next_state(table, *state, c[blk]);
        next--;
    }
    int first_mode2 = next;
    for (int p = first_sig; p > first_mode2; p--) {
        int blk = s->raster[p];
        if (c[blk] >= 4) {
            uint32_t k = derive_rice(s, c, p, 4);
            c[blk] += (int32_t)(dec_rem_abs_ep(d, k, 5, 15) << 1);
        }
    }
    for (int p = first_mode2; p >= min_sub; p--) {
        uint32_t k = derive_rice(s, c, p, 0);
        int pos0 = (*state < 2 ? 1 : 2) << k;
        int rem = (int)dec_rem_abs_ep(d, k, 5, 15);
        int level = rem == pos0 ? 0 : rem < pos0 ? rem + 1 : rem;
        *state = next_state(table, *state, level);
        if (level) {
            int blk = s->raster[p];
            sig_blk[nnz++] = blk;
            first_nz = p;
            if (p > last_nz) last_nz = p;
            c[blk] = level;
        }
    }
    int hide = sign_hiding && (last_nz - first_nz > 3) && nnz > 0;
    int nsigns = hide ? nnz - 1 : nnz;
    uint32_t pattern = nsigns > 0 ? dec_bins_ep(d, (uint32_t)nsigns) << (32 - nsigns) : 0;
    int64_t sum_abs = 0;
    for (int i = 0; i < nsigns; i++) {
        int32_t a = c[sig_blk[i]];
        if (a < -32768 || a > 32767) return -1;
        sum_abs += a;
        c[sig_blk[i]] = (pattern & 0x80000000u) ? -a : a;
        pattern <<= 1;
    }
    if (nnz > nsigns) {
        int32_t a = c[sig_blk[nsigns]];
        if (a < -32768 || a > 32767) return -1;
        sum_abs += a;
        c[sig_blk[nsigns]] = (sum_abs & 1) ? -a : a;
    }
    return 0;
}

static int decode_residual(BinDecoder* d, orc_model* m, int w, int h, int luma, int dq, int sdh, int32_t* c) {
    Scan s;
    make_scan(&s, w, h);
    memset(c, 0, sizeof(int32_t) * (size_t)w * h);
    int last = decode_last(d, m, &s, luma);
    if (last < 0) return -1;
    int gsize = 1 << s.log2_cg;
    int table = dq ? 32040 : 0;
    int state = 0, remaining = (s.n * 28) >> 4;
    Tmpl t = {-1, 0};
    uint8_t sig_group[1024];
    memset(sig_group, 0, sizeof sig_group);
    for (int g = last >> s.log2_cg; g >= 0; g--) {
        int min_sub = g * gsize, cr = s.cg_raster[g];
        int sig = (g == (last >> s.log2_cg)) || min_sub == 0;
        if (!sig) {
            int cx = cr % s.wg, cy = cr / s.wg;
            int right = cx + 1 < s.wg && sig_group[cr + 1];
            int lower = cy + 1 < s.hg && sig_group[cr + s.wg];
            sig = dec_bin(d, &m[(luma ? C_SIG_COEFF_GROUP_LUMA : C_SIG_COEFF_GROUP_CHROMA) + (right || lower)]);
        }
        if (!sig) continue;
        sig_group[cr] = 1;
        if (decode_subblock(d, m, &s, c, min_sub, min_sub + gsize - 1, last, luma, table, &state, sdh && !dq, &t,
                            &remaining))
            return -1;
    }
    return 0;
}

/* A substream of n_blocks TUs of w x h coded by rmh_cabac_residual_stream:
 * each TU's residual_coding, then end_of_slice_segment_flag (1 after the
 * last). Contexts are initialised for an I slice at qp. levels receives
 * n_blocks * w * h values in raster order. 0 on success, -1 on a malformed
 * stream. */
int orc_cabac_decode_residuals(const uint8_t* data, size_t n, int w, int h, int n_blocks, int luma, int dq, int sdh,
                               int qp, int32_t* levels) {
    orc_model m[ORC_NUM_CONTEXTS];
    orc_ctx_store(m, 0, qp);
    BinDecoder d;
    dec_init(&d, data, n);
    for (int b = 0; b < n_blocks; b++) {
        if (decode_residual(&d, m, w, h, luma, dq, sdh, levels + (size_t)b * w * h)) return -1;
        if (dec_bin_trm(&d) != (b == n_blocks - 1)) return -1;
    }
    return dec_finish(&d);
}

/* Known-answer hooks for the reference's own BinEncoder tests
 * (vvc/cabac.rs:1016-1120): decode n bins with one model at (qp, init 25,
 * window 13) or in bypass, then the terminating bin. */
int orc_cabac_decode_bins(const uint8_t* data, size_t n, int qp, int init_id, int ws, int n_ctx, int n_ep,
                          uint8_t* bins) {
    orc_model m = orc_ctx_init(qp, init_id, ws);
    BinDecoder d;
    dec_init(&d, data, n);
    for (int i = 0; i < n_ctx; i++) bins[i] = (uint8_t)dec_bin(&d, &m);
    for (int i = 0; i < n_ep; i++) bins[n_ctx + i] = (uint8_t)dec_bin_ep(&d);
    if (!dec_bin_trm(&d)) return -1;
    return dec_finish(&d);
}

/* ---- coding-tree split syntax ---- */

typedef struct {
    int no, qt, bh, bv, th, tv;
} CanSplit;

static CanSplit can_split_tree(int x, int y, int w, int h, int pw, int ph, int mtt, int min_qt, int max_bt, int max_tt,
                               int max_mtt, int min_cb, int last_split) {
    int right_out = x + w > pw, bottom_out = y + h > ph;
    int qt = w == h && mtt == 0 && w > (min_qt > min_cb ? min_qt : min_cb) && w > 4;
    CanSplit c = {!right_out && !bottom_out, qt, 0, 0, 0, 0};
    if (right_out && bottom_out) {
        c.qt = qt || (w == h && w > min_cb);
        c.bh = !qt && h > min_cb;
        return c;
    }
    if (right_out) {
        c.bv = w > min_cb;
        return c;
    }
    if (bottom_out) {
        c.bh = h > min_cb;
        return c;
    }
    if (mtt >= max_mtt) return c;
    c.bh = h > min_cb && w <= max_bt && h <= max_bt;
    c.bv = w > min_cb && w <= max_bt && h <= max_bt;
    c.th = h >= 4 * min_cb && w <= max_tt && h <= max_tt && last_split != 4;
    c.tv = w >= 4 * min_cb && w <= max_tt && h <= max_tt && last_split != 5;
    return c;
}

typedef struct {
    int split, qt, hv, hor_bt, ver_bt;
} SplitCtx;

/* left_h / above_w of 0: unavailable. */
static SplitCtx split_ctx(int w, int h, int qt_depth, int mtt, int left_h, int left_qt, int above_w, int above_qt,
                          CanSplit can) {
    SplitCtx r;
    int cs = 0;
    if (left_h > 0 && left_h < h) cs++;
    if (above_w > 0 && above_w < w) cs++;
    int ns = (can.qt ? 2 : 0) + can.bh + can.bv + can.th + can.tv;
    if (ns > 0) ns--;
    cs += 3 * (ns >> 1);
    int cq = 0;
    if (left_h > 0 && left_qt > qt_depth) cq++;
    if (above_w > 0 && above_qt > qt_depth) cq++;
    if (qt_depth >= 2) cq += 3;
    int nh = can.bh + can.th, nv = can.bv + can.tv, chv;
    if (nv == nh) {
        if (left_h > 0 && above_w > 0) {
            int da = w / above_w, dl = h / left_h;
            chv = da == dl ? 0 : da < dl ? 1 : 2;
        } else {
            chv = 0;
        }
    } else {
        chv = nv < nh ? 3 : 4;
    }
    r.split = cs < 8 ? cs : 8;
    r.qt = cq < 5 ? cq : 5;
    r.hv = chv < 4 ? chv : 4;
    r.hor_bt = mtt <= 1 ? 1 : 0;
    r.ver_bt = mtt <= 1 ? 3 : 2;
    return r;
}

/* Returns the CodingTreeSplit order: 0 none, 1 quad, 2 BH, 3 BV, 4 TH, 5 TV. */
static int decode_split(BinDecoder* d, orc_model* m, CanSplit can, SplitCtx sc) {
    int any = can.qt || can.bh || can.bv || can.th || can.tv;
    int any_btt = can.bh || can.bv || can.th || can.tv;
    int is_split = any;
    if (can.no && is_split) is_split = dec_bin(d, &m[C_SPLIT_FLAG + sc.split]);
    if (!is_split) return 0;
    int is_qt = can.qt;
    if (is_qt && any_btt) is_qt = dec_bin(d, &m[C_SPLIT_QT_FLAG + sc.qt]);
    if (is_qt) return 1;
    int can_hor = can.bh || can.th;
    int is_ver = can.bv || can.tv;
    if (is_ver && can_hor) is_ver = dec_bin(d, &m[C_SPLIT_HV_FLAG + sc.hv]);
    int can14 = is_ver ? can.tv : can.th;
    int is12 = is_ver ? can.bv : can.bh;
    if (is12 && can14) is12 = dec_bin(d, &m[C_SPLIT_12_FLAG + (is_ver ? sc.ver_bt : sc.hor_bt)]);
    return is_ver ? (is12 ? 3 : 5) : (is12 ? 2 : 4);
}

/* ---- the luma substream of the RD search ---- */

typedef struct {
    int W, H;
    uint8_t* rec;
    uint8_t* decoded;                  /* per sample */
    uint8_t *cu_w, *cu_h, *cu_qt;      /* per 4x4 cell; cu_w 0 = not decoded */
} Pic;

static void neighbour(const Pic* p, int x, int y, int* w, int* h, int* qt) {
    *w = *h = *qt = 0;
    if (x < 0 || y < 0 || x >= p->W || y >= p->H) return;
    size_t c = (size_t)(y >> 2) * (p->W >> 2) + (x >> 2);
    *w = p->cu_w[c];
    *h = p->cu_h[c];
    *qt = p->cu_qt[c];
}

static int read_split(BinDecoder* d, orc_model* m, const Pic* p, int x, int y, int w, int h, int qt_depth, int mtt,
                      int last_split) {
    CanSplit can = can_split_tree(x, y, w, h, p->W, p->H, mtt, 8, 32, 32, 1, 4, last_split);
    int lw, lh, lq, aw, ah, aq;
    neighbour(p, x - 1, y, &lw, &lh, &lq);
    neighbour(p, x, y - 1, &aw, &ah, &aq);
    return decode_split(d, m, can, split_ctx(w, h, qt_depth, mtt, lh, lq, aw, aq, can));
}

/* PLANAR (0) / DC (1) from the decoded neighbours: left column bottom-up
 * (h + 1 samples), the corner, the top row (w + 1); unavailable samples take
 * the previous available one in that order, the first available one before
 * it, or 128 when none is. */
static void predict(const Pic* p, int x0, int y0, int w, int h, int mode, int* pred) {
    int line[2 * 65 + 1];
    int nl = h + 1, nt = w + 1, tot = nl + 1 + nt;
    for (int i = 0; i < tot; i++) {
        int x, y;
        if (i < nl) {
            x = x0 - 1;
            y = y0 + nl - 1 - i;
        } else if (i == nl) {
            x = x0 - 1;
            y = y0 - 1;
        } else {
            x = x0 + (i - nl - 1);
            y = y0 - 1;
        }
        int ok = x >= 0 && y >= 0 && x < p->W && y < p->H && p->decoded[(size_t)y * p->W + x];
        line[i] = ok ? p->rec[(size_t)y * p->W + x] : -1;
    }
    int first = -1;
    for (int i = 0; i < tot && first < 0; i++) first = line[i];
    int prev = first < 0 ? 128 : first;
    for (int i = 0; i < tot; i++) {
        if (line[i] < 0)
            line[i] = prev;
        else
            prev = line[i];
    }
    int left[66], top[66];
    for (int j = 0; j <= h; j++) left[j] = line[nl - 1 - j];
    for (int i = 0; i <= w; i++) top[i] = line[nl + 1 + i];
    int lw = 0, lh = 0;
    while ((1 << lw) < w) lw++;
    while ((1 << lh) < h) lh++;
    if (mode == 0) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int pv = ((h - 1 - y) * top[x] + (y + 1) * left[h]) << lw;
                int ph = ((w - 1 - x) * left[y] + (x + 1) * top[w]) << lh;
                pred[y * w + x] = (pv + ph + w * h) >> (lw + lh + 1);
            }
        return;
    }
    int s = 0, dc;
    if (w == h) {
        for (int i = 0; i < w; i++) s += top[i] + left[i];
        dc = (s + w) >> (lw + 1);
    } else if (w > h) {
        for (int i = 0; i < w; i++) s += top[i];
        dc = (s + (w >> 1)) >> lw;
    } else {
        for (int i = 0; i < h; i++) s += left[i];
        dc = (s + (h >> 1)) >> lh;
    }
    for (int i = 0; i < w * h; i++) pred[i] = dc;
}

static int decode_leaf(BinDecoder* d, orc_model* m, Pic* p, int x0, int y0, int w, int h, int qt_depth, int qp) {
    int mode;
    if (!dec_bin(d, &m[C_INTRA_LUMA_MPM_FLAG])) return -1; /* only MPM modes occur */
    if (!dec_bin(d, &m[C_INTRA_LUMA_PLANAR_FLAG + 1])) {
        mode = 0;
    } else {
        if (dec_unary_max_ep(d, 4) != 0) return -1; /* MPM 1 is DC beside PLANAR/DC neighbours */
        mode = 1;
    }
    int pred[32 * 32];
    int32_t lv[32 * 32], coef[32 * 32], res[32 * 32];
    predict(p, x0, y0, w, h, mode, pred);
    memset(res, 0, sizeof res);
    if (dec_bin(d, &m[C_QT_CBF_Y])) {
        if (decode_residual(d, m, w, h, 1, 0, 0, lv)) return -1;
        orc_dequant_rect(w, h, lv, qp, 8, coef);
        orc_inverse_transform(w, h, coef, RMH_TR_DCT2, RMH_TR_DCT2, 0, 0, 8, res);
    }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int v = pred[y * w + x] + res[y * w + x];
            v = v < 0 ? 0 : v > 255 ? 255 : v;
            size_t i = (size_t)(y0 + y) * p->W + x0 + x;
            p->rec[i] = (uint8_t)v;
            p->decoded[i] = 1;
        }
    for (int cy = y0 >> 2; cy < (y0 + h) >> 2; cy++)
        for (int cx = x0 >> 2; cx < (x0 + w) >> 2; cx++) {
            size_t c = (size_t)cy * (p->W >> 2) + cx;
            p->cu_w[c] = (uint8_t)w;
            p->cu_h[c] = (uint8_t)h;
            p->cu_qt[c] = (uint8_t)qt_depth;
        }
    return 0;
}

/* Children of a split of a w x h block (split_coding_tree_node,
 * vvc/transform.rs:363-409): {x, y, w, h} relative to the block. */
static int split_children(int split, int w, int h, int (*c)[4]) {
    int n = 0;
#define CH(X, Y, W, H) (c[n][0] = (X), c[n][1] = (Y), c[n][2] = (W), c[n][3] = (H), n++)
    switch (split) {
        case 1: CH(0, 0, w / 2, h / 2); CH(w / 2, 0, w / 2, h / 2); CH(0, h / 2, w / 2, h / 2); CH(w / 2, h / 2, w / 2, h / 2); break;
        case 2: CH(0, 0, w, h / 2); CH(0, h / 2, w, h / 2); break;
        case 3: CH(0, 0, w / 2, h); CH(w / 2, 0, w / 2, h); break;
        case 4: CH(0, 0, w, h / 4); CH(0, h / 4, w, h / 2); CH(0, 3 * h / 4, w, h / 4); break;
        default: CH(0, 0, w / 4, h); CH(w / 4, 0, w / 2, h); CH(3 * w / 4, 0, w / 4, h); break;
    }
#undef CH
    return n;
}

/* coding_tree(): the split, then each child in order, down to the leaf CUs. */
static int decode_tree(BinDecoder* d, orc_model* m, Pic* p, int x, int y, int w, int h, int qt_depth, int mtt,
                       int last_split, int qp, int* split_out) {
    int split = read_split(d, m, p, x, y, w, h, qt_depth, mtt, last_split);
    if (split_out) *split_out = split;
    if (split == 0) return decode_leaf(d, m, p, x, y, w, h, qt_depth, qp);
    int ch[4][4];
    int n = split_children(split, w, h, ch);
    for (int c = 0; c < n; c++) {
        int rc = decode_tree(d, m, p, x + ch[c][0], y + ch[c][1], ch[c][2], ch[c][3], qt_depth + (split == 1),
                             mtt + (split >= 2), split, qp, NULL);
        if (rc) return rc;
    }
    return 0;
}

/* Decodes rmh_rd_encode_luma_stream's substream of a width x height picture
 * (multiples of 32) at qp into recon (width * height), the split of each
 * 32x32 CTU and, when cu_wh is not null, the width and height of the CU
 * covering each 4x4 cell (2 bytes per cell, raster order). 0 on success; -1
 * on a malformed stream or a syntax element this subset does not produce. */
int orc_decode_luma_stream(const uint8_t* data, size_t n, int width, int height, int qp, uint8_t* recon,
                           uint8_t* splits, uint8_t* cu_wh) {
    if (width <= 0 || height <= 0 || width % 32 || height % 32) return -1;
    orc_model m[ORC_NUM_CONTEXTS];
    orc_ctx_store(m, 0, qp);
    BinDecoder d;
    dec_init(&d, data, n);
    size_t cells = (size_t)(width / 4) * (height / 4);
    Pic p = {width, height, recon, calloc((size_t)width * height, 1), calloc(cells, 1), calloc(cells, 1),
             calloc(cells, 1)};
    int rc = 0;
    const int nbx = width / 32, nby = height / 32;
    for (int b = 0; b < nbx * nby && rc == 0; b++) {
        int split = 0;
        rc = decode_tree(&d, m, &p, (b % nbx) * 32, (b / nbx) * 32, 32, 32, 0, 0, -1, qp, &split);
        splits[b] = (uint8_t)split;
        if (rc == 0 && dec_bin_trm(&d) != (b == nbx * nby - 1)) rc = -1;
    }
    if (rc == 0) rc = dec_finish(&d);
    if (cu_wh)
        for (size_t c = 0; c < cells; c++) {
            cu_wh[2 * c] = p.cu_w[c];
            cu_wh[2 * c + 1] = p.cu_h[c];
        }
    free(p.decoded);
    free(p.cu_w);
    free(p.cu_h);
    free(p.cu_qt);
    return rc;
}

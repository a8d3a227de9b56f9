// cabac.cpp - the CABAC coder of cabac.h (host code).
//
// Reference functions restated, each at its cited lines:
//   Model / Contexts::init          CtxModel::init, update, get_lps (vvc/cabac.rs:41-99),
//                                   CtxStore::from_table (vvc/cabac.rs:269-278)
//   Encoder                          BinEncoder (vvc/cabac.rs:477-745)
//   encode_residual                  encode_residual_with_options, encode_last_sig_coeff,
//                                    encode_subblock and their context helpers
//                                    (vvc/residual.rs:85-240, 280-357, 448-743)
//   can_split / split_contexts       can_split_coding_tree, derive_split_contexts
//                                    (vvc/transform.rs:537-715)
//   encode_split                     encode_coding_tree_split (vvc/slice.rs:3850-3893)
//   Decoder                          BinDecoder (vvc/cabac.rs:761-1007)
//   decode_residual                  decode_residual_with_options, decode_last_sig_coeff,
//                                    decode_subblock (vvc/residual.rs:359-433, 537-887)
//   decode_split                     decode_coding_tree_split (vvc/slice.rs:3895-3946)
// The coefficient scan is CoeffScan::new (vvc/transform.rs:37-79, 100-212).
#include "cabac.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "../../include/rmh.h"
#include "cabac_data.h"

namespace rmh {
namespace cabac {
namespace {

constexpr uint32_t kMask0 = 0x7FE0, kMask1 = 0x7FFE;

// LPS renormalisation shift by lps_range >> 3 (H.266 Table 127).
constexpr uint8_t kRenormLps[32] = {6, 5, 4, 4, 3, 3, 3, 3, 2, 2, 2, 2, 2, 2, 2, 2,
                                    1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};

inline int model_state(const Model& m) { return (m.s0 + m.s1) >> 8; }

inline uint32_t model_lps(const Model& m, uint32_t range) {
    int q = model_state(m);
    if (q & 0x80) q ^= 0xFF;
    return ((((uint32_t)q >> 2) * (range >> 5)) >> 1) + 4;
}

inline void model_update(Model& m, int bin) {
    uint32_t n0 = m.s0 - ((m.s0 >> m.r0) & kMask0);
    uint32_t n1 = m.s1 - ((m.s1 >> m.r1) & kMask1);
    if (bin) {
        n0 += (0x7FFFu >> m.r0) & kMask0;
        n1 += (0x7FFFu >> m.r1) & kMask1;
    }
    m.s0 = (uint16_t)n0;
    m.s1 = (uint16_t)n1;
}

Model model_init(int qp, int init_id, int log2_window) {
    const int slope = (init_id >> 3) - 4;
    const int offset = (init_id & 7) * 18 + 1;
    const int st = std::clamp(((slope * (qp - 16)) >> 1) + offset, 1, 127);
    const uint32_t p = (uint32_t)st << 8;
    Model m;
    m.r0 = (uint8_t)(2 + ((log2_window >> 2) & 3));
    m.r1 = (uint8_t)(3 + m.r0 + (log2_window & 3));
    m.s0 = (uint16_t)(p & kMask0);
    m.s1 = (uint16_t)(p & kMask1);
    return m;
}

// ---- coefficient scan ---------------------------------------------------

struct Scan {
    int w = 0, h = 0, log2_cg = 0, wg = 0, hg = 0, n = 0;
    std::vector<uint16_t> raster;   // scan position -> y * w + x
    std::vector<uint8_t> px, py;    // scan position -> x, y
    std::vector<uint16_t> cg;       // group index (scan order) -> group raster y * wg + x
};

void diag_order(int w, int h, std::vector<int>& xs, std::vector<int>& ys) {
    int line = 0, col = 0;
    for (int i = 0; i < w * h; i++) {
        xs.push_back(col);
        ys.push_back(line);
        if (col == w - 1 || line == 0) {
            line += col + 1;
            col = 0;
            if (line >= h) {
                col += line - (h - 1);
                line = h - 1;
            }
        } else {
            col++;
            line--;
        }
    }
}

void group_log2(int lx, int ly, int* a, int* b) {
    if (lx == 0) {
        *a = 0;
        *b = std::min(ly, 4);
    } else if (ly == 0) {
        *a = std::min(lx, 4);
        *b = 0;
    } else if (lx == 1) {
        *a = 1;
        *b = ly <= 2 ? 1 : 3;
    } else if (ly == 1) {
        *a = lx == 2 ? 1 : 3;
        *b = 1;
    } else {
        *a = 2;
        *b = 2;
    }
}

Scan make_scan(int lx, int ly) {
    Scan s;
    s.w = 1 << lx;
    s.h = 1 << ly;
    int a, b;
    group_log2(lx, ly, &a, &b);
    s.log2_cg = a + b;
    s.wg = std::min(s.w, 32) >> a;
    s.hg = std::min(s.h, 32) >> b;
    std::vector<int> gx, gy, ix, iy;
    diag_order(s.wg, s.hg, gx, gy);
    diag_order(1 << a, 1 << b, ix, iy);
    for (size_t g = 0; g < gx.size(); g++) {
        s.cg.push_back((uint16_t)(gy[g] * s.wg + gx[g]));
        for (size_t i = 0; i < ix.size(); i++) {
            const int x = (gx[g] << a) + ix[i], y = (gy[g] << b) + iy[i];
            s.raster.push_back((uint16_t)(y * s.w + x));
            s.px.push_back((uint8_t)x);
            s.py.push_back((uint8_t)y);
        }
    }
    s.n = (int)s.raster.size();
    return s;
}

const Scan& scan_for(int w, int h) {
    struct All {
        Scan s[7][7];
        All() {
            for (int a = 0; a < 7; a++)
                for (int b = 0; b < 7; b++) s[a][b] = make_scan(a, b);
        }
    };
    static const All all;
    int lx = 0, ly = 0;
    while ((1 << lx) < w) lx++;
    while ((1 << ly) < h) ly++;
    return all.s[lx][ly];
}

// ---- residual_coding ----------------------------------------------------

constexpr int kPrefixCtx[8] = {0, 0, 0, 3, 6, 10, 15, 21};
constexpr int kMinInGroup[14] = {0, 1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96};
constexpr uint8_t kGroupIdx[64] = {0,  1,  2,  3,  4,  4,  5,  5,  6,  6,  6,  6,  7,  7,  7,  7,
                                   8,  8,  8,  8,  8,  8,  8,  8,  9,  9,  9,  9,  9,  9,  9,  9,
                                   10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10,
                                   11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11};
constexpr uint8_t kRiceByTemplate[32] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 2, 2,
                                         2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3};
constexpr uint32_t kRemainBinReduction = 5, kMaxLog2TrRange = 15;
constexpr int kRiceBaseLevel = 4;
constexpr int kDepQuantTransitions = 32040;  // four 4-bit rows: next state by (state, parity)

inline int next_state(int table, int state, int level) { return (table >> ((state << 2) + ((level & 1) << 1))) & 3; }

// Template over the right, two-right, below-right, below and two-below neighbours.
struct Template {
    int diag = -1, sum1 = 0;
};

int sig_ctx(const Scan& s, const int32_t* lv, int pos, bool luma, int state, Template& t) {
    const int x = s.px[pos], y = s.py[pos], w = s.w, h = s.h, base = y * w + x;
    int sum_abs = 0, num_pos = 0;
    auto add = [&](int i) {
        const int a = std::abs(lv[i]);
        sum_abs += std::min(4 + (a & 1), a);
        num_pos += a != 0;
    };
    if (x + 1 < w) {
        add(base + 1);
        if (x + 2 < w) add(base + 2);
        if (y + 1 < h) add(base + w + 1);
    }
    if (y + 1 < h) {
        add(base + w);
        if (y + 2 < h) add(base + 2 * w);
    }
    const int d = x + y;
    int ofs = std::min((sum_abs + 1) >> 1, 3) + (d < 2 ? 4 : 0);
    if (luma) ofs += d < 5 ? 4 : 0;
    t.diag = d;
    t.sum1 = sum_abs - num_pos;
    const int set = std::min(std::max(state - 1, 0), 2);
    static constexpr int kLuma[3] = {SIG_FLAG_LUMA, SIG_FLAG_LUMA_1, SIG_FLAG_LUMA_2};
    static constexpr int kChroma[3] = {SIG_FLAG_CHROMA, SIG_FLAG_CHROMA_1, SIG_FLAG_CHROMA_2};
    return (luma ? kLuma[set] : kChroma[set]) + ofs;
}

int level_ctx_offset(const Template& t, bool luma) {
    if (t.diag < 0) return 0;
    int o = std::min(t.sum1, 4) + 1;
    if (t.diag == 0)
        o += luma ? 15 : 5;
    else if (luma)
        o += t.diag < 3 ? 10 : t.diag < 10 ? 5 : 0;
    return o;
}

uint32_t rice_param(const Scan& s, const int32_t* lv, int pos, int base_level) {
    const int x = s.px[pos], y = s.py[pos], w = s.w, h = s.h, base = y * w + x;
    int sum = 0;
    if (x + 1 < w) {
        sum += std::abs(lv[base + 1]);
        if (x + 2 < w) sum += std::abs(lv[base + 2]);
        if (y + 1 < h) sum += std::abs(lv[base + w + 1]);
    }
    if (y + 1 < h) {
        sum += std::abs(lv[base + w]);
        if (y + 2 < h) sum += std::abs(lv[base + 2 * w]);
    }
    return kRiceByTemplate[std::clamp(sum - 5 * base_level, 0, 31)];
}

void encode_last(Encoder& e, Contexts& c, const Scan& s, int last, bool luma) {
    const int x = s.px[last], y = s.py[last];
    const int gx = kGroupIdx[x], gy = kGroupIdx[y];
    const int maxx = kGroupIdx[std::min(s.w, 32) - 1], maxy = kGroupIdx[std::min(s.h, 32) - 1];
    int lw = 0, lh = 0;
    while ((1 << lw) < s.w) lw++;
    while ((1 << lh) < s.h) lh++;
    const int bx = luma ? LAST_X_LUMA : LAST_X_CHROMA, by = luma ? LAST_Y_LUMA : LAST_Y_CHROMA;
    const int ox = luma ? kPrefixCtx[lw] : 0, oy = luma ? kPrefixCtx[lh] : 0;
    const int sx = luma ? (lw + 1) >> 2 : std::min(s.w >> 3, 2);
    const int sy = luma ? (lh + 1) >> 2 : std::min(s.h >> 3, 2);
    for (int i = 0; i < gx; i++) e.bin(1, c.m[bx + ox + (i >> sx)]);
    if (gx < maxx) e.bin(0, c.m[bx + ox + (gx >> sx)]);
    for (int i = 0; i < gy; i++) e.bin(1, c.m[by + oy + (i >> sy)]);
    if (gy < maxy) e.bin(0, c.m[by + oy + (gy >> sy)]);
    if (gx > 3) {
        const int v = x - kMinInGroup[gx], cnt = ((gx - 2) >> 1) - 1;
        for (int i = cnt; i >= 0; i--) e.bypass((v >> i) & 1);
    }
    if (gy > 3) {
        const int v = y - kMinInGroup[gy], cnt = ((gy - 2) >> 1) - 1;
        for (int i = cnt; i >= 0; i--) e.bypass((v >> i) & 1);
    }
}

void encode_group(Encoder& e, Contexts& c, const Scan& s, const int32_t* lv, int min_sub, int last, bool luma,
                  int table, int& state, bool sign_hiding, Template& t, int& remaining) {
    const int max_sub = min_sub + (1 << s.log2_cg) - 1;
    const bool has_last = (last >> s.log2_cg) == (min_sub >> s.log2_cg);
    int next = has_last ? last : max_sub;
    const int infer = next != last ? ((min_sub >> s.log2_cg) != 0 ? min_sub : -1) : next;
    const int gt1_base = luma ? GTX1_LUMA : GTX1_CHROMA;
    const int par_base = luma ? PAR_FLAG_LUMA : PAR_FLAG_CHROMA;
    const int gt2_base = luma ? GTX2_LUMA : GTX2_CHROMA;
    int first_nz = next, last_nz = -1, nnz = 0, first_2nd = min_sub - 1;
    uint32_t signs = 0;

    // Pass 1: significance, greater-than-1, parity and greater-than-3 flags.
    for (; next >= min_sub && remaining >= 4; next--) {
        const int32_t level = lv[s.raster[next]];
        const int sig = level != 0;
        if (nnz != 0 || next != infer) {
            e.bin(sig, c.m[sig_ctx(s, lv, next, luma, state, t)]);
            remaining--;
        } else if (next != last) {
            sig_ctx(s, lv, next, luma, state, t);
        }
        if (sig) {
            const int o = level_ctx_offset(t, luma);
            nnz++;
            first_nz = next;
            last_nz = std::max(last_nz, next);
            const int a = std::abs(level);
            if (next != last) signs <<= 1;
            if (level < 0) signs += 1;
            e.bin(a > 1, c.m[gt1_base + o]);
            remaining--;
            if (a > 1) {
                e.bin(a & 1, c.m[par_base + o]);
                e.bin(a > 3, c.m[gt2_base + o]);
                remaining -= 2;
                if (a > 3) first_2nd = std::max(first_2nd, next);
            }
        }
        state = next_state(table, state, level);
    }
    const int min_2nd = next;

    // Pass 2: Rice remainders of the levels above 3.
    for (int p = first_2nd; p > min_2nd; p--) {
        const int a = std::abs(lv[s.raster[p]]);
        if (a >= 4) e.rem_abs((uint32_t)((a - 4) >> 1), rice_param(s, lv, p, kRiceBaseLevel), kRemainBinReduction,
                              kMaxLog2TrRange);
    }

    // Pass 3: whole levels in bypass for the positions pass 1 did not reach.
    for (int p = min_2nd; p >= min_sub; p--) {
        const int32_t level = lv[s.raster[p]];
        const int a = std::abs(level);
        const uint32_t k = rice_param(s, lv, p, 0);
        const int pos0 = (state < 2 ? 1 : 2) << k;
        const int rem = a == 0 ? pos0 : a <= pos0 ? a - 1 : a;
        e.rem_abs((uint32_t)rem, k, kRemainBinReduction, kMaxLog2TrRange);
        state = next_state(table, state, a);
        if (a) {
            nnz++;
            first_nz = p;
            last_nz = std::max(last_nz, p);
            signs <<= 1;
            if (level < 0) signs += 1;
        }
    }

    int nsigns = nnz;
    if (sign_hiding && last_nz - first_nz > 3 && nsigns > 0) {
        nsigns--;
        signs >>= 1;
    }
    e.bypass_bins(signs, nsigns);
}

}  // namespace

void Contexts::init(int slice, int qp) {
    qp = std::clamp(qp, 0, 63);
    for (int i = 0; i < NUM_CONTEXTS; i++) m[i] = model_init(qp, RMH_CTX_INIT[slice][i][0], RMH_CTX_INIT[slice][i][1]);
}

// ---- Encoder ------------------------------------------------------------

void Encoder::put_bits(uint32_t v, int n) {
    written_ += (uint64_t)n;
    for (int i = n - 1; i >= 0; i--) {
        acc_ = (acc_ << 1) | ((v >> i) & 1);
        if (++acc_n_ == 8) {
            if (sink_) sink_->push_back((uint8_t)acc_);
            acc_ = 0;
            acc_n_ = 0;
        }
    }
}

// Moves the top byte of `low` out, holding 0xFF bytes until a later byte shows
// whether a carry reaches them.
void Encoder::out_byte_() {
    const uint32_t lead = low_ >> (24 - bits_left_);
    bits_left_ += 8;
    low_ &= bits_left_ >= 32 ? 0u : 0xFFFFFFFFu >> bits_left_;
    if (lead == 0xFF) {
        held_++;
    } else if (held_ > 0) {
        const uint32_t carry = lead >> 8;
        put_bits(held_byte_ + carry, 8);
        held_byte_ = lead & 0xFF;
        const uint32_t fill = (0xFF + carry) & 0xFF;
        for (; held_ > 1; held_--) put_bits(fill, 8);
    } else {
        held_ = 1;
        held_byte_ = lead;
    }
}

void Encoder::bin(int b, Model& m) {
    const uint32_t lps = model_lps(m, range_);
    const int mps = model_state(m) >= 128;
    range_ -= lps;
    if (b != mps) {
        const int nb = kRenormLps[lps >> 3];
        bits_left_ -= nb;
        low_ = (low_ + range_) << nb;
        range_ = lps << nb;
        if (bits_left_ < 12) out_byte_();
    } else if (range_ < 256) {
        bits_left_ -= 1;
        low_ <<= 1;
        range_ <<= 1;
        if (bits_left_ < 12) out_byte_();
    }
    model_update(m, b);
}

void Encoder::bypass(int b) {
    low_ <<= 1;
    if (b) low_ += range_;
    if (--bits_left_ < 12) out_byte_();
}

void Encoder::bypass_bins(uint32_t v, int n) {
    if (n <= 0) return;
    if (range_ == 256) {
        // range 256: each bin is one bit of low, so whole groups go in at once
        for (int rem = n; rem > 0;) {
            const int k = std::min(rem, 8);
            const uint32_t bits = (v >> (rem - k)) & ((1u << k) - 1);
            low_ = (low_ << k) + (bits << 8);
            rem -= k;
            bits_left_ -= k;
            if (bits_left_ < 12) out_byte_();
        }
        return;
    }
    int rem = n;
    while (rem > 8) {
        rem -= 8;
        const uint32_t pat = (v >> rem) & 0xFF;
        low_ = (low_ << 8) + range_ * pat;
        v -= pat << rem;
        bits_left_ -= 8;
        if (bits_left_ < 12) out_byte_();
    }
    low_ = (low_ << rem) + range_ * v;
    bits_left_ -= rem;
    if (bits_left_ < 12) out_byte_();
}

void Encoder::terminate(int b) {
    range_ -= 2;
    if (b) {
        low_ = (low_ + range_) << 7;
        range_ = 2 << 7;
        bits_left_ -= 7;
    } else if (range_ >= 256) {
        return;
    } else {
        low_ <<= 1;
        range_ <<= 1;
        bits_left_ -= 1;
    }
    if (bits_left_ < 12) out_byte_();
}

void Encoder::unary_max(uint32_t sym, uint32_t max_sym) {
    if (max_sym == 0) return;
    for (uint32_t i = 0; i < std::min(sym, max_sym); i++) bypass(1);
    if (sym < max_sym) bypass(0);
}

// Rice prefix (unary up to `cutoff`, then Exp-Golomb escape) and suffix.
void Encoder::rem_abs(uint32_t v, uint32_t k, uint32_t cutoff, uint32_t max_log2) {
    if (v < (cutoff << k)) {
        const uint32_t len = (v >> k) + 1;
        bypass_bins((1u << len) - 2, (int)len);
        bypass_bins(v & ((1u << k) - 1), (int)k);
        return;
    }
    const uint32_t max_prefix = 32 - cutoff - max_log2;
    const uint32_t code = (v >> k) - cutoff;
    uint32_t plen, slen;
    if (code >= (1u << max_prefix) - 1) {
        plen = max_prefix;
        slen = max_log2;
    } else {
        plen = 0;
        while (code > (2u << plen) - 2) plen++;
        slen = plen + k + 1;
    }
    const uint32_t total = plen + cutoff;
    const uint32_t suffix = ((code - ((1u << plen) - 1)) << k) | (v & ((1u << k) - 1));
    bypass_bins((1u << total) - 1, (int)total);
    bypass_bins(suffix, (int)slen);
}

void Encoder::finish() {
    if (bits_left_ < 32 && (low_ >> (32 - bits_left_)) != 0) {
        put_bits(held_byte_ + 1, 8);
        for (; held_ > 1; held_--) put_bits(0x00, 8);
        low_ -= 1u << (32 - bits_left_);
    } else {
        if (held_ > 0) put_bits(held_byte_, 8);
        for (; held_ > 1; held_--) put_bits(0xFF, 8);
    }
    put_bits(low_ >> 8, 24 - bits_left_);
    put_bits(1, 1);
    if (acc_n_) put_bits(0, 8 - acc_n_);
    held_ = 0;
    bits_left_ = 23;
    low_ = 0;
    range_ = 510;
}

// ---- residual_coding ----------------------------------------------------

void encode_residual(Encoder& e, Contexts& c, int w, int h, const int32_t* lv, bool luma, bool dep_quant,
                     bool sign_hiding) {
    const Scan& s = scan_for(w, h);
    int last = -1;
    for (int p = s.n - 1; p >= 0; p--)
        if (lv[s.raster[p]]) {
            last = p;
            break;
        }
    if (last < 0) return;  // callers code cbf = 0 instead
    encode_last(e, c, s, last, luma);

    const int gsize = 1 << s.log2_cg;
    const int ngroups = s.wg * s.hg;
    uint8_t coded[1024], sig_raster[1024];
    std::fill(coded, coded + ngroups, 0);
    std::fill(sig_raster, sig_raster + ngroups, 0);
    for (int p = 0; p < s.n; p++)
        if (lv[s.raster[p]]) coded[p >> s.log2_cg] = 1;

    const int table = dep_quant ? kDepQuantTransitions : 0;
    int state = 0, remaining = (s.n * 28) >> 4;
    Template t;
    const int sg_base = luma ? SIG_COEFF_GROUP_LUMA : SIG_COEFF_GROUP_CHROMA;
    for (int g = last >> s.log2_cg; g >= 0; g--) {
        const int cr = s.cg[g];
        sig_raster[cr] = coded[g];
        if (g != 0 && g != (last >> s.log2_cg)) {
            const int cx = cr % s.wg, cy = cr / s.wg;
            const bool right = cx + 1 < s.wg && sig_raster[cr + 1];
            const bool below = cy + 1 < s.hg && sig_raster[cr + s.wg];
            e.bin(coded[g], c.m[sg_base + (right || below)]);
            if (!coded[g]) continue;
        }
        encode_group(e, c, s, lv, g * gsize, last, luma, table, state, sign_hiding && !dep_quant, t, remaining);
    }
}

// ---- coding-tree split syntax -------------------------------------------

CanSplit can_split(int x, int y, int w, int h, int pic_w, int pic_h, int mtt_depth, const SplitLimits& lim,
                   int last_split) {
    const bool right_out = x + w > pic_w, bottom_out = y + h > pic_h;
    const bool qt = w == h && mtt_depth == 0 && w > std::max(lim.min_qt, lim.min_cb) && w > 4;
    if (right_out && bottom_out) return {false, qt || (w == h && w > lim.min_cb), !qt && h > lim.min_cb, false, false, false};
    if (right_out) return {false, qt, false, w > lim.min_cb, false, false};
    if (bottom_out) return {false, qt, h > lim.min_cb, false, false, false};
    if (mtt_depth >= lim.max_mtt_depth) return {true, qt, false, false, false, false};
    const bool bt_fits = w <= lim.max_bt && h <= lim.max_bt, tt_fits = w <= lim.max_tt && h <= lim.max_tt;
    return {true,
            qt,
            h > lim.min_cb && bt_fits,
            w > lim.min_cb && bt_fits,
            h >= 4 * lim.min_cb && tt_fits && last_split != 4,
            w >= 4 * lim.min_cb && tt_fits && last_split != 5};
}

SplitCtx split_contexts(int w, int h, int qt_depth, int mtt_depth, int left_h, int left_qt, int above_w, int above_qt,
                        const CanSplit& can) {
    SplitCtx r{};
    const bool has_left = left_h > 0, has_above = above_w > 0;
    int cs = (has_left && left_h < h) + (has_above && above_w < w);
    int nsplit = (can.qt ? 2 : 0) + can.bh + can.bv + can.th + can.tv;
    if (nsplit > 0) nsplit--;
    cs += 3 * (nsplit >> 1);
    int cq = (has_left && left_qt > qt_depth) + (has_above && above_qt > qt_depth) + (qt_depth >= 2 ? 3 : 0);
    const int nh = can.bh + can.th, nv = can.bv + can.tv;
    int chv;
    if (nv == nh) {
        if (has_left && has_above) {
            const int dep_above = w / std::max(above_w, 1), dep_left = h / std::max(left_h, 1);
            chv = dep_above == dep_left ? 0 : dep_above < dep_left ? 1 : 2;
        } else {
            chv = 0;
        }
    } else {
        chv = nv < nh ? 3 : 4;
    }
    r.split = std::min(cs, 8);
    r.qt = std::min(cq, 5);
    r.hv = std::min(chv, 4);
    r.hor_bt = mtt_depth <= 1 ? 1 : 0;
    r.ver_bt = mtt_depth <= 1 ? 3 : 2;
    return r;
}

void encode_split(Encoder& e, Contexts& c, int split, const CanSplit& can, const SplitCtx& sc) {
    const bool is_split = split != 0;
    if (can.no && can.any()) e.bin(is_split, c.m[SPLIT_FLAG + sc.split]);
    if (!is_split) return;
    const bool is_qt = split == 1;
    if (can.qt && can.any_mtt()) e.bin(is_qt, c.m[SPLIT_QT_FLAG + sc.qt]);
    if (is_qt) return;
    const bool can_hor = can.bh || can.th, can_ver = can.bv || can.tv;
    const bool is_ver = split == 3 || split == 5;
    if (can_ver && can_hor) e.bin(is_ver, c.m[SPLIT_HV_FLAG + sc.hv]);
    const bool can12 = is_ver ? can.bv : can.bh, can14 = is_ver ? can.tv : can.th;
    if (can12 && can14) e.bin(split == 2 || split == 3, c.m[SPLIT_12_FLAG + (is_ver ? sc.ver_bt : sc.hor_bt)]);
}

// ---- decoding side -----------------------------------------------------

Decoder::Decoder(const uint8_t* data, size_t n) : data_(data), n_(n), pos_(n < 2 ? n : 2) {
    const uint32_t b0 = n > 0 ? data[0] : 0, b1 = n > 1 ? data[1] : 0;
    value_ = (b0 << 8) + b1;
    last_byte_ = (uint8_t)b1;
}

uint32_t Decoder::read_byte_() {
    if (pos_ >= n_) return 0;
    last_byte_ = data_[pos_++];
    return last_byte_;
}

int Decoder::bin(Model& m) {
    const int mps = model_state(m) >= 128;
    const uint32_t lps = model_lps(m, range_);
    range_ -= lps;
    const uint32_t scaled = range_ << 7;
    int b;
    if (value_ < scaled) {
        b = mps;
        if (range_ < 256) {
            range_ <<= 1;
            value_ <<= 1;
            if (++bits_needed_ >= 0) {
                value_ += read_byte_() << bits_needed_;
                bits_needed_ -= 8;
            }
        }
    } else {
        b = !mps;
        const int nb = kRenormLps[lps >> 3];
        value_ = (value_ - scaled) << nb;
        range_ = lps << nb;
        bits_needed_ += nb;
        if (bits_needed_ >= 0) {
            value_ += read_byte_() << bits_needed_;
            bits_needed_ -= 8;
        }
    }
    model_update(m, b);
    return b;
}

int Decoder::bypass() {
    value_ <<= 1;
    if (++bits_needed_ >= 0) {
        value_ += read_byte_();
        bits_needed_ = -8;
    }
    const uint32_t scaled = range_ << 7;
    if (value_ < scaled) return 0;
    value_ -= scaled;
    return 1;
}

uint32_t Decoder::bypass_bins(int n) {
    uint32_t v = 0;
    for (int i = 0; i < n; i++) v = (v << 1) | (uint32_t)bypass();
    return v;
}

int Decoder::terminate() {
    range_ -= 2;
    if (value_ >= (range_ << 7)) return 1;
    if (range_ < 256) {
        range_ <<= 1;
        value_ <<= 1;
        if (++bits_needed_ == 0) {
            value_ += read_byte_();
            bits_needed_ = -8;
        }
    }
    return 0;
}

uint32_t Decoder::unary_max(uint32_t max_sym) {
    uint32_t s = 0;
    while (s < max_sym && bypass()) s++;
    return s;
}

uint32_t Decoder::rem_abs(uint32_t k, uint32_t cutoff, uint32_t max_log2) {
    const uint32_t max_prefix = 32 - max_log2;
    uint32_t prefix = 0, bit;
    do {
        prefix++;
        bit = (uint32_t)bypass();
    } while (bit && prefix < max_prefix);
    prefix -= 1 - bit;
    uint32_t len = k, offset;
    if (prefix < cutoff) {
        offset = prefix << k;
    } else {
        offset = ((1u << (prefix - cutoff)) + cutoff - 1) << k;
        len += prefix == max_prefix ? max_log2 - k : prefix - cutoff;
    }
    return offset + bypass_bins((int)len);
}

bool Decoder::finish_ok() const {
    const int shift = 8 + bits_needed_;
    return shift < 0 || (((uint32_t)last_byte_ << shift) & 0xFF) == 0x80;
}

namespace {

int decode_last(Decoder& d, Contexts& c, const Scan& s, bool luma) {
    const int maxx = kGroupIdx[std::min(s.w, 32) - 1], maxy = kGroupIdx[std::min(s.h, 32) - 1];
    int lw = 0, lh = 0;
    while ((1 << lw) < s.w) lw++;
    while ((1 << lh) < s.h) lh++;
    const int bx = luma ? LAST_X_LUMA : LAST_X_CHROMA, by = luma ? LAST_Y_LUMA : LAST_Y_CHROMA;
    const int ox = luma ? kPrefixCtx[lw] : 0, oy = luma ? kPrefixCtx[lh] : 0;
    const int sx = luma ? (lw + 1) >> 2 : std::min(s.w >> 3, 2);
    const int sy = luma ? (lh + 1) >> 2 : std::min(s.h >> 3, 2);
    int x = 0, y = 0;
    while (x < maxx && d.bin(c.m[bx + ox + (x >> sx)])) x++;
    while (y < maxy && d.bin(c.m[by + oy + (y >> sy)])) y++;
    if (x > 3) {
        const int cnt = ((x - 2) >> 1) - 1;
        x = kMinInGroup[x] + (int)d.bypass_bins(cnt + 1);
    }
    if (y > 3) {
        const int cnt = ((y - 2) >> 1) - 1;
        y = kMinInGroup[y] + (int)d.bypass_bins(cnt + 1);
    }
    if (x >= s.w || y >= s.h) return -1;
    for (int p = 0; p < s.n; p++)
        if (s.raster[p] == y * s.w + x) return p;
    return -1;
}

bool decode_group(Decoder& d, Contexts& c, const Scan& s, int32_t* lv, int min_sub, int last, bool luma, int table,
                  int& state, bool sign_hiding, Template& t, int& remaining) {
    const int max_sub = min_sub + (1 << s.log2_cg) - 1;
    const bool has_last = (last >> s.log2_cg) == (min_sub >> s.log2_cg);
    const int first_sig = has_last ? last : max_sub;
    int next = first_sig;
    const int infer = next != last ? ((min_sub >> s.log2_cg) != 0 ? min_sub : -1) : next;
    const int gt1_base = luma ? GTX1_LUMA : GTX1_CHROMA;
    const int par_base = luma ? PAR_FLAG_LUMA : PAR_FLAG_CHROMA;
    const int gt2_base = luma ? GTX2_LUMA : GTX2_CHROMA;
    int first_nz = next, last_nz = -1, nnz = 0;
    int sig_pos[16];
    for (; next >= min_sub && remaining >= 4; next--) {
        const int blk = s.raster[next];
        int sig = nnz == 0 && next == infer;
        if (!sig) {
            sig = d.bin(c.m[sig_ctx(s, lv, next, luma, state, t)]);
            remaining--;
        } else if (next != last) {
            sig_ctx(s, lv, next, luma, state, t);
        }
        if (sig) {
            const int o = level_ctx_offset(t, luma);
            sig_pos[nnz++] = blk;
            first_nz = next;
            last_nz = std::max(last_nz, next);
            const int gt1 = d.bin(c.m[gt1_base + o]);
            remaining--;
            int par = 0, gt2 = 0;
            if (gt1) {
                par = d.bin(c.m[par_base + o]);
                gt2 = d.bin(c.m[gt2_base + o]);
                remaining -= 2;
            }
            lv[blk] += 1 + par + gt1 + (gt2 << 1);
        }
        state = next_state(table, state, lv[blk]);
    }
    const int first_mode2 = next;
    for (int p = first_sig; p > first_mode2; p--) {
        const int blk = s.raster[p];
        if (lv[blk] >= 4)
            lv[blk] += (int32_t)(d.rem_abs(rice_param(s, lv, p, kRiceBaseLevel), kRemainBinReduction, kMaxLog2TrRange)
                                 << 1);
    }
    for (int p = first_mode2; p >= min_sub; p--) {
        const uint32_t k = rice_param(s, lv, p, 0);
        const int pos0 = (state < 2 ? 1 : 2) << k;
        const int rem = (int)d.rem_abs(k, kRemainBinReduction, kMaxLog2TrRange);
        const int level = rem == pos0 ? 0 : rem < pos0 ? rem + 1 : rem;
        state = next_state(table, state, level);
        if (level) {
            sig_pos[nnz++] = s.raster[p];
            first_nz = p;
            last_nz = std::max(last_nz, p);
            lv[s.raster[p]] = level;
        }
    }
    const bool hide = sign_hiding && last_nz - first_nz > 3 && nnz > 0;
    const int nsigns = hide ? nnz - 1 : nnz;
    uint32_t pattern = nsigns > 0 ? d.bypass_bins(nsigns) << (32 - nsigns) : 0;
    int64_t sum = 0;
    for (int i = 0; i < nnz; i++) {
        const int32_t a = lv[sig_pos[i]];
        if (a > 32767) return false;
        sum += a;
        if (i < nsigns) {
            if (pattern & 0x80000000u) lv[sig_pos[i]] = -a;
            pattern <<= 1;
        } else if (sum & 1) {
            lv[sig_pos[i]] = -a;
        }
    }
    return true;
}

}  // namespace

bool decode_residual(Decoder& d, Contexts& c, int w, int h, int32_t* lv, bool luma, bool dep_quant,
                     bool sign_hiding) {
    const Scan& s = scan_for(w, h);
    std::fill(lv, lv + (size_t)w * h, 0);
    const int last = decode_last(d, c, s, luma);
    if (last < 0) return false;
    const int gsize = 1 << s.log2_cg;
    uint8_t sig_raster[1024];
    std::fill(sig_raster, sig_raster + s.wg * s.hg, 0);
    const int table = dep_quant ? kDepQuantTransitions : 0;
    int state = 0, remaining = (s.n * 28) >> 4;
    Template t;
    const int sg_base = luma ? SIG_COEFF_GROUP_LUMA : SIG_COEFF_GROUP_CHROMA;
    for (int g = last >> s.log2_cg; g >= 0; g--) {
        const int cr = s.cg[g];
        int sig = g == (last >> s.log2_cg) || g == 0;
        if (!sig) {
            const int cx = cr % s.wg, cy = cr / s.wg;
            const bool right = cx + 1 < s.wg && sig_raster[cr + 1];
            const bool below = cy + 1 < s.hg && sig_raster[cr + s.wg];
            sig = d.bin(c.m[sg_base + (right || below)]);
        }
        if (!sig) continue;
        sig_raster[cr] = 1;
        if (!decode_group(d, c, s, lv, g * gsize, last, luma, table, state, sign_hiding && !dep_quant, t, remaining))
            return false;
    }
    return true;
}

int decode_split(Decoder& d, Contexts& c, const CanSplit& can, const SplitCtx& sc) {
    int is_split = can.any();
    if (can.no && is_split) is_split = d.bin(c.m[SPLIT_FLAG + sc.split]);
    if (!is_split) return 0;
    int is_qt = can.qt;
    if (is_qt && can.any_mtt()) is_qt = d.bin(c.m[SPLIT_QT_FLAG + sc.qt]);
    if (is_qt) return 1;
    const bool can_hor = can.bh || can.th;
    int is_ver = can.bv || can.tv;
    if (is_ver && can_hor) is_ver = d.bin(c.m[SPLIT_HV_FLAG + sc.hv]);
    const bool can14 = is_ver ? can.tv : can.th;
    int is12 = is_ver ? can.bv : can.bh;
    if (is12 && can14) is12 = d.bin(c.m[SPLIT_12_FLAG + (is_ver ? sc.ver_bt : sc.hor_bt)]);
    return is_ver ? (is12 ? 3 : 5) : (is12 ? 2 : 4);
}

}  // namespace cabac
}  // namespace rmh

extern "C" __attribute__((visibility("default"))) rmh_status rmh_cabac_residual_stream(
    const int32_t* levels, uint32_t width, uint32_t height, uint32_t n_blocks, uint32_t flags, uint32_t qp,
    uint8_t* stream, size_t stream_capacity, size_t* stream_len) {
    auto pow2 = [](uint32_t v) { return v >= 1 && v <= 64 && (v & (v - 1)) == 0; };
    if (!levels || !stream_len || !pow2(width) || !pow2(height) || n_blocks == 0 || qp > 63 ||
        (flags & ~7u) != 0 || (stream_capacity && !stream))
        return RMH_INVALID_INPUT;
    const size_t area = (size_t)width * height;
    for (uint32_t b = 0; b < n_blocks; b++) {
        bool any = false;
        for (size_t i = 0; i < area; i++) {
            const int32_t v = levels[b * area + i];
            if (v == 0) continue;
            if (i % width >= 32 || i / width >= 32 || v < -32768 || v > 32767) return RMH_INVALID_INPUT;
            any = true;
        }
        if (!any) return RMH_INVALID_INPUT;
    }
    std::vector<uint8_t> bytes;
    rmh::cabac::Encoder e(&bytes);
    rmh::cabac::Contexts c;
    c.init(0, (int)qp);
    for (uint32_t b = 0; b < n_blocks; b++) {
        rmh::cabac::encode_residual(e, c, (int)width, (int)height, levels + b * area, flags & RMH_CABAC_LUMA,
                                    flags & RMH_CABAC_DEP_QUANT, flags & RMH_CABAC_SIGN_HIDING);
        e.terminate(b == n_blocks - 1);
    }
    e.finish();
    *stream_len = bytes.size();
    if (bytes.size() > stream_capacity) return RMH_INVALID_INPUT;
    std::memcpy(stream, bytes.data(), bytes.size());
    return RMH_OK;
}

// rdsearch.cpp - rate-distortion search over the coding-tree split of each
// 32x32 luma node (encoder side, host code).
//
// The reference has no full RD search over splits: its only chooser is the
// quadrant-mean heuristic choose_qt_mtt_partition_rdo
// (crates/rustmpeg-video/src/vvc/slice.rs:4127-4167). This file writes the RD
// search from the reference's building blocks:
//   - the six splits and their children: CodingTreeSplit and
//     split_coding_tree_node (vvc/transform.rs:346-409), in that order
//     (0 none, 1 quad, 2 binary hor, 3 binary ver, 4 ternary hor, 5 ternary ver);
//   - each child is a leaf CU coded with PLANAR or DC chosen by SAD, PLANAR on
//     ties (choose_intra_mode, vvc/predict.rs:2706), DCT-II forward transform
//     (forward_transform_rect_bd, vvc/dct.rs:865-991), intra quantisation
//     (quantize_rect_bd, vvc/dct.rs:1113-1147), dequantisation and inverse
//     transform (vvc/dct.rs:1151-1187, 996-1109), clip to 8 bits;
//   - J = SSD + lambda * R, lambda = 0.57 * 2^((QP - 12) / 3) (the VTM intra
//     form) and R from a static bit model of the levels (below). Both are this
//     search's own choices: the reference has neither.
//
// Prediction is PLANAR / DC over rectangular blocks from the reconstructed
// neighbours (top row w + 1, left column h + 1 samples) with nearest-available
// substitution; there is no reference smoothing or PDPC in this search.
//
// Modes: full (all six splits), pruned (the splits whose predicted probability
// is >= p_prune, or all six when the largest is < p_fallback) and heuristic
// (the reference heuristic's one split).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/rmh.h"
#include "cabac.h"
#include "tables.h"

namespace {

constexpr int kNode = 32;

int lg2(int v) {
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

struct Dct2 {
    int m[6][32 * 32];  // log2 n = 2..5: m[k * n + j]
    Dct2() {
        for (int l = 2; l <= 5; l++) {
            const int n = 1 << l;
            bool ok = true;
            for (int k = 0; k < n; k++)
                for (int j = 0; j < n; j++) m[l][k * n + j] = rmh::transform_entry(RMH_TR_DCT2, n, k, j, &ok);
        }
    }
};

const Dct2& dct2() {
    static const Dct2 t;
    return t;
}

const int kQuantScale[2][6] = {{26214, 23302, 20560, 18396, 16384, 14564}, {18396, 16384, 14564, 13107, 11651, 10280}};
const int kInvScale[2][6] = {{40, 45, 51, 57, 64, 72}, {57, 64, 72, 80, 90, 102}};

int clip16(int64_t v) { return (int)(v < -32768 ? -32768 : v > 32767 ? 32767 : v); }

// forward_transform_rect_bd at bd = 8, DCT-II both ways, no zero-out (w, h <= 32).
void forward(int w, int h, const int* r, int* out) {
    const int lw = lg2(w), lh = lg2(h);
    const int* mh = dct2().m[lw];
    const int* mv = dct2().m[lh];
    const int s1 = lw + 8 + 6 - 15, s2 = lh + 6;
    const int64_t a1 = s1 > 0 ? (int64_t)1 << (s1 - 1) : 0, a2 = (int64_t)1 << (s2 - 1);
    int tmp[32 * 32];
    for (int y = 0; y < h; y++)
        for (int k = 0; k < w; k++) {
            int64_t s = 0;
            for (int n = 0; n < w; n++) s += (int64_t)mh[k * w + n] * r[y * w + n];
            tmp[k * h + y] = (int)(s1 >= 0 ? (s + a1) >> s1 : s << -s1);
        }
    for (int v = 0; v < h; v++)
        for (int k = 0; k < w; k++) {
            int64_t s = 0;
            for (int y = 0; y < h; y++) s += (int64_t)mv[v * h + y] * tmp[k * h + y];
            out[v * w + k] = clip16((s + a2) >> s2);
        }
}

// quantize_rect_bd, intra rounding offset, bd = 8.
void quantise(int w, int h, const int* c, int qp, int* lv) {
    const int lsum = lg2(w) + lg2(h), odd = lsum & 1;
    const int tr_shift = 15 - 8 - (lsum >> 1) - odd;
    const int qbits = 14 + qp / 6 + tr_shift;
    const int64_t add = (int64_t)171 << (qbits - 9 > 0 ? qbits - 9 : 0);
    const int64_t sc = kQuantScale[odd][qp % 6];
    for (int i = 0; i < w * h; i++) {
        const int64_t a = c[i] < 0 ? -(int64_t)c[i] : c[i];
        const int64_t l = (a * sc + add) >> qbits;
        lv[i] = clip16(c[i] < 0 ? -l : l);
    }
}

// dequantize_rect_bd then inverse_transform_rect_bd at bd = 8 (DCT-II).
void reconstruct_residual(int w, int h, const int* lv, int qp, int* res) {
    const int lw = lg2(w), lh = lg2(h), lsum = lw + lh, odd = lsum & 1;
    const int tr_shift = 15 - 8 - (lsum >> 1) - odd;
    const int rs = 6 - (tr_shift + qp / 6);
    const int64_t sc = kInvScale[odd][qp % 6];
    int c[32 * 32];
    for (int i = 0; i < w * h; i++) {
        const int64_t v = (int64_t)lv[i] * sc;
        c[i] = clip16(rs > 0 ? (v + ((int64_t)1 << (rs - 1))) >> rs : v << -rs);
    }
    const int* mh = dct2().m[lw];
    const int* mv = dct2().m[lh];
    int tmp[32 * 32];
    for (int i = 0; i < w; i++)
        for (int j = 0; j < h; j++) {
            int64_t s = 0;
            for (int k = 0; k < h; k++) s += (int64_t)mv[k * h + j] * c[k * w + i];
            tmp[i * h + j] = clip16((s + 64) >> 7);
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int64_t s = 0;
            for (int k = 0; k < w; k++) s += (int64_t)mh[k * w + x] * tmp[k * h + y];
            res[y * w + x] = clip16((s + 2048) >> 12);
        }
}

// Static bit model of a leaf: cbf, the last position, and per coded position
// 0.5 bit for a zero or 2 bits (significance, sign) plus an Exp-Golomb-like
// 2*floor(log2 |l|) + 1 for |l| > 1, in diagonal order (x + y, then y).
double leaf_bits(int w, int h, const int* lv) {
    double bits = 1.0;
    int last = -1;
    int order[32 * 32];
    int n = 0;
    for (int d = 0; d < w + h - 1; d++)
        for (int y = std::min(d, h - 1); y >= 0 && d - y < w; y--) order[n++] = y * w + (d - y);
    for (int i = n - 1; i >= 0; i--)
        if (lv[order[i]]) {
            last = i;
            break;
        }
    if (last < 0) return bits;
    bits += lg2(w) + lg2(h);
    for (int i = 0; i <= last; i++) {
        const int a = std::abs(lv[order[i]]);
        if (!a)
            bits += 0.5;
        else
            bits += 2.0 + (a > 1 ? 2.0 * std::floor(std::log2((double)a)) + 1.0 : 0.0);
    }
    return bits;
}

const double kSplitBits[6] = {1.0, 2.0, 4.0, 4.0, 4.0, 4.0};

// Children of a split of a w x h block (split_coding_tree_node,
// vvc/transform.rs:363-409), relative to the block: {x, y, w, h}.
int children(int split, int w, int h, int (*c)[4]) {
    auto set = [&](int k, int x, int y, int cw, int ch) {
        c[k][0] = x;
        c[k][1] = y;
        c[k][2] = cw;
        c[k][3] = ch;
    };
    switch (split) {
        case 0: set(0, 0, 0, w, h); return 1;
        case 1:
            for (int i = 0; i < 4; i++) set(i, (i & 1) * (w / 2), (i >> 1) * (h / 2), w / 2, h / 2);
            return 4;
        case 2: set(0, 0, 0, w, h / 2); set(1, 0, h / 2, w, h / 2); return 2;
        case 3: set(0, 0, 0, w / 2, h); set(1, w / 2, 0, w / 2, h); return 2;
        case 4: set(0, 0, 0, w, h / 4); set(1, 0, h / 4, w, h / 2); set(2, 0, 3 * h / 4, w, h / 4); return 3;
        default: set(0, 0, 0, w / 4, h); set(1, w / 4, 0, w / 2, h); set(2, 3 * w / 4, 0, w / 4, h); return 3;
    }
}

struct Frame {
    const uint8_t* src;
    int W, H, stride;
    uint8_t* rec;  // W x H, row pitch W
};

// One node's candidate: leaves coded in order into a local 32x32 recon.
struct Node {
    uint8_t rec[kNode * kNode];
    uint8_t done[kNode * kNode];
    uint8_t cu_w[64], cu_h[64], cu_qt[64];  // per 4x4 cell of the node: the CU covering it (0: not coded)
};

// Coding-tree limits of the luma stream: CTU = node = 32, QT down to 16, one
// MTT level of binary or ternary splits, CUs down to 4.
constexpr rmh::cabac::SplitLimits kLimits{8, 32, 32, 1, 4};

// CABAC state of the stream and the picture's CU map (per 4x4 cell) that the
// split contexts read across node boundaries.
struct Syntax {
    rmh::cabac::Encoder enc;
    rmh::cabac::Contexts ctx;
    std::vector<uint8_t> cu_w, cu_h, cu_qt;
    int cells_w = 0;
};

// Reference sample at picture position (x, y) for a leaf of the node at
// (bx, by): inside the node from the candidate's own recon when that sample
// is coded; outside from the picture recon when it precedes the node in
// raster order of 32x32 nodes. Returns -1 when unavailable.
int ref_sample(const Frame& f, const Node& nd, int bx, int by, int x, int y) {
    if (x < 0 || y < 0 || x >= f.W || y >= f.H) return -1;
    if (x >= bx && x < bx + kNode && y >= by && y < by + kNode) {
        const int i = (y - by) * kNode + (x - bx);
        return nd.done[i] ? nd.rec[i] : -1;
    }
    if (y < by) return f.rec[(size_t)y * f.W + x];
    if (y < by + kNode && x < bx) return f.rec[(size_t)y * f.W + x];
    return -1;
}

// PLANAR (0) / DC (1) prediction of a w x h block from its reference line in
// substitution order (left column bottom-up, h + 1 samples; the corner; the top
// row, w + 1 samples), -1 marking unavailable samples: each takes the previous
// available one, the first available one when none precedes it, or 128.
void planar_dc(int* line, int w, int h, int mode, int* pred) {
    const int nl = h + 1, nt = w + 1, tot = nl + 1 + nt;
    int first = -1;
    for (int i = 0; i < tot; i++)
        if (line[i] >= 0) {
            first = line[i];
            break;
        }
    if (first < 0) {
        for (int i = 0; i < tot; i++) line[i] = 128;
    } else {
        int prev = first;
        for (int i = 0; i < tot; i++) {
            if (line[i] < 0)
                line[i] = prev;
            else
                prev = line[i];
        }
    }
    int left[33], top[33];
    for (int j = 0; j <= h; j++) left[j] = line[nl - 1 - j];
    for (int i = 0; i <= w; i++) top[i] = line[nl + 1 + i];
    const int lw = lg2(w), lh = lg2(h);
    if (mode == 0) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                const int pv = ((h - 1 - y) * top[x] + (y + 1) * left[h]) << lw;
                const int ph = ((w - 1 - x) * left[y] + (x + 1) * top[w]) << lh;
                pred[y * w + x] = (pv + ph + w * h) >> (lw + lh + 1);
            }
    } else {
        int dc;
        if (w == h) {
            int s = 0;
            for (int i = 0; i < w; i++) s += top[i] + left[i];
            dc = (s + w) >> (lw + 1);
        } else if (w > h) {
            int s = 0;
            for (int i = 0; i < w; i++) s += top[i];
            dc = (s + (w >> 1)) >> lw;
        } else {
            int s = 0;
            for (int i = 0; i < h; i++) s += left[i];
            dc = (s + (h >> 1)) >> lh;
        }
        for (int i = 0; i < w * h; i++) pred[i] = dc;
    }
}


// PLANAR (0) / DC (1) prediction of the w x h leaf at picture (x0, y0).
void predict(const Frame& f, const Node& nd, int bx, int by, int x0, int y0, int w, int h, int mode, int* pred) {
    int line[2 * 33 + 1];
    const int nl = h + 1, nt = w + 1;
    for (int i = 0; i < nl; i++) line[i] = ref_sample(f, nd, bx, by, x0 - 1, y0 + nl - 1 - i);
    line[nl] = ref_sample(f, nd, bx, by, x0 - 1, y0 - 1);
    for (int i = 0; i < nt; i++) line[nl + 1 + i] = ref_sample(f, nd, bx, by, x0 + i, y0 - 1);
    planar_dc(line, w, h, mode, pred);
}

// CU covering picture sample (x, y) as seen from the node at (bx, by): inside
// the node from the candidate's own cells, outside from the picture map where
// that sample precedes the node in raster order of nodes. False when none.
bool cu_at(const Frame& f, const Syntax& sx, const Node& nd, int bx, int by, int x, int y, int* w, int* h, int* qt) {
    if (x < 0 || y < 0 || x >= f.W || y >= f.H) return false;
    if (x >= bx && x < bx + kNode && y >= by && y < by + kNode) {
        const int c = ((y - by) >> 2) * 8 + ((x - bx) >> 2);
        if (!nd.cu_w[c]) return false;
        *w = nd.cu_w[c];
        *h = nd.cu_h[c];
        *qt = nd.cu_qt[c];
        return true;
    }
    if (!(y < by || (y < by + kNode && x < bx))) return false;
    const size_t c = (size_t)(y >> 2) * sx.cells_w + (x >> 2);
    *w = sx.cu_w[c];
    *h = sx.cu_h[c];
    *qt = sx.cu_qt[c];
    return true;
}

// coding_tree split syntax of a node or leaf at (x, y), contexts from its left
// and above neighbours (derive_split_contexts, vvc/transform.rs:624-715).
void code_split(Syntax& sx, const Frame& f, const Node& nd, int bx, int by, int x, int y, int w, int h, int qt_depth,
                int mtt_depth, int last_split, int split) {
    namespace cb = rmh::cabac;
    const cb::CanSplit can = cb::can_split(x, y, w, h, f.W, f.H, mtt_depth, kLimits, last_split);
    int lw = 0, lh = 0, lq = 0, aw = 0, ah = 0, aq = 0;
    if (!cu_at(f, sx, nd, bx, by, x - 1, y, &lw, &lh, &lq)) lh = 0;
    if (!cu_at(f, sx, nd, bx, by, x, y - 1, &aw, &ah, &aq)) aw = 0;
    const cb::SplitCtx sc = cb::split_contexts(w, h, qt_depth, mtt_depth, lh, lq, aw, aq, can);
    cb::encode_split(sx.enc, sx.ctx, split, can, sc);
}

// The syntax of one leaf CU after its split flag: the luma mode (PLANAR is
// MPM 0; DC is MPM 1 whenever the neighbours are PLANAR or DC, which is every
// CU here), the luma cbf and residual_coding (vvc/slice.rs:1656-1687, 399).
void code_leaf(Syntax& sx, int mode, int w, int h, const int* lv) {
    namespace cb = rmh::cabac;
    sx.enc.bin(1, sx.ctx.m[cb::INTRA_LUMA_MPM_FLAG]);
    sx.enc.bin(mode == 1, sx.ctx.m[cb::INTRA_LUMA_PLANAR_FLAG + 1]);
    if (mode == 1) sx.enc.unary_max(0, 4);
    bool cbf = false;
    for (int i = 0; i < w * h && !cbf; i++) cbf = lv[i] != 0;
    sx.enc.bin(cbf, sx.ctx.m[cb::QT_CBF_Y]);
    if (cbf) cb::encode_residual(sx.enc, sx.ctx, w, h, lv, true, false, false);
}

// Codes the leaf CU at node-relative (lx, ly), w x h, into nd: PLANAR or DC
// by SAD, forward transform, quantisation, reconstruction. Its syntax (split
// flag 0, mode, cbf, residual) goes to sx; without sx the static model's bits
// are added to *bits.
void code_leaf_cu(const Frame& f, int bx, int by, int lx, int ly, int w, int h, int qt_depth, int mtt_depth,
                  int last_split, int qp, Node& nd, Syntax* sx, double* bits, uint64_t* sse) {
    int pred[2][32 * 32], res[32 * 32], coef[32 * 32], lv[32 * 32], src[32 * 32];
    const int x0 = bx + lx, y0 = by + ly;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) src[y * w + x] = f.src[(size_t)(y0 + y) * f.stride + x0 + x];
    int sad[2];
    for (int m = 0; m < 2; m++) {
        predict(f, nd, bx, by, x0, y0, w, h, m, pred[m]);
        int s = 0;
        for (int i = 0; i < w * h; i++) s += std::abs(src[i] - pred[m][i]);
        sad[m] = s;
    }
    const int m = sad[1] < sad[0] ? 1 : 0;
    for (int i = 0; i < w * h; i++) res[i] = src[i] - pred[m][i];
    forward(w, h, res, coef);
    quantise(w, h, coef, qp, lv);
    if (sx) {
        code_split(*sx, f, nd, bx, by, x0, y0, w, h, qt_depth, mtt_depth, last_split, 0);
        code_leaf(*sx, m, w, h, lv);
    } else {
        *bits += 1.0 + leaf_bits(w, h, lv);
    }
    for (int cy = ly >> 2; cy < (ly + h) >> 2; cy++)
        for (int cx = lx >> 2; cx < (lx + w) >> 2; cx++) {
            nd.cu_w[cy * 8 + cx] = (uint8_t)w;
            nd.cu_h[cy * 8 + cx] = (uint8_t)h;
            nd.cu_qt[cy * 8 + cx] = (uint8_t)qt_depth;
        }
    reconstruct_residual(w, h, lv, qp, res);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int v = pred[m][y * w + x] + res[y * w + x];
            v = v < 0 ? 0 : v > 255 ? 255 : v;
            const int d = v - src[y * w + x];
            *sse += (uint64_t)(d * d);
            nd.rec[(ly + y) * kNode + lx + x] = (uint8_t)v;
            nd.done[(ly + y) * kNode + lx + x] = 1;
        }
}

// A coded candidate held for adoption: its node state, stream state and the
// bytes it wrote.
struct Trial {
    Node nd;
    rmh::cabac::Encoder enc;
    rmh::cabac::Contexts ctx;
    std::vector<uint8_t> bytes;
    uint64_t sse = 0;
    double j = 0.0;
    int split = -1;
};

// Makes the stream state the trial's: its bytes follow what sx has written.
void adopt(Syntax& sx, const Trial& t) {
    std::vector<uint8_t>* out = sx.enc.sink();
    if (out) out->insert(out->end(), t.bytes.begin(), t.bytes.end());
    sx.enc = t.enc.with_sink(out);
    sx.ctx = t.ctx;
}

bool split_allowed(const rmh::cabac::CanSplit& c, int s) {
    const bool a[6] = {c.no, c.qt, c.bh, c.bv, c.th, c.tv};
    return a[s];
}

// Codes the block at node-relative (lx, ly), w x h, with `split`. Its children
// are leaves, except that with `search` the four children of a quad split
// each choose their own split (leaf children) by J on counting copies of the
// stream state: the search's second level.
void code_tree(const Frame& f, int bx, int by, int lx, int ly, int w, int h, int qt_depth, int mtt_depth,
               int last_split, int split, int qp, double lambda, Node& nd, Syntax* sx, bool search, double* bits,
               uint64_t* sse) {
    if (split == 0) {
        code_leaf_cu(f, bx, by, lx, ly, w, h, qt_depth, mtt_depth, last_split, qp, nd, sx, bits, sse);
        return;
    }
    if (sx) code_split(*sx, f, nd, bx, by, bx + lx, by + ly, w, h, qt_depth, mtt_depth, last_split, split);
    int ch[4][4];
    const int n = children(split, w, h, ch);
    const int cqt = qt_depth + (split == 1), cmtt = mtt_depth + (split >= 2);
    for (int c = 0; c < n; c++) {
        const int cx = lx + ch[c][0], cy = ly + ch[c][1], cw = ch[c][2], chh = ch[c][3];
        if (!(search && split == 1 && sx)) {
            code_leaf_cu(f, bx, by, cx, cy, cw, chh, cqt, cmtt, split, qp, nd, sx, bits, sse);
            continue;
        }
        const rmh::cabac::CanSplit can =
            rmh::cabac::can_split(bx + cx, by + cy, cw, chh, f.W, f.H, cmtt, kLimits, split);
        Trial best;
        for (int s = 0; s < 6; s++) {
            if (!split_allowed(can, s)) continue;
            Trial t;
            t.nd = nd;
            Syntax probe{sx->enc.with_sink(sx->enc.sink() ? &t.bytes : nullptr), sx->ctx, {}, {}, {}, sx->cells_w};
            std::swap(probe.cu_w, sx->cu_w);
            std::swap(probe.cu_h, sx->cu_h);
            std::swap(probe.cu_qt, sx->cu_qt);
            const uint64_t b0 = probe.enc.bits();
            double unused = 0.0;
            code_tree(f, bx, by, cx, cy, cw, chh, cqt, cmtt, split, s, qp, lambda, t.nd, &probe, false, &unused, &t.sse);
            t.j = (double)t.sse + lambda * (double)(probe.enc.bits() - b0);
            std::swap(probe.cu_w, sx->cu_w);
            std::swap(probe.cu_h, sx->cu_h);
            std::swap(probe.cu_qt, sx->cu_qt);
            if (best.split < 0 || t.j < best.j) {
                t.split = s;
                t.enc = probe.enc;
                t.ctx = probe.ctx;
                best = std::move(t);
            }
        }
        adopt(*sx, best);
        nd = best.nd;
        *sse += best.sse;
    }
}

// Codes one candidate split of the node at (bx, by); returns J and fills nd.
// With sx the rate is the CABAC bits the candidate adds to sx's stream (and
// sx advances) and the children of a quad split search their own splits;
// without it, the static bit model over leaf children.
double code_candidate(const Frame& f, int bx, int by, int split, int qp, double lambda, Node& nd, double* bits_out,
                      uint64_t* sse_out, Syntax* sx) {
    std::memset(nd.done, 0, sizeof nd.done);
    std::memset(nd.cu_w, 0, sizeof nd.cu_w);
    const uint64_t bits0 = sx ? sx->enc.bits() : 0;
    double bits = sx ? 0.0 : kSplitBits[split];
    uint64_t sse = 0;
    code_tree(f, bx, by, 0, 0, kNode, kNode, 0, 0, -1, split, qp, lambda, nd, sx, sx != nullptr, &bits, &sse);
    if (sx) bits = (double)(sx->enc.bits() - bits0);
    *bits_out = bits;
    *sse_out = sse;
    return (double)sse + lambda * bits;
}

// choose_qt_mtt_partition_rdo (vvc/slice.rs:4127-4167) for a 32x32 node.
int heuristic_split(const Frame& f, int bx, int by, int qp) {
    const int hw = kNode / 2;
    auto mean = [&](int qx, int qy) {
        int64_t s = 0;
        for (int y = qy; y < qy + hw; y++)
            for (int x = qx; x < qx + hw; x++) s += f.src[(size_t)y * f.stride + x];
        return s / (hw * hw);
    };
    const int64_t m00 = mean(bx, by), m10 = mean(bx + hw, by), m01 = mean(bx, by + hw), m11 = mean(bx + hw, by + hw);
    const int64_t hor = std::llabs(m00 - m01) + std::llabs(m10 - m11);
    const int64_t ver = std::llabs(m00 - m10) + std::llabs(m01 - m11);
    const int64_t thr = (qp > 4 ? qp : 4) / 2 + 8;
    if (hor > thr && ver > thr) return 1;
    if (hor > ver && hor > thr) return 2;
    if (ver > thr) return 3;
    return 0;
}

// Shared body of both entry points. With cabac set the rate is CABAC bits and
// the coded luma substream is appended to *stream_out (when not null).
rmh_status rd_encode(const uint8_t* src, uint32_t width, uint32_t height, uint32_t stride, uint32_t qp, uint32_t mode,
                     const float* probs, float p_prune, float p_fallback, uint8_t* recon, uint8_t* splits,
                     float* costs, rmh_rd_stats* stats, bool cabac, std::vector<uint8_t>* stream_out) {
    if (!src || !recon || !splits || !stats) return RMH_INVALID_INPUT;
    if (width == 0 || height == 0 || width % kNode || height % kNode || stride < width || width > 16384 ||
        height > 16384)
        return RMH_INVALID_INPUT;
    if (qp > 63 || mode > RMH_RD_HEURISTIC || (mode == RMH_RD_PRUNED && !probs)) return RMH_INVALID_INPUT;
    const int nbx = (int)width / kNode, nby = (int)height / kNode;
    Frame f{src, (int)width, (int)height, (int)stride, recon};
    const double lambda = 0.57 * std::pow(2.0, ((double)qp - 12.0) / 3.0);
    rmh_rd_stats st{};
    Node best, cand;
    std::unique_ptr<Syntax> sx;
    if (cabac) {
        sx.reset(new Syntax{rmh::cabac::Encoder(stream_out), {}, {}, {}, {}, (int)width / 4});
        sx->ctx.init(0, (int)qp);
        const size_t cells = (size_t)(width / 4) * (height / 4);
        sx->cu_w.assign(cells, 0);
        sx->cu_h.assign(cells, 0);
        sx->cu_qt.assign(cells, 0);
    }
    for (int by = 0; by < nby; by++)
        for (int bx = 0; bx < nbx; bx++) {
            const int b = by * nbx + bx;
            bool try_split[6] = {true, true, true, true, true, true};
            if (mode == RMH_RD_PRUNED) {
                const float* p = probs + 6 * (size_t)b;
                float mx = 0.0f;
                for (int s = 0; s < 6; s++) mx = std::max(mx, p[s]);
                if (mx >= p_fallback) {
                    for (int s = 0; s < 6; s++) try_split[s] = p[s] >= p_prune;
                } else {
                    st.fallbacks++;
                }
            } else if (mode == RMH_RD_HEURISTIC) {
                const int h = heuristic_split(f, bx * kNode, by * kNode, (int)qp);
                for (int s = 0; s < 6; s++) try_split[s] = s == h;
            }
            double best_j = 0.0, best_bits = 0.0;
            uint64_t best_sse = 0;
            int best_s = -1;
            Trial kept;  // CABAC: the best candidate's stream state and bytes
            for (int s = 0; s < 6; s++) {
                if (costs) costs[6 * (size_t)b + s] = -1.0f;
                if (!try_split[s]) continue;
                double bits;
                uint64_t sse;
                double j;
                Trial t;
                if (sx) {
                    // code the candidate on a copy of the stream state with its own bytes
                    Syntax probe{sx->enc.with_sink(&t.bytes), sx->ctx, {}, {}, {}, sx->cells_w};
                    std::swap(probe.cu_w, sx->cu_w);
                    std::swap(probe.cu_h, sx->cu_h);
                    std::swap(probe.cu_qt, sx->cu_qt);
                    j = code_candidate(f, bx * kNode, by * kNode, s, (int)qp, lambda, cand, &bits, &sse, &probe);
                    std::swap(probe.cu_w, sx->cu_w);
                    std::swap(probe.cu_h, sx->cu_h);
                    std::swap(probe.cu_qt, sx->cu_qt);
                    t.enc = probe.enc;
                    t.ctx = probe.ctx;
                } else {
                    j = code_candidate(f, bx * kNode, by * kNode, s, (int)qp, lambda, cand, &bits, &sse, nullptr);
                }
                st.candidates++;
                if (costs) costs[6 * (size_t)b + s] = (float)j;
                if (best_s < 0 || j < best_j) {
                    best_j = j;
                    best_s = s;
                    best_bits = bits;
                    best_sse = sse;
                    best = cand;
                    if (sx) kept = std::move(t);
                }
            }
            splits[b] = (uint8_t)best_s;
            st.sse += best_sse;
            if (sx) {
                adopt(*sx, kept);
                sx->enc.terminate(b == nbx * nby - 1);  // end_of_slice_segment_flag after each CTU
                for (int c = 0; c < 64; c++) {
                    const size_t pc = (size_t)(by * 8 + (c >> 3)) * sx->cells_w + bx * 8 + (c & 7);
                    sx->cu_w[pc] = best.cu_w[c];
                    sx->cu_h[pc] = best.cu_h[c];
                    sx->cu_qt[pc] = best.cu_qt[c];
                }
            } else {
                st.bits += best_bits;
            }
            for (int y = 0; y < kNode; y++)
                std::memcpy(recon + (size_t)(by * kNode + y) * width + bx * kNode, best.rec + y * kNode, kNode);
        }
    if (sx) {
        sx->enc.finish();
        st.bits = (double)sx->enc.bits();
    }
    st.blocks = (uint32_t)(nbx * nby);
    *stats = st;
    return RMH_OK;
}

}  // namespace

extern "C" __attribute__((visibility("default"))) rmh_status rmh_rd_encode_luma(
    const uint8_t* src, uint32_t width, uint32_t height, uint32_t stride, uint32_t qp, uint32_t mode,
    const float* probs, float p_prune, float p_fallback, uint8_t* recon, uint8_t* splits, float* costs,
    rmh_rd_stats* stats) {
    return rd_encode(src, width, height, stride, qp, mode, probs, p_prune, p_fallback, recon, splits, costs, stats,
                     false, nullptr);
}

extern "C" __attribute__((visibility("default"))) rmh_status rmh_rd_encode_luma_stream(
    const uint8_t* src, uint32_t width, uint32_t height, uint32_t stride, uint32_t qp, uint32_t mode,
    const float* probs, float p_prune, float p_fallback, uint8_t* recon, uint8_t* splits, float* costs,
    rmh_rd_stats* stats, uint8_t* stream, size_t stream_capacity, size_t* stream_len) {
    if (!stream_len || (stream_capacity && !stream)) return RMH_INVALID_INPUT;
    std::vector<uint8_t> bytes;
    const rmh_status s = rd_encode(src, width, height, stride, qp, mode, probs, p_prune, p_fallback, recon, splits,
                                   costs, stats, true, &bytes);
    if (s != RMH_OK) return s;
    *stream_len = bytes.size();
    if (bytes.size() > stream_capacity) return RMH_INVALID_INPUT;
    if (!bytes.empty()) std::memcpy(stream, bytes.data(), bytes.size());
    return RMH_OK;
}

namespace {

// A leaf CU of a parsed luma substream, in coding order; tu < 0 when cbf = 0.
struct ParsedCu {
    int x, y, w, h, mode, tu;
};

struct StreamParser {
    rmh::cabac::Decoder dec;
    rmh::cabac::Contexts ctx;
    int W, H, qp;
    std::vector<uint8_t> cu_w, cu_h, cu_qt;  // per 4x4 cell, 0: not parsed yet
    std::vector<ParsedCu> cus;
    std::vector<rmh_tu_desc> tus;
    std::vector<int16_t> coeff;

    void neighbour(int x, int y, int* w, int* h, int* qt) const {
        *w = *h = *qt = 0;
        if (x < 0 || y < 0 || x >= W || y >= H) return;
        const size_t c = (size_t)(y >> 2) * (W >> 2) + (x >> 2);
        *w = cu_w[c];
        *h = cu_h[c];
        *qt = cu_qt[c];
    }

    // coding_tree() down to the leaf CUs; false on a malformed stream.
    bool tree(int x, int y, int w, int h, int qt_depth, int mtt_depth, int last_split, int* split_out) {
        namespace cb = rmh::cabac;
        const cb::CanSplit can = cb::can_split(x, y, w, h, W, H, mtt_depth, kLimits, last_split);
        int lw, lh, lq, aw, ah, aq;
        neighbour(x - 1, y, &lw, &lh, &lq);
        neighbour(x, y - 1, &aw, &ah, &aq);
        const int split = cb::decode_split(dec, ctx, can, cb::split_contexts(w, h, qt_depth, mtt_depth, lh, lq, aw, aq, can));
        if (split_out) *split_out = split;
        if (split == 0) return leaf(x, y, w, h, qt_depth);
        int ch[4][4];
        const int n = children(split, w, h, ch);
        for (int c = 0; c < n; c++)
            if (!tree(x + ch[c][0], y + ch[c][1], ch[c][2], ch[c][3], qt_depth + (split == 1),
                      mtt_depth + (split >= 2), split, nullptr))
                return false;
        return true;
    }

    bool leaf(int x, int y, int w, int h, int qt_depth) {
        namespace cb = rmh::cabac;
        if (w > kNode || h > kNode) return false;
        if (!dec.bin(ctx.m[cb::INTRA_LUMA_MPM_FLAG])) return false;  // the subset codes MPM modes only
        int mode = 0;
        if (dec.bin(ctx.m[cb::INTRA_LUMA_PLANAR_FLAG + 1])) {
            if (dec.unary_max(4) != 0) return false;  // MPM 1 is DC beside PLANAR/DC neighbours
            mode = 1;
        }
        int tu = -1;
        if (dec.bin(ctx.m[cb::QT_CBF_Y])) {
            int32_t lv[32 * 32];
            if (!cb::decode_residual(dec, ctx, w, h, lv, true, false, false)) return false;
            rmh_tu_desc t{};
            t.offset = (uint32_t)coeff.size();
            t.log2_w = (uint8_t)lg2(w);
            t.log2_h = (uint8_t)lg2(h);
            t.qp = (uint8_t)qp;
            t.bit_depth = 8;
            t.tr_hor = t.tr_ver = RMH_TR_DCT2;
            for (int i = 0; i < w * h; i++) coeff.push_back((int16_t)lv[i]);
            tu = (int)tus.size();
            tus.push_back(t);
        }
        cus.push_back({x, y, w, h, mode, tu});
        for (int cy = y >> 2; cy < (y + h) >> 2; cy++)
            for (int cx = x >> 2; cx < (x + w) >> 2; cx++) {
                const size_t c = (size_t)cy * (W >> 2) + cx;
                cu_w[c] = (uint8_t)w;
                cu_h[c] = (uint8_t)h;
                cu_qt[c] = (uint8_t)qt_depth;
            }
        return true;
    }
};

}  // namespace

extern "C" __attribute__((visibility("default"))) rmh_status rmh_luma_stream_decode(
    rmh_ctx* ctx, const uint8_t* stream, size_t stream_len, uint32_t width, uint32_t height, uint32_t qp,
    uint8_t* recon, uint8_t* splits) {
    if (!ctx || !stream || !recon || !splits || width == 0 || height == 0 || width % kNode || height % kNode ||
        width > 16384 || height > 16384 || qp > 63)
        return RMH_INVALID_INPUT;
    const int W = (int)width, H = (int)height, nbx = W / kNode, nby = H / kNode;
    const size_t cells = (size_t)(W / 4) * (H / 4);
    StreamParser p{rmh::cabac::Decoder(stream, stream_len), {}, W, H, (int)qp, std::vector<uint8_t>(cells, 0),
                   std::vector<uint8_t>(cells, 0), std::vector<uint8_t>(cells, 0), {}, {}, {}};
    p.ctx.init(0, (int)qp);
    // pass 1, host: parse the whole picture (parsing needs no reconstruction)
    for (int b = 0; b < nbx * nby; b++) {
        int split = 0;
        if (!p.tree((b % nbx) * kNode, (b / nbx) * kNode, kNode, kNode, 0, 0, -1, &split)) return RMH_INVALID_INPUT;
        splits[b] = (uint8_t)split;
        if (p.dec.terminate() != (b == nbx * nby - 1)) return RMH_INVALID_INPUT;
    }
    if (!p.dec.finish_ok()) return RMH_INVALID_INPUT;
    // pass 2, device: every residual of the picture in one batch
    std::vector<int16_t> res(p.coeff.size());
    if (!p.tus.empty()) {
        const rmh_status s = rmh_residual_batch(ctx, p.tus.data(), p.tus.size(), p.coeff.data(), p.coeff.size(),
                                                res.data());
        if (s != RMH_OK) return s;
    }
    // pass 3, host: prediction and reconstruction in coding order
    std::vector<uint8_t> done((size_t)W * H, 0);
    int line[2 * 33 + 1], pred[32 * 32];
    for (const ParsedCu& cu : p.cus) {
        const int nl = cu.h + 1, nt = cu.w + 1;
        auto sample = [&](int x, int y) -> int {
            if (x < 0 || y < 0 || x >= W || y >= H || !done[(size_t)y * W + x]) return -1;
            return recon[(size_t)y * W + x];
        };
        for (int i = 0; i < nl; i++) line[i] = sample(cu.x - 1, cu.y + nl - 1 - i);
        line[nl] = sample(cu.x - 1, cu.y - 1);
        for (int i = 0; i < nt; i++) line[nl + 1 + i] = sample(cu.x + i, cu.y - 1);
        planar_dc(line, cu.w, cu.h, cu.mode, pred);
        const int16_t* r = cu.tu >= 0 ? res.data() + p.tus[cu.tu].offset : nullptr;
        for (int y = 0; y < cu.h; y++)
            for (int x = 0; x < cu.w; x++) {
                int v = pred[y * cu.w + x] + (r ? r[y * cu.w + x] : 0);
                v = v < 0 ? 0 : v > 255 ? 255 : v;
                const size_t i = (size_t)(cu.y + y) * W + cu.x + x;
                recon[i] = (uint8_t)v;
                done[i] = 1;
            }
    }
    return RMH_OK;
}

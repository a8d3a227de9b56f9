// intra_subset.cpp - caller-side driver of the intra-profile subset decode:
// one device batch for every residual of the picture, then PLANAR / DC
// prediction and reconstruction on the host in coding order.
//
// Prediction follows predict_block for PLANAR and DC
// (crates/rustmpeg-video/src/vvc/predict.rs:2619-2646): reference samples
// gathered and padded as gather_references (predict.rs:1084-1139), the
// [1 2 1] reference filter for luma PLANAR blocks larger than 32 samples
// (predict.rs:1233-1296), predict_planar / predict_dc (predict.rs:1688-1736)
// and PDPC (predict.rs:1739-1754); reconstruction as reconstruct_block
// (slice.rs:1050-1097).
#include <cstring>
#include <vector>

#include "../../include/rmh.h"

namespace {

struct Refs {
    int above[64];
    int left[64];
    int tl;
};

int lg2(int v) {
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

void gather(const uint8_t* p, int pw, int ph, int bx, int by, int n, Refs& r) {
    const int an = by == 0 ? 0 : (pw - bx < 2 * n ? pw - bx : 2 * n);
    const int ln = bx == 0 ? 0 : (ph - by < n ? ph - by : n);
    for (int i = 0; i < an; i++) r.above[i] = p[(by - 1) * pw + bx + i];
    for (int i = 0; i < ln; i++) r.left[i] = p[(by + i) * pw + bx - 1];
    r.tl = an && ln ? p[(by - 1) * pw + bx - 1] : an ? r.above[0] : ln ? r.left[0] : 128;
    if (!an) {
        for (int i = 0; i < 2 * n; i++) r.above[i] = ln ? r.left[0] : 128;
    } else {
        for (int i = an; i < 2 * n; i++) r.above[i] = r.above[an - 1];
    }
    if (!ln) {
        for (int i = 0; i < 2 * n; i++) r.left[i] = an ? r.above[0] : 128;
    } else {
        for (int i = ln; i < 2 * n; i++) r.left[i] = r.left[ln - 1];
    }
}

void smooth(const Refs& s, int n, Refs& o) {
    o.tl = (2 * s.tl + s.above[0] + s.left[0] + 2) >> 2;
    for (int j = 0; j + 1 < 2 * n; j++) {
        o.above[j] = ((j ? s.above[j - 1] : s.tl) + 2 * s.above[j] + s.above[j + 1] + 2) >> 2;
        o.left[j] = ((j ? s.left[j - 1] : s.tl) + 2 * s.left[j] + s.left[j + 1] + 2) >> 2;
    }
    o.above[2 * n - 1] = s.above[2 * n - 1];
    o.left[2 * n - 1] = s.left[2 * n - 1];
}

// PLANAR (mode 0) or DC (mode 1) prediction of an n x n block.
void predict(const uint8_t* p, int pw, int ph, int bx, int by, int n, bool luma, int mode, int* pred) {
    Refs raw, filt;
    gather(p, pw, ph, bx, by, n, raw);
    const Refs* r = &raw;
    if (mode == 0 && luma && n * n > 32) {
        smooth(raw, n, filt);
        r = &filt;
    }
    const int l = lg2(n);
    if (mode == 0) {
        int top[65], left[65], bottom[64], right[64];
        for (int i = 0; i <= n; i++) {
            top[i] = r->above[i];
            left[i] = r->left[i];
        }
        for (int k = 0; k < n; k++) {
            bottom[k] = left[n] - top[k];
            right[k] = top[n] - left[k];
        }
        for (int i = 0; i <= n; i++) {
            top[i] <<= l;
            left[i] <<= l;
        }
        for (int y = 0; y < n; y++) {
            int hor = left[y];
            for (int x = 0; x < n; x++) {
                hor += right[y];
                top[x] += bottom[x];
                pred[y * n + x] = ((hor << l) + (top[x] << l) + (1 << (2 * l))) >> (1 + 2 * l);
            }
        }
    } else {
        int sum = 0;
        for (int i = 0; i < n; i++) sum += r->above[i] + r->left[i];
        const int dc = (sum + n) >> (l + 1);
        for (int i = 0; i < n * n; i++) pred[i] = dc;
    }
    const int scale = (2 * l - 2) >> 2;
    for (int y = 0; y < n; y++) {
        const int sy = (2 * y) >> scale, wt = 32 >> (sy < 31 ? sy : 31);
        for (int x = 0; x < n; x++) {
            const int sx = (2 * x) >> scale, wl = 32 >> (sx < 31 ? sx : 31);
            const int v = pred[y * n + x];
            pred[y * n + x] = v + ((wl * (r->left[y] - v) + wt * (r->above[x] - v) + 32) >> 6);
        }
    }
}

}  // namespace

extern "C" __attribute__((visibility("default"))) rmh_status rmh_intra_subset_decode(
    rmh_ctx* ctx, uint32_t width, uint32_t height, const uint8_t* modes, const rmh_tu_desc* tus, size_t n_tu,
    const int16_t* coeff, size_t n_coeff, uint8_t* y, uint8_t* u, uint8_t* v) {
    if (!ctx) return RMH_INVALID_INPUT;
    if (!modes || !tus || !coeff || !y || !u || !v || width == 0 || height == 0 || width % 32 || height % 32 ||
        width > 16384 || height > 16384)
        return RMH_INVALID_INPUT;
    const size_t ctus = (size_t)(width / 32) * (height / 32);
    if (n_tu != 3 * ctus) return RMH_INVALID_INPUT;
    for (size_t i = 0; i < ctus; i++) {
        if (modes[i] > 1) return RMH_INVALID_INPUT;
        for (int c = 0; c < 3; c++) {
            const rmh_tu_desc& t = tus[3 * i + c];
            const int l = c ? 4 : 5;
            if (t.log2_w != l || t.log2_h != l || t.bit_depth != 8) return RMH_INVALID_INPUT;
        }
    }
    std::vector<int16_t> res(n_coeff);
    rmh_status s = rmh_residual_batch(ctx, tus, n_tu, coeff, n_coeff, res.data());
    if (s != RMH_OK) return s;
    const int w = (int)width, h = (int)height, cw = w / 2, ch = h / 2;
    std::vector<uint8_t> ry((size_t)w * h), ru((size_t)cw * ch), rv((size_t)cw * ch);
    int pred[1024];
    size_t k = 0;
    for (int cy0 = 0; cy0 < h; cy0 += 32)
        for (int cx0 = 0; cx0 < w; cx0 += 32, k++)
            for (int c = 0; c < 3; c++) {
                const int n = c ? 16 : 32, pw = c ? cw : w, ph = c ? ch : h;
                const int bx = c ? cx0 / 2 : cx0, by = c ? cy0 / 2 : cy0;
                uint8_t* p = c == 0 ? ry.data() : c == 1 ? ru.data() : rv.data();
                predict(p, pw, ph, bx, by, n, c == 0, modes[k], pred);
                const int16_t* r = res.data() + tus[3 * k + c].offset;
                for (int yy = 0; yy < n; yy++)
                    for (int xx = 0; xx < n; xx++) {
                        const int val = pred[yy * n + xx] + r[yy * n + xx];
                        p[(by + yy) * pw + bx + xx] = (uint8_t)(val < 0 ? 0 : val > 255 ? 255 : val);
                    }
            }
    memcpy(y, ry.data(), ry.size());
    memcpy(u, ru.data(), ru.size());
    memcpy(v, rv.data(), rv.size());
    return RMH_OK;
}

extern "C" __attribute__((visibility("default"))) rmh_status rmh_intra_predict(const uint8_t* plane, uint32_t width,
                                                                              uint32_t height, uint32_t bx,
                                                                              uint32_t by, uint32_t n, int is_luma,
                                                                              int mode, int32_t* pred) {
    if (!plane || !pred || (n != 4 && n != 8 && n != 16 && n != 32) || mode < 0 || mode > 1 || width > 16384 ||
        height > 16384 || bx + n > width || by + n > height)
        return RMH_INVALID_INPUT;
    predict(plane, (int)width, (int)height, (int)bx, (int)by, (int)n, is_luma != 0, mode, pred);
    return RMH_OK;
}

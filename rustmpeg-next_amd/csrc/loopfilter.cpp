// loopfilter.cpp - host-side tables of the LMCS and ALF plans: the LMCS
// lookup tables and the per-picture ALF filter banks the kernels read.
#include "loopfilter.h"

#include <algorithm>
#include <cstring>

#include "alf_data.h"

#define RMH_EXPORT extern "C" __attribute__((visibility("default")))

namespace rmh {

namespace {

constexpr int kLmcsPrec = 11;  // LMCS_FP_PREC, vvc/filter.rs:22

// The 7x7 diamond's 12 symmetric tap pairs as (dx, dy) with dy >= 0, in
// coefficient order (the partner of each is (-dx, -dy)).
constexpr int kTap[12][2] = {{0, 3}, {1, 2}, {0, 2}, {-1, 2}, {2, 1}, {1, 1},
                             {0, 1}, {-1, 1}, {-2, 1}, {3, 0}, {2, 0}, {1, 0}};

// Index of the tap pair holding (dx, dy) or (-dx, -dy).
int tap_index(int dx, int dy) {
    for (int j = 0; j < 12; j++)
        if ((kTap[j][0] == dx && kTap[j][1] == dy) || (kTap[j][0] == -dx && kTap[j][1] == -dy)) return j;
    return -1;
}

// The coefficient that tap k takes under transposition t (H.266 8.8.5.2):
// t = 1 swaps the axes, 2 mirrors x, 3 swaps then mirrors x. Filtering tap k
// with coefficient perm[t][k] equals filtering the transposed block.
struct Perm {
    int p[4][12];
    Perm() {
        for (int k = 0; k < 12; k++) {
            const int dx = kTap[k][0], dy = kTap[k][1];
            p[0][k] = k;
            p[1][k] = tap_index(dy, dx);
            p[2][k] = tap_index(-dx, dy);
            p[3][k] = tap_index(-dy, dx);
        }
    }
};

// Clip values of the 2-bit clip indices: 2^(B - {0, 3, 5, 7}) with B = 8 below
// 10-bit, 10 below 12-bit, else 12 (ALF_CLIP_VALS_{8,10,12}BIT,
// vvc/filter.rs:1509, 1872, 1876, chosen by bit depth as decode_general.rs:5428-5434 does).
int clip_value(int bd, int idx) {
    const int base = bd >= 12 ? 12 : bd >= 10 ? 10 : 8;
    static const int drop[4] = {0, 3, 5, 7};
    return 1 << (base - drop[idx & 3]);
}

}  // namespace

bool lmcs_lut(const rmh_lmcs_aps& aps, int bit_depth, uint16_t* fwd, uint16_t* inv, int32_t* chroma_adj,
              int32_t* pivot_out) {
    if (bit_depth < 8 || bit_depth > 12) return false;
    const int max_val = (1 << bit_depth) - 1;
    const int log2_cw = bit_depth - 4, org_cw = 1 << log2_cw;
    const int lo = std::min<int>(aps.min_bin_idx, 15);
    const int hi = std::max(std::min<int>(aps.max_bin_idx, 15), lo);
    // codewords per bin and the piecewise-linear pivots of both domains
    int cw[16] = {}, in_piv[17] = {}, out_piv[17] = {};
    for (int i = lo; i <= hi; i++) cw[i] = std::max(org_cw + aps.bin_cw_delta[i], 0);
    for (int i = 0; i < 16; i++) {
        out_piv[i + 1] = out_piv[i] + cw[i];
        in_piv[i + 1] = in_piv[i] + org_cw;
    }
    const int half = 1 << (kLmcsPrec - 1);
    for (int s = 0; s <= max_val; s++) {
        if (fwd) {
            const int i = std::min(s >> log2_cw, 15);
            const int scale = ((cw[i] << kLmcsPrec) + (1 << (log2_cw - 1))) >> log2_cw;
            fwd[s] = (uint16_t)std::clamp(out_piv[i] + ((scale * (s - in_piv[i]) + half) >> kLmcsPrec), 0, max_val);
        }
        if (inv) {
            // the reshaped-domain bin: first bin from lo whose upper pivot exceeds s
            int i = lo;
            while (i <= hi && s >= out_piv[i + 1]) i++;
            i = std::min(i, 15);
            const int scale = cw[i] ? (org_cw << kLmcsPrec) / cw[i] : 0;
            // 32-bit wrapping product, as the reference's i32 arithmetic in a
            // release build: it overflows only past the last pivot of an APS
            // whose codewords fall short of the sample range
            const int32_t t = (int32_t)((uint32_t)scale * (uint32_t)(s - out_piv[i]) + (uint32_t)half);
            inv[s] = (uint16_t)std::clamp(in_piv[i] + (t >> kLmcsPrec), 0, max_val);
        }
    }
    if (chroma_adj)
        for (int i = 0; i < 16; i++)
            chroma_adj[i] = cw[i] ? (org_cw << kLmcsPrec) / std::max(cw[i] + aps.chroma_res_scaling_offset, 1)
                                  : 1 << kLmcsPrec;
    if (pivot_out) memcpy(pivot_out, out_piv, sizeof out_piv);
    return true;
}

void alf_luma_bank(const rmh_alf_aps* aps, int bit_depth, int16_t* out) {
    static const Perm perm;
    int16_t coef[12];
    uint8_t clip[12];
    for (int bank = 0; bank < 17; bank++)
        for (int cls = 0; cls < 25; cls++) {
            if (bank < 16) {
                const int f = RMH_ALF_CLASS_MAP[bank][cls];
                for (int k = 0; k < 12; k++) {
                    coef[k] = RMH_ALF_FIXED[f][k];
                    clip[k] = 0;
                }
            } else if (aps && aps->n_luma) {
                const int f = std::min<int>(aps->class_to_filter[cls], aps->n_luma - 1);
                memcpy(coef, aps->luma_coeff[f], sizeof coef);
                memcpy(clip, aps->luma_clip[f], sizeof clip);
            } else {
                memset(coef, 0, sizeof coef);
                memset(clip, 0, sizeof clip);
            }
            for (int t = 0; t < 4; t++) {
                int16_t* e = out + ((bank * 25 + cls) * 4 + t) * 24;
                for (int k = 0; k < 12; k++) {
                    e[k] = coef[perm.p[t][k]];
                    e[12 + k] = (int16_t)clip_value(bit_depth, clip[perm.p[t][k]]);
                }
            }
        }
}

void alf_chroma_bank(const rmh_alf_aps& aps, int bit_depth, int16_t* out) {
    for (int a = 0; a < 8; a++)
        for (int k = 0; k < 6; k++) {
            out[a * 12 + k] = aps.chroma_coeff[a][k];
            out[a * 12 + 6 + k] = (int16_t)clip_value(bit_depth, aps.chroma_clip[a][k]);
        }
}

}  // namespace rmh

RMH_EXPORT rmh_status rmh_lmcs_tables(const rmh_lmcs_aps* aps, uint32_t bit_depth, uint16_t* fwd_lut,
                                      uint16_t* inv_lut, int32_t* chroma_adj_lut, int32_t* reshape_pivot) {
    if (!aps || aps->reserved) return RMH_INVALID_INPUT;
    return rmh::lmcs_lut(*aps, (int)bit_depth, fwd_lut, inv_lut, chroma_adj_lut, reshape_pivot) ? RMH_OK
                                                                                                 : RMH_INVALID_INPUT;
}

RMH_EXPORT rmh_status rmh_alf_filter_bank(const rmh_alf_aps* aps, uint32_t bit_depth, int16_t* out) {
    if (!out || bit_depth < 8 || bit_depth > 12) return RMH_INVALID_INPUT;
    rmh::alf_luma_bank(aps, (int)bit_depth, out);
    return RMH_OK;
}

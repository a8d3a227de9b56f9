// residual.hip - batched dequantisation + inverse LFNST + inverse DCT-II /
// DST-VII / DCT-VIII over every TU of a batch, for gfx950.
//
// Semantics follow the reference TU driver inverse_residual / residual_of
// (crates/rustmpeg-video/src/vvc/decode_general.rs:4100-4236) over
// dequantize_rect_bd / dequantize_dep_quant_bd / dequantize_ts
// (vvc/dct.rs:1151-1248, vvc/transform.rs:229-246), inverse_lfnst_rect
// (vvc/lfnst.rs:219-297) and inverse_transform_rect_bd (vvc/dct.rs:996-1109).
//
// Mapping: one 64-lane wavefront (one workgroup) per TU.
//   1. the coded region (after the default, LFNST and caller zero-outs) is read
//      once from HBM, dequantised and stored column-major in LDS as int16;
//   2. LFNST (if any) rewrites the top-left 4x4 / 8x8 region in LDS;
//   3. the bounding box of non-zero coefficients shrinks both passes (an exact
//      optimisation: zero coefficients contribute nothing);
//   4. the vertical pass writes int16 intermediates row-major to LDS, clipped
//      exactly like vvc/dct.rs:1094; the horizontal pass writes int16 residuals
//      to HBM. Both passes are int16 x int16 dot products with int32
//      accumulation (v_dot2_i32_i16): |coefficient| <= 2^15, |matrix| <= 91 and
//      at most 32 taps keep every sum below 2^27, so int32 is exact.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "gmem.h"
#include "kernels.h"

namespace rmh {
namespace {

typedef short short2v __attribute__((ext_vector_type(2)));

__constant__ int16_t kInvQuant[2][6] = {{40, 45, 51, 57, 64, 72}, {57, 64, 72, 80, 90, 102}};

__device__ __forceinline__ int clip16(int64_t v) { return (int)(v < -32768 ? -32768 : v > 32767 ? 32767 : v); }
__device__ __forceinline__ int clip16i(int v) { return v < -32768 ? -32768 : v > 32767 ? 32767 : v; }

__device__ __forceinline__ int dot2(uint32_t a, uint32_t b, int acc) {
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(short2v, a), __builtin_bit_cast(short2v, b), acc, false);
}

// The first product of a chain: the clamping form is encoded with acc as a
// separate operand, so no copy of the start value into the accumulator is
// needed. It changes nothing: |acc| < 2^12 here and each product is below
// 2^22 in magnitude.
__device__ __forceinline__ int dot2_first(uint32_t a, uint32_t b, int acc) {
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(short2v, a), __builtin_bit_cast(short2v, b), acc, true);
}

// Two int32 values saturated to int16 and packed (lo = a, hi = b).
__device__ __forceinline__ uint32_t pack_sat16(int a, int b) {
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pk_i16(a, b));
}

__device__ __forceinline__ int dequant_flat(int level, int scale, int rshift) {
    int64_t v = (int64_t)level * scale;
    if (rshift > 0) return clip16((v + ((int64_t)1 << (rshift - 1))) >> rshift);
    return clip16((int64_t)((uint64_t)v << (-rshift)));
}

__device__ __forceinline__ int dequant_ts(int level, int qp) {
    int per = qp / 6, rem = qp % 6;
    int rs = 6 - per;
    int64_t v = (int64_t)level * kInvQuant[0][rem];
    int64_t s = rs > 0 ? (v + ((int64_t)1 << (rs - 1))) >> rs : (int64_t)((uint64_t)v << (-rs));
    int n = (int)(uint32_t)(uint64_t)s;
    return clip16i(n);
}

// LDS row pitch in dwords for both the coefficient columns and the
// intermediate rows: 16 int16 pairs plus 4 dwords of padding, keeping every
// row 16-byte aligned for 128-bit reads.
constexpr int kPitch = 20;

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        int o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}

// acc + sum over pairs kk < 16 of dot2(src[kk], m[kk]); m[kk] is zero past the
// live pairs, so the 128-bit reads may run past them.
__device__ __forceinline__ int dot_pairs(const uint32_t* __restrict__ src, const uint32_t (&m)[16], int npairs) {
    const uint4* v4 = reinterpret_cast<const uint4*>(src);
    int acc = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (4 * q < npairs) {
            uint4 v = v4[q];
            acc = dot2(v.x, m[4 * q + 0], acc);
            acc = dot2(v.y, m[4 * q + 1], acc);
            acc = dot2(v.z, m[4 * q + 2], acc);
            acc = dot2(v.w, m[4 * q + 3], acc);
        }
    }
    return acc;
}


// Recon target of one TU: the address of sample (0, 0) of its block in its
// plane, the row pitch (samples), the sample size and the maximum value.
struct ReconDst {
    char* base;
    int stride;
    int bps;
    int maxv;
};

__device__ __forceinline__ ReconDst recon_dst(const rmh_tu_place* __restrict__ place,
                                              const rmh_recon_frame* __restrict__ frames, uint32_t tu_index) {
    const rmh_tu_place pl = place[tu_index];
    const rmh_plane p = frames[pl.frame].planes[pl.plane];
    ReconDst d;
    d.bps = p.bytes_per_sample;
    d.stride = (int)p.stride;
    d.maxv = (1 << p.bit_depth) - 1;
    d.base = (char*)p.data + ((size_t)pl.y * p.stride + pl.x) * d.bps;
    return d;
}

// pred + res clipped to [0, maxv], in place (reconstruct_tu add and clip).
__device__ __forceinline__ void recon_add(const ReconDst& d, int x, int y, int res) {
    if (d.bps == 1) {
        RMH_GLOBAL uint8_t* q = gdst<uint8_t>(d.base) + (size_t)y * d.stride + x;
        const int v = (int)*q + res;
        *q = (uint8_t)(v < 0 ? 0 : v > d.maxv ? d.maxv : v);
    } else {
        RMH_GLOBAL uint16_t* q = gdst<uint16_t>(d.base) + (size_t)y * d.stride + x;
        const int v = (int)*q + res;
        *q = (uint16_t)(v < 0 ? 0 : v > d.maxv ? d.maxv : v);
    }
}

// The prediction sample (x, y) of a recon target, and pred + res clipped to
// [0, maxv] stored back: the add of recon_add split so that a lane can fetch
// all its prediction samples before the first store.
__device__ __forceinline__ int recon_load(const ReconDst& d, int x, int y) {
    const size_t o = (size_t)y * d.stride + x;
    return d.bps == 1 ? (int)gsrc<uint8_t>(d.base)[o] : (int)gsrc<uint16_t>(d.base)[o];
}
__device__ __forceinline__ void recon_store(const ReconDst& d, int x, int y, int v) {
    const size_t o = (size_t)y * d.stride + x;
    v = v < 0 ? 0 : v > d.maxv ? d.maxv : v;
    if (d.bps == 1)
        gdst<uint8_t>(d.base)[o] = (uint8_t)v;
    else
        gdst<uint16_t>(d.base)[o] = (uint16_t)v;
}

// PREP: the prepass (a.prep_out set): stops after dequantisation and LFNST
// and writes the coded region for the tiled kernel. An instantiation of its
// own, so the full path keeps its register budget.
template <bool PREP>
__global__ __launch_bounds__(64) void residual_kernel(ResidualArgs a) {
    __shared__ __align__(16) uint32_t ct[32 * kPitch];  // coefficients: column i, int16 pairs along row k
    __shared__ __align__(16) uint32_t tt[64 * kPitch];  // intermediates: row y, int16 pairs along column i
    // Raw levels for dependent quantisation share the intermediates' space:
    // they are last read before the barrier that precedes the vertical pass.
    static_assert(32 * 32 * sizeof(int16_t) <= sizeof(tt), "levels fit in the intermediates");
    int16_t* lv = reinterpret_cast<int16_t*>(tt);

    const int lane = threadIdx.x;
    const uint32_t tu_index = a.order ? a.order[a.first_tu + blockIdx.x] : a.first_tu + blockIdx.x;
    const rmh_tu_desc d = a.tus[tu_index];
    const int lw = d.log2_w, lh = d.log2_h;
    const int w = 1 << lw, h = 1 << lh, n = w * h;
    const int bd = d.bit_depth;
    const int qp = d.qp < 4 ? 4 : d.qp;
    const int16_t* __restrict__ cin = a.coeff + d.offset;
    int16_t* __restrict__ rout = a.place ? nullptr : a.residual + d.offset;
    ReconDst dst{}, jdst{};
    // joint Cb-Cr (reconstruct_joint_cbcr, decode_general.rs:3908-3981): the
    // other chroma component's residual is derived from this one
    const int jmode = d.joint & 3, jneg = (d.joint >> 2) & 1, jshift = jmode == 3 ? 0 : 1;
    if (a.place) {
        dst = recon_dst(a.place, a.frames, tu_index);
        if (jmode) {
            const rmh_tu_place pl = a.place[tu_index];
            const rmh_plane p = a.frames[pl.frame].planes[3 - pl.plane];
            jdst.bps = p.bytes_per_sample;
            jdst.stride = (int)p.stride;
            jdst.maxv = (1 << p.bit_depth) - 1;
            jdst.base = (char*)p.data + ((size_t)pl.y * p.stride + pl.x) * jdst.bps;
        }
    }
    // writes residual e (raster index in the TU) to the residual buffer or adds it to the picture
    auto put = [&](int e, int v) {
        if (rout) {
            rout[e] = (int16_t)v;
        } else {
            recon_add(dst, e & (w - 1), e >> lw, v);
            if (jmode) recon_add(jdst, e & (w - 1), e >> lw, (jneg ? -v : v) >> jshift);
        }
    };

    if (d.flags & RMH_TU_TRANSFORM_SKIP) {
        for (int e = lane; e < n; e += 64) put(e, dequant_ts(cin[e], qp));
        return;
    }

    // Effective zero-out (vvc/dct.rs:1063-1077, decode_general.rs:4191-4236).
    const int th = d.tr_hor, tv = d.tr_ver;
    int sw = (th != RMH_TR_DCT2 && w == 32) ? 16 : (w > 32 ? w - 32 : 0);
    int sh = (tv != RMH_TR_DCT2 && h == 32) ? 16 : (h > 32 ? h - 32 : 0);
    sw = d.skip_w > sw ? d.skip_w : sw;
    sh = d.skip_h > sh ? d.skip_h : sh;
    const bool lfnst = d.lfnst_idx != 0;
    if (lfnst) {
        int nz = (w == 4 || h == 4) ? 4 : 8;
        int lsw = w - (nz < w ? nz : w), lsh = h - (nz < h ? nz : h);
        sw = lsw > sw ? lsw : sw;
        sh = lsh > sh ? lsh : sh;
    }
    const int cw = w - sw, ch = h - sh;  // coded columns / rows that feed the transform
    // Region held in LDS: the coded region, widened to 4x4 when LFNST reads it,
    // with rows padded to a multiple of 8 (zeros) for the 128-bit reads.
    const int rw = lfnst && cw < 4 ? 4 : cw;
    const int rh = lfnst && ch < 4 ? 4 : ch;
    const int rh_pad = (rh + 7) & ~7;
    const bool dq = (d.flags & RMH_TU_DEP_QUANT) != 0;

    // Dequantisation parameters (vvc/dct.rs:1164-1175, 1218-1234).
    const int lsum = lw + lh;
    const int sqrt2 = lsum & 1;
    const int tr_shift = 15 - bd - (lsum >> 1) - sqrt2;

    int16_t* ct16 = reinterpret_cast<int16_t*>(ct);
    int my_r = -1, my_c = -1;
    int lrw = 0;
    while ((1 << lrw) < rw) lrw++;
    if (!dq) {
        const int per = qp / 6, rem = qp % 6;
        const int rshift = 6 - (tr_shift + per);
        const int scale = kInvQuant[sqrt2][rem];
        const int total = rh_pad << lrw;
        for (int e = lane; e < total; e += 64) {
            const int y = e >> lrw, x = e & ((1 << lrw) - 1);
            if (x < rw) {
                int v = y < rh ? dequant_flat(cin[y * w + x], scale, rshift) : 0;
                ct16[x * 2 * kPitch + y] = (int16_t)v;
                if (v != 0 && x < cw && y < ch) {
                    my_r = y > my_r ? y : my_r;
                    my_c = x > my_c ? x : my_c;
                }
            }
        }
    } else {
        const int total = rh_pad << lrw;
        for (int e = lane; e < total; e += 64) {
            const int y = e >> lrw, x = e & ((1 << lrw) - 1);
            if (x < rw) ct16[x * 2 * kPitch + y] = 0;
        }
        const int lw_c = w < 32 ? w : 32, lh_c = h < 32 ? h : 32;
        for (int e = lane; e < lw_c * lh_c; e += 64) {
            int y = e / lw_c, x = e - y * lw_c;
            lv[e] = cin[y * w + x];
        }
        __syncthreads();
        // dequantize_dep_quant_bd (vvc/dct.rs:1191-1248): from the last
        // non-zero position down to 0 in scan order, a 4-state automaton
        // driven by the parity of each level (zeros included) sets the
        // reconstruction offset. The state sequence depends only on the
        // parities, and each step is one of two maps of {0..3}, so it is a
        // prefix of map compositions: every lane composes the maps of its
        // chunk of the sequence, a wave scan gives each lane its start
        // state, and the lanes then dequantise their chunks in parallel.
        const uint16_t* scan = a.scans + a.scan_off[lw][lh];
        const int len = a.scan_len[lw][lh];
        auto level_at = [&](int s) {
            const int r = scan[s];
            return (int)lv[(r >> lw) * lw_c + (r & (w - 1))];
        };
        int last = -1;
        for (int s = lane; s < len; s += 64)
            if (level_at(s) != 0) last = s;
        last = wave_max(last);
        const int L = last + 1, chunk = (L + 63) >> 6;
        const int t0 = lane * chunk, t1 = min(t0 + chunk, L);  // positions s = last - t, t in [t0, t1)
        // maps packed as 4 x 2 bits, entry x = image of state x; step maps
        // T0 = (0, 2, 1, 3) and T1 = (2, 0, 3, 1), from 32040 >> (4 x + 2 p)
        constexpr uint32_t kT0 = 0u | 2u << 2 | 1u << 4 | 3u << 6, kT1 = 2u | 0u << 2 | 3u << 4 | 1u << 6;
        auto apply = [](uint32_t g, uint32_t f) {  // g o f
            uint32_t h = 0;
#pragma unroll
            for (int x = 0; x < 4; x++) h |= ((g >> (2 * ((f >> (2 * x)) & 3))) & 3) << (2 * x);
            return h;
        };
        uint32_t f = 0xE4u;  // identity (0, 1, 2, 3)
        for (int t = t0; t < t1; t++) f = apply((level_at(last - t) & 1) ? kT1 : kT0, f);
        // inclusive scan of the chunk maps over the lanes (lane j: f_j o ... o f_0)
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)f, d);
            if (lane >= d) f = apply(f, o);
        }
        const uint32_t prev = (uint32_t)__shfl_up((int)f, 1);
        int state = lane ? (int)(prev & 3) : 0;  // lanes 0 .. lane - 1's maps applied to the start state 0
        const int qdq = qp + 1, per = qdq / 6, rem = qdq % 6;
        const int shift = 7 - per - tr_shift;
        int64_t inv = kInvQuant[sqrt2][rem];
        if (shift < 0) inv = (int64_t)((uint64_t)inv << (-shift));
        const int64_t add = shift > 0 ? ((int64_t)1 << (shift - 1)) : 0;
        const int rs = shift > 0 ? shift : 0;
        for (int t = t0; t < t1; t++) {
            const int r = scan[last - t];
            const int x = r & (w - 1), y = r >> lw;
            const int level = lv[y * lw_c + x];
            if (level != 0) {
                const int64_t q = 2 * (int64_t)level + (level > 0 ? -(state >> 1) : (state >> 1));
                if (x < rw && y < rh) ct16[x * 2 * kPitch + y] = (int16_t)clip16((q * inv + add) >> rs);
            }
            state = (32040 >> ((state << 2) + ((level & 1) << 1))) & 3;
        }
    }
    __syncthreads();

    if (lfnst) {
        // inverse_lfnst_rect (vvc/lfnst.rs:219-297): inputs are the first 16
        // positions of the 4x4 diagonal scan.
        const int sb = (w >= 8 && h >= 8) ? 8 : 4;
        const int tr_size = sb == 4 ? 16 : 48;
        const int zero_out = ((w == 4 && h == 4) || (w == 8 && h == 8)) ? 8 : 16;
        const int set = d.lfnst_mode <= 1 ? 0 : d.lfnst_mode <= 12 ? 1 : d.lfnst_mode <= 23 ? 2
                      : d.lfnst_mode <= 44 ? 3 : d.lfnst_mode <= 55 ? 2 : 1;
        const bool tp = (d.lfnst_mode >= 35 && d.lfnst_mode < 67) || d.lfnst_mode >= 81;
        const int8_t* m = sb == 8 ? a.lfnst8 + ((set * 2 + (d.lfnst_idx - 1)) * 16) * 48
                                  : a.lfnst4 + ((set * 2 + (d.lfnst_idx - 1)) * 16) * 16;
        const int mstride = sb == 8 ? 48 : 16;
        // 4x4 diagonal scan positions (x, y) packed as y*4+x
        const uint8_t scan4[16] = {0, 4, 1, 8, 5, 2, 12, 9, 6, 3, 13, 10, 7, 14, 11, 15};
        int src[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            int x = scan4[i] & 3, y = scan4[i] >> 2;
            src[i] = ct16[x * 2 * kPitch + y];
        }
        int outv = 0, ox = -1, oy = -1;
        if (lane < tr_size) {
            int sum = 0;
            for (int i = 0; i < zero_out; i++) sum += src[i] * (int)m[i * mstride + lane];
            outv = clip16i((sum + 64) >> 7);
            const int j = lane;
            if (tp) {
                if (sb == 4) {
                    oy = j & 3;
                    ox = j >> 2;
                } else if (j < 32) {
                    oy = j & 7;
                    ox = j >> 3;
                } else {
                    oy = (j - 32) & 3;
                    ox = 4 + ((j - 32) >> 2);
                }
            } else {
                if (sb == 4) {
                    oy = j >> 2;
                    ox = j & 3;
                } else if (j < 32) {
                    oy = j >> 3;
                    ox = j & 7;
                } else {
                    oy = 4 + ((j - 32) >> 2);
                    ox = (j - 32) & 3;
                }
            }
        }
        __syncthreads();
        if (lane < tr_size && ox < rw && oy < rh) ct16[ox * 2 * kPitch + oy] = (int16_t)outv;
        __syncthreads();
    }
    if (lfnst || dq) {
        // A caller zero-out narrower than the 4x4 LFNST input leaves values
        // outside the coded region in LDS; they must not reach the transform.
        // The bounding box is taken after LFNST / dependent dequantisation.
        const int total = rh_pad << lrw;
        for (int e = lane; e < total; e += 64) {
            const int y = e >> lrw, x = e & ((1 << lrw) - 1);
            if (x < rw) {
                if (x >= cw || y >= ch) {
                    ct16[x * 2 * kPitch + y] = 0;
                } else if (ct16[x * 2 * kPitch + y] != 0) {
                    my_r = y > my_r ? y : my_r;
                    my_c = x > my_c ? x : my_c;
                }
            }
        }
        __syncthreads();
    }
    if constexpr (PREP) {
        // prepass: the dequantised (and LFNST-ed) coded region, for the tiled kernel
        int16_t* po = a.prep_out + a.prep_off[blockIdx.x];
        for (int e = lane; e < cw * ch; e += 64) {
            const int y = e / cw, x = e - y * cw;
            po[y * w + x] = ct16[x * 2 * kPitch + y];
        }
        return;
    }
    const int er = __builtin_amdgcn_readfirstlane(wave_max(my_r)) + 1;  // effective rows
    const int ec = __builtin_amdgcn_readfirstlane(wave_max(my_c)) + 1;  // effective columns
    if (er <= 0) {
        for (int e = lane; e < n; e += 64) put(e, 0);
        return;
    }

    const int bdc = bd < 8 ? 8 : bd > 12 ? 12 : bd;
    if (w == 1 || h == 1) {
        // 1-D transform (vvc/dct.rs:1010-1057): one pass, shift 21 - bd.
        const int len = w == 1 ? h : w;
        const int ltype = w == 1 ? tv : th;
        const int lg = w == 1 ? lh : lw;
        const uint32_t* mp = a.mat + a.mat_off[ltype][lg];
        const int shift = 6 + 15 - bdc;
        const int taps = w == 1 ? er : ec;
        const int npairs = (taps + 1) >> 1;
        for (int j = lane; j < len; j += 64) {
            int acc = 0;
            for (int kk = 0; kk < npairs; kk++) {
                int c0, c1;
                if (w == 1) {
                    c0 = ct16[2 * kk];
                    c1 = ct16[2 * kk + 1];
                } else {
                    c0 = ct16[(2 * kk) * 2 * kPitch];
                    c1 = 2 * kk + 1 < ec ? ct16[(2 * kk + 1) * 2 * kPitch] : 0;
                }
                uint32_t pc = (uint32_t)(uint16_t)c0 | ((uint32_t)(uint16_t)c1 << 16);
                acc = dot2(pc, mp[kk * len + j], acc);
            }
            put(j, clip16i((acc + (1 << (shift - 1))) >> shift));
        }
        return;
    }

    // Stage 1 (vertical): tt[y][i] = clip16((sum_k Mv[k][y] * C[k][i] + 64) >> 7)
    // for i < ec, zero for the padding column. Lane y holds column y of Mv.
    {
        const int k1 = (er + 1) >> 1;
        const uint32_t* mv = a.mat + a.mat_off[tv][lh];
        const int y = lane & (h - 1), g = lane >> lh, G = 64 >> lh;
        uint32_t m[16];
#pragma unroll
        for (int kk = 0; kk < 16; kk++) m[kk] = kk < k1 ? mv[kk * h + y] : 0u;
        const int ecp = (ec + 1) & ~1;
        int16_t* tt16 = reinterpret_cast<int16_t*>(tt);
        for (int i = g; i < ecp; i += G) {
            int v = 0;
            if (i < ec) v = clip16i((dot_pairs(ct + i * kPitch, m, k1) + 64) >> 7);
            tt16[y * 2 * kPitch + i] = (int16_t)v;
        }
    }
    __syncthreads();

    // Stage 2 (horizontal): out[y][x] = clip16((sum_k Mh[k][x] * tt[y][k] + rnd) >> (20 - bd)).
    {
        const int k2 = (ec + 1) >> 1;
        const uint32_t* mh = a.mat + a.mat_off[th][lw];
        const int x = lane & (w - 1), g = lane >> lw, G = 64 >> lw;
        uint32_t m[16];
#pragma unroll
        for (int kk = 0; kk < 16; kk++) m[kk] = kk < k2 ? mh[kk * w + x] : 0u;
        const int shift = 6 + 15 - 1 - bdc;
        const int rnd = 1 << (shift - 1);
        for (int y = g; y < h; y += G) put(y * w + x, clip16i((dot_pairs(tt + y * kPitch, m, k2) + rnd) >> shift));
    }
}


// ---------------------------------------------------------------------------
// Tiled kernel: one kTileThreads-lane workgroup (one wave) per tile of
// same-size TUs, 16 samples per lane.
//
//   phase 0  per-TU parameters (offset, coded region, dequant scale/shift,
//            matrix offsets) into LDS; bounding box cleared
//   phase A  every coefficient of the tile's coded regions read with 8-byte
//            loads, dequantised, stored column-major in LDS; bounding box of
//            the non-zero levels
//   phase B  vertical pass: lane (TU, row y) holds column y of Mv in registers
//            and produces its row of intermediates for the live columns
//   phase C  horizontal pass: lane (TU, columns x, x + 1) holds those columns
//            of Mh and produces their residuals, stored to HBM (or added into
//            the picture under RECON, one column per lane)
// Many TUs per workgroup keep many loads in flight, which a one-TU workgroup
// cannot. Arithmetic and clipping are identical to residual_kernel.
// ---------------------------------------------------------------------------

struct TuParams {
    uint32_t off;    // first residual sample (offset + TU size <= n_coeff < 2^32)
    uint32_t src;    // first level: in a.coeff, or in a.prep when cwch bit 16 is set
    int32_t cwch;    // cw | ch << 8 | prepped << 16 | LFNST code (matrix | transposed << 3 | applies << 4) << 17 | dq << 22
    int32_t scale;
    int32_t shifts;  // dequant left shift | right shift << 8 | second-stage shift (20 - bd) << 16
    int32_t mh;      // dword offset of Mh pairs
    int32_t mv;      // dword offset of Mv pairs
    int32_t stride;  // recon: picture row pitch (samples)
    int32_t bpsmax;  // recon: bytes per sample | (2^bd - 1) << 8
    int32_t dq;      // dependent quantisation: S | rs << 24 | (left shift > 0) << 30 (scale and shifts are then 1, 0)
    char* base;      // recon: sample (0, 0) of the block
};
static_assert(sizeof(TuParams) == 48, "residual_tile_lds reserves 48 bytes per TU");

// dequantize_rect_bd for one TU of the tiled kernel: with ls = 0, S = scale;
// with 0 < ls <= 9, S = scale << ls (|level * S| <= 2^15 * 102 * 2^9 < 2^31,
// exact); with ls >= 10, S = 2^16, since every non-zero level saturates
// (|level * scale| << ls >= 40 * 2^10) and level * 2^16 saturates with the
// same sign. rs = 0 whenever ls > 0. S < 2^17 and |level| <= 2^15 fit the
// 24-bit multiplier.
__device__ __forceinline__ int dequant_tile(int level, int S, int rnd, int rs) {
    return min(max((__mul24(level, S & 0x1ffff) + rnd) >> rs, -32768), 32767);
}

// acc + the dot product of NQ 8-byte groups (2 pairs each) of src with
// m[0 .. 2 NQ); every read is issued before the first product.
template <int NQ>
__device__ __forceinline__ int dot_groups(const uint2* __restrict__ src, const uint32_t (&m)[16], int acc) {
    if constexpr (NQ == 0) return acc;
    uint2 v[NQ > 0 ? NQ : 1];
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = src[q];
    acc = dot2_first(v[0].x, m[0], acc);
    acc = dot2(v[0].y, m[1], acc);
#pragma unroll
    for (int q = 1; q < NQ; q++) {
        acc = dot2(v[q].x, m[2 * q], acc);
        acc = dot2(v[q].y, m[2 * q + 1], acc);
    }
    return acc;
}

// Stage-1 outputs of one lane: columns c0 .. c0 + nc - 1 (c0 and nc even)
// of one row, clip16((sum + 64) >> 7), stored as int16 pairs. src
// (coefficient columns) and dst (the intermediate row) are disjoint LDS
// regions, which lets reads of later columns be issued before earlier
// results are stored.
template <int NQ, int CSTRIDE>
__device__ __forceinline__ void stage1_row(const uint2* __restrict__ src, uint32_t* __restrict__ dst,
                                           const uint32_t (&m)[16], int c0, int nc) {
    for (int j = 0; j < nc; j += 2) {
        const int i = c0 + j;
        const int a0 = dot_groups<NQ>(src + i * CSTRIDE, m, 64) >> 7;
        const int a1 = dot_groups<NQ>(src + (i + 1) * CSTRIDE, m, 64) >> 7;
        dst[i >> 1] = pack_sat16(a0, a1);
    }
}

// The same for two matrix columns over one read of src: a0 and a1 start at
// acc.
template <int NQ>
__device__ __forceinline__ void dot_groups2(const uint2* __restrict__ src, const uint32_t (&m0)[16],
                                            const uint32_t (&m1)[16], int acc, int& a0, int& a1) {
    a0 = a1 = acc;
    if constexpr (NQ > 0) {
        uint2 v[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) v[q] = src[q];
        a0 = dot2_first(v[0].x, m0[0], acc);
        a1 = dot2_first(v[0].x, m1[0], acc);
        a0 = dot2(v[0].y, m0[1], a0);
        a1 = dot2(v[0].y, m1[1], a1);
#pragma unroll
        for (int q = 1; q < NQ; q++) {
            a0 = dot2(v[q].x, m0[2 * q], a0);
            a1 = dot2(v[q].x, m1[2 * q], a1);
            a0 = dot2(v[q].y, m0[2 * q + 1], a0);
            a1 = dot2(v[q].y, m1[2 * q + 1], a1);
        }
    }
}

// f(std::integral_constant<int, Q>) for the smallest Q in [Q0, QMAX] with
// nq <= Q (QMAX when none; nq never exceeds QMAX here): one wave-uniform branch selects a fully unrolled
// body instead of a branch per group inside it.
template <int Q0, int QMAX, typename F>
__device__ __forceinline__ void with_groups(int nq, F& f) {
    if constexpr (Q0 >= QMAX) {
        f(std::integral_constant<int, QMAX>{});
    } else {
        if (nq <= Q0)
            f(std::integral_constant<int, Q0>{});
        else
            with_groups<Q0 + 1, QMAX>(nq, f);
    }
}

__device__ __forceinline__ int wave_max_u(int v) { return __builtin_amdgcn_readfirstlane(wave_max(v)); }

// The first NP pairs (2, 4, 8 or 16) of one column of the padded table;
// the rest of m is left unset and never read.
template <int NP>
__device__ __forceinline__ void load_col(const uint32_t* __restrict__ src, uint32_t (&m)[16]) {
    if constexpr (NP <= 2) {
        const uint2 v = *reinterpret_cast<const uint2*>(src);
        m[0] = v.x;
        m[1] = v.y;
    } else {
        const uint4* s4 = reinterpret_cast<const uint4*>(src);
#pragma unroll
        for (int q = 0; q < NP / 4; q++) {
            const uint4 v = s4[q];
            m[4 * q + 0] = v.x;
            m[4 * q + 1] = v.y;
            m[4 * q + 2] = v.z;
            m[4 * q + 3] = v.w;
        }
    }
}

// load_col into m[0 .. NP).
template <int NP>
__device__ __forceinline__ void load_colp(const uint32_t* __restrict__ src, uint32_t* m) {
    if constexpr (NP <= 2) {
        const uint2 v = *reinterpret_cast<const uint2*>(src);
        m[0] = v.x;
        m[1] = v.y;
    } else {
        const uint4* s4 = reinterpret_cast<const uint4*>(src);
#pragma unroll
        for (int q = 0; q < NP / 4; q++) {
            const uint4 v = s4[q];
            m[4 * q + 0] = v.x;
            m[4 * q + 1] = v.y;
            m[4 * q + 2] = v.z;
            m[4 * q + 3] = v.w;
        }
    }
}

// Phases B and C of a plain tile of a uniform class (residual_uniform_class:
// sides of 8 and 16, and 32 x 32, 32 x 16, 16 x 32) whose TUs share one
// transform pair. (4-sample sides measured slower this way: one lane then
// holds a whole TU, and the lanes' LDS accesses fall on few banks.) The matrix column a step needs is then the same in every
// lane: it comes through the scalar cache into SGPRs (the constant address
// space), one v_dot2 operand, instead of every lane fetching it.
//   phase B  lane = coefficient column(s) (TU, x): the column's live rows in
//            registers, one intermediate per output row y, for every y (the
//            intermediates overwrite the coefficients: a barrier separates
//            the last column read from the first store)
//   phase C  lane = intermediate row(s) (TU, y): the row's live pairs in
//            registers, one residual per output column x, for every x; the
//            row goes out with 8-byte stores
// Arithmetic and clipping are those of the other passes.
typedef const __attribute__((address_space(4))) uint32_t* MatPtr;

// The even and the odd samples of a group of 4 (int16 pairs d0 = (v0, v1),
// d1 = (v2, v3)): (v0, v2) and (v1, v3), the operands of a butterfly.
__device__ __forceinline__ uint32_t even_pair(uint32_t d0, uint32_t d1) { return __builtin_amdgcn_perm(d1, d0, 0x05040100u); }
__device__ __forceinline__ uint32_t odd_pair(uint32_t d0, uint32_t d1) { return __builtin_amdgcn_perm(d1, d0, 0x07060302u); }

constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
// The largest power of two <= v (v >= 1): block sizes that divide the
// power-of-two loop ranges they step through.
constexpr int pow2_floor(int v) { return v >= 2 ? 2 * pow2_floor(v / 2) : 1; }

// acc + sum over the NQ groups of dot2(v[q], m[q]) (one pair per group).
template <int NQ>
__device__ __forceinline__ int dot_half(const uint32_t* v, const uint32_t* m, int acc) {
    acc = dot2_first(v[0], m[0], acc);
#pragma unroll
    for (int q = 1; q < NQ; q++) acc = dot2(v[q], m[q], acc);
    return acc;
}

template <int LW, int LH, int GMAX, int CP, int TP>
__device__ __forceinline__ void uniform_passes(const ResidualArgs& a, uint32_t* C, uint32_t* T, const TuParams* P,
                                               int G, int ec, int k1, int k2) {
    constexpr int W = 1 << LW, H = 1 << LH;
    constexpr int CPL = GMAX * W / kTileThreads;  // coefficient columns per lane: 1, 2 or 4
    constexpr int RPL = GMAX * H / kTileThreads;  // intermediate rows per lane: 1, 2 or 4
    static_assert(CPL >= 1 && RPL >= 1 && GMAX * W == CPL * kTileThreads && GMAX * H == RPL * kTileThreads,
                  "a tile's columns and rows spread evenly over the lanes");
    const int tid = threadIdx.x;
    const MatPtr mat = (MatPtr)a.col16;
    const int mv = __builtin_amdgcn_readfirstlane(P[0].mv), mh = __builtin_amdgcn_readfirstlane(P[0].mh);
    // DCT-II in a direction: its output pairs (n, N - 1 - n) come from one
    // even and one odd half-length sum (M[k][N - 1 - n] = (-1)^k M[k][n]),
    // which halves the products; the sums are exact, so the results equal
    // the full products'
    const bool bf_v = mv == a.col16_off[RMH_TR_DCT2][LH], bf_h = mh == a.col16_off[RMH_TR_DCT2][LW];
    const int eo_v = a.eo_off[LH], eo_h = a.eo_off[LW];

    // phase B: lane columns j = tid CPL + c (adjacent, so pairs pack)
    if (ec > 0) {
        auto vpass = [&](auto nq) {
            constexpr int NQ = decltype(nq)::value;
            uint2 cv[CPL][NQ];
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                const int j = tid * CPL + c, g = j >> LW, x = j & (W - 1);
                const uint2* src = reinterpret_cast<const uint2*>(C + (g * W + x) * CP);
#pragma unroll
                for (int q = 0; q < NQ; q++) cv[c][q] = src[q];
            }
            __syncthreads();
            const int j0 = tid * CPL, g = j0 >> LW, x0 = j0 & (W - 1);
            // one intermediate per lane column: clip16((sum + 64) >> 7), stored
            auto put = [&](int y, const int (&v)[CPL]) {
                uint32_t* row = T + (g * H + y) * TP;
                if constexpr (CPL == 1) {
                    reinterpret_cast<int16_t*>(row)[x0] = (int16_t)(pack_sat16(v[0], 0) & 0xffff);
                } else if constexpr (CPL == 2) {
                    row[x0 >> 1] = pack_sat16(v[0], v[1]);
                } else {
                    *reinterpret_cast<uint2*>(row + (x0 >> 1)) = make_uint2(pack_sat16(v[0], v[1]), pack_sat16(v[2], v[3]));
                }
            };
            if (bf_v) {
                uint32_t ev[CPL][NQ], od[CPL][NQ];
#pragma unroll
                for (int c = 0; c < CPL; c++)
#pragma unroll
                    for (int q = 0; q < NQ; q++) {
                        ev[c][q] = even_pair(cv[c][q].x, cv[c][q].y);
                        od[c][q] = odd_pair(cv[c][q].x, cv[c][q].y);
                    }
                // matrix columns in blocks of about 32 SGPRs: one wait per
                // block on the scalar loads instead of one per column
                constexpr int YB = cmin(H / 2, pow2_floor(16 / NQ));
                static_assert((H / 2) % YB == 0, "blocks tile the rows");
#pragma unroll
                for (int y0 = 0; y0 < H / 2; y0 += YB) {
                    uint32_t me[YB][NQ], mo[YB][NQ];
#pragma unroll
                    for (int b = 0; b < YB; b++)
#pragma unroll
                        for (int q = 0; q < NQ; q++) {
                            me[b][q] = mat[eo_v + (y0 + b) * 16 + q];
                            mo[b][q] = mat[eo_v + (y0 + b) * 16 + 8 + q];
                        }
#pragma unroll
                    for (int b = 0; b < YB; b++) {
                        int lo[CPL], hi[CPL];
#pragma unroll
                        for (int c = 0; c < CPL; c++) {
                            const int e = dot_half<NQ>(ev[c], me[b], 64), o = dot_half<NQ>(od[c], mo[b], 0);
                            lo[c] = (e + o) >> 7;
                            hi[c] = (e - o) >> 7;
                        }
                        put(y0 + b, lo);
                        put(H - 1 - y0 - b, hi);
                    }
                }
            } else {
                constexpr int YB = cmin(H, pow2_floor(16 / NQ));
                static_assert(H % YB == 0, "blocks tile the rows");
#pragma unroll
                for (int y0 = 0; y0 < H; y0 += YB) {
                    uint32_t m[YB][2 * NQ];
#pragma unroll
                    for (int b = 0; b < YB; b++)
#pragma unroll
                        for (int k = 0; k < 2 * NQ; k++) m[b][k] = mat[mv + (y0 + b) * 16 + k];
#pragma unroll
                    for (int b = 0; b < YB; b++) {
                        int v[CPL];
#pragma unroll
                        for (int c = 0; c < CPL; c++) {
                            int acc = dot2_first(cv[c][0].x, m[b][0], 64);
                            acc = dot2(cv[c][0].y, m[b][1], acc);
#pragma unroll
                            for (int q = 1; q < NQ; q++) {
                                acc = dot2(cv[c][q].x, m[b][2 * q], acc);
                                acc = dot2(cv[c][q].y, m[b][2 * q + 1], acc);
                            }
                            v[c] = acc >> 7;
                        }
                        put(y0 + b, v);
                    }
                }
            }
        };
        with_groups<1, (H / 4 > 0 ? H / 4 : 1)>((k1 + 1) >> 1, vpass);
    }
    __syncthreads();

    // phase C: lane rows r = tid RPL + c
    auto hpass = [&](auto nq) {
        constexpr int NQ = decltype(nq)::value;
        constexpr int NQ1 = NQ > 0 ? NQ : 1;
        uint2 tv[RPL][NQ1];
        int sh[RPL], rnd[RPL];
#pragma unroll
        for (int c = 0; c < RPL; c++) {
            const int r = tid * RPL + c, g = r >> LH, y = r & (H - 1);
            const uint2* src = reinterpret_cast<const uint2*>(T + (g * H + y) * TP);
#pragma unroll
            for (int q = 0; q < NQ; q++) tv[c][q] = src[q];
            sh[c] = P[g < G ? g : 0].shifts >> 16;
            rnd[c] = 1 << (sh[c] - 1);
        }
        uint32_t o[RPL][W / 2];
        if (NQ > 0 && bf_h) {
            uint32_t ev[RPL][NQ1], od[RPL][NQ1];
#pragma unroll
            for (int c = 0; c < RPL; c++)
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    ev[c][q] = even_pair(tv[c][q].x, tv[c][q].y);
                    od[c][q] = odd_pair(tv[c][q].x, tv[c][q].y);
                }
            // columns x, x + 1 and their mirrors W - 1 - x, W - 2 - x
            constexpr int XB = cmin(W / 2, cmax(2, pow2_floor(16 / NQ1)));  // columns per block of scalar loads (even)
            static_assert((W / 2) % XB == 0 && XB % 2 == 0, "blocks tile the half row in pairs");
#pragma unroll
            for (int x0 = 0; x0 < W / 2; x0 += XB) {
                uint32_t me[XB][NQ1], mo[XB][NQ1];
#pragma unroll
                for (int b = 0; b < XB; b++)
#pragma unroll
                    for (int q = 0; q < NQ; q++) {
                        me[b][q] = mat[eo_h + (x0 + b) * 16 + q];
                        mo[b][q] = mat[eo_h + (x0 + b) * 16 + 8 + q];
                    }
#pragma unroll
                for (int b = 0; b < XB; b += 2) {
                    const int x = x0 + b;
#pragma unroll
                    for (int c = 0; c < RPL; c++) {
                        const int e0 = dot_half<NQ1>(ev[c], me[b], rnd[c]), o0 = dot_half<NQ1>(od[c], mo[b], 0);
                        const int e1 = dot_half<NQ1>(ev[c], me[b + 1], rnd[c]), o1 = dot_half<NQ1>(od[c], mo[b + 1], 0);
                        o[c][x >> 1] = pack_sat16((e0 + o0) >> sh[c], (e1 + o1) >> sh[c]);
                        o[c][(W - 2 - x) >> 1] = pack_sat16((e1 - o1) >> sh[c], (e0 - o0) >> sh[c]);
                    }
                }
            }
        } else {
            constexpr int XB = cmin(W, cmax(2, pow2_floor(16 / NQ1)));  // columns per block of scalar loads (even)
            static_assert(W % XB == 0 && XB % 2 == 0, "blocks tile the row in pairs");
#pragma unroll
            for (int x0 = 0; x0 < W; x0 += XB) {
              uint32_t mb[XB][2 * NQ1];
#pragma unroll
              for (int b = 0; b < XB; b++)
#pragma unroll
                  for (int k = 0; k < 2 * NQ; k++) mb[b][k] = mat[mh + (x0 + b) * 16 + k];
#pragma unroll
              for (int b = 0; b < XB; b += 2) {
                const int x = x0 + b;
                const uint32_t* m0 = mb[b];
                const uint32_t* m1 = mb[b + 1];
#pragma unroll
                for (int c = 0; c < RPL; c++) {
                    int a0 = rnd[c], a1 = rnd[c];
                    if constexpr (NQ > 0) {
                        a0 = dot2_first(tv[c][0].x, m0[0], rnd[c]);
                        a1 = dot2_first(tv[c][0].x, m1[0], rnd[c]);
                        a0 = dot2(tv[c][0].y, m0[1], a0);
                        a1 = dot2(tv[c][0].y, m1[1], a1);
#pragma unroll
                        for (int q = 1; q < NQ; q++) {
                            a0 = dot2(tv[c][q].x, m0[2 * q], a0);
                            a1 = dot2(tv[c][q].x, m1[2 * q], a1);
                            a0 = dot2(tv[c][q].y, m0[2 * q + 1], a0);
                            a1 = dot2(tv[c][q].y, m1[2 * q + 1], a1);
                        }
                    }
                    o[c][x >> 1] = pack_sat16(a0 >> sh[c], a1 >> sh[c]);
                }
              }
            }
        }
        // each row is W contiguous int16 residuals; TU offsets are multiples
        // of 4 samples, so every row start is 8-byte aligned
#pragma unroll
        for (int c = 0; c < RPL; c++) {
            const int r = tid * RPL + c, g = r >> LH, y = r & (H - 1);
            if (g < G) {
                RMH_GLOBAL u32x2* dst = gdst<u32x2>(a.residual + P[g].off + (uint32_t)(y * W));
#pragma unroll
                for (int k = 0; k < W / 4; k++) dst[k] = u32x2{o[c][2 * k], o[c][2 * k + 1]};
            }
        }
    };
    with_groups<0, (W / 4 > 0 ? W / 4 : 1)>((k2 + 1) >> 1, hpass);
}

// Rows per prediction-fetch batch of a recon lane (phase C): 8 keeps the
// kernel at 6 waves per SIMD, where 16 would need 128 VGPRs.
constexpr int kReconBatch = 8;

template <int LW, int LH, bool RECON, bool LFNST>
__device__ __forceinline__ void tile_body(const ResidualArgs& a, const TileItem& it, uint32_t* smem) {
    constexpr int W = 1 << LW, H = 1 << LH;
    constexpr int LWC = LW < 5 ? LW : 5, LHC = LH < 5 ? LH : 5;
    constexpr int Wc = 1 << LWC, Hc = 1 << LHC;
    // dword pitch of a coefficient column / an intermediate row: the live
    // pairs, padded to 2 mod 4 dwords so that the 8 column quads of a phase A
    // store, or 32 consecutive rows of a phase B store, fall in distinct
    // banks; even, for 8-byte reads
    constexpr int CP = residual_tile_pitch(LHC), TP = residual_tile_pitch(LWC);
    // Plain tiles of the uniform classes take the uniform-matrix passes
    // (uniform_passes): every TU of such a tile has the same transform pair.
    constexpr bool UNI = !RECON && !LFNST && residual_uniform_class(LW, LH);
    constexpr int GMAX = residual_tile_tus(LW, LH, UNI);
    const int G = it.count;
    // LDS: the coefficient columns, the intermediate rows (over the
    // coefficients in a uniform tile, whose lanes hold their columns in
    // registers before the first intermediate is stored), the per-TU
    // parameters, the tile bounding box; residual_tile_lds has the size.
    constexpr int CW = GMAX * Wc * CP, TW = GMAX * H * TP;
    uint32_t* C = smem;
    uint32_t* T = UNI ? C : C + CW;
    TuParams* P = reinterpret_cast<TuParams*>(C + (UNI ? (CW > TW ? CW : TW) : CW + TW));
    int* TB = reinterpret_cast<int*>(P + GMAX);  // tile bounding box: max row, max column of non-zeros
    const int tid = threadIdx.x;

    if (tid < 2) TB[tid] = -1;
    bool dq_lane = false;
    for (int g = tid; g < G; g += kTileThreads) {
        const TileTu d = a.tile_tus[it.first + g];
        TuParams p;
        p.off = d.off;
        p.src = d.src;
        p.cwch = d.cw | (d.ch << 8) | (d.prepped << 16) | (d.lfnst << 17) | (d.dq << 22);
        // dequantisation as level * S + rnd >> rs (see dequant_tile)
        const int ls = d.shifts & 15;
        p.scale = ls >= 10 ? 65536 : (int)d.scale << ls;
        p.shifts = (ls ? 0 : (d.shifts >> 4)) | (d.shift2 << 16);
        p.dq = 0;
        if (d.dq) {
            // phase A keeps the raw levels; the dependent-quantisation pass
            // dequantises them with these parameters
            p.dq = p.scale | (((p.shifts & 255) | (ls > 0 ? 64 : 0)) << 24);
            p.scale = 1;
            p.shifts = d.shift2 << 16;
            dq_lane = true;
        }
        p.mh = d.mh;
        p.mv = d.mv;
        p.stride = 0;
        p.bpsmax = 0;
        p.base = nullptr;
        if (RECON) {
            const TileDst t = a.tile_dst[it.first + g];
            p.stride = t.stride;
            p.base = t.base;
            p.bpsmax = t.bps | (t.maxv << 8);
        }
        P[g] = p;
    }
    // GMAX <= kTileThreads, so one lane per TU: a ballot covers the tile.
    // Only tool tiles (the LFNST instantiation) hold such TUs.
    static_assert(GMAX <= kTileThreads, "one phase-0 lane per TU");
    const bool any_dq = LFNST && __builtin_amdgcn_ballot_w64(dq_lane) != 0;
    __syncthreads();

    // phase B's lane items: (TU, row y, column chunk), NCH_B chunks, in
    // RB rounds of kTileThreads; its matrix columns (NPB pairs each) of every round
    // are fetched with the coefficients when they fit 16 registers
    constexpr int NCH_B = (kTileThreads / (GMAX * H)) > 0 ? kTileThreads / (GMAX * H) : 1;
    constexpr int RB = (GMAX * H * NCH_B + kTileThreads - 1) / kTileThreads;
    constexpr int NPB = Hc / 2;  // matrix pairs per column
    constexpr bool B_PRE = !UNI && RB * NPB <= 16;
    uint32_t mB[16];

    // coefficient groups (4 x 4) per TU of the dependent-quantisation pass:
    // every lane owns at most one, since a tile's coded regions hold at most
    // 16 kTileThreads samples
    constexpr int NCG = (Wc / 4) * (Hc / 4);
    static_assert(!LFNST || GMAX * NCG <= kTileThreads, "one coefficient group per lane (tool tiles)");
    int cg_org = 0;

    // phase A: a lane takes 2 rows x 4 columns and stores 4 row pairs; the
    // whole Hc x Wc region is written (zeros outside the coded region), so
    // rows and columns past the tile bounding box read as zero later.
    {
        // a fixed number of items per lane (at most 2); every load of the
        // lane is issued before the first dequantisation
        constexpr int Q = Wc / 4, RP = Hc / 2, NA = (GMAX * RP * Q + kTileThreads - 1) / kTileThreads;
        const int total = G * RP * Q;
        int mr = -1, mc = -1;
        uint64_t raw[NA][2];
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int e = tid + kTileThreads * i;
            raw[i][0] = raw[i][1] = 0;
            if (e < total) {
                const int x0 = (e % Q) * 4, rp = (e / Q) % RP, g = e / (Q * RP);
                const uint32_t off = P[g].src;
                const int cwch = P[g].cwch, cw = cwch & 255, ch = (cwch >> 8) & 255, r0 = 2 * rp;
                const int16_t* __restrict__ src = ((cwch >> 16) & 1) ? a.prep : a.coeff;
                if (x0 < cw) {
                    if (r0 < ch) raw[i][0] = *reinterpret_cast<const uint64_t*>(src + (off + (uint32_t)(r0 * W + x0)));
                    if (r0 + 1 < ch)
                        raw[i][1] = *reinterpret_cast<const uint64_t*>(src + (off + (uint32_t)((r0 + 1) * W + x0)));
                }
            }
        }
        // the dependent-quantisation pass's coefficient group (one per lane,
        // in scan order within each TU): its top-left raster position, from
        // the scan (whose first entry of every 16 is a group's origin)
        if (any_dq) cg_org = a.scans[a.scan_off[LW][LH] + 16 * (tid & (NCG - 1))];
        // phase B's matrix columns: loaded with the coefficients instead of
        // after the barrier
        if constexpr (B_PRE) {
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                const int qb = rb * kTileThreads + tid;
                const int gb = qb / (H * NCH_B);
                if (gb < G) load_colp<NPB>(a.col16 + P[gb].mv + (qb & (H - 1)) * 16, mB + rb * NPB);
            }
        }
        // S fits int16 in every TU of the wave's items (every QP short of the
        // top corner of the range): each level then dequantises with one dot2
        // against (S, 0) or (0, S), and the saturating pack is the int16 clip
        bool wide = false;
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int e = tid + kTileThreads * i;
            if (e < total && P[e / (Q * RP)].scale > 32767) wide = true;
        }
        const bool fast = __builtin_amdgcn_ballot_w64(wide) == 0;
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int e = tid + kTileThreads * i;
            if (e >= total) continue;
            const int x0 = (e % Q) * 4, rp = (e / Q) % RP, g = e / (Q * RP);
            const int cwch = P[g].cwch, S = P[g].scale, rs = P[g].shifts & 31;
            const int cw = cwch & 255, rnd = (1 << rs) >> 1;
            const int r0 = 2 * rp;
            uint64_t raw0 = raw[i][0], raw1 = raw[i][1];
            // columns past the coded region read as zero levels
            const int nv = cw - x0;
            if (nv < 4) {
                const uint64_t keep = nv > 0 ? (1ull << (16 * nv)) - 1 : 0;
                raw0 &= keep;
                raw1 &= keep;
            }
            // bounding box of the non-zero levels (a zero level dequantises
            // to zero) in whole row pairs and column quads, the granularity at
            // which the passes below use it
            if (raw0 | raw1) {
                mr = max(mr, r0 + 1);
                mc = max(mc, x0 + 3);
            }
            const uint32_t w0[2] = {(uint32_t)raw0, (uint32_t)(raw0 >> 32)};
            const uint32_t w1[2] = {(uint32_t)raw1, (uint32_t)(raw1 >> 32)};
            uint32_t* dst = C + (g * Wc + x0) * CP + rp;
            if (fast) {
                const uint32_t mlo = (uint32_t)S, mhi = (uint32_t)S << 16;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t mk = (k & 1) ? mhi : mlo;
                    const int a0 = dot2_first(w0[k >> 1], mk, rnd) >> rs;
                    const int a1 = dot2_first(w1[k >> 1], mk, rnd) >> rs;
                    dst[k * CP] = pack_sat16(a0, a1);
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t u0 = w0[k >> 1], u1 = w1[k >> 1];
                    const int l0 = (k & 1) ? (int)u0 >> 16 : (int)(int16_t)(u0 & 0xffff);
                    const int l1 = (k & 1) ? (int)u1 >> 16 : (int)(int16_t)(u1 & 0xffff);
                    const int a0 = dequant_tile(l0, S, rnd, rs), a1 = dequant_tile(l1, S, rnd, rs);
                    dst[k * CP] = __builtin_amdgcn_perm((uint32_t)a1, (uint32_t)a0, 0x05040100u);
                }
            }
        }
        mr = wave_max(mr);
        mc = wave_max(mc);
        if ((tid & 63) == 0) {
            atomicMax(&TB[0], mr);
            atomicMax(&TB[1], mc);
        }
    }
    __syncthreads();
    if (LFNST && any_dq) {
        // dequantize_dep_quant_bd (vvc/dct.rs:1191-1248) on the raw levels
        // phase A left in LDS. The automaton's step from state x on a level of
        // parity p is x' = swap_bits(x) ^ 2p (the table 32040 read as 2-bit
        // entries), an affine map over GF(2). Run from the end of the scan
        // (zeros keep state 0, so this equals starting at the last non-zero
        // level), bit 1 of the state on arrival at scan position s, the only
        // bit the reconstruction offset uses, is therefore the XOR of the
        // parities at positions s + 1, s + 3, s + 5, ... A lane takes one 4 x 4
        // coefficient group (16 consecutive scan positions): the XORs over the
        // later groups of its TU come from two ballots, the rest from a walk
        // over its own 16 levels.
        const int g = tid / NCG, c = tid & (NCG - 1);
        const bool act = g < G && ((P[g].cwch >> 22) & 1);
        const int x0 = cg_org & (W - 1), y0 = cg_org >> LW;
        uint2* col = reinterpret_cast<uint2*>(C + (g * Wc + x0) * CP + (y0 >> 1));
        uint2 v[4] = {};
        if (act) {
#pragma unroll
            for (int x = 0; x < 4; x++) v[x] = col[x * (CP / 2)];
        }
        // level (x, y) of the group, int16 pairs along the rows of a column
        auto lev = [&](int x, int y) {
            const uint32_t u = y < 2 ? v[x].x : v[x].y;
            return (y & 1) ? (int)u >> 16 : (int)(int16_t)(u & 0xffff);
        };
        // raster positions y * 4 + x of the group's diagonal scan
        constexpr uint8_t kScan4[16] = {0, 4, 1, 8, 5, 2, 12, 9, 6, 3, 13, 10, 7, 14, 11, 15};
        int xe = 0, xo = 0;  // parity XORs over the group's even / odd scan positions
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int par = lev(kScan4[i] & 3, kScan4[i] >> 2) & 1;
            if (i & 1) xo ^= par; else xe ^= par;
        }
        // the later groups of the TU: lanes tid + 1 .. (g + 1) NCG - 1
        const int seg_end = (g + 1) * NCG;
        const uint64_t upto = seg_end >= 64 ? ~0ull : (1ull << seg_end) - 1;
        const uint64_t later = upto & ~((2ull << tid) - 1);
        const uint64_t be = __builtin_amdgcn_ballot_w64(act && xe), bo = __builtin_amdgcn_ballot_w64(act && xo);
        int se = __popcll(be & later) & 1, so = __popcll(bo & later) & 1;
        if (act) {
            const int dqp = P[g].dq;
            const int S = dqp & 0x1ffff, rs = (dqp >> 24) & 63, wide_q = (dqp >> 30) & 1;
            const int rnd = (1 << rs) >> 1;
            int out[16];
#pragma unroll
            for (int i = 15; i >= 0; i--) {
                const int pos = kScan4[i];
                const int level = lev(pos & 3, pos >> 2);
                // bit 1 of the state: the XOR over positions s + 1, s + 3, ...,
                // the even-index positions after s when i is odd
                const int A = (i & 1) ? se : so;
                int val = 0;
                if (level != 0) {
                    int q = 2 * level + (level > 0 ? -A : A);
                    // with a left shift every |q| >= 2^15 saturates: the clamp
                    // keeps q * S within int32
                    if (wide_q) q = min(max(q, -32768), 32767);
                    val = min(max((__mul24(q, S) + rnd) >> rs, -32768), 32767);
                }
                out[pos] = val;
                if (i & 1) so ^= level & 1; else se ^= level & 1;
            }
#pragma unroll
            for (int x = 0; x < 4; x++) {
                uint2 w;
                w.x = __builtin_amdgcn_perm((uint32_t)out[4 + x], (uint32_t)out[x], 0x05040100u);
                w.y = __builtin_amdgcn_perm((uint32_t)out[12 + x], (uint32_t)out[8 + x], 0x05040100u);
                col[x * (CP / 2)] = w;
            }
        }
        __syncthreads();
    }
    // tool tiles: every TU is LFNST or dependent quantisation; the tile flag
    // says whether any applies LFNST
    if (LFNST && (it.lw & kTileLfnst)) {
        // inverse_lfnst_rect (vvc/lfnst.rs:219-297) on the dequantised block in
        // LDS, for the TUs that apply it: the first NZ positions of the 4x4 diagonal scan times the
        // kernel matrix give TR outputs, (sum + 64) >> 7 clipped to int16,
        // placed (transposed or not) in the top-left 4x4 / 8x8 region; the
        // coded region is 4..8 on a side, so it holds the whole input and
        // clips the output. One item per (TU, output); the outputs are
        // staged in the intermediate-row area (unused until phase B, and
        // large enough for GMAX * TR values) before they overwrite the inputs.
        constexpr int SB = (W >= 8 && H >= 8) ? 8 : 4;
        constexpr int TR = SB == 8 ? 48 : 16;
        constexpr int NZ = ((W == 4 && H == 4) || (W == 8 && H == 8)) ? 8 : 16;
        static_assert(H * residual_tile_pitch(LWC) * 2 >= TR, "LFNST outputs fit the intermediate rows");
        constexpr uint32_t kScan4Lo = 0x9c258140u, kScan4Hi = 0xfbe7ad36u;  // y * 4 + x of the 4x4 diagonal scan, 4 bits each
        const int8_t* __restrict__ lt = SB == 8 ? a.lfnst8 : a.lfnst4;
        int16_t* stage = reinterpret_cast<int16_t*>(T);
#pragma unroll 1
        for (int q = tid; q < G * TR; q += kTileThreads) {
            const int g = q / TR, j = q - g * TR;
            const int code = (P[g].cwch >> 17) & 31;
            if (!(code & 16)) continue;
            const int8_t* __restrict__ m = lt + (code & 7) * 16 * TR + j;
            const int16_t* c16 = reinterpret_cast<const int16_t*>(C + g * Wc * CP);
            int sum = 0;
#pragma unroll
            for (int i = 0; i < NZ; i++) {
                const int pos = ((i < 8 ? kScan4Lo >> (4 * i) : kScan4Hi >> (4 * (i - 8)))) & 15;
                sum += (int)c16[(pos & 3) * 2 * CP + (pos >> 2)] * (int)m[i * TR];
            }
            stage[q] = (int16_t)clip16i((sum + 64) >> 7);
        }
        __syncthreads();
#pragma unroll 1
        for (int q = tid; q < G * TR; q += kTileThreads) {
            const int g = q / TR, j = q - g * TR;
            const int cwch = P[g].cwch, cw = cwch & 255, ch = (cwch >> 8) & 255;
            if (!((cwch >> 21) & 1)) continue;  // no LFNST: its block stays as dequantised
            int ox, oy;
            if (SB == 4) {
                ox = j & 3;
                oy = j >> 2;
            } else if (j < 32) {
                ox = j & 7;
                oy = j >> 3;
            } else {
                ox = (j - 32) & 3;
                oy = 4 + ((j - 32) >> 2);
            }
            if ((cwch >> 20) & 1) {  // transposed
                const int t = ox;
                ox = oy;
                oy = t;
            }
            if (ox < cw && oy < ch) reinterpret_cast<int16_t*>(C + g * Wc * CP)[ox * 2 * CP + oy] = stage[q];
        }
        // the outputs reach the whole coded region of every LFNST TU (rows
        // and columns 0 .. 3 or 0 .. 7): cover it in the bounding box
        if (tid < G && ((P[tid].cwch >> 21) & 1)) {
            const int cwch = P[tid].cwch;
            atomicMax(&TB[0], ((cwch >> 8) & 255) - 1);
            atomicMax(&TB[1], (cwch & 255) - 1);
        }
        __syncthreads();
    }
    const int er = __builtin_amdgcn_readfirstlane(TB[0]) + 1;
    const int ec = __builtin_amdgcn_readfirstlane(TB[1]) + 1;
    const int k1 = (er + 1) >> 1, k2 = (ec + 1) >> 1;
    if constexpr (UNI) {
        uniform_passes<LW, LH, GMAX, CP, TP>(a, C, T, P, G, ec, k1, k2);
        return;
    }

    // phase B: vertical pass, lane = (TU, row y, column chunk); rows past
    // the bounding box are read in groups of 4 (zeros, from zero coefficients)
    if (ec > 0) {
        constexpr int NCH = NCH_B;
        constexpr int ITEMS = GMAX * H * NCH;
        // the horizontal pass reads intermediates two pairs at a time: the
        // columns are produced up to a multiple of 4 (zeros past the box,
        // from zero coefficients), in NCH chunks of an even number of columns
        constexpr int CM = NCH > 2 ? 2 * NCH : 4;
        const int ecn = (ec + CM - 1) & ~(CM - 1), nc = ecn / NCH;
        auto vpass = [&](auto nq) {
            constexpr int NQ = decltype(nq)::value;
#pragma unroll
            for (int base = 0; base < ITEMS; base += kTileThreads) {
                const int q = base + tid;
                const int y = q & (H - 1);
                const int c0 = (q / H) % NCH;
                const int g = q / (H * NCH);
                if (g < G) {
                    uint32_t m[16];
                    if constexpr (B_PRE) {
#pragma unroll
                        for (int k = 0; k < NPB; k++) m[k] = mB[base / kTileThreads * NPB + k];
                    } else {
                        load_col<NPB>(a.col16 + P[g].mv + y * 16, m);
                    }
                    // the same trip count in every lane
                    stage1_row<NQ, CP / 2>(reinterpret_cast<const uint2*>(C + g * Wc * CP), T + (g * H + y) * TP, m,
                                           c0 * nc, nc);
                }
            }
        };
        with_groups<1, (Hc / 4 > 0 ? Hc / 4 : 1)>((k1 + 1) >> 1, vpass);
    }
    __syncthreads();

    if constexpr (!RECON) {
        // phase C: horizontal pass, lane = (TU, column pair, row group); each
        // lane produces RL rows (8, or all rows of a TU of fewer) of two
        // adjacent columns per round and stores them as saturated int16
        // pairs straight to HBM. Long rounds beat longer store runs: with
        // 8 x 8 TUs, 4 rounds of 2 rows (64-byte runs) took 49.5 us per
        // 26 M samples, 1 round of 8 rows (16-byte runs) 36.1 us.
        constexpr int W2 = W / 2;
        constexpr int NRR0 = H / 8;
        constexpr int NRR = NRR0 < 1 ? 1 : (NRR0 > H ? H : NRR0);
        constexpr int RL = H / NRR;        // rows per lane per round (<= 8)
        constexpr int LPT = W2 * NRR;      // lanes per TU
        constexpr int ITEMS = GMAX * LPT;  // 8 kTileThreads / RL, a multiple of kTileThreads
        constexpr int NP = Wc / 2;         // matrix pairs per column
        constexpr int RC = ITEMS / kTileThreads;  // rounds
        // the two matrix columns of every round, fetched together before the
        // first round when they fit 32 registers (one wait instead of RC)
        constexpr bool C_PRE = RC > 1 && RC * 2 * NP <= 32;
        const int x2 = tid % W2;
        uint32_t mC[C_PRE ? RC * 2 * NP : 1];
        if constexpr (C_PRE) {
#pragma unroll
            for (int rc = 0; rc < RC; rc++) {
                const int g = (rc * kTileThreads + tid) / LPT;
                if (g < G) {
                    const int mh = P[g].mh;
                    load_colp<NP>(a.col16 + mh + (2 * x2) * 16, mC + rc * 2 * NP);
                    load_colp<NP>(a.col16 + mh + (2 * x2 + 1) * 16, mC + rc * 2 * NP + NP);
                }
            }
        }
        auto hpass = [&](auto nq) {
            constexpr int NQ = decltype(nq)::value;
#pragma unroll
            for (int base = 0; base < ITEMS; base += kTileThreads) {
                const int q = base + tid;
                const int rr = (q / W2) % NRR;
                const int g = q / LPT;
                if (g < G) {
                    const TuParams p = P[g];
                    uint32_t m0[16], m1[16];
                    if constexpr (C_PRE) {
#pragma unroll
                        for (int k = 0; k < NP; k++) {
                            m0[k] = mC[base / kTileThreads * 2 * NP + k];
                            m1[k] = mC[base / kTileThreads * 2 * NP + NP + k];
                        }
                    } else {
                        load_col<NP>(a.col16 + p.mh + (2 * x2) * 16, m0);
                        load_col<NP>(a.col16 + p.mh + (2 * x2 + 1) * 16, m1);
                    }
                    const int shift2 = p.shifts >> 16;
                    const int rnd = 1 << (shift2 - 1);
                    // row rr of the TU; rows rr + j NRR are at constant offsets
                    const uint2* src = reinterpret_cast<const uint2*>(T + (g * H + rr) * TP);
                    RMH_GLOBAL uint32_t* dst = gdst<uint32_t>(a.residual + p.off) + (rr * W2 + x2);
#pragma unroll
                    for (int j = 0; j < RL; j++) {
                        int a0, a1;
                        dot_groups2<NQ>(src + j * NRR * (TP / 2), m0, m1, rnd, a0, a1);
                        dst[j * NRR * W2] = pack_sat16(a0 >> shift2, a1 >> shift2);
                    }
                }
            }
        };
        // NQ = 0 for a tile without non-zero coefficients (phase B wrote no
        // intermediates; every residual is zero)
        with_groups<0, (Wc / 4 > 0 ? Wc / 4 : 1)>((k2 + 1) >> 1, hpass);
    } else {
        // phase C (recon): horizontal pass, lane = (TU, column x, row chunk);
        // picture samples are updated in place
        constexpr int NR0 = kTileThreads / (GMAX * W);
        constexpr int NR = NR0 > 0 ? NR0 : 1;
        constexpr int ITEMS = GMAX * W * NR;
        auto hpass = [&](auto nq) {
            constexpr int NQ = decltype(nq)::value;
#pragma unroll
            for (int base = 0; base < ITEMS; base += kTileThreads) {
                const int q = base + tid;
                const int x = q & (W - 1);
                const int r0 = (q / W) % NR;
                const int g = q / (W * NR);
                if (g < G) {
                    const TuParams p = P[g];
                    uint32_t m[16];
                    load_col<Wc / 2>(a.col16 + p.mh + x * 16, m);
                    const int shift2 = p.shifts >> 16;
                    const int rnd = 1 << (shift2 - 1);
                    const ReconDst rd{p.base, p.stride, p.bpsmax & 255, p.bpsmax >> 8};
                    const uint2* src = reinterpret_cast<const uint2*>(T + g * H * TP);
                    {
                        // H / NR <= 16 rows per lane, in batches of RB: the
                        // prediction samples of a batch are requested before its first add
                        constexpr int NY = H / NR, RB = NY < kReconBatch ? NY : kReconBatch;
#pragma unroll
                        for (int j0 = 0; j0 < NY; j0 += RB) {
                            int pred[RB];
#pragma unroll
                            for (int j = 0; j < RB; j++) pred[j] = recon_load(rd, x, r0 + (j0 + j) * NR);
#pragma unroll
                            for (int j = 0; j < RB; j++) {
                                const int y = r0 + (j0 + j) * NR;
                                const int acc = dot_groups<NQ>(src + y * (TP / 2), m, rnd);
                                recon_store(rd, x, y, pred[j] + min(max(acc >> shift2, -32768), 32767));
                            }
                        }
                    }
                }
            }
        };
        // no group at all for a tile without non-zero coefficients (phase B
        // wrote no intermediates; every residual is zero)
        with_groups<0, (Wc / 4 > 0 ? Wc / 4 : 1)>((k2 + 1) >> 1, hpass);
    }
}

// RECON: add into the picture planes (a.place / a.frames) instead of writing
// int16 residuals. LFNST: tool tiles, of TUs that apply the inverse LFNST or
// dependent quantisation (an instantiation of its own, so the plain tiles
// keep their register budget and code size).
template <bool RECON, bool LFNST>
__global__ __launch_bounds__(kTileThreads) void residual_tile_kernel(ResidualArgs a, const TileItem* __restrict__ items) {
    extern __shared__ __align__(16) uint32_t smem[];
    const TileItem it = items[blockIdx.x];
    switch ((it.lw & 7) * 8 + it.lh) {
#define RMH_TILE(LW, LH) \
    case LW * 8 + LH:    \
        tile_body<LW, LH, RECON, LFNST>(a, it, smem); \
        break;
        RMH_TILE(2, 2) RMH_TILE(2, 3) RMH_TILE(2, 4) RMH_TILE(2, 5) RMH_TILE(2, 6)
        RMH_TILE(3, 2) RMH_TILE(3, 3) RMH_TILE(3, 4) RMH_TILE(3, 5) RMH_TILE(3, 6)
        RMH_TILE(4, 2) RMH_TILE(4, 3) RMH_TILE(4, 4) RMH_TILE(4, 5) RMH_TILE(4, 6)
        RMH_TILE(5, 2) RMH_TILE(5, 3) RMH_TILE(5, 4) RMH_TILE(5, 5) RMH_TILE(5, 6)
        RMH_TILE(6, 2) RMH_TILE(6, 3) RMH_TILE(6, 4) RMH_TILE(6, 5) RMH_TILE(6, 6)
#undef RMH_TILE
        default:
            break;
    }
}

}  // namespace

hipError_t launch_residual(const ResidualArgs& a, hipStream_t stream) {
    if (a.n_tu == 0) return hipSuccess;
    if (a.prep_out)
        hipLaunchKernelGGL(residual_kernel<true>, dim3(a.n_tu), dim3(64), 0, stream, a);
    else
        hipLaunchKernelGGL(residual_kernel<false>, dim3(a.n_tu), dim3(64), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_residual_tiles(const ResidualArgs& a, const TileItem* items, uint32_t n_items, uint32_t n_plain,
                                 uint32_t lds_bytes, hipStream_t stream) {
    if (n_plain > 0) {
        if (a.place)
            hipLaunchKernelGGL((residual_tile_kernel<true, false>), dim3(n_plain), dim3(kTileThreads), lds_bytes, stream, a,
                               items);
        else
            hipLaunchKernelGGL((residual_tile_kernel<false, false>), dim3(n_plain), dim3(kTileThreads), lds_bytes, stream,
                               a, items);
    }
    if (n_items > n_plain) {
        if (a.place)
            hipLaunchKernelGGL((residual_tile_kernel<true, true>), dim3(n_items - n_plain), dim3(kTileThreads), lds_bytes,
                               stream, a, items + n_plain);
        else
            hipLaunchKernelGGL((residual_tile_kernel<false, true>), dim3(n_items - n_plain), dim3(kTileThreads), lds_bytes,
                               stream, a, items + n_plain);
    }
    return hipGetLastError();
}

}  // namespace rmh

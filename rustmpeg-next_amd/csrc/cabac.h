// cabac.h - host-side CABAC coder for the encoder's rate: the arithmetic
// coder, the adaptive context models, residual_coding and the syntax the
// RD search writes (coding-tree splits, PLANAR/DC luma modes, luma cbf).
//
// Arithmetic and binarisations follow the reference's BinEncoder and
// CtxModel (crates/rustmpeg-video/src/vvc/cabac.rs:32-99, 477-745), its
// residual coding (vvc/residual.rs:142-240, 280-357, 448-743) and its split
// syntax (vvc/slice.rs:3850-3893, vvc/transform.rs:537-715). These are
// normative (H.266 clause 9.3): a decoder reads the stream only when every
// interval update, context selection and binarisation is the standard's.
#pragma once
#include <cstdint>
#include <vector>

namespace rmh {
namespace cabac {

// Context-set offsets into the 379 models (vvc/cabac.rs:307-398).
enum : int {
    SPLIT_FLAG = 0,
    SPLIT_QT_FLAG = 9,
    SPLIT_HV_FLAG = 15,
    SPLIT_12_FLAG = 20,
    INTRA_LUMA_MPM_FLAG = 40,
    INTRA_LUMA_PLANAR_FLAG = 41,
    QT_CBF_Y = 77,
    SIG_COEFF_GROUP_LUMA = 86,
    SIG_COEFF_GROUP_CHROMA = 88,
    SIG_FLAG_LUMA = 90,
    SIG_FLAG_CHROMA = 102,
    SIG_FLAG_LUMA_1 = 110,
    SIG_FLAG_CHROMA_1 = 122,
    SIG_FLAG_LUMA_2 = 130,
    SIG_FLAG_CHROMA_2 = 142,
    PAR_FLAG_LUMA = 150,
    PAR_FLAG_CHROMA = 171,
    GTX2_LUMA = 182,
    GTX2_CHROMA = 203,
    GTX1_LUMA = 214,
    GTX1_CHROMA = 235,
    LAST_X_LUMA = 246,
    LAST_X_CHROMA = 266,
    LAST_Y_LUMA = 269,
    LAST_Y_CHROMA = 289,
    NUM_CONTEXTS = 379,
};

// One adaptive binary model: two probability estimates with their own
// adaptation rates, averaged (H.266 9.3.2.2, 9.3.4.3.2).
struct Model {
    uint16_t s0, s1;
    uint8_t r0, r1;
};

struct Contexts {
    Model m[NUM_CONTEXTS];
    // slice 0 = I, 1 = P, 2 = B initialisation table at slice QP qp.
    void init(int slice, int qp);
};

// Arithmetic encoder. With a sink the coded bytes are appended to it; without
// one it only counts, which is how the RD search prices a candidate.
class Encoder {
public:
    explicit Encoder(std::vector<uint8_t>* sink = nullptr) : sink_(sink) {}
    // A copy of the current state writing to another sink (null: counting
    // only). The bytes it writes are those this encoder would write from here,
    // so a candidate coded on a copy can be adopted without coding it again.
    Encoder with_sink(std::vector<uint8_t>* sink) const {
        Encoder e = *this;
        e.sink_ = sink;
        return e;
    }
    std::vector<uint8_t>* sink() const { return sink_; }
    void bin(int b, Model& m);
    void bypass(int b);
    void bypass_bins(uint32_t v, int n);
    void terminate(int b);
    void unary_max(uint32_t sym, uint32_t max_sym);
    void rem_abs(uint32_t v, uint32_t k, uint32_t cutoff, uint32_t max_log2);
    // Bits the stream holds so far, counting the bytes still held for carry
    // propagation and the bits of `low` already committed.
    uint64_t bits() const { return written_ + 8ull * held_ + (uint64_t)(23 - bits_left_); }
    // Flushes the interval and byte-aligns (a 1 bit, then zeros).
    void finish();

private:
    void put_bits(uint32_t v, int n);
    void out_byte_();
    uint32_t low_ = 0, range_ = 510;
    uint32_t held_byte_ = 0xFF, held_ = 0;
    int bits_left_ = 23;
    uint64_t written_ = 0;  // bits emitted to the sink (or that would have been)
    uint32_t acc_ = 0;
    int acc_n_ = 0;
    std::vector<uint8_t>* sink_;
};

struct CanSplit;
struct SplitCtx;

// Arithmetic decoder over one substream (bytes past the end read as zero).
class Decoder {
public:
    Decoder(const uint8_t* data, size_t n);
    int bin(Model& m);
    int bypass();
    uint32_t bypass_bins(int n);
    int terminate();
    uint32_t unary_max(uint32_t max_sym);
    uint32_t rem_abs(uint32_t k, uint32_t cutoff, uint32_t max_log2);
    // True when the substream ends in the stop bit and alignment zeros.
    bool finish_ok() const;

private:
    uint32_t read_byte_();
    const uint8_t* data_;
    size_t n_, pos_;
    uint32_t range_ = 510, value_ = 0;
    int bits_needed_ = -8;
    uint8_t last_byte_ = 0;
};

// residual_coding() of one TU into levels (w*h, raster order); false on a
// position outside the block or a level outside int16.
bool decode_residual(Decoder& d, Contexts& c, int w, int h, int32_t* levels, bool luma, bool dep_quant,
                     bool sign_hiding);
int decode_split(Decoder& d, Contexts& c, const CanSplit& can, const SplitCtx& sc);

// residual_coding() of one TU (regular, not transform skip) with the levels in
// raster order, w x h of 1..64 (only the top-left 32x32 are coded). At least one
// level must be non-zero.
void encode_residual(Encoder& e, Contexts& c, int w, int h, const int32_t* levels, bool luma, bool dep_quant,
                     bool sign_hiding);

struct CanSplit {
    bool no, qt, bh, bv, th, tv;
    bool any() const { return qt || bh || bv || th || tv; }
    bool any_mtt() const { return bh || bv || th || tv; }
};
struct SplitCtx {
    int split, qt, hv, hor_bt, ver_bt;
};

// Split order is CodingTreeSplit's: 0 none, 1 quad, 2 binary hor, 3 binary
// ver, 4 ternary hor, 5 ternary ver.
struct SplitLimits {
    int min_qt, max_bt, max_tt, max_mtt_depth, min_cb;
};
CanSplit can_split(int x, int y, int w, int h, int pic_w, int pic_h, int mtt_depth, const SplitLimits& lim,
                   int last_split);
// left_h / above_w: height and width of the CU left of / above the node's
// top-left sample, with their QT depths; a size of 0 means unavailable.
SplitCtx split_contexts(int w, int h, int qt_depth, int mtt_depth, int left_h, int left_qt, int above_w, int above_qt,
                        const CanSplit& can);
void encode_split(Encoder& e, Contexts& c, int split, const CanSplit& can, const SplitCtx& sc);

}  // namespace cabac
}  // namespace rmh

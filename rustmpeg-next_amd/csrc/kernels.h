// kernels.h - launch interface between the C-ABI host layer (rmh.cpp) and the
// HIP kernels (residual.hip, mc.hip, deblock.hip). Internal to librmh.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rmh.h"

namespace rmh {

// Device-resident constant tables, built once per context.
struct DeviceTables {
    const uint32_t* mat = nullptr;   // transform matrices as int16 pairs: [kk][j] = (M[2kk][j], M[2kk+1][j])
    const uint32_t* col16 = nullptr; // per column j: 16 pairs (M[2kk][j], M[2kk+1][j]), zero past the last row
    int32_t col16_off[3][7];         // dword offset of (type, log2 n); -1 if unsupported
    // DCT-II of 8, 16 and 32 points split for the even/odd butterfly, in
    // col16: per column j, 8 even pairs (M[4q][j], M[4q + 2][j]) then 8 odd
    // pairs (M[4q + 1][j], M[4q + 3][j]), zero past the last row; -1 if absent
    int32_t eo_off[7];
    const int8_t* lfnst8 = nullptr;  // [4][2][16][48]
    const int8_t* lfnst4 = nullptr;  // [4][2][16][16]
    const uint16_t* scans = nullptr; // grouped diagonal scans, raster positions
    int32_t mat_off[3][7];           // dword offset of (type, log2 n); -1 if unsupported
    int32_t scan_off[7][7];          // element offset of scan (log2 w, log2 h)
    int32_t scan_len[7][7];
};

// Per tile-eligible TU, in tile order, built at plan creation: what the tiled
// kernel needs about the TU in one 20-byte record.
struct TileTu {
    uint32_t off;     // element offset of the TU block in the residual output
    uint32_t src;     // element offset of its levels (in ResidualArgs::prep when prepped)
    uint8_t cw, ch;   // coded region after zero-out (<= 32)
    uint8_t scale;    // dequantisation scale (1 when prepped: the values are already dequantised;
                      // of QP + 1 under dependent quantisation)
    uint8_t shifts;   // dequant left shift | right shift << 4
    uint16_t mh, mv;  // dword offsets of the column-major matrices (col16)
    uint8_t shift2;   // second-stage shift, 20 - bit depth
    uint8_t prepped;  // 1: read the prepass output (ResidualArgs::prep), not the levels
    uint8_t lfnst;    // tool tiles: LFNST matrix (set * 2 + idx - 1) | transposed << 3 | 16 when the TU applies LFNST
    uint8_t dq;       // 1: dependent quantisation of the whole coded region, applied in the kernel
};
// Recon target of a tile-eligible TU (recon plans only).
struct TileDst {
    char* base;       // picture sample (0, 0) of the block
    int32_t stride;   // row pitch, samples
    uint16_t bps;     // bytes per sample
    uint16_t maxv;    // 2^bd - 1
};

struct ResidualArgs {
    const rmh_tu_desc* tus;
    const int16_t* coeff;
    int16_t* residual;
    const uint32_t* mat;
    const int8_t* lfnst8;
    const int8_t* lfnst4;
    const uint16_t* scans;
    const uint32_t* order;  // optional TU index list (nullptr: descriptor order)
    const uint32_t* col16;  // column-major padded matrix pairs (tiled kernel)
    int32_t col16_off[3][7];
    int32_t eo_off[7];                 // DeviceTables::eo_off
    const rmh_tu_place* place;         // recon mode when non-null: add into frames' planes
    const rmh_recon_frame* frames;
    const TileTu* tile_tus;            // tiled kernel: per TU of order[0 .. n_fast)
    const TileDst* tile_dst;           // tiled kernel, recon mode
    // Prepass of TUs with LFNST or dependent quantisation that the tiled
    // kernel then transforms: residual_kernel with prep_out set writes each
    // TU's dequantised (and LFNST-ed) coefficients, zero outside the coded
    // region, as a raster w x h block at prep_out + prep_off[k] (k = the TU's
    // position in its launch) and stops before the transform; the tiled
    // kernel reads them from `prep`.
    int16_t* prep_out;
    const uint32_t* prep_off;
    const int16_t* prep;
    uint32_t first_tu;
    uint32_t n_tu;
    int32_t mat_off[3][7];
    int32_t scan_off[7][7];
    int32_t scan_len[7][7];
};

hipError_t launch_residual(const ResidualArgs& a, hipStream_t stream);

// A tile of TUs of one size class (log2 w, log2 h in 2..6) processed by one
// kTileThreads-lane workgroup of the tiled residual kernel: TUs order[first .. first+count).
struct TileItem {
    uint32_t first;
    uint16_t count;
    uint8_t lw;  // | kTileLfnst: a tool tile, whose TUs apply the inverse LFNST or dependent quantisation in the kernel
    uint8_t lh;
};
constexpr uint8_t kTileLfnst = 0x80;

// Tiled residual kernel for 2-D TUs with both sides >= 4 and no transform
// skip: plain TUs from their levels, TUs with LFNST or dependent
// quantisation from the prepass output. `a.order` is required.
// items[0 .. n_plain) are plain tiles, items[n_plain .. n_items) tool tiles (LFNST or dependent quantisation).
hipError_t launch_residual_tiles(const ResidualArgs& a, const TileItem* items, uint32_t n_items, uint32_t n_plain,
                                 uint32_t lds_bytes, hipStream_t stream);

// Size classes whose plain tiles (residual output, no LFNST or dependent
// quantisation in the kernel) run the uniform-matrix passes: every TU of such
// a tile has the same transform pair, a lane owns coefficient columns and then
// intermediate rows, and the intermediates overlay the coefficients in LDS.
// The host orders and cuts tiles by this same predicate.
constexpr bool residual_uniform_class(int lw, int lh) {
    return ((lw == 3 || lw == 4) && (lh == 3 || lh == 4)) || (lw == 5 && (lh == 4 || lh == 5)) || (lw == 4 && lh == 5);
}

// Lanes of a tiled-kernel workgroup; a full tile holds 16 samples per lane
// (a uniform tile with a 32-sample side twice as many, so that its rows and
// its columns each cover the lanes).
constexpr int kTileThreads = 64;

// TUs per full tile of class (lw, lh), uniform (`uni`) or not.
constexpr int residual_tile_tus(int lw, int lh, bool uni = false) {
    int g = ((kTileThreads * 16) >> (lw + lh)) > 0 ? (kTileThreads * 16) >> (lw + lh) : 1;
    if (uni) {
        if ((g << lh) < kTileThreads) g = kTileThreads >> lh;
        if ((g << lw) < kTileThreads) g = kTileThreads >> lw;
    }
    return g;
}
// Dword pitch of a coefficient column (lc = log2 of the coded height) or an
// intermediate row (lc = log2 of the coded width) in the tiled kernel's LDS:
// 2^lc / 2 int16 pairs, padded to 2 mod 4 dwords.
constexpr int residual_tile_pitch(int lc) { return lc <= 2 ? 2 : (1 << lc) / 2 + 2; }

// The LDS bytes a full tile needs (the kernel lays LDS out for a full tile
// whatever the tile's count).
constexpr uint32_t residual_tile_lds(int lw, int lh, bool uni = false) {
    const uint32_t count = (uint32_t)residual_tile_tus(lw, lh, uni);
    const int lwc = lw < 5 ? lw : 5, lhc = lh < 5 ? lh : 5;
    const uint32_t c = count * (1u << lwc) * residual_tile_pitch(lhc) * 4;  // coefficient columns
    const uint32_t t = count * (1u << lh) * residual_tile_pitch(lwc) * 4;   // intermediate rows
    return (uni ? (c > t ? c : t) : c + t) + count * 48 + 16;               // + per-TU parameters + tile bounding box
}

// One inter work item: a 16x16 luma tile (and its 8x8 chroma tiles) of one
// PU, with its sizes already clipped to the picture (built at plan creation).
struct McItem {
    uint16_t x0, y0;  // luma tile origin in the picture
    uint8_t tw, th;   // luma tile size, 1..16
    uint8_t cw, ch;   // chroma tile size, 0..8
    uint16_t frame;
    uint16_t flags;   // bit 0: alternative half-sample luma filter
    int32_t mv_x, mv_y;
    uint32_t reserved;
};

// Items [0, n_items8) read and write 8-bit planes, the rest 16-bit ones.
// waves: work items (one wave each) per workgroup, 1, 2 or 4.
hipError_t launch_mc(const rmh_mc_frame* frames, const McItem* items, uint32_t n_items, uint32_t n_items8,
                     hipStream_t stream, int waves = 4);

// One bi-prediction work item: the tile geometry of McItem, both MVs and the
// BCW weight of L1 (the L0 weight is 8 minus it).
struct McBiItem {
    uint16_t x0, y0;  // luma tile origin in the picture
    uint8_t tw, th;   // luma tile size, 1..16
    uint8_t cw, ch;   // chroma tile size, 0..8
    uint16_t frame;
    uint8_t flags;    // bit 0: alternative half-sample luma filter
    int8_t w1;        // L1 weight, one of -2, 3, 4, 5, 10
    int32_t mv0_x, mv0_y, mv1_x, mv1_y;
    uint32_t reserved;
};

hipError_t launch_mc_bi(const rmh_mc_bi_frame* frames, const McBiItem* items, uint32_t n_items, uint32_t n_items8,
                        hipStream_t stream, int waves = 4);

// DMVR of one bi PU (8-bit frames): the signalled MV pair, refined into
// slot `slot` of the plan's MV array (4 int32 per slot).
struct DmvrJob {
    uint16_t x, y;    // PU origin (luma)
    uint8_t lw, lh;   // log2 PU size (3..7)
    uint16_t frame;
    uint32_t slot;
    int32_t mv[4];    // mv0_x, mv0_y, mv1_x, mv1_y
};

// One 16x16 luma tile (with its 8x8 chroma tiles) of a bi PU that DMVR or
// BDOF refines; its MV pair is read from slot `slot` of the MV array.
struct McRefItem {
    uint16_t px, py;  // PU origin
    uint8_t lw, lh;   // log2 PU size
    uint8_t tx, ty;   // tile index in the PU (16-sample units)
    uint16_t frame;
    uint8_t flags;    // bit 0: alternative half-sample luma filter; bit 1: BDOF
    int8_t w1;        // L1 weight
    uint32_t slot;
};

// One GPM work item: a 16x16 luma tile of a GPM PU and its 8x8 chroma
// tiles, sizes clipped to the picture, with the PU's split-line parameters
// (predict_gpm_block): the GEO_DIS indices of the angle and of the angle + 8,
// and the offsets of the PU origin from the split line on the luma grid, for
// the luma and the chroma planes.
struct GpmItem {
    uint16_t x0, y0;   // luma tile origin (picture)
    uint16_t px, py;   // PU origin
    uint8_t tw, th;    // luma tile size, 0..16
    uint8_t cw, ch;    // chroma tile size, 0..8
    uint16_t frame;
    uint8_t dx, dy;
    int16_t lox, loy;  // luma plane
    int16_t cox, coy;  // chroma planes
};

hipError_t launch_gpm(const rmh_gpm_frame* frames, const GpmItem* items, uint32_t n_items, uint32_t luma_scale,
                      uint32_t chroma_scale, hipStream_t stream);

// lds_bytes: 2 (w + 4) (h + 4) for the largest PU of the jobs.
hipError_t launch_dmvr(const rmh_mc_bi_frame* frames, const DmvrJob* jobs, uint32_t n_jobs, uint32_t lds_bytes,
                       int32_t* mvs, hipStream_t stream);
hipError_t launch_mc_bi_refined(const rmh_mc_bi_frame* frames, const McRefItem* items, uint32_t n_items,
                                const int32_t* mvs, hipStream_t stream);

// Deblocking: a copy of every plane into `shadows` (3 per frame, same sizes,
// device memory), then one tiled launch that reads the shadows and writes the
// pictures. The frames are taken in the order frame_of[n_frames] (device
// memory), 8-bit frames first; tile_start[n_frames + 1] holds the running
// sum of deblock_tiles() in that order (tile_start[0] = 0, tile_start[n_frames]
// = total_tiles) and tiles8 is the number of tiles on 8-bit frames;
// max_rows = largest luma height.
uint32_t deblock_tiles(uint32_t width, uint32_t height);
// With copy_shadows, the frames' planes are first copied into `shadows` and
// filtered in place; without it, `shadows` are the source planes and the
// frames' planes receive the output (every sample written once).
hipError_t launch_deblock(const rmh_dbk_frame* frames, const rmh_plane* shadows, const uint32_t* tile_start,
                          const uint32_t* frame_of, uint32_t n_frames, uint32_t total_tiles, uint32_t tiles8,
                          uint32_t max_rows, bool copy_shadows, hipStream_t stream);

// Edge-map builder over `n_frames` frames (device frame table), one thread per
// cell; cell_start[f] is the running sum of grid_w * grid_h.
hipError_t launch_edge_map(const rmh_edge_frame* frames, const uint32_t* cell_start, uint32_t n_frames,
                           uint32_t total_cells, hipStream_t stream);

// One plane of one picture of an SAO batch (built at plan creation).
struct SaoJob {
    const void* src;               // deblocked plane
    void* dst;                     // output plane, same size and pitch
    const rmh_sao_param* params;   // the frame's CTB parameters
    uint32_t width, height, stride;  // samples
    uint32_t vis_w, vis_h;         // visible extent of this component
    uint32_t ctb;                  // CTB size of this component
    uint32_t ctb_cols;             // params row pitch in CTBs
    uint32_t comp;                 // 0 Y, 1 Cb, 2 Cr
    uint32_t bd;                   // bit depth
    uint32_t blocks_per_row;       // row form: 256-lane segments per group of kSaoRows rows
    uint32_t first_block;          // first workgroup of the job
    uint32_t vec;                  // rows 8-byte aligned: whole-lane vector access
    uint32_t lanes_across;         // CTB form: lanes across a wave's block (4, 8, 16); 0: row form
    uint32_t tiles_x;              // CTB form: blocks across the plane
    uint32_t tiles;                // CTB form: blocks of the plane (4 kSaoBlocksPerWave per workgroup)
    uint32_t ctb_log2;             // log2 of ctb
    double inv_tiles_x;            // CTB form: 1 / tiles_x
};

// Rows of a plane per workgroup of the SAO row form; blocks per wave of the
// CTB form.
constexpr int kSaoRows = 4;
constexpr int kSaoBlocksPerWave = 1;

// block_job[b] is the job of workgroup b. Workgroups [range[i], range[i+1])
// hold, in order, row-form 8-bit, CTB-form 8-bit, row-form 16-bit and
// CTB-form 16-bit planes. The CTB form serves planes whose CTBs are at least
// 16 samples on a side.
hipError_t launch_sao(const SaoJob* jobs, const uint32_t* block_job, const uint32_t (&range)[5], hipStream_t stream);

// Encoder forward path: levels of every TU from its residual block.
struct FwdArgs {
    const rmh_tu_desc* tus;
    const int16_t* residual;
    int16_t* levels;
    const uint32_t* mat;
    const int8_t* lfnst8;
    const int8_t* lfnst4;
    const uint16_t* scans;
    int32_t mat_off[3][7];
    int32_t scan_off[7][7];
    int32_t scan_len[7][7];
    uint32_t n_tu;
};
hipError_t launch_forward(const FwdArgs& a, hipStream_t stream);

// Per-TU reconstruction SSE of a candidate batch (pred may be null).
struct SseArgs {
    const rmh_tu_desc* tus;
    const int16_t* residual;
    const int16_t* pred;
    const int16_t* recon;
    uint64_t* sse;
    uint32_t n_tu;
};
hipError_t launch_sse(const SseArgs& a, hipStream_t stream);

// One luma plane of an LMCS batch: samples mapped through lut[min(s, cap)].
struct LmcsJob {
    const void* src;
    void* dst;
    const uint16_t* lut;        // device, cap + 1 entries
    uint32_t width, height;
    uint32_t src_stride, dst_stride;  // samples
    uint32_t bps;               // bytes per sample
    uint32_t cap;               // 2^bd - 1
    uint32_t first_block;       // first workgroup of the job
    uint32_t reserved;
};
// Rows of a plane per LMCS workgroup. Workgroups [range[0], range[1]) map
// 8-bit planes, [range[1], range[2]) 16-bit ones.
constexpr int kLmcsRows = 8;
hipError_t launch_lmcs(const LmcsJob* jobs, const uint32_t* block_job, const uint32_t (&range)[3], hipStream_t stream);

// ALF luma tiles are 64 x 16 samples, one 256-lane workgroup each; chroma
// tiles 64 x 16 chroma samples.
constexpr int kAlfTileW = 64, kAlfTileH = 16;
// Filter bank of one luma APS and the 16 fixed sets: [bank 0..16][class
// 0..24][transpose 0..3][coefficient 0..11, clip value 0..11], with the
// transposition already applied (bank 16 is the APS).
constexpr int kAlfBankEntries = 17 * 25 * 4 * 24;

struct AlfLumaJob {
    const void* src;
    void* dst;
    const rmh_alf_ctb* ctbs;
    const int16_t* bank;          // kAlfBankEntries
    uint32_t src_stride, dst_stride;
    uint32_t width, height;       // plane
    uint32_t vis_w, vis_h;
    uint32_t bw, bh;              // min(visible, plane): the clamp of every read
    uint32_t bps, bd;
    uint32_t ctb_log2, ctb_cols, ctb_xn, ctb_yn;  // CTBs of the loop: ceil(vis / ctb)
    uint32_t has_aps;
    uint32_t tiles_x, first_block;
    uint32_t reserved;
};

struct AlfChromaJob {
    const void* src;
    void* dst;
    const void* luma;             // pre-ALF luma (CCALF input)
    const rmh_alf_ctb* ctbs;
    const int16_t* chroma_bank;   // [alternative][6 coefficients, 6 clip values]; null: chroma ALF off
    const int16_t* cc_bank;       // [filter][7]; null: CCALF off
    uint32_t src_stride, dst_stride, luma_stride;
    uint32_t width, height;       // chroma plane
    uint32_t cbw, cbh;            // clamp of chroma reads: min(ceil(vis / 2), plane)
    uint32_t lw, lh;              // clamp of luma reads: min(vis, luma plane)
    uint32_t cvis_w, cvis_h, vis_w, vis_h;
    uint32_t bps, bd;
    uint32_t ctb_log2, ctb_cols, ctb_xn, ctb_yn;
    uint32_t comp;                // 1 Cb, 2 Cr
    uint32_t n_chroma, n_cc;
    uint32_t tiles_x, first_block;
    uint32_t reserved;
};

// Workgroups [range[0], range[1]) of each kind serve 8-bit planes,
// [range[1], range[2]) 16-bit ones.
hipError_t launch_alf(const AlfLumaJob* luma, const uint32_t* luma_block_job, const uint32_t (&luma_range)[3],
                      const AlfChromaJob* chroma, const uint32_t* chroma_block_job, const uint32_t (&chroma_range)[3],
                      hipStream_t stream);

}  // namespace rmh

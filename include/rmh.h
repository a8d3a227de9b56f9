/*
 * rmh.h - C ABI of librmh, the MI355X (gfx950) accelerator for the data-parallel
 * stages of VVC (H.266) reconstruction.
 *
 * The library replaces three CPU stages of rustmpeg-video's VVC decoder:
 *
 *   residual  dequantisation (flat, dependent, transform skip) + inverse LFNST +
 *             inverse DCT-II / DST-VII / DCT-VIII over every transform unit (TU)
 *   inter MC  8-tap luma / 4-tap chroma fractional-sample interpolation
 *   deblock   luma and chroma deblocking of a whole picture
 *
 * Every entry point takes plain pointers and sizes. Host-memory entry points
 * copy in, run and copy out; on any error the caller's outputs are left
 * untouched, so a caller can fall back to its own CPU path (the contract of
 * crates/rustmpeg-video/src/av2/deblock.rs:30-35). Plan entry points keep the
 * work description resident on the device and run on a caller-supplied HIP
 * stream against device buffers; they are what a frame-step hipGraph captures.
 *
 * Status codes follow AcceleratorError (crates/rustmpeg-video/src/accelerator.rs:18-32).
 */
#ifndef RMH_H
#define RMH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RMH_ABI_VERSION 1u

/* ------------------------------------------------------------------------ */
/* Status and context                                                        */
/* ------------------------------------------------------------------------ */

/* Mirrors AcceleratorError {Unavailable, InvalidInput, LimitExceeded, Execution}
 * (crates/rustmpeg-video/src/accelerator.rs:18-32). RMH_TIMEOUT is reported by
 * bounded waits and maps to Execution, as the Vulkan read-back timeout does
 * (crates/rustmpeg-vulkan/src/compute/mod.rs:416-418, 833-847). */
typedef enum rmh_status {
    RMH_OK = 0,
    RMH_UNAVAILABLE = 1,
    RMH_INVALID_INPUT = 2,
    RMH_LIMIT_EXCEEDED = 3,
    RMH_EXECUTION = 4,
    RMH_TIMEOUT = 5
} rmh_status;

typedef struct rmh_ctx rmh_ctx;
typedef struct rmh_plan rmh_plan;
typedef struct rmh_graph rmh_graph;

typedef struct rmh_config {
    int32_t device;       /* HIP device ordinal */
    uint32_t timeout_ms;  /* bound on every blocking wait; 0 selects 30000 ms
                             (DEFAULT_TIMEOUT, crates/rustmpeg-vulkan/src/compute/mod.rs:39) */
    uint32_t flags;       /* reserved, must be 0 */
} rmh_config;

typedef struct rmh_device_info {
    char name[64];
    char arch[32];             /* e.g. "gfx950" */
    int32_t compute_units;
    int32_t max_clock_khz;
    uint64_t global_mem_bytes;
    uint32_t abi_version;
    uint32_t reserved;
} rmh_device_info;

/* Replaces Compute::new (crates/rustmpeg-vulkan/src/compute/mod.rs:112-160):
 * opens the device and compiles nothing at run time (kernels are prebuilt). */
rmh_status rmh_create(const rmh_config* cfg, rmh_ctx** out);
void rmh_destroy(rmh_ctx* ctx);
/* Text of the last error on this context (thread-local per context). */
const char* rmh_last_error(const rmh_ctx* ctx);
rmh_status rmh_info(rmh_ctx* ctx, rmh_device_info* out);
/* Sticky failure latch, as Accelerator::disable_gpu / gpu_failed
 * (crates/rustmpeg-vulkan/src/accelerator.rs:38,106-109,163-171): once set,
 * every call returns RMH_UNAVAILABLE. */
void rmh_disable(rmh_ctx* ctx);
int rmh_is_disabled(const rmh_ctx* ctx);
/* Test hook: the next `count` device calls fail with RMH_EXECUTION before any
 * device work is enqueued, so callers can test their fallback without a fault. */
void rmh_inject_failures(rmh_ctx* ctx, uint32_t count);

/* ------------------------------------------------------------------------ */
/* Residual stage                                                            */
/* ------------------------------------------------------------------------ */

enum { RMH_TR_DCT2 = 0, RMH_TR_DST7 = 1, RMH_TR_DCT8 = 2 };  /* TrType, vvc/dct.rs:27-34 */

enum {
    RMH_TU_TRANSFORM_SKIP = 1u, /* dequantize_ts, no transform (vvc/transform.rs:229-246) */
    RMH_TU_DEP_QUANT = 2u,      /* dependent quantisation (vvc/dct.rs:1191-1248) */
    RMH_TU_QUANT_INTER = 4u     /* forward plans only: the inter rounding offset of quantisation
                                   (85 instead of 171, vvc/dct.rs:1134, transform.rs:422) */
};

/* One transform unit. 16 bytes; the descriptor fields of TuData / CuPlan
 * (crates/rustmpeg-video/src/vvc/decode_general.rs:773-811) that the residual
 * driver inverse_residual (decode_general.rs:4125-4199) consumes.
 *
 * The TU's coefficients are a raster w*h block of int16 levels starting at
 * coeff[offset]; its residual is written as a raster w*h int16 block at
 * residual[offset]. Columns >= w - skip or rows >= h - skip (after the
 * default zero-out of vvc/dct.rs:1063-1077 and, with LFNST, the LFNST
 * zero-out of decode_general.rs:4233-4236) are ignored. */
typedef struct rmh_tu_desc {
    uint32_t offset;     /* element offset of the TU block in coeff[] / residual[] */
    uint8_t log2_w;      /* 0..6 (widths 1..64; w*h >= 2) */
    uint8_t log2_h;      /* 0..6 */
    uint8_t qp;          /* Qp' = QP + 6*(bit_depth-8), already clamped as in decode_general.rs:4135-4138 */
    uint8_t bit_depth;   /* 8..12 */
    uint8_t tr_hor;      /* RMH_TR_* */
    uint8_t tr_ver;      /* RMH_TR_* */
    uint8_t skip_w;      /* extra zero-out width (the extra_skip_width of vvc/dct.rs:1002) */
    uint8_t skip_h;      /* extra zero-out height */
    uint8_t lfnst_idx;   /* 0 (off), 1, 2 */
    uint8_t lfnst_mode;  /* LFNST intra-mode index 0..94 (vvc/lfnst.rs:85-95) */
    uint8_t flags;       /* RMH_TU_* */
    uint8_t joint;       /* joint Cb-Cr (ICT) of a chroma TU pair, recon plans only: 0 none;
                          * else TuCResMode 1..3 (TuData::joint, decode_general.rs:773-783) in
                          * bits 0-1 and ph_joint_cbcr_sign_flag in bit 2. The TU carries the coded
                          * component (Cr for mode 1, Cb for modes 2 and 3, its placement's plane);
                          * the other component's residual is derived from it as
                          * reconstruct_joint_cbcr does (decode_general.rs:3908-3981): mode 1 Cb =
                          * (+-Cr) >> 1, mode 2 Cr = (+-Cb) >> 1, mode 3 Cr = +-Cb, and added to
                          * the same block of the other chroma plane */
} rmh_tu_desc;

/* Host memory, synchronous. `coeff` and `residual` hold n_coeff int16 each.
 * The new VVC trait VvcResidualBatch::residuals maps onto this call. */
rmh_status rmh_residual_batch(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu,
                              const int16_t* coeff, size_t n_coeff, int16_t* residual);

/* Plan: validates and uploads the descriptors once. */
rmh_status rmh_residual_plan_create(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu,
                                    size_t n_coeff, rmh_plan** out);
/* Device memory, asynchronous on `stream` (a hipStream_t, NULL = default stream). */
rmh_status rmh_residual_plan_run(rmh_plan* plan, const int16_t* d_coeff, int16_t* d_residual,
                                 void* stream);
/* Unbatched baseline: one launch per TU range [first_tu, first_tu + count) in
 * descriptor order (the same kernel, a smaller grid). */
rmh_status rmh_residual_plan_run_range(rmh_plan* plan, size_t first_tu, size_t count,
                                       const int16_t* d_coeff, int16_t* d_residual, void* stream);

/* Autotuner (no reference counterpart; part of batching the frame step): for
 * each TU size class of the plan (largest first), times the plan's tiled
 * launch with that class's tiles holding 1, 1/2, 1/4 and 1/8 of the default
 * TUs per workgroup on the given buffers (the fastest of three batches of
 * reps launches, 0 selects 5) and keeps a change that saves at least 2 %. Results are
 * unchanged (tiling does not change the arithmetic); the buffers are
 * overwritten. Synchronous. split_out (49 bytes, [log2 w][log2 h], may be
 * NULL) receives the chosen split per class. */
rmh_status rmh_residual_plan_autotune(rmh_plan* plan, const int16_t* d_coeff, int16_t* d_residual,
                                      void* stream, uint32_t reps, uint8_t* split_out);

/* ------------------------------------------------------------------------ */
/* Planes                                                                    */
/* ------------------------------------------------------------------------ */

/* A sample plane: `bytes_per_sample` 1 (8-bit content) or 2 (up to 12-bit),
 * row-major with a row pitch of `stride` samples, like Plane / Plane16
 * (crates/rustmpeg-video/src/vvc/predict.rs:987-1000, 2866-2945). */
typedef struct rmh_plane {
    void* data;
    uint32_t width;
    uint32_t height;
    uint32_t stride;            /* samples */
    uint8_t bytes_per_sample;   /* 1 or 2 */
    uint8_t bit_depth;          /* 8..12 */
    uint16_t reserved;
} rmh_plane;

/* Inter reconstruction (the add of reconstruct_tus / reconstruct_component_full,
 * crates/rustmpeg-video/src/vvc/decode_general.rs:3885-4096, slice.rs:1920-1977):
 * instead of writing int16 residuals, a recon plan adds each TU's residual to
 * the prediction already in a picture plane and clips to [0, 2^bd - 1], in
 * place. TU i lands at (x, y) of plane `plane` of frame `frame`. */
typedef struct rmh_tu_place {
    uint16_t frame;
    uint8_t plane;     /* 0 Y, 1 Cb, 2 Cr */
    uint8_t reserved;  /* must be 0 */
    uint16_t x;
    uint16_t y;
} rmh_tu_place;

typedef struct rmh_recon_frame {
    rmh_plane planes[3];  /* device addresses; bit_depth must match the TUs */
} rmh_recon_frame;

/* Recon plan; run it with rmh_residual_plan_run(plan, d_coeff, NULL, stream). */
rmh_status rmh_residual_recon_plan_create(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu, size_t n_coeff,
                                          const rmh_tu_place* place, const rmh_recon_frame* frames,
                                          size_t n_frames, rmh_plan** out);

/* ------------------------------------------------------------------------ */
/* Inter prediction (uni-prediction, 4:2:0)                                  */
/* ------------------------------------------------------------------------ */

enum { RMH_PU_ALT_HPEL = 0x10u };  /* use LUMA_ALT_HPEL_FILTER_8TAP at frac 8 (vvc/inter.rs:80) */

/* One prediction unit, 16 bytes. Luma block (x, y, 1<<log2_w, 1<<log2_h) is
 * predicted from frame `frame`'s reference with the luma MV in 1/16 sample;
 * chroma uses the same MV at 1/32 (vvc/inter.rs:1746-1749, 1826-1827). */
typedef struct rmh_pu_desc {
    uint16_t x;
    uint16_t y;
    uint8_t log2_w;      /* low 4 bits: log2 width (3..7); high bits: RMH_PU_* flags */
    uint8_t log2_h;      /* log2 height (3..7) */
    uint16_t frame;      /* index into the frame table */
    int32_t mv_x;        /* 1/16 luma sample */
    int32_t mv_y;
} rmh_pu_desc;

typedef struct rmh_mc_frame {
    rmh_plane ref[3];    /* Y, Cb, Cr reference (4:2:0) */
    rmh_plane dst[3];    /* prediction output */
} rmh_mc_frame;

rmh_status rmh_mc_batch(rmh_ctx* ctx, const rmh_mc_frame* frames, size_t n_frames,
                        const rmh_pu_desc* pus, size_t n_pu);
/* Planes in `frames` hold device addresses. */
rmh_status rmh_mc_plan_create(rmh_ctx* ctx, const rmh_mc_frame* frames, size_t n_frames,
                              const rmh_pu_desc* pus, size_t n_pu, rmh_plan** out);
rmh_status rmh_mc_plan_run(rmh_plan* plan, void* stream);
/* Autotuner of an inter-prediction plan, uni or bi (no reference
 * counterpart; part of batching the frame step, as the residual autotuner):
 * times the plan's interpolation launch with 4, 2 and 1 work items (one wave
 * each) per workgroup, the fastest of three batches of five launches, and
 * keeps the fastest shape (a change needs a 2 % gain); *waves_out (may be
 * NULL) receives it. Outputs do not depend on the shape; the runs write the
 * plan's output planes. Refused while a graph holds the plan. */
rmh_status rmh_mc_plan_autotune(rmh_plan* plan, void* stream, uint32_t* waves_out);

/* Bi-prediction (inter_dir 3) with CU-level weights, replacing the bi branch
 * of predict_inter_pu (vvc/inter.rs:1753-1801; DMVR and BDOF per PU, below): both
 * lists are interpolated to 14-bit intermediates (interpolate_luma_block /
 * interpolate_chroma_block_420 with for_bi_pred, vvc/inter.rs:886-1182; the
 * 8..12-bit forms vvc/inter.rs:2010-2134, 2543-2670) and combined by
 * weighted_average_bi_16 (vvc/inter.rs:2676-2696; 8-bit weighted_average_bi,
 * vvc/inter.rs:1189-1201). Every plane of a frame shares one bit depth. */
typedef struct rmh_mc_bi_frame {
    rmh_plane ref0[3];   /* L0 reference (Y, Cb, Cr, 4:2:0) */
    rmh_plane ref1[3];   /* L1 reference, same sizes */
    rmh_plane dst[3];    /* prediction output */
} rmh_mc_bi_frame;

/* One bi-predicted PU, 28 bytes. */
typedef struct rmh_bi_pu_desc {
    uint16_t x;
    uint16_t y;
    uint8_t log2_w;      /* low 4 bits: log2 width (3..7); high bits: RMH_PU_* flags */
    uint8_t log2_h;      /* log2 height (3..7) */
    uint16_t frame;      /* index into the bi frame table */
    int32_t mv0_x;       /* L0 MV, 1/16 luma sample */
    int32_t mv0_y;
    int32_t mv1_x;       /* L1 MV */
    int32_t mv1_y;
    uint8_t bcw_idx;     /* L1 weight BCW_WEIGHTS[min(bcw_idx, 4)] / 8 of {-2, 3, 4, 5, 10}
                          * (vvc/inter.rs:128); L0 weight 8 minus that; 2 = equal weights */
    uint8_t refine;      /* RMH_BI_DMVR | RMH_BI_BDOF, 8-bit frames only; 0 otherwise */
    uint8_t reserved[2]; /* zero */
} rmh_bi_pu_desc;

/* Decoder-side refinements of the 8-bit bi branch of predict_inter_pu
 * (vvc/inter.rs:1763-1800). The caller sets a flag when the conditions that
 * need its state hold: RMH_BI_DMVR for the SPS / picture-header enables, a
 * merge CU and references at symmetric POC distances (sym_poc); RMH_BI_BDOF
 * for the enables and references on opposite sides (opp_poc). The library
 * then applies, as the reference does:
 *   DMVR (refine_dmvr_subblock, inter.rs:1371-1459) over the whole PU when
 *        bcw_idx is 2 and the PU is at least 8x8 with 128 samples: the MV pair
 *        is refined by a bilateral search over the 5x5 integer offsets (L0 +
 *        offset, L1 - offset) on bilinear predictions, every second row, and a
 *        parabolic sub-sample step; both luma and chroma use the refined pair;
 *   BDOF (apply_bdof_luma, inter.rs:1218-1339) on luma when bcw_idx is 2, in
 *        place of the weighted average. */
enum { RMH_BI_DMVR = 1u, RMH_BI_BDOF = 2u };

/* Host memory, synchronous (as rmh_mc_batch, which it mirrors). */
rmh_status rmh_mc_bi_batch(rmh_ctx* ctx, const rmh_mc_bi_frame* frames, size_t n_frames,
                           const rmh_bi_pu_desc* pus, size_t n_pu);
/* Planes in `frames` hold device addresses; run with rmh_mc_plan_run. */
rmh_status rmh_mc_bi_plan_create(rmh_ctx* ctx, const rmh_mc_bi_frame* frames, size_t n_frames,
                                 const rmh_bi_pu_desc* pus, size_t n_pu, rmh_plan** out);

/* Geometric partitioning (GPM): the blend of predict_gpm_block
 * (vvc/inter.rs:1673-1730) that the general decoder applies to a GPM coding
 * unit's two merge-candidate predictions (decode_general.rs:2716-2741). The
 * caller predicts the two candidates of every GPM PU into the planes pred0
 * and pred1 (MC plans writing into those planes); this plan blends them into
 * dst with the mask of each PU's split direction (derive_gpm_params_table,
 * inter.rs:1641-1671, over GEO_DIS / GEO_ANGLE_TO_MASK, inter.rs:131-140):
 * per sample w0 = clip3(0, 8, (32 + idx + 4) >> 3) from the signed distance
 * idx of the sample to the split line, out = clip((w0 p0 + (8 - w0) p1 + 4)
 * >> 3, 0, 255). Positions go onto the luma grid scaled by 1 << luma_scale
 * (luma) and 1 << chroma_scale (chroma, whose blocks are the PU's halves):
 * the reference's decoder passes 1 and 2 (RMH_GPM_*_SCALE below). Samples
 * are 8-bit; every plane of a frame has the same size as its dst. */
typedef struct rmh_gpm_frame {
    rmh_plane pred0[3];  /* candidate 0's prediction (Y, Cb, Cr) */
    rmh_plane pred1[3];  /* candidate 1's prediction */
    rmh_plane dst[3];    /* the blended prediction */
} rmh_gpm_frame;

typedef struct rmh_gpm_pu {
    uint16_t x;
    uint16_t y;
    uint8_t log2_w;      /* 3..7 */
    uint8_t log2_h;      /* 3..7 */
    uint16_t frame;
    uint8_t split_dir;   /* merge_gpm_partition_idx, 0..63 */
    uint8_t reserved[3]; /* zero */
} rmh_gpm_pu;

enum { RMH_GPM_LUMA_SCALE = 1, RMH_GPM_CHROMA_SCALE = 2 };

/* Device planes; luma_scale and chroma_scale 0..3. Run with rmh_gpm_plan_run
 * or in a graph. */
rmh_status rmh_gpm_plan_create(rmh_ctx* ctx, const rmh_gpm_frame* frames, size_t n_frames, const rmh_gpm_pu* pus,
                               size_t n_pu, uint32_t luma_scale, uint32_t chroma_scale, rmh_plan** out);
rmh_status rmh_gpm_plan_run(rmh_plan* plan, void* stream);

/* ------------------------------------------------------------------------ */
/* Deblocking                                                                */
/* ------------------------------------------------------------------------ */

/* Edge byte of one 4-sample edge segment:
 *   bits 0-1  bS (0 = not filtered, 1, 2)
 *   bits 2-3  maxFilterLengthP code c, length = 2c+1 (1, 3, 5, 7)
 *   bits 4-5  maxFilterLengthQ code
 *   bit  6    luma horizontal edge on a CTB row boundary (vvc/filter.rs:338)
 * The map holds 4 bytes per 4x4 luma cell (cell (i, j) covers luma x 4i..4i+3,
 * y 4j..4j+3):
 *   [0] luma vertical edge at x = 4i, rows 4j..4j+3
 *   [1] luma horizontal edge at y = 4j, columns 4i..4i+3
 *   [2] chroma vertical edge at chroma x = 2i, chroma rows 2j..2j+1
 *   [3] chroma horizontal edge at chroma y = 2j, chroma columns 2i..2i+1
 * (one decision group per segment, as deblock_luma_edge_bd, vvc/filter.rs:638,
 * and deblock_chroma_edge with two lines, vvc/filter.rs:935). */
enum {
    RMH_EDGE_BS_MASK = 0x03u,
    RMH_EDGE_LENP_SHIFT = 2,
    RMH_EDGE_LENQ_SHIFT = 4,
    RMH_EDGE_CTB_ROW = 0x40u
};

typedef struct rmh_dbk_frame {
    rmh_plane planes[3];         /* Y, Cb, Cr, filtered in place */
    const uint8_t* edges;        /* grid_w * grid_h * 4 bytes */
    uint32_t grid_w;             /* ceil(luma width / 4) */
    uint32_t grid_h;             /* ceil(luma height / 4) */
    uint8_t qp[3];               /* QpY, QpCb, QpCr (tC/beta index base, vvc/filter.rs:653) */
    int8_t beta_offset_div2[3];
    int8_t tc_offset_div2[3];
    uint8_t reserved[3];
} rmh_dbk_frame;

rmh_status rmh_deblock(rmh_ctx* ctx, const rmh_dbk_frame* frames, size_t n_frames);
/* Planes and `edges` in `frames` hold device addresses. */
rmh_status rmh_deblock_plan_create(rmh_ctx* ctx, const rmh_dbk_frame* frames, size_t n_frames,
                                   rmh_plan** out);
rmh_status rmh_deblock_plan_run(rmh_plan* plan, void* stream);
/* Out-of-place deblocking plan (no counterpart in the reference, whose
 * DeblockFilter filters in place; for frame-step graphs that chain deblocking
 * into SAO): frames[f].planes are read only, dst[3 f + c] (same size, bit
 * depth and sample size, any pitch, not overlapping a source plane) receives
 * every sample of the deblocked component c of frame f. It saves the copy of
 * each plane that the in-place form makes first. Run with rmh_deblock_plan_run.
 * Every sample of the source planes reaches dst, filtered or not, so a frame
 * step that builds the picture in scratch planes must write every sample of
 * them first (including samples no inter PU or TU covers). */
rmh_status rmh_deblock_to_plan_create(rmh_ctx* ctx, const rmh_dbk_frame* frames, size_t n_frames,
                                      const rmh_plane* dst, rmh_plan** out);

/* Edge-map builder: the edge bytes above, derived on the device from the
 * coded picture's CU / TU maps as the general decoder derives them
 * (apply_deblock, vvc/decode_general.rs:4977-5187, with deblock_bs
 * 4939-4975, luma_lens / chroma_lens 4916-4935 and derive_boundary_strength,
 * vvc/filter.rs:202-226):
 *   luma   an edge wherever the TU ids of two 4x4 cells differ (both inside a
 *          TU): bS 2 if either CU is intra, else 1 if either cell's luma cbf
 *          is set, else 1 if the reference POCs differ or an MV component
 *          differs by >= 8 (1/16 sample), else 0; lengths luma_lens of the
 *          two TU widths (vertical) or heights (horizontal); horizontal edges
 *          on y % ctb_size == 0 carry RMH_EDGE_CTB_ROW;
 *   chroma edges on the 8-sample chroma grid where the chroma TU ids differ,
 *          bS 2 (the general decoder always passes 2), lengths chroma_lens
 *          of the chroma TU sizes, (1, 3) on horizontal CTB-row boundaries.
 * Edges are taken inside the visible picture only, as apply_deblock's loops
 * run. One cell per 4x4 luma block, row-major, grid_w x grid_h: */
typedef struct rmh_edge_cell {
    uint32_t tu_y;     /* luma TU id, UINT32_MAX outside any TU */
    uint32_t tu_c;     /* chroma TU id of the co-located 2x2 chroma cell, UINT32_MAX if none */
    uint8_t tu_wy;     /* luma TU width and height (4..64) */
    uint8_t tu_hy;
    uint8_t tu_wc;     /* chroma TU width and height in chroma samples (2..32) */
    uint8_t tu_hc;
    uint8_t flags;     /* RMH_CELL_* */
    uint8_t reserved[3];
    int32_t mv_x;      /* the MV and reference POC of the first list with a reference */
    int32_t mv_y;      /* (deblock_bs's `field`, decode_general.rs:4951-4958), 1/16 sample */
    int32_t ref_poc;
} rmh_edge_cell;

enum { RMH_CELL_INTRA = 1u, RMH_CELL_CBF = 2u };

typedef struct rmh_edge_frame {
    const rmh_edge_cell* cells;  /* grid_w * grid_h cells */
    uint8_t* edges;              /* output: grid_w * grid_h * 4 edge bytes (rmh_dbk_frame.edges) */
    uint32_t grid_w;
    uint32_t grid_h;
    uint32_t vis_w;              /* visible luma size: 4 * grid - 3 .. 4 * grid */
    uint32_t vis_h;
    uint32_t ctb_size;           /* luma CTB size (power of two, 16..128) */
    uint32_t reserved;
} rmh_edge_frame;

/* Host cells in, host edge bytes out (written only on RMH_OK). */
rmh_status rmh_edge_map(rmh_ctx* ctx, const rmh_edge_frame* frames, size_t n_frames);
/* Device cells and edge buffers; run with rmh_edge_map_plan_run. */
rmh_status rmh_edge_map_plan_create(rmh_ctx* ctx, const rmh_edge_frame* frames, size_t n_frames, rmh_plan** out);
rmh_status rmh_edge_map_plan_run(rmh_plan* plan, void* stream);

/* In-place deblocking that allocates nothing. No counterpart in the
 * reference: it serves callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same frames, results and validation as rmh_deblock_plan_create
 * followed by rmh_deblock_plan_run, enqueued on `stream`, but the frame tables
 * and the copies of the planes that the filter reads live in one
 * library-owned device buffer per GPU of RMH_NOALLOC_SCRATCH_BYTES, held until
 * the process ends: at most RMH_NOALLOC_PLANE_BYTES of plane copies (each
 * plane's rows padded to 16 bytes, each plane to 256: one 3840x2160 10-bit
 * frame, or five 1920x1080 10-bit frames) and 64 KiB of tables (a few hundred
 * frames). Larger batches return RMH_LIMIT_EXCEEDED before any device work.
 * Every call on a GPU shares that GPU's buffer: calls that overlap on one GPU
 * (on different streams, or from threads on different streams) race and must
 * be serialised by the caller; calls on one stream are ordered. */
#define RMH_NOALLOC_SCRATCH_BYTES (32u << 20)
#define RMH_NOALLOC_PLANE_BYTES ((32u << 20) - (64u << 10))
rmh_status rmh_deblock_noalloc(rmh_ctx* ctx, const rmh_dbk_frame* frames, size_t n_frames, void* stream);

/* ------------------------------------------------------------------------ */
/* Sample adaptive offset (after deblocking)                                 */
/* ------------------------------------------------------------------------ */

/* SAO of one component of one CTB, the SaoType enum
 * (crates/rustmpeg-video/src/vvc/filter.rs:1225-1248):
 *   RMH_SAO_OFF   nothing;
 *   RMH_SAO_BAND  band offset: band_or_class = band_pos (0..31), offsets for
 *                 bands band_pos .. band_pos + 3 (mod 32);
 *   RMH_SAO_EDGE  edge offset: band_or_class = eo_class (0: horizontal,
 *                 1: vertical, 2: 135 degrees, 3: 45 degrees), offsets for
 *                 edge categories 1..4.
 * Offsets are in sample units at every bit depth, as the reference applies
 * them (filter.rs:1968-1975). */
enum { RMH_SAO_OFF = 0, RMH_SAO_BAND = 1, RMH_SAO_EDGE = 2 };

typedef struct rmh_sao_param {
    uint8_t type;
    uint8_t band_or_class;
    uint8_t reserved[2];
    int32_t offsets[4];
} rmh_sao_param;

/* One picture of apply_sao (crates/rustmpeg-video/src/vvc/decode_general.rs:
 * 5189-5222), each CTB through apply_sao_region_bounded_16 (vvc/filter.rs:
 * 1954-2043): dst = src with the CTB's offsets applied inside the visible
 * region (vis_w x vis_h luma, ceil(vis/2) chroma); every other sample, and
 * every edge-offset sample with a neighbour outside the visible region or the
 * plane, is copied unchanged. src is read only; dst must not overlap it.
 * params[(cy * ctb_cols + cx) * 3 + c] is the SAO of component c (Y, Cb, Cr)
 * of CTB (cx, cy); chroma CTBs are ctb_size / 2 on a side (4:2:0). A decoder
 * whose slice has sao_chroma_used = 0 passes RMH_SAO_OFF for Cb and Cr. */
typedef struct rmh_sao_frame {
    rmh_plane src[3];            /* deblocked Y, Cb, Cr */
    rmh_plane dst[3];            /* SAO output, same sizes and pitch */
    const rmh_sao_param* params; /* ctb_cols * ctb_rows * 3 entries */
    uint32_t ctb_size;           /* luma CTB size: 16, 32, 64 or 128 */
    uint32_t ctb_cols;           /* row pitch of params in CTBs, >= ceil(vis_w / ctb_size) */
    uint32_t ctb_rows;           /* >= ceil(vis_h / ctb_size) */
    uint32_t vis_w;              /* visible luma size (<= the Y plane's) */
    uint32_t vis_h;
    uint32_t reserved;
} rmh_sao_frame;

/* Host planes and params: uploads, runs and writes dst (written only on
 * RMH_OK). Replaces the decoder's apply_sao loop over apply_sao_region_bounded_16. */
rmh_status rmh_sao(rmh_ctx* ctx, const rmh_sao_frame* frames, size_t n_frames);
/* Planes and params in `frames` hold device addresses; a plan keeps the frame
 * table and the work index on the device and runs on any stream. */
rmh_status rmh_sao_plan_create(rmh_ctx* ctx, const rmh_sao_frame* frames, size_t n_frames,
                               rmh_plan** out);
rmh_status rmh_sao_plan_run(rmh_plan* plan, void* stream);

/* ------------------------------------------------------------------------ */
/* Encoder side: forward transform and quantisation                          */
/* ------------------------------------------------------------------------ */

/* The encoder's forward path over a batch of TUs, in the order of
 * plan_primary (crates/rustmpeg-video/src/vvc/slice.rs:718-761) and the 8..12-bit
 * encoder (slice.rs:4279-4296): each TU's residual block (a raster w*h int16
 * block at residual[offset], |value| < 2^15) goes through
 * forward_transform_rect_bd (vvc/dct.rs:865-991) with the TU's zero-out (widened
 * to the LFNST region when lfnst_idx, as on the decoder side), forward_lfnst_rect
 * (vvc/lfnst.rs:131-204), then quantize_rect_bd (dct.rs:1113-1147; inter
 * rounding with RMH_TU_QUANT_INTER) or quantize_dep_quant_bd (dct.rs:1252-1301)
 * with RMH_TU_DEP_QUANT; transform-skip TUs through quantize_ts_with_mode
 * (vvc/transform.rs:415-427). The levels are written as a raster w*h int16
 * block at levels[offset]. The descriptors follow rmh_tu_desc (joint must be 0);
 * the plan also holds the decoder-side residual plan of the same TUs. */
rmh_status rmh_forward_plan_create(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu, size_t n_coeff,
                                   rmh_plan** out);
/* Device buffers of n_coeff int16; asynchronous on `stream`. */
rmh_status rmh_forward_plan_run(rmh_plan* plan, const int16_t* d_residual, int16_t* d_levels, void* stream);
/* Candidate evaluation, as encode_luma scores a transform choice
 * (slice.rs:817-866): the levels as above, the reconstructed residual the
 * decoder derives from them (the residual plan) into d_recon, and per TU i
 * d_sse[i] = sum over the block of (pred + residual - clip(pred + recon, 0,
 * 2^bd - 1))^2, the block_sse of the clipped reconstruction against the source
 * pred + residual; with d_pred NULL, sum of (residual - recon)^2. */
rmh_status rmh_forward_plan_evaluate(rmh_plan* plan, const int16_t* d_residual, const int16_t* d_pred,
                                     int16_t* d_levels, int16_t* d_recon, uint64_t* d_sse, void* stream);

/* ------------------------------------------------------------------------ */
/* Luma mapping with chroma scaling (LMCS)                                   */
/* ------------------------------------------------------------------------ */

/* The LMCS APS, as VvcLmcsAps (crates/rustmpeg-video/src/vvc/hls.rs:2723-2732). */
typedef struct rmh_lmcs_aps {
    uint8_t min_bin_idx;              /* lmcs_min_bin_idx, 0..15 */
    uint8_t max_bin_idx;              /* 15 - lmcs_delta_max_bin_idx */
    int8_t chroma_res_scaling_offset; /* LmcsDeltaCrs */
    uint8_t reserved;                 /* must be 0 */
    int16_t bin_cw_delta[16];         /* LmcsDeltaCW[i] */
} rmh_lmcs_aps;

/* The tables of LmcsReshaper16::from_aps (vvc/filter.rs:2628-2716) at
 * bit_depth 8..12: fwd_lut and inv_lut receive 2^bit_depth entries,
 * chroma_adj_lut 16 and reshape_pivot 17; any output may be NULL. Host
 * only; needs no context. */
rmh_status rmh_lmcs_tables(const rmh_lmcs_aps* aps, uint32_t bit_depth, uint16_t* fwd_lut, uint16_t* inv_lut,
                           int32_t* chroma_adj_lut, int32_t* reshape_pivot);

enum { RMH_LMCS_INVERSE = 0, RMH_LMCS_FORWARD = 1 };

/* One luma plane mapped sample by sample through the LUT of `aps`:
 * apply_inverse_plane_16 (vvc/filter.rs:2715-2722; the decoder applies it to
 * the whole luma plane after the slice, before deblocking,
 * decode_general.rs:1134-1138) or apply_forward_plane_16 (filter.rs:2707-2713).
 * A sample above 2^bd - 1 reads the last entry, as there. `dst` has the
 * size of `src`; dst.data == src.data maps in place, any other overlap is
 * rejected. */
typedef struct rmh_lmcs_frame {
    rmh_plane src;
    rmh_plane dst;
    rmh_lmcs_aps aps;
    uint32_t direction;   /* RMH_LMCS_INVERSE or RMH_LMCS_FORWARD */
} rmh_lmcs_frame;

/* Device planes; the LUTs are built and uploaded at plan creation. */
rmh_status rmh_lmcs_plan_create(rmh_ctx* ctx, const rmh_lmcs_frame* frames, size_t n_frames, rmh_plan** out);
rmh_status rmh_lmcs_plan_run(rmh_plan* plan, void* stream);

/* ------------------------------------------------------------------------ */
/* Adaptive loop filter (ALF) and cross-component ALF (CCALF)                */
/* ------------------------------------------------------------------------ */

/* Per-CTB switches, the decoder's alf_luma / alf_cb / alf_cr / ccalf_cb /
 * ccalf_cr entries (crates/rustmpeg-video/src/vvc/decode_general.rs:1399-1515). */
enum { RMH_ALF_LUMA_APS = 17 };
typedef struct rmh_alf_ctb {
    uint8_t luma;    /* 0 off; 1..16 fixed filter set luma - 1 (AlfLumaMode::Fixed);
                      * RMH_ALF_LUMA_APS the luma APS (AlfLumaMode::Aps); other values off */
    uint8_t cb;      /* 0 off; else chroma APS alternative cb - 1 (clamped to the last, as there) */
    uint8_t cr;
    uint8_t cc_cb;   /* CCALF filter idc: 0 off; 1..n_cc_cb; larger values off (decode_general.rs:5317) */
    uint8_t cc_cr;
    uint8_t reserved[3];
} rmh_alf_ctb;

/* The parts of an ALF APS (VvcAlfAps, vvc/hls.rs:2699-2718) that the filters
 * read. Clip indices are the ones the decoder applies: luma_clip all zero when
 * alf_luma_clip_flag is 0 (decode_general.rs:5466-5475); chroma_clip zero for
 * an alternative without signalled indices (5588). */
typedef struct rmh_alf_aps {
    int16_t luma_coeff[25][12];
    uint8_t luma_clip[25][12];      /* 2-bit clipping indices */
    uint8_t class_to_filter[25];    /* filter_coeff_delta_idx */
    uint8_t n_luma;                 /* luma_filters.len(), 1..25 when the APS serves luma */
    uint8_t n_chroma;               /* chroma_filters.len(), 0..8 */
    uint8_t n_cc_cb;                /* cc_cb_filters.len(), 0..4 */
    uint8_t n_cc_cr;
    uint8_t reserved;
    int16_t chroma_coeff[8][6];
    uint8_t chroma_clip[8][6];
    int16_t cc_cb[4][7];
    int16_t cc_cr[4][7];
} rmh_alf_aps;

/* One picture of apply_alf (decode_general.rs:5225-5333): per CTB the 7x7
 * luma ALF with 4x4 classification (alf_luma_block, 5411-5566,
 * derive_alf_class_4x4_vb_16, filter.rs:1558-1691), the 5x5 chroma ALF
 * (alf_chroma_block, 5568-5660) and CCALF from the pre-ALF luma added to the
 * chroma result (ccalf_block, 5335-5409), with the ALF virtual boundary 4
 * luma rows above each CTB row. src is read only (the SAO output); dst
 * receives every sample: filtered where the CTB's switches and the visible
 * region say, src elsewhere. The APS pointers are host memory read at plan
 * creation; NULL disables what that APS serves (a CTB asking for the luma APS
 * without one is left unfiltered, where the reference would stop). */
typedef struct rmh_alf_frame {
    rmh_plane src[3];
    rmh_plane dst[3];                /* same sizes as src; must not overlap src */
    const rmh_alf_ctb* ctbs;         /* device: ctb_cols * ctb_rows entries, row-major */
    const rmh_alf_aps* luma_aps;     /* the slice's first luma APS (alf_aps_ids_luma[0]) */
    const rmh_alf_aps* chroma_aps;   /* alf_aps_id_chroma */
    const rmh_alf_aps* cc_cb_aps;    /* cc_alf_cb_aps_id */
    const rmh_alf_aps* cc_cr_aps;    /* cc_alf_cr_aps_id */
    uint32_t ctb_size;               /* luma CTB size: 16, 32, 64 or 128 */
    uint32_t ctb_cols;               /* >= ceil(vis_w / ctb_size) */
    uint32_t ctb_rows;               /* >= ceil(vis_h / ctb_size) */
    uint32_t vis_w;                  /* visible luma size (<= the Y plane's) */
    uint32_t vis_h;
    uint32_t reserved;
} rmh_alf_frame;

/* The luma filter bank a plan builds from `aps` (NULL: none) at bit_depth
 * 8..12 and the kernels read: out[((bank * 25 + class) * 4 + transpose) * 24
 * + k], banks 0..15 the fixed filter sets and 16 the APS, k < 12 the
 * coefficient filtering tap k after the block's transposition
 * (ALF_TRANSPOSE_PERM, vvc/filter.rs:2036-2041) and k >= 12 that tap's clip
 * value. 17 * 25 * 4 * 24 entries; host only, needs no context. */
rmh_status rmh_alf_filter_bank(const rmh_alf_aps* aps, uint32_t bit_depth, int16_t* out);

/* Device planes and CTB switches; the filter banks are built at creation.
 * APS coefficients must lie in [-128, 128] (ALF) and [-64, 64] (CCALF). */
rmh_status rmh_alf_plan_create(rmh_ctx* ctx, const rmh_alf_frame* frames, size_t n_frames, rmh_plan** out);
rmh_status rmh_alf_plan_run(rmh_plan* plan, void* stream);

/* ------------------------------------------------------------------------ */
/* Plans and frame-step graphs                                               */
/* ------------------------------------------------------------------------ */

void rmh_plan_destroy(rmh_plan* plan);
/* Algorithmic HBM bytes one run of the plan moves (the roofline numerator). */
uint64_t rmh_plan_bytes(const rmh_plan* plan);

/* Captures the given plans, in order, into one hipGraph (one frame step). For
 * residual plans the device buffers are bound at capture time. */
typedef struct rmh_graph_node {
    rmh_plan* plan;
    const int16_t* d_coeff;      /* residual plans only */
    int16_t* d_residual;         /* residual plans only */
} rmh_graph_node;

rmh_status rmh_graph_create(rmh_ctx* ctx, const rmh_graph_node* nodes, size_t n_nodes,
                            rmh_graph** out);
rmh_status rmh_graph_launch(rmh_graph* graph, void* stream);
void rmh_graph_destroy(rmh_graph* graph);

/* ------------------------------------------------------------------------ */
/* Caller-side driver: intra-profile subset decode                           */
/* ------------------------------------------------------------------------ */

/* Two-pass reconstruction of one 8-bit 4:2:0 picture coded with the intra
 * subset (CTU = CU = TU = 32 luma / 16 chroma, PLANAR or DC per CTU, chroma
 * DM), the decode path of slice::decode_picture -> reconstruct_block
 * (crates/rustmpeg-video/src/vvc/slice.rs:458-632, 1050-1097) split as
 *   (i)  the caller parses the picture into a TU batch (three TUs per CTU in
 *        raster order: Y 32x32, Cb 16x16, Cr 16x16) and one mode per CTU
 *        (0 PLANAR, 1 DC);
 *   (ii) every residual of the picture is computed in one device batch;
 *   (iii) PLANAR / DC prediction (reference filter, PDPC) and the add run on
 *        the host in coding order, since each CTU predicts from its
 *        reconstructed neighbours.
 * width and height are multiples of 32; y/u/v receive the reconstruction
 * (row pitch = width, width / 2) and are untouched on error. */
rmh_status rmh_intra_subset_decode(rmh_ctx* ctx, uint32_t width, uint32_t height, const uint8_t* modes,
                                   const rmh_tu_desc* tus, size_t n_tu, const int16_t* coeff, size_t n_coeff,
                                   uint8_t* y, uint8_t* u, uint8_t* v);

/* ------------------------------------------------------------------------ */
/* Introspection (tests)                                                     */
/* ------------------------------------------------------------------------ */

/* The host PLANAR (mode 0) / DC (mode 1) predictor that rmh_intra_subset_decode
 * uses, for an n x n block (n = 4..32) at (bx, by) of an 8-bit plane with row
 * pitch `width`: predict_block's PLANAR / DC path (vvc/predict.rs:2619-2646:
 * gather_references, the [1 2 1] reference filter for luma PLANAR above 32
 * samples, predict_planar / predict_dc, apply_pdpc). Host code only. */
rmh_status rmh_intra_predict(const uint8_t* plane, uint32_t width, uint32_t height, uint32_t bx, uint32_t by,
                             uint32_t n, int is_luma, int mode, int32_t* pred);

/* Writes the n*n row-major integer matrix the kernels use for (type, n);
 * returns RMH_INVALID_INPUT for an unsupported pair (matrix_for_any,
 * vvc/dct.rs:251-256). */
rmh_status rmh_transform_matrix(int32_t type, int32_t n, int32_t* out);

/* ------------------------------------------------------------------------ */
/* Coding-tree split search (encoder side, host code)                        */
/* ------------------------------------------------------------------------ */

/* Split modes in the order of CodingTreeSplit (vvc/transform.rs:346-359). */
enum {
    RMH_SPLIT_NONE = 0,
    RMH_SPLIT_QT = 1,
    RMH_SPLIT_BT_HOR = 2,
    RMH_SPLIT_BT_VER = 3,
    RMH_SPLIT_TT_HOR = 4,
    RMH_SPLIT_TT_VER = 5
};

enum {
    RMH_RD_FULL = 0,      /* every split of every 32x32 node */
    RMH_RD_PRUNED = 1,    /* splits with probs >= p_prune; all six when max prob < p_fallback */
    RMH_RD_HEURISTIC = 2  /* the one split of choose_qt_mtt_partition_rdo (vvc/slice.rs:4127-4167) */
};

typedef struct rmh_rd_stats {
    double bits;          /* bits of the chosen splits: modelled (rmh_rd_encode_luma) or the
                             coded stream's length (rmh_rd_encode_luma_stream) */
    uint64_t sse;         /* luma SSE of the reconstruction */
    uint32_t blocks;      /* 32x32 nodes */
    uint32_t candidates;  /* splits coded */
    uint32_t fallbacks;   /* nodes searched fully because max prob < p_fallback */
    uint32_t reserved;
} rmh_rd_stats;

/* RD search over the six splits of each 32x32 node of an 8-bit luma picture
 * (width, height multiples of 32), nodes in raster order; each child is a
 * leaf CU (PLANAR/DC by SAD, DCT-II, intra quantisation at `qp`).
 * J = SSD + lambda * bits with lambda = 0.57 * 2^((qp - 12) / 3). The
 * reference has no full RD search; its only chooser is the heuristic above,
 * so this entry point has no reference counterpart. It replaces the split
 * choice of encode_picture's coding-tree loop (vvc/slice.rs:4127-4167).
 * `probs` (6 per node, RMH_RD_PRUNED only) come from rmh_cnn_forward.
 * Outputs: recon (width*height), splits (one per node), optional costs
 * (6 per node: J, or -1 for splits not coded). Runs on the calling thread. */
rmh_status rmh_rd_encode_luma(const uint8_t* src, uint32_t width, uint32_t height, uint32_t stride,
                              uint32_t qp, uint32_t mode, const float* probs, float p_prune,
                              float p_fallback, uint8_t* recon, uint8_t* splits, float* costs,
                              rmh_rd_stats* stats);

/* residual_coding of n_blocks TUs of width x height (powers of two, 1..64;
 * levels in raster order, w*h per TU, each TU with a non-zero level and
 * nothing outside its top-left 32x32) into one CABAC substream, each TU
 * followed by end_of_slice_segment_flag (1 after the last), contexts
 * initialised for an I slice at qp. Replaces encode_residual_with_options
 * (vvc/residual.rs:448-522) over BinEncoder (vvc/cabac.rs:477-745).
 * RMH_INVALID_INPUT when the stream exceeds stream_capacity (*stream_len then
 * holds the size needed). Host code. */
enum {
    RMH_CABAC_LUMA = 1,         /* luma contexts (else chroma) */
    RMH_CABAC_DEP_QUANT = 2,    /* dependent-quantisation state selects the sig contexts */
    RMH_CABAC_SIGN_HIDING = 4   /* sign data hiding (ignored with RMH_CABAC_DEP_QUANT) */
};
rmh_status rmh_cabac_residual_stream(const int32_t* levels, uint32_t width, uint32_t height, uint32_t n_blocks,
                                     uint32_t flags, uint32_t qp, uint8_t* stream, size_t stream_capacity,
                                     size_t* stream_len);

/* The same search with CABAC as the rate: every candidate is priced by the
 * bits it adds to the picture's CABAC substream (context state carried in
 * coding order), and the chosen candidates are written to `stream`. The
 * stream is a luma-only (4:0:0) I-slice substream with 32x32 CTUs: per CTU the
 * coding-tree split (QT to 16, one binary/ternary level, CUs down to 4), per
 * leaf CU the PLANAR/DC mode as MPM 0/1, the luma cbf and residual_coding,
 * then end_of_slice_segment_flag; contexts initialised for an I slice at qp.
 * Syntax and arithmetic follow the reference's encoder: BinEncoder
 * (vvc/cabac.rs:477-745), encode_residual (vvc/residual.rs:435-743),
 * encode_coding_tree_split (vvc/slice.rs:3850-3893), the MPM luma mode coding
 * (vvc/slice.rs:1656-1687) and the luma cbf (vvc/slice.rs:399). Reconstruction
 * uses this search's PLANAR/DC predictor above (no reference-sample filter or
 * PDPC), so the stream is decodable with that predictor. stats->bits is the
 * stream length in bits. RMH_INVALID_INPUT when the stream exceeds
 * stream_capacity (*stream_len then holds the size needed). */
rmh_status rmh_rd_encode_luma_stream(const uint8_t* src, uint32_t width, uint32_t height, uint32_t stride,
                                     uint32_t qp, uint32_t mode, const float* probs, float p_prune,
                                     float p_fallback, uint8_t* recon, uint8_t* splits, float* costs,
                                     rmh_rd_stats* stats, uint8_t* stream, size_t stream_capacity,
                                     size_t* stream_len);

/* Decodes a substream of rmh_rd_encode_luma_stream (width, height multiples
 * of 32, the same qp) into recon (width * height) and the split of each 32x32
 * CTU. Two passes around one device batch: the host parses the whole
 * picture (coding trees, modes, levels; parsing needs no reconstruction), every
 * TU's dequantisation and inverse transform runs in one rmh_residual_batch,
 * then the host predicts (the search's PLANAR/DC) and adds in coding order.
 * The parse follows the reference's decoder side: BinDecoder
 * (vvc/cabac.rs:761-1007), decode_coding_tree_split (vvc/slice.rs:3895-3946),
 * decode_residual_with_options (vvc/residual.rs:537-887), as the general-tree
 * intra decoder drives them (decode_general_picture, vvc/decode_general.rs:822).
 * RMH_INVALID_INPUT on a malformed stream. Synchronous. */
rmh_status rmh_luma_stream_decode(rmh_ctx* ctx, const uint8_t* stream, size_t stream_len, uint32_t width,
                                  uint32_t height, uint32_t qp, uint8_t* recon, uint8_t* splits);

/* ------------------------------------------------------------------------ */
/* CU-split predictor: BF16 CNN on MFMA (no reference counterpart: the      */
/* reference has no learned component; it serves RMH_RD_PRUNED above)        */
/* ------------------------------------------------------------------------ */

/* Input: 32x32 8-bit luma blocks (raster, 1024 bytes each) and QP / 51 per
 * block. conv3x3(1->32), conv3x3(32->32) x2, ReLU, global average pool,
 * MLP 33->64->6, softmax over the six split modes (RMH_SPLIT_*). Weights and
 * activations are stored as bf16, sums accumulate in fp32 on MFMA. */
typedef struct rmh_cnn rmh_cnn;

size_t rmh_cnn_param_count(void);
/* Seeded He-normal initialisation; buffers sized for batches of up to max_batch. */
rmh_status rmh_cnn_create(rmh_ctx* ctx, uint32_t max_batch, uint64_t seed, rmh_cnn** out);
void rmh_cnn_destroy(rmh_cnn* cnn);
/* fp32 master parameters, rmh_cnn_param_count() floats, host memory. Setting
 * them resets the optimiser state. */
rmh_status rmh_cnn_get_params(rmh_cnn* cnn, float* out);
rmh_status rmh_cnn_set_params(rmh_cnn* cnn, const float* in);
/* Device buffers: blocks n*1024 bytes, qp n floats, probs n*6 floats. */
rmh_status rmh_cnn_forward(rmh_cnn* cnn, const uint8_t* d_blocks, const float* d_qp, uint32_t n,
                           float* d_probs, void* stream);
/* Host buffers, synchronous (stages through device memory). Runs on the
 * network's own stream: calls on one network serialise, calls on different
 * networks of one context may overlap on the device. */
rmh_status rmh_cnn_predict(rmh_cnn* cnn, const uint8_t* blocks, const float* qp, uint32_t n, float* probs);
/* Gradient of the mean softmax cross-entropy against labels (0..5, device
 * bytes) into d_grads (rmh_cnn_param_count() floats) and the per-block
 * losses into d_loss (n floats), without changing the parameters. */
rmh_status rmh_cnn_gradients(rmh_cnn* cnn, const uint8_t* d_blocks, const float* d_qp, const uint8_t* d_labels,
                             uint32_t n, float* d_grads, float* d_loss, void* stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) at learning rate lr on the batch. */
rmh_status rmh_cnn_train_step(rmh_cnn* cnn, const uint8_t* d_blocks, const float* d_qp,
                              const uint8_t* d_labels, uint32_t n, float lr, float* d_loss, void* stream);
/* The same against target distributions over the six splits (d_targets: n*6
 * floats, each row summing to 1), e.g. a softmax of negated RD costs. */
rmh_status rmh_cnn_train_step_soft(rmh_cnn* cnn, const uint8_t* d_blocks, const float* d_qp,
                                   const float* d_targets, uint32_t n, float lr, float* d_loss, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RMH_H */

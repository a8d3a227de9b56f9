/*
 * vvc_oracle.h - CPU oracle for the VVC reconstruction stages.
 *
 * TEST INFRASTRUCTURE ONLY. This is a scalar C restatement of the reference
 * Rust arithmetic (crates/rustmpeg-video/src/vvc/), used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product (librmh) never links, loads or calls it.
 *
 * Every function names the reference function it restates. Arithmetic is in
 * int64 where the reference uses i64, so results are exact.
 */
#ifndef VVC_ORACLE_H
#define VVC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/rmh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tables (vvc/dct.rs:58-256, vvc/tr_tables.rs, vvc/lfnst.rs) ---- */
int orc_transform_matrix(int type, int n, int32_t* out);        /* matrix_for_any */
int orc_lfnst_set(int mode);                                    /* lfnst_set */
int orc_lfnst_transposes(int mode);                              /* transposes */
int orc_lfnst_coeff(int set, int idx, int big, int i, int j);    /* LFNST_8X8/LFNST_4X4 entry */
/* Grouped 4x4 diagonal scan raster positions (CoeffScan::new, vvc/transform.rs:37-79). */
int orc_coeff_scan(int w, int h, int32_t* raster);
void orc_cg_log2(int w, int h, int* lw, int* lh);               /* coefficient_group_log2 */

/* ---- residual stage ---- */
int orc_dequant_rect(int w, int h, const int32_t* levels, int qp, int bd, int32_t* out);
int orc_dequant_dep_quant(int w, int h, const int32_t* levels, int qp, int bd, int32_t* out);
int32_t orc_dequant_ts(int32_t level, int qp);
int orc_inverse_lfnst(int w, int h, int32_t* coeffs, int lfnst_idx, int mode);
int orc_forward_lfnst(int w, int h, int32_t* coeffs, int lfnst_idx, int mode);
int orc_inverse_transform(int w, int h, const int32_t* coeffs, int tr_hor, int tr_ver,
                          int extra_skip_w, int extra_skip_h, int bd, int32_t* out);
int orc_forward_transform(int w, int h, const int32_t* residual, int tr_hor, int tr_ver,
                          int extra_skip_w, int extra_skip_h, int bd, int32_t* out);
/* Square-only forward (vvc/dct.rs:296-342): no clip after the second stage. */
int orc_forward_transform_square(int n, const int32_t* residual, int tr_hor, int tr_ver,
                                 int skip_w, int skip_h, int32_t* out);
int orc_quant_rect(int w, int h, const int32_t* coeffs, int qp, int bd, int is_intra, int32_t* out);
int orc_quant_dep_quant(int w, int h, const int32_t* coeffs, int qp, int bd, int32_t* out);
int32_t orc_quant_ts(int32_t residual, int qp);
/* dequantize_dct / quantize_dct for square DCT-2 (vvc/dct.rs:563-613). */
int orc_quant_dct(int n, const int32_t* coeffs, int qp, int32_t* out);
int orc_dequant_dct(int n, const int32_t* levels, int qp, int32_t* out);
/* One TU through the inverse_residual order (decode_general.rs:4125-4199). */
int orc_residual_tu(const rmh_tu_desc* tu, const int16_t* coeff, int16_t* residual);
/* Every TU of a batch; returns the index+1 of the first rejected TU, 0 on success. */
size_t orc_residual_batch(const rmh_tu_desc* tus, size_t n_tu, const int16_t* coeff,
                          size_t n_coeff, int16_t* residual);
/* Encoder forward path (plan_primary order, vvc/slice.rs:718-761). */
int32_t orc_quant_ts_mode(int32_t residual, int qp, int is_intra);
int orc_forward_tu(const rmh_tu_desc* tu, const int16_t* residual, int16_t* levels);
size_t orc_forward_batch(const rmh_tu_desc* tus, size_t n_tu, const int16_t* residual, size_t n_coeff,
                         int16_t* levels);
/* Residual + prediction, clipped, in place in host planes (reconstruct_tu's add). */
size_t orc_recon_batch(const rmh_tu_desc* tus, size_t n_tu, const int16_t* coeff, size_t n_coeff,
                       const rmh_tu_place* place, const rmh_recon_frame* frames, size_t n_frames);

/* ---- inter MC stage ---- */
/* Luma interpolation (interpolate_luma_block, vvc/inter.rs:896-1051, generalised
 * to bit depth bd with head_room = 14 - bd). Samples are read as uint16. */
int orc_interp_luma(const uint16_t* plane, int stride, int pw, int ph, int bd, int bx, int by,
                    int w, int h, int mv_x, int mv_y, int bi, int alt_hpel, int affine,
                    int32_t* out);
/* Chroma 4:2:0 interpolation (interpolate_chroma_block_420, vvc/inter.rs:1056-1182). */
int orc_interp_chroma(const uint16_t* plane, int stride, int pw, int ph, int bd, int cx, int cy,
                      int cw, int ch, int mv_x, int mv_y, int bi, int32_t* out);
/* interpolate_luma_block_16 exactly as written in the reference
 * (vvc/inter.rs:2010-2134), including its uni-prediction 2-D shift. */
int orc_interp_luma_16_as_written(const uint16_t* plane, int stride, int pw, int ph, int bd,
                                  int bx, int by, int w, int h, int mv_x, int mv_y, int bi,
                                  int32_t* out);
/* weighted_average_bi generalised to bd (vvc/inter.rs:1189-1201, 2676-2696). */
void orc_weighted_average_bi(const int32_t* p0, const int32_t* p1, size_t n, int bcw_idx, int bd,
                             int32_t* out);
/* predict_inter_pu uni-prediction + write_prediction (vvc/inter.rs:1734-1848,
 * decode_general.rs:2641-2667) over a PU list. */
size_t orc_mc_batch(const rmh_mc_frame* frames, size_t n_frames, const rmh_pu_desc* pus,
                    size_t n_pu);
/* predict_inter_pu bi-prediction without DMVR and BDOF (vvc/inter.rs:1753-1801):
 * both lists with for_bi_pred, weighted_average_bi_16 (vvc/inter.rs:2676-2696),
 * then write_prediction, over a bi PU list. */
/* apply_bdof_luma (8-bit): dst = BDOF of the 14-bit intermediates of a w x h block. */
int orc_bdof_luma(const int32_t* pred0, const int32_t* pred1, int w, int h, int32_t* dst);
/* refine_dmvr_subblock over a dx x dy block; mv = {mv0_x, mv0_y, mv1_x, mv1_y} in and out. */
void orc_dmvr_refine(const uint16_t* r0, const uint16_t* r1, int stride, int pw, int ph, int bx, int by, int dx,
                     int dy, int32_t* mv);
/* derive_gpm_params_table and predict_gpm_block (8-bit). */
void orc_gpm_params(uint8_t* angle, uint8_t* distance);
int orc_gpm_blend(const int32_t* p0, const int32_t* p1, int w, int h, int split_dir, int scale_x, int scale_y,
                  int32_t* out);
size_t orc_gpm_batch(const rmh_gpm_frame* frames, size_t n_frames, const rmh_gpm_pu* pus, size_t n_pu, int luma_scale,
                     int chroma_scale);
size_t orc_mc_bi_batch(const rmh_mc_bi_frame* frames, size_t n_frames, const rmh_bi_pu_desc* pus,
                       size_t n_pu);

/* ---- deblock stage ---- */
/* deblock_luma_edge_bd (vvc/filter.rs:638-679) on lines[4][16] = p7..p0 q0..q7. */
void orc_deblock_luma_edge(int32_t* lines, int qp_p, int qp_q, int bs, int beta_off, int tc_off,
                           int bd, int len_p, int len_q, int at_ctb);
/* deblock_chroma_edge (vvc/filter.rs:935-1007) on n lines[8] = p3..p0 q0..q3. */
void orc_deblock_chroma_edge(int32_t* lines, int n, int qp, int bs, int beta_off, int tc_off,
                             int len_p, int len_q, int bd);
/* Whole-picture deblocking in the reference edge order (apply_deblock,
 * decode_general.rs:4977-5187) driven by an edge map. */
size_t orc_deblock_frames(const rmh_dbk_frame* frames, size_t n_frames);
/* deblock_luma_plane_16 / deblock_chroma_plane_16 (vvc/filter.rs:2344-2527). */
void orc_deblock_luma_plane(uint16_t* plane, int w, int h, int bd, int cu_step, int ctb, int qp,
                            int bs, int beta_off, int tc_off);
void orc_deblock_chroma_plane(uint16_t* plane, int w, int h, int bd, int cu_step, int ctb_c,
                              int qp, int bs, int beta_off, int tc_off);

/* ---- intra-profile subset codec (harness for the residual stage) ---- */
void orc_intra_predict(const uint8_t* plane, int pw, int ph, int bx, int by, int w, int h, int is_luma, int mode,
                       int32_t* pred);
/* choose_intra_mode for an n x n luma block (vvc/predict.rs:2706-2730): 0 PLANAR, 1 DC. */
int orc_intra_choose_mode(const uint8_t* rec, const uint8_t* src, int pw, int ph, int bx, int by, int n);
int orc_intra_encode(const uint8_t* sy, const uint8_t* su, const uint8_t* sv, int w, int h, int qp, uint8_t* ry,
                     uint8_t* ru, uint8_t* rv, uint8_t* modes, rmh_tu_desc* tus, int16_t* coeff);
int orc_intra_decode(int w, int h, const uint8_t* modes, const rmh_tu_desc* tus, const int16_t* coeff, uint8_t* ry,
                     uint8_t* ru, uint8_t* rv);

/* Threads of the batch drivers (1, the default: serial in the reference's
 * order; results never depend on it). */
void orc_set_threads(int n);

/* Per-line deblocking filters and sample decisions (vvc/filter.rs:276-304,
 * 477-631, 907-919, 1010-1057). Luma lines are [p7..p0, q0..q7], chroma
 * lines [p3..p0, q0..q3]. */
void orc_dbk_long_line(int32_t* line, int len_p, int len_q, int tc, int max_val);
void orc_dbk_short_line(int32_t* line, int de, int dep, int deq, int tc, int max_val);
void orc_dbk_chroma_line(int32_t* line, int len_p, int len_q, int tc, int max_val);
int orc_dbk_luma_sample_decision(int p0, int p3, int q0, int q3, int dpq, int sp, int sq, int spq, int side_p_large,
                                 int side_q_large, int beta, int tc);
int orc_dbk_chroma_sample_decision(int p0, int p3, int q0, int q3, int dpq, int beta, int tc);

/* The edge map of apply_deblock's edge loops (vvc/decode_general.rs:4977-5187)
 * from a cell grid: bS by deblock_bs / derive_boundary_strength, lengths by
 * luma_lens / chroma_lens. Host buffers. */
size_t orc_edge_map(const rmh_edge_frame* frames, size_t n_frames);

/* tC and beta after bit-depth scaling (deblock_luma_edge_bd, vvc/filter.rs:653-666). */
void orc_tc_beta(int qp, int bs, int beta_off, int tc_off, int bd, int* tc, int* beta);

/* ---- SAO (vvc/filter.rs:1276-1345, 1954-2043; decode_general.rs:5189-5222) ---- */
int orc_sao_region_bounded16(const rmh_plane* src, rmh_plane* dst, int bx, int by, int w, int h, int bw, int bh,
                             const rmh_sao_param* sao);
int orc_sao_region_bounded8(const rmh_plane* src, rmh_plane* dst, int bx, int by, int w, int h, int bw, int bh,
                            const rmh_sao_param* sao);
int orc_sao_frame(const rmh_sao_frame* f);

/* ---- LMCS and ALF (vvc/filter.rs:1580-1691, 2628-2750; decode_general.rs:5225-5660) ---- */
int orc_lmcs_tables(const rmh_lmcs_aps* aps, int bit_depth, uint16_t* fwd_lut, uint16_t* inv_lut,
                    int32_t* chroma_adj_lut, int32_t* reshape_pivot);
int orc_lmcs_map(const rmh_plane* src, const rmh_plane* dst, const uint16_t* lut, int bit_depth);
/* Scales residual[0..n) in place; returns the scale (chroma_adj_lut entry) used. */
int orc_lmcs_scale_chroma(const rmh_lmcs_aps* aps, int bit_depth, int avg_luma, int32_t* residual, int n);
/* Class index 0..24 of the 4x4 block at (bx, by); *transpose gets 0..3. */
int orc_alf_classify(const rmh_plane* plane, int bx, int by, int bw, int bh, int vb_h, int vb_pos, int* transpose);
void orc_alf_luma_filter(int mode, int class_idx, int transpose, const rmh_alf_aps* aps, int bit_depth, int32_t* fc,
                         int32_t* fclip);
int orc_alf_frame(const rmh_alf_frame* f);

#ifdef __cplusplus
}
#endif

#endif

"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, as the checker. The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the same sources, e.g. the
# sanitizer build of `make -C oracle asan`.
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_lib = None

I32P = ctypes.POINTER(ctypes.c_int32)
I16P = ctypes.POINTER(ctypes.c_int16)
U16P = ctypes.POINTER(ctypes.c_uint16)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _setup(_lib)
    return _lib


def _setup(L):
    c_int, c_size = ctypes.c_int, ctypes.c_size_t
    L.orc_transform_matrix.argtypes = [c_int, c_int, I32P]
    L.orc_lfnst_set.argtypes = [c_int]
    L.orc_lfnst_transposes.argtypes = [c_int]
    L.orc_lfnst_coeff.argtypes = [c_int] * 5
    L.orc_coeff_scan.argtypes = [c_int, c_int, I32P]
    L.orc_cabac_decode_residuals.argtypes = [ctypes.c_char_p, ctypes.c_size_t, c_int, c_int, c_int, c_int, c_int,
                                             c_int, c_int, I32P]
    L.orc_cabac_decode_bins.argtypes = [ctypes.c_char_p, ctypes.c_size_t, c_int, c_int, c_int, c_int, c_int,
                                        ctypes.POINTER(ctypes.c_uint8)]
    L.orc_decode_luma_stream.argtypes = [ctypes.c_char_p, ctypes.c_size_t, c_int, c_int, c_int,
                                         ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint8),
                                         ctypes.POINTER(ctypes.c_uint8)]
    L.orc_cabac_ctx_offset.argtypes = [c_int]
    L.orc_dequant_rect.argtypes = [c_int, c_int, I32P, c_int, c_int, I32P]
    L.orc_dequant_dep_quant.argtypes = [c_int, c_int, I32P, c_int, c_int, I32P]
    L.orc_dequant_ts.argtypes = [ctypes.c_int32, c_int]
    L.orc_dequant_ts.restype = ctypes.c_int32
    L.orc_quant_ts.argtypes = [ctypes.c_int32, c_int]
    L.orc_quant_ts.restype = ctypes.c_int32
    L.orc_inverse_lfnst.argtypes = [c_int, c_int, I32P, c_int, c_int]
    L.orc_forward_lfnst.argtypes = [c_int, c_int, I32P, c_int, c_int]
    L.orc_inverse_transform.argtypes = [c_int, c_int, I32P, c_int, c_int, c_int, c_int, c_int, I32P]
    L.orc_forward_transform.argtypes = [c_int, c_int, I32P, c_int, c_int, c_int, c_int, c_int, I32P]
    L.orc_forward_transform_square.argtypes = [c_int, I32P, c_int, c_int, c_int, c_int, I32P]
    L.orc_quant_rect.argtypes = [c_int, c_int, I32P, c_int, c_int, c_int, I32P]
    L.orc_quant_dep_quant.argtypes = [c_int, c_int, I32P, c_int, c_int, I32P]
    L.orc_quant_dct.argtypes = [c_int, I32P, c_int, I32P]
    L.orc_dequant_dct.argtypes = [c_int, I32P, c_int, I32P]
    L.orc_residual_batch.argtypes = [ctypes.c_void_p, c_size, I16P, c_size, I16P]
    L.orc_residual_batch.restype = c_size
    L.orc_interp_luma.argtypes = [U16P] + [c_int] * 13 + [I32P]
    L.orc_interp_chroma.argtypes = [U16P] + [c_int] * 11 + [I32P]
    L.orc_interp_luma_16_as_written.argtypes = [U16P] + [c_int] * 11 + [I32P]
    L.orc_weighted_average_bi.argtypes = [I32P, I32P, c_size, c_int, c_int, I32P]
    L.orc_bdof_luma.argtypes = [I32P, I32P, c_int, c_int, I32P]
    L.orc_dmvr_refine.argtypes = [U16P, U16P] + [c_int] * 7 + [I32P]
    L.orc_gpm_params.argtypes = [ctypes.POINTER(ctypes.c_uint8)] * 2
    L.orc_gpm_blend.argtypes = [I32P, I32P] + [c_int] * 5 + [I32P]
    L.orc_gpm_batch.argtypes = [ctypes.c_void_p, c_size, ctypes.c_void_p, c_size, c_int, c_int]
    L.orc_gpm_batch.restype = c_size
    L.orc_recon_batch.argtypes = [ctypes.c_void_p, c_size, I16P, c_size, ctypes.c_void_p, ctypes.c_void_p, c_size]
    L.orc_recon_batch.restype = c_size
    L.orc_mc_batch.argtypes = [ctypes.c_void_p, c_size, ctypes.c_void_p, c_size]
    L.orc_mc_batch.restype = c_size
    L.orc_mc_bi_batch.argtypes = [ctypes.c_void_p, c_size, ctypes.c_void_p, c_size]
    L.orc_mc_bi_batch.restype = c_size
    L.orc_deblock_luma_edge.argtypes = [I32P] + [c_int] * 9
    L.orc_deblock_chroma_edge.argtypes = [I32P] + [c_int] * 8
    L.orc_deblock_frames.argtypes = [ctypes.c_void_p, c_size]
    L.orc_deblock_frames.restype = c_size
    L.orc_deblock_luma_plane.argtypes = [U16P] + [c_int] * 9
    L.orc_deblock_chroma_plane.argtypes = [U16P] + [c_int] * 9
    L.orc_tc_beta.argtypes = [c_int] * 5 + [I32P, I32P]
    L.orc_set_threads.argtypes = [c_int]
    L.orc_edge_map.argtypes = [ctypes.c_void_p, c_size]
    L.orc_edge_map.restype = c_size
    L.orc_dbk_long_line.argtypes = [I32P] + [c_int] * 4
    L.orc_dbk_short_line.argtypes = [I32P] + [c_int] * 5
    L.orc_dbk_chroma_line.argtypes = [I32P] + [c_int] * 4
    L.orc_dbk_luma_sample_decision.argtypes = [c_int] * 12
    L.orc_dbk_chroma_sample_decision.argtypes = [c_int] * 7
    U8P = ctypes.POINTER(ctypes.c_uint8)
    L.orc_intra_predict.argtypes = [U8P] + [c_int] * 8 + [I32P]
    L.orc_intra_choose_mode.argtypes = [U8P, U8P] + [c_int] * 5
    L.orc_intra_encode.argtypes = [U8P, U8P, U8P, c_int, c_int, c_int, U8P, U8P, U8P, U8P, ctypes.c_void_p, I16P]
    L.orc_intra_decode.argtypes = [c_int, c_int, U8P, ctypes.c_void_p, I16P, U8P, U8P, U8P]
    L.orc_sao_region_bounded16.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [c_int] * 6 + [ctypes.c_void_p]
    L.orc_sao_region_bounded8.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [c_int] * 6 + [ctypes.c_void_p]
    L.orc_sao_frame.argtypes = [ctypes.c_void_p]
    L.orc_lmcs_tables.argtypes = [ctypes.c_void_p, c_int, U16P, U16P, I32P, I32P]
    L.orc_lmcs_map.argtypes = [ctypes.c_void_p, ctypes.c_void_p, U16P, c_int]
    L.orc_lmcs_scale_chroma.argtypes = [ctypes.c_void_p, c_int, c_int, I32P, c_int]
    L.orc_alf_classify.argtypes = [ctypes.c_void_p] + [c_int] * 6 + [ctypes.POINTER(c_int)]
    L.orc_alf_frame.argtypes = [ctypes.c_void_p]
    L.orc_quant_ts_mode.argtypes = [ctypes.c_int32, c_int, c_int]
    L.orc_quant_ts_mode.restype = ctypes.c_int32
    L.orc_forward_batch.argtypes = [ctypes.c_void_p, c_size, I16P, c_size, I16P]
    L.orc_forward_batch.restype = c_size
    L.orc_alf_luma_filter.argtypes = [c_int] * 3 + [ctypes.c_void_p, c_int, I32P, I32P]
    L.orc_alf_luma_filter.restype = None


def _p(a, t):
    return a.ctypes.data_as(t)


def transform_matrix(tr_type, n):
    out = np.zeros(n * n, np.int32)
    if lib().orc_transform_matrix(tr_type, n, _p(out, I32P)):
        return None
    return out.reshape(n, n)


def coeff_scan(w, h):
    out = np.zeros(4096, np.int32)
    k = lib().orc_coeff_scan(w, h, _p(out, I32P))
    return out[:k].copy()


def _blk(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))


def dequant_rect(w, h, levels, qp, bd):
    lv = _blk(levels)
    out = np.zeros(w * h, np.int32)
    assert lib().orc_dequant_rect(w, h, _p(lv, I32P), qp, bd, _p(out, I32P)) == 0
    return out.reshape(h, w)


def dequant_dep_quant(w, h, levels, qp, bd):
    lv = _blk(levels)
    out = np.zeros(w * h, np.int32)
    assert lib().orc_dequant_dep_quant(w, h, _p(lv, I32P), qp, bd, _p(out, I32P)) == 0
    return out.reshape(h, w)


def dequant_ts(level, qp):
    return lib().orc_dequant_ts(int(level), qp)


def quant_ts(res, qp):
    return lib().orc_quant_ts(int(res), qp)


def inverse_lfnst(w, h, coeffs, idx, mode):
    c = _blk(coeffs).copy()
    rc = lib().orc_inverse_lfnst(w, h, _p(c, I32P), idx, mode)
    return None if rc else c.reshape(h, w)


def forward_lfnst(w, h, coeffs, idx, mode):
    c = _blk(coeffs).copy()
    rc = lib().orc_forward_lfnst(w, h, _p(c, I32P), idx, mode)
    return None if rc else c.reshape(h, w)


def inverse_transform(w, h, coeffs, th, tv, skip_w=0, skip_h=0, bd=8):
    c = _blk(coeffs)
    out = np.zeros(w * h, np.int32)
    rc = lib().orc_inverse_transform(w, h, _p(c, I32P), th, tv, skip_w, skip_h, bd, _p(out, I32P))
    return None if rc else out.reshape(h, w)


def forward_transform(w, h, res, th, tv, skip_w=0, skip_h=0, bd=8):
    r = _blk(res)
    out = np.zeros(w * h, np.int32)
    rc = lib().orc_forward_transform(w, h, _p(r, I32P), th, tv, skip_w, skip_h, bd, _p(out, I32P))
    return None if rc else out.reshape(h, w)


def forward_transform_square(n, res, th, tv, skip_w=0, skip_h=0):
    r = _blk(res)
    out = np.zeros(n * n, np.int32)
    rc = lib().orc_forward_transform_square(n, _p(r, I32P), th, tv, skip_w, skip_h, _p(out, I32P))
    return None if rc else out.reshape(n, n)


def quant_rect(w, h, coeffs, qp, bd, is_intra=True):
    c = _blk(coeffs)
    out = np.zeros(w * h, np.int32)
    assert lib().orc_quant_rect(w, h, _p(c, I32P), qp, bd, int(is_intra), _p(out, I32P)) == 0
    return out.reshape(h, w)


def quant_dep_quant(w, h, coeffs, qp, bd):
    c = _blk(coeffs)
    out = np.zeros(w * h, np.int32)
    assert lib().orc_quant_dep_quant(w, h, _p(c, I32P), qp, bd, _p(out, I32P)) == 0
    return out.reshape(h, w)


def quant_dct(n, coeffs, qp):
    c = _blk(coeffs)
    out = np.zeros(n * n, np.int32)
    rc = lib().orc_quant_dct(n, _p(c, I32P), qp, _p(out, I32P))
    return None if rc else out.reshape(n, n)


def dequant_dct(n, levels, qp):
    c = _blk(levels)
    out = np.zeros(n * n, np.int32)
    rc = lib().orc_dequant_dct(n, _p(c, I32P), qp, _p(out, I32P))
    return None if rc else out.reshape(n, n)


def residual_batch(tus, coeff):
    """tus: numpy structured array with the rmh_tu_desc dtype; coeff: int16."""
    coeff = np.ascontiguousarray(coeff, dtype=np.int16)
    out = np.zeros_like(coeff)
    bad = lib().orc_residual_batch(tus.ctypes.data, len(tus), _p(coeff, I16P), coeff.size, _p(out, I16P))
    if bad:
        raise ValueError("oracle rejected TU %d" % (bad - 1))
    return out


def interp_luma(plane, bd, bx, by, w, h, mv_x, mv_y, bi=False, alt_hpel=False):
    p = np.ascontiguousarray(plane, dtype=np.uint16)
    out = np.zeros(w * h, np.int32)
    assert lib().orc_interp_luma(_p(p, U16P), p.shape[1], p.shape[1], p.shape[0], bd, bx, by, w, h,
                                 mv_x, mv_y, int(bi), int(alt_hpel), 0, _p(out, I32P)) == 0
    return out.reshape(h, w)


def interp_chroma(plane, bd, cx, cy, cw, ch, mv_x, mv_y, bi=False):
    p = np.ascontiguousarray(plane, dtype=np.uint16)
    out = np.zeros(cw * ch, np.int32)
    assert lib().orc_interp_chroma(_p(p, U16P), p.shape[1], p.shape[1], p.shape[0], bd, cx, cy, cw, ch,
                                   mv_x, mv_y, int(bi), _p(out, I32P)) == 0
    return out.reshape(ch, cw)


def interp_luma_16_as_written(plane, bd, bx, by, w, h, mv_x, mv_y, bi=False):
    p = np.ascontiguousarray(plane, dtype=np.uint16)
    out = np.zeros(w * h, np.int32)
    assert lib().orc_interp_luma_16_as_written(_p(p, U16P), p.shape[1], p.shape[1], p.shape[0], bd, bx,
                                               by, w, h, mv_x, mv_y, int(bi), _p(out, I32P)) == 0
    return out.reshape(h, w)


def weighted_average_bi(p0, p1, bcw_idx, bd):
    a = _blk(p0)
    b = _blk(p1)
    out = np.zeros_like(a)
    lib().orc_weighted_average_bi(_p(a, I32P), _p(b, I32P), a.size, bcw_idx, bd, _p(out, I32P))
    return out


def bdof_luma(p0, p1, w, h):
    """apply_bdof_luma of two (h, w) 14-bit intermediate blocks."""
    a = np.ascontiguousarray(np.asarray(p0, np.int32).reshape(-1))
    b = np.ascontiguousarray(np.asarray(p1, np.int32).reshape(-1))
    out = np.zeros(w * h, np.int32)
    assert lib().orc_bdof_luma(_p(a, I32P), _p(b, I32P), w, h, _p(out, I32P)) == 0
    return out.reshape(h, w)


def dmvr_refine(ref0, ref1, bx, by, dx, dy, mv0, mv1):
    """refine_dmvr_subblock over a dx x dy block: the refined (mv0, mv1)."""
    r0 = np.ascontiguousarray(ref0, dtype=np.uint16)
    r1 = np.ascontiguousarray(ref1, dtype=np.uint16)
    assert r0.shape == r1.shape
    mv = np.array([mv0[0], mv0[1], mv1[0], mv1[1]], np.int32)
    lib().orc_dmvr_refine(_p(r0, U16P), _p(r1, U16P), r0.shape[1], r0.shape[1], r0.shape[0], bx, by, dx, dy,
                          _p(mv, I32P))
    return (int(mv[0]), int(mv[1])), (int(mv[2]), int(mv[3]))


def gpm_params():
    """derive_gpm_params_table: [(angle, distance)] of the 64 split directions."""
    a = np.zeros(64, np.uint8)
    d = np.zeros(64, np.uint8)
    lib().orc_gpm_params(_p(a, U8P), _p(d, U8P))
    return list(zip(a.tolist(), d.tolist()))


def gpm_blend(p0, p1, w, h, split_dir, scale_x, scale_y):
    """predict_gpm_block of two (h, w) 8-bit predictions."""
    a = np.ascontiguousarray(np.asarray(p0, np.int32).reshape(-1))
    b = np.ascontiguousarray(np.asarray(p1, np.int32).reshape(-1))
    out = np.zeros(w * h, np.int32)
    assert lib().orc_gpm_blend(_p(a, I32P), _p(b, I32P), w, h, split_dir, scale_x, scale_y, _p(out, I32P)) == 0
    return out.reshape(h, w)


def gpm_batch(frames, pus, luma_scale, chroma_scale):
    """orc_gpm_batch over a list of GpmFrame structs (host planes)."""
    arr = (type(frames[0]) * len(frames))(*frames)
    pus = np.ascontiguousarray(pus)
    return lib().orc_gpm_batch(ctypes.addressof(arr), len(frames), pus.ctypes.data, len(pus), luma_scale,
                               chroma_scale)


def deblock_luma_edge(lines, qp_p, qp_q, bs, beta_off, tc_off, bd, len_p, len_q, at_ctb=False):
    l = np.ascontiguousarray(np.asarray(lines, dtype=np.int32).reshape(4, 16)).copy()
    lib().orc_deblock_luma_edge(_p(l, I32P), qp_p, qp_q, bs, beta_off, tc_off, bd, len_p, len_q, int(at_ctb))
    return l


def deblock_chroma_edge(lines, qp, bs, beta_off, tc_off, len_p, len_q, bd):
    l = np.ascontiguousarray(np.asarray(lines, dtype=np.int32).reshape(-1, 8)).copy()
    lib().orc_deblock_chroma_edge(_p(l, I32P), l.shape[0], qp, bs, beta_off, tc_off, len_p, len_q, bd)
    return l


def deblock_luma_plane(plane, bd, cu_step, ctb, qp, bs, beta_off=0, tc_off=0):
    p = np.ascontiguousarray(plane, dtype=np.uint16).copy()
    lib().orc_deblock_luma_plane(_p(p, U16P), p.shape[1], p.shape[0], bd, cu_step, ctb, qp, bs, beta_off, tc_off)
    return p


def deblock_chroma_plane(plane, bd, cu_step, ctb_c, qp, bs, beta_off=0, tc_off=0):
    p = np.ascontiguousarray(plane, dtype=np.uint16).copy()
    lib().orc_deblock_chroma_plane(_p(p, U16P), p.shape[1], p.shape[0], bd, cu_step, ctb_c, qp, bs, beta_off, tc_off)
    return p


def dbk_long_line(line, len_p, len_q, tc, max_val):
    a = np.array(line, np.int32)
    lib().orc_dbk_long_line(_p(a, I32P), len_p, len_q, tc, max_val)
    return a.tolist()


def dbk_short_line(line, de, dep, deq, tc, max_val):
    a = np.array(line, np.int32)
    lib().orc_dbk_short_line(_p(a, I32P), de, int(dep), int(deq), tc, max_val)
    return a.tolist()


def dbk_chroma_line(line, len_p, len_q, tc, max_val):
    a = np.array(line, np.int32)
    lib().orc_dbk_chroma_line(_p(a, I32P), len_p, len_q, tc, max_val)
    return a.tolist()


def tc_beta(qp, bs, beta_off, tc_off, bd):
    tc = ctypes.c_int32()
    beta = ctypes.c_int32()
    lib().orc_tc_beta(qp, bs, beta_off, tc_off, bd, ctypes.byref(tc), ctypes.byref(beta))
    return tc.value, beta.value


U8P = ctypes.POINTER(ctypes.c_uint8)


def intra_encode(y, u, v, qp, tu_dtype):
    """Subset intra encoder decisions for one 8-bit 4:2:0 frame; returns
    (recon (y, u, v), modes, tus, coeff)."""
    h, w = y.shape
    n_ctu = (w // 32) * (h // 32)
    ry, ru, rv = np.zeros_like(y), np.zeros_like(u), np.zeros_like(v)
    modes = np.zeros(n_ctu, np.uint8)
    tus = np.zeros(3 * n_ctu, tu_dtype)
    coeff = np.zeros(n_ctu * (1024 + 512), np.int16)
    args = [np.ascontiguousarray(a, dtype=np.uint8) for a in (y, u, v)]
    n = lib().orc_intra_encode(_p(args[0], U8P), _p(args[1], U8P), _p(args[2], U8P), w, h, qp, _p(ry, U8P),
                               _p(ru, U8P), _p(rv, U8P), _p(modes, U8P), tus.ctypes.data, _p(coeff, I16P))
    assert n == n_ctu
    return (ry, ru, rv), modes, tus, coeff


def intra_decode(w, h, modes, tus, coeff):
    ry = np.zeros((h, w), np.uint8)
    ru = np.zeros((h // 2, w // 2), np.uint8)
    rv = np.zeros((h // 2, w // 2), np.uint8)
    n = lib().orc_intra_decode(w, h, _p(np.ascontiguousarray(modes, dtype=np.uint8), U8P), tus.ctypes.data,
                               _p(np.ascontiguousarray(coeff, dtype=np.int16), I16P), _p(ry, U8P), _p(ru, U8P),
                               _p(rv, U8P))
    assert n == len(modes)
    return ry, ru, rv


def mc_bi_batch(frame, pus):
    """orc_mc_bi_batch over one abi.McBiFrame (host planes) and a BI_PU_DTYPE array."""
    pus = np.ascontiguousarray(pus)
    arr = (type(frame) * 1)(frame)
    return lib().orc_mc_bi_batch(ctypes.addressof(arr), 1, pus.ctypes.data, len(pus))


def mc_batch(frame, pus):
    """orc_mc_batch over one abi.McFrame (host planes) and a PU_DTYPE array."""
    pus = np.ascontiguousarray(pus)
    arr = (type(frame) * 1)(frame)
    return lib().orc_mc_batch(ctypes.addressof(arr), 1, pus.ctypes.data, len(pus))


_abi_mod = None


def _abi():
    """The ABI struct definitions (rustmpeg-next_amd/abi.py: ctypes layouts
    only, no library calls), loaded from the file."""
    global _abi_mod
    if _abi_mod is None:
        import importlib.util
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "rustmpeg-next_amd", "abi.py")
        spec = importlib.util.spec_from_file_location("rmh_abi_layouts", path)
        _abi_mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(_abi_mod)
    return _abi_mod


def _sao_param(p):
    """p: a SAO_PARAM_DTYPE record, or (type, band_or_class, offsets)."""
    if isinstance(p, tuple):
        r = np.zeros(1, _abi().SAO_PARAM_DTYPE)
        r["type"], r["band_or_class"], r["offsets"] = p[0], p[1], p[2]
        return r
    return np.ascontiguousarray(np.atleast_1d(p), dtype=_abi().SAO_PARAM_DTYPE)


def sao_region(src, bd, bx, by, w, h, bw, bh, param, eight_bit_form=False):
    """apply_sao_region_bounded_16 (or the 8-bit apply_sao_region_bounded) of
    one region: returns a copy of src with the region filtered."""
    A = _abi()
    src = np.ascontiguousarray(src)
    dst = src.copy()
    prm = _sao_param(param)
    fn = lib().orc_sao_region_bounded8 if eight_bit_form else lib().orc_sao_region_bounded16
    ps, pd = A.plane_from_array(src, bd), A.plane_from_array(dst, bd)
    if fn(ctypes.addressof(ps), ctypes.addressof(pd), bx, by, w, h, bw, bh, prm.ctypes.data):
        raise ValueError("invalid SAO parameters")
    return dst


def sao_frame(planes, bd, params, ctb_size, ctb_cols, ctb_rows, vis_w, vis_h):
    """apply_sao over one picture: planes [Y, Cb, Cr] (not modified); params a
    SAO_PARAM_DTYPE array of ctb_rows * ctb_cols * 3 records. Returns the three
    output planes."""
    A = _abi()
    planes = [np.ascontiguousarray(p) for p in planes]
    outs = [np.zeros_like(p) for p in planes]
    params = np.ascontiguousarray(params, dtype=A.SAO_PARAM_DTYPE)
    f = A.sao_frame([A.plane_from_array(p, bd) for p in planes], [A.plane_from_array(o, bd) for o in outs],
                    params.ctypes.data, ctb_size, ctb_cols, ctb_rows, vis_w, vis_h)
    if lib().orc_sao_frame(ctypes.addressof(f)):
        raise ValueError("invalid SAO parameters")
    return outs


def lmcs_tables(aps, bd):
    """LmcsReshaper16::from_aps: (fwd_lut, inv_lut, chroma_adj_lut, reshape_pivot) for an abi.LmcsAps."""
    n = 1 << min(max(bd, 8), 12)
    fwd, inv = np.zeros(n, np.uint16), np.zeros(n, np.uint16)
    cadj, piv = np.zeros(16, np.int32), np.zeros(17, np.int32)
    lib().orc_lmcs_tables(ctypes.addressof(aps), bd, _p(fwd, U16P), _p(inv, U16P), _p(cadj, I32P), _p(piv, I32P))
    return fwd, inv, cadj, piv


def lmcs_map(plane, lut, bd):
    """apply_{forward,inverse}_plane_16 of a uint8/uint16 plane: returns the mapped copy."""
    A = _abi()
    src = np.ascontiguousarray(plane)
    out = np.zeros_like(src)
    lut = np.ascontiguousarray(lut, dtype=np.uint16)
    ps, po = A.plane_from_array(src, bd), A.plane_from_array(out, bd)
    lib().orc_lmcs_map(ctypes.byref(ps), ctypes.byref(po), _p(lut, U16P), bd)
    return out


def lmcs_scale_chroma(aps, bd, avg_luma, residual):
    """chroma_scale_for_luma + scale_chroma_residual: (scaled copy, scale)."""
    r = np.ascontiguousarray(residual, dtype=np.int32).copy()
    scale = lib().orc_lmcs_scale_chroma(ctypes.addressof(aps), bd, avg_luma, _p(r, I32P), r.size)
    return r, scale


def alf_classify(plane, bd, bx, by, bw, bh, vb_h, vb_pos):
    """derive_alf_class_4x4_vb_16 of the 4x4 block at (bx, by): (class, transpose)."""
    A = _abi()
    src = np.ascontiguousarray(plane)
    p = A.plane_from_array(src, bd)
    t = ctypes.c_int(0)
    c = lib().orc_alf_classify(ctypes.byref(p), bx, by, bw, bh, vb_h, vb_pos, ctypes.byref(t))
    return c, t.value


def alf_frame(planes, bd, ctbs, ctb_size, ctb_cols, ctb_rows, vis_w, vis_h, luma_aps=None, chroma_aps=None,
              cc_cb_aps=None, cc_cr_aps=None):
    """apply_alf over one picture: planes [Y, Cb, Cr] (not modified), ctbs an
    ALF_CTB_DTYPE array of ctb_rows * ctb_cols records, APS arguments abi.AlfAps.
    Returns the three output planes."""
    A = _abi()
    planes = [np.ascontiguousarray(p) for p in planes]
    outs = [np.zeros_like(p) for p in planes]
    ctbs = np.ascontiguousarray(ctbs, dtype=A.ALF_CTB_DTYPE)
    f = A.alf_frame([A.plane_from_array(p, bd) for p in planes], [A.plane_from_array(o, bd) for o in outs],
                    ctbs.ctypes.data, ctb_size, ctb_cols, ctb_rows, vis_w, vis_h, luma_aps, chroma_aps, cc_cb_aps,
                    cc_cr_aps)
    lib().orc_alf_frame(ctypes.addressof(f))
    return outs


def alf_luma_filter(mode, class_idx, transpose, aps, bd):
    """(coefficients, clip values) alf_luma_block applies to a block (mode 1..16 fixed, 17 APS)."""
    fc, cl = np.zeros(12, np.int32), np.zeros(12, np.int32)
    lib().orc_alf_luma_filter(mode, class_idx, transpose, ctypes.addressof(aps) if aps is not None else None, bd,
                              _p(fc, I32P), _p(cl, I32P))
    return fc, cl


def quant_ts_mode(res, qp, is_intra):
    return lib().orc_quant_ts_mode(int(res), qp, int(is_intra))


def forward_batch(tus, residual):
    """The encoder forward path (orc_forward_batch) over TU_DTYPE descriptors: int16 levels."""
    residual = np.ascontiguousarray(residual, dtype=np.int16)
    out = np.zeros_like(residual)
    tus = np.ascontiguousarray(tus)
    bad = lib().orc_forward_batch(tus.ctypes.data, len(tus), _p(residual, I16P), residual.size, _p(out, I16P))
    if bad:
        raise ValueError("TU %d rejected" % (bad - 1))
    return out


# ---- CABAC decoding side (orc_cabac.c) ----

def cabac_decode_residuals(data, w, h, n_blocks, qp=26, luma=True, dep_quant=False, sign_hiding=False):
    """Levels [n_blocks, h, w] read back from a residual substream, or None
    when the stream is malformed."""
    out = np.zeros(n_blocks * w * h, np.int32)
    r = lib().orc_cabac_decode_residuals(bytes(data), len(data), w, h, n_blocks, int(luma), int(dep_quant),
                                         int(sign_hiding), qp, _p(out, I32P))
    return None if r != 0 else out.reshape(n_blocks, h, w)


def cabac_decode_bins(data, n_ctx, n_ep, qp=26, init_id=25, ws=13):
    """n_ctx context-coded bins (one model) then n_ep bypass bins, then the
    terminating bin; None when the stream does not end as it should."""
    out = np.zeros(n_ctx + n_ep, np.uint8)
    r = lib().orc_cabac_decode_bins(bytes(data), len(data), qp, init_id, ws, n_ctx, n_ep,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return None if r != 0 else out


def decode_luma_stream(data, w, h, qp):
    """(recon[h, w], splits[h/32, w/32], cu sizes[h/4, w/4, 2] as (width,
    height)) of an rmh_rd_encode_luma_stream substream, or None when it is
    malformed."""
    rec = np.zeros((h, w), np.uint8)
    sp = np.zeros((h // 32) * (w // 32), np.uint8)
    cu = np.zeros((h // 4, w // 4, 2), np.uint8)
    u8 = ctypes.POINTER(ctypes.c_uint8)
    r = lib().orc_decode_luma_stream(bytes(data), len(data), w, h, qp, rec.ctypes.data_as(u8),
                                     sp.ctypes.data_as(u8), cu.ctypes.data_as(u8))
    return None if r != 0 else (rec, sp.reshape(h // 32, w // 32), cu)


def cabac_ctx_offset(i):
    return lib().orc_cabac_ctx_offset(i)

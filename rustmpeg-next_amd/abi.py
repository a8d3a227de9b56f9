"""ctypes mirrors of the structs in include/rmh.h (no library is loaded here)."""
import ctypes

import numpy as np

RMH_OK, RMH_UNAVAILABLE, RMH_INVALID_INPUT, RMH_LIMIT_EXCEEDED, RMH_EXECUTION, RMH_TIMEOUT = range(6)
STATUS_NAMES = {0: "OK", 1: "Unavailable", 2: "InvalidInput", 3: "LimitExceeded", 4: "Execution",
                5: "Timeout"}

TR_DCT2, TR_DST7, TR_DCT8 = 0, 1, 2
TU_TRANSFORM_SKIP, TU_DEP_QUANT, TU_QUANT_INTER = 1, 2, 4
PU_ALT_HPEL = 0x10
EDGE_BS_MASK, EDGE_LENP_SHIFT, EDGE_LENQ_SHIFT, EDGE_CTB_ROW = 0x03, 2, 4, 0x40

TU_DTYPE = np.dtype([
    ("offset", "<u4"), ("log2_w", "u1"), ("log2_h", "u1"), ("qp", "u1"), ("bit_depth", "u1"),
    ("tr_hor", "u1"), ("tr_ver", "u1"), ("skip_w", "u1"), ("skip_h", "u1"),
    ("lfnst_idx", "u1"), ("lfnst_mode", "u1"), ("flags", "u1"), ("joint", "u1")])
PU_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("log2_w", "u1"), ("log2_h", "u1"),
                     ("frame", "<u2"), ("mv_x", "<i4"), ("mv_y", "<i4")])


class Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("timeout_ms", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32), ("compute_units", ctypes.c_int32),
                ("max_clock_khz", ctypes.c_int32), ("global_mem_bytes", ctypes.c_uint64),
                ("abi_version", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Plane(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("stride", ctypes.c_uint32), ("bytes_per_sample", ctypes.c_uint8), ("bit_depth", ctypes.c_uint8),
                ("reserved", ctypes.c_uint16)]


class McFrame(ctypes.Structure):
    _fields_ = [("ref", Plane * 3), ("dst", Plane * 3)]


class McBiFrame(ctypes.Structure):
    _fields_ = [("ref0", Plane * 3), ("ref1", Plane * 3), ("dst", Plane * 3)]


class GpmFrame(ctypes.Structure):
    _fields_ = [("pred0", Plane * 3), ("pred1", Plane * 3), ("dst", Plane * 3)]


# rmh_gpm_pu (12 bytes) and the scales the reference's decoder passes
GPM_PU_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("log2_w", "u1"), ("log2_h", "u1"), ("frame", "<u2"),
                         ("split_dir", "u1"), ("reserved", "u1", (3,))])
assert GPM_PU_DTYPE.itemsize == 12
GPM_LUMA_SCALE, GPM_CHROMA_SCALE = 1, 2


# rmh_bi_pu_desc (28 bytes); bcw_idx 0..4 selects the L1 weight of {-2, 3, 4, 5, 10} / 8;
# refine: BI_DMVR | BI_BDOF (8-bit frames)
BI_DMVR, BI_BDOF = 1, 2
BI_PU_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("log2_w", "u1"), ("log2_h", "u1"), ("frame", "<u2"),
                        ("mv0_x", "<i4"), ("mv0_y", "<i4"), ("mv1_x", "<i4"), ("mv1_y", "<i4"),
                        ("bcw_idx", "u1"), ("refine", "u1"), ("reserved", "u1", (2,))])
assert BI_PU_DTYPE.itemsize == 28
BCW_DEFAULT = 2


class ReconFrame(ctypes.Structure):
    _fields_ = [("planes", Plane * 3)]


# rmh_tu_place
PLACE_DTYPE = np.dtype([("frame", "<u2"), ("plane", "u1"), ("reserved", "u1"), ("x", "<u2"), ("y", "<u2")])


class RdStats(ctypes.Structure):
    _fields_ = [("bits", ctypes.c_double), ("sse", ctypes.c_uint64), ("blocks", ctypes.c_uint32),
                ("candidates", ctypes.c_uint32), ("fallbacks", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


RD_FULL, RD_PRUNED, RD_HEURISTIC = 0, 1, 2
CABAC_LUMA, CABAC_DEP_QUANT, CABAC_SIGN_HIDING = 1, 2, 4
SPLIT_NAMES = ("none", "quad", "binary_hor", "binary_ver", "ternary_hor", "ternary_ver")


class DbkFrame(ctypes.Structure):
    _fields_ = [("planes", Plane * 3), ("edges", ctypes.c_void_p), ("grid_w", ctypes.c_uint32),
                ("grid_h", ctypes.c_uint32), ("qp", ctypes.c_uint8 * 3), ("beta_offset_div2", ctypes.c_int8 * 3),
                ("tc_offset_div2", ctypes.c_int8 * 3), ("reserved", ctypes.c_uint8 * 3)]


SAO_OFF, SAO_BAND, SAO_EDGE = 0, 1, 2
# rmh_sao_param: type, band_pos or eo_class, 2 reserved bytes, 4 int32 offsets
SAO_PARAM_DTYPE = np.dtype([("type", "u1"), ("band_or_class", "u1"), ("reserved", "u1", (2,)),
                            ("offsets", "<i4", (4,))])


class SaoFrame(ctypes.Structure):
    _fields_ = [("src", Plane * 3), ("dst", Plane * 3), ("params", ctypes.c_void_p), ("ctb_size", ctypes.c_uint32),
                ("ctb_cols", ctypes.c_uint32), ("ctb_rows", ctypes.c_uint32), ("vis_w", ctypes.c_uint32),
                ("vis_h", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


def sao_frame(srcs, dsts, params_addr, ctb_size, ctb_cols, ctb_rows, vis_w, vis_h):
    f = SaoFrame()
    for i in range(3):
        f.src[i] = srcs[i]
        f.dst[i] = dsts[i]
    f.params = params_addr
    f.ctb_size, f.ctb_cols, f.ctb_rows, f.vis_w, f.vis_h = ctb_size, ctb_cols, ctb_rows, vis_w, vis_h
    return f


class LmcsAps(ctypes.Structure):
    """rmh_lmcs_aps: VvcLmcsAps (vvc/hls.rs:2723-2732)."""
    _fields_ = [("min_bin_idx", ctypes.c_uint8), ("max_bin_idx", ctypes.c_uint8),
                ("chroma_res_scaling_offset", ctypes.c_int8), ("reserved", ctypes.c_uint8),
                ("bin_cw_delta", ctypes.c_int16 * 16)]


def lmcs_aps(min_bin_idx, max_bin_idx, bin_cw_delta, chroma_res_scaling_offset=0):
    return LmcsAps(min_bin_idx, max_bin_idx, chroma_res_scaling_offset, 0, (ctypes.c_int16 * 16)(*bin_cw_delta))


LMCS_INVERSE, LMCS_FORWARD = 0, 1


class LmcsFrame(ctypes.Structure):
    _fields_ = [("src", Plane), ("dst", Plane), ("aps", LmcsAps), ("direction", ctypes.c_uint32)]


class AlfAps(ctypes.Structure):
    """rmh_alf_aps: the filter data of VvcAlfAps (vvc/hls.rs:2699-2718) as the decoder applies it."""
    _fields_ = [("luma_coeff", (ctypes.c_int16 * 12) * 25), ("luma_clip", (ctypes.c_uint8 * 12) * 25),
                ("class_to_filter", ctypes.c_uint8 * 25), ("n_luma", ctypes.c_uint8), ("n_chroma", ctypes.c_uint8),
                ("n_cc_cb", ctypes.c_uint8), ("n_cc_cr", ctypes.c_uint8), ("reserved", ctypes.c_uint8),
                ("chroma_coeff", (ctypes.c_int16 * 6) * 8), ("chroma_clip", (ctypes.c_uint8 * 6) * 8),
                ("cc_cb", (ctypes.c_int16 * 7) * 4), ("cc_cr", (ctypes.c_int16 * 7) * 4)]


assert ctypes.sizeof(AlfAps) == 1186


def alf_aps(luma_filters=(), luma_clip_idx=(), luma_clip=False, filter_coeff_delta_idx=None, chroma_filters=(),
            chroma_clip_idx=(), cc_cb_filters=(), cc_cr_filters=()):
    """An AlfAps from VvcAlfAps fields, with the clip indices the decoder reads:
    luma ones only under alf_luma_clip_flag (decode_general.rs:5466-5475), chroma
    ones where signalled, else zero (5588)."""
    a = AlfAps()
    if len(luma_filters) > 25 or len(chroma_filters) > 8 or len(cc_cb_filters) > 4 or len(cc_cr_filters) > 4:
        raise ValueError("too many filters")
    for i, f in enumerate(luma_filters):
        a.luma_coeff[i][:] = list(f)
        if luma_clip and i < len(luma_clip_idx):
            a.luma_clip[i][:] = list(luma_clip_idx[i])
    a.class_to_filter[:] = list(filter_coeff_delta_idx if filter_coeff_delta_idx is not None else [0] * 25)
    a.n_luma, a.n_chroma = len(luma_filters), len(chroma_filters)
    for i, f in enumerate(chroma_filters):
        a.chroma_coeff[i][:] = list(f)
        if i < len(chroma_clip_idx):
            a.chroma_clip[i][:] = list(chroma_clip_idx[i])
    a.n_cc_cb, a.n_cc_cr = len(cc_cb_filters), len(cc_cr_filters)
    for i, f in enumerate(cc_cb_filters):
        a.cc_cb[i][:] = list(f)
    for i, f in enumerate(cc_cr_filters):
        a.cc_cr[i][:] = list(f)
    return a


# rmh_alf_ctb (8 bytes): luma 0 off / 1..16 fixed set + 1 / 17 APS; cb, cr alternative + 1; CCALF idc
ALF_CTB_DTYPE = np.dtype([("luma", "u1"), ("cb", "u1"), ("cr", "u1"), ("cc_cb", "u1"), ("cc_cr", "u1"),
                          ("reserved", "u1", (3,))])
assert ALF_CTB_DTYPE.itemsize == 8
ALF_LUMA_APS = 17


class AlfFrame(ctypes.Structure):
    _fields_ = [("src", Plane * 3), ("dst", Plane * 3), ("ctbs", ctypes.c_void_p), ("luma_aps", ctypes.c_void_p),
                ("chroma_aps", ctypes.c_void_p), ("cc_cb_aps", ctypes.c_void_p), ("cc_cr_aps", ctypes.c_void_p),
                ("ctb_size", ctypes.c_uint32), ("ctb_cols", ctypes.c_uint32), ("ctb_rows", ctypes.c_uint32),
                ("vis_w", ctypes.c_uint32), ("vis_h", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


def alf_frame(srcs, dsts, ctbs_addr, ctb_size, ctb_cols, ctb_rows, vis_w, vis_h, luma_aps=None, chroma_aps=None,
              cc_cb_aps=None, cc_cr_aps=None):
    """An AlfFrame; the AlfAps arguments must outlive the call that reads them."""
    f = AlfFrame()
    for i in range(3):
        f.src[i] = srcs[i]
        f.dst[i] = dsts[i]
    f.ctbs = ctbs_addr
    for name, aps in (("luma_aps", luma_aps), ("chroma_aps", chroma_aps), ("cc_cb_aps", cc_cb_aps),
                      ("cc_cr_aps", cc_cr_aps)):
        setattr(f, name, ctypes.addressof(aps) if aps is not None else None)
    f.ctb_size, f.ctb_cols, f.ctb_rows, f.vis_w, f.vis_h = ctb_size, ctb_cols, ctb_rows, vis_w, vis_h
    return f


class GraphNode(ctypes.Structure):
    _fields_ = [("plan", ctypes.c_void_p), ("d_coeff", ctypes.c_void_p), ("d_residual", ctypes.c_void_p)]


def plane_from_address(addr, width, height, stride, bytes_per_sample, bit_depth):
    return Plane(ctypes.c_void_p(addr), width, height, stride, bytes_per_sample, bit_depth, 0)


def plane_from_array(a, bit_depth):
    """A host plane over a C-contiguous 2-D uint8/uint16 numpy array."""
    assert a.ndim == 2 and a.flags["C_CONTIGUOUS"] and a.dtype in (np.uint8, np.uint16)
    return plane_from_address(a.ctypes.data, a.shape[1], a.shape[0], a.shape[1], a.dtype.itemsize, bit_depth)


def dbk_frame(planes, edges_addr, grid_w, grid_h, qps, beta_offsets=(0, 0, 0), tc_offsets=(0, 0, 0)):
    f = DbkFrame()
    for i in range(3):
        f.planes[i] = planes[i]
    f.edges = edges_addr
    f.grid_w, f.grid_h = grid_w, grid_h
    for i in range(3):
        f.qp[i] = qps[i]
        f.beta_offset_div2[i] = beta_offsets[i]
        f.tc_offset_div2[i] = tc_offsets[i]
    return f


def mc_frame(refs, dsts):
    f = McFrame()
    for i in range(3):
        f.ref[i] = refs[i]
        f.dst[i] = dsts[i]
    return f


# rmh_edge_cell: one 4x4 luma cell of the coded picture for the edge-map builder
EDGE_CELL_DTYPE = np.dtype([("tu_y", "<u4"), ("tu_c", "<u4"), ("tu_wy", "u1"), ("tu_hy", "u1"), ("tu_wc", "u1"),
                            ("tu_hc", "u1"), ("flags", "u1"), ("reserved", "u1", (3,)), ("mv_x", "<i4"),
                            ("mv_y", "<i4"), ("ref_poc", "<i4")])
assert EDGE_CELL_DTYPE.itemsize == 28
CELL_INTRA, CELL_CBF = 1, 2


class EdgeFrame(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_void_p), ("edges", ctypes.c_void_p), ("grid_w", ctypes.c_uint32),
                ("grid_h", ctypes.c_uint32), ("vis_w", ctypes.c_uint32), ("vis_h", ctypes.c_uint32),
                ("ctb_size", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]

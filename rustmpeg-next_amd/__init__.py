"""rustmpeg-next_amd: MI355X (gfx950) accelerator for the data-parallel stages of
VVC reconstruction, over librmh (include/rmh.h).

This module is the Python mirror of the accelerator surface a rustmpeg caller
sees (crates/rustmpeg-video/src/accelerator.rs:18-32 for the error type,
crates/rustmpeg-vulkan/src/accelerator.rs for the façade): an ``Accelerator``
that owns one device context, and three provider objects with the shape of the
VVC traits the drop-in adds next to ``DeblockFilter<P>``
(crates/rustmpeg-video/src/av2/deblock.rs:36-48):

    VvcResidualBatch.residuals(tus, coeffs)      dequant + LFNST + inverse transform
    VvcInterPredBatch.predict(frames, pus)       uni-prediction 8/4-tap interpolation
    VvcDeblockFilter.deblock(frames)             whole-picture deblocking, in place

There is no CPU fallback in this package: when librmh.so is missing or the
device is absent every call raises ``Unavailable`` and the caller runs its own
CPU path, as the reference's codecs do when no accelerator is installed.
"""
import ctypes
import os

import numpy as np

from . import abi, hip  # noqa: F401
from .abi import PU_DTYPE, TU_DTYPE  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librmh.so")


class AcceleratorError(Exception):
    """AcceleratorError (crates/rustmpeg-video/src/accelerator.rs:18-32)."""
    status = abi.RMH_EXECUTION

    def __init__(self, message, status=None):
        super().__init__(message)
        if status is not None:
            self.status = status


class Unavailable(AcceleratorError):
    status = abi.RMH_UNAVAILABLE


class InvalidInput(AcceleratorError):
    status = abi.RMH_INVALID_INPUT


class LimitExceeded(AcceleratorError):
    status = abi.RMH_LIMIT_EXCEEDED


class Execution(AcceleratorError):
    status = abi.RMH_EXECUTION


class Timeout(Execution):
    status = abi.RMH_TIMEOUT


_ERRORS = {abi.RMH_UNAVAILABLE: Unavailable, abi.RMH_INVALID_INPUT: InvalidInput,
           abi.RMH_LIMIT_EXCEEDED: LimitExceeded, abi.RMH_EXECUTION: Execution, abi.RMH_TIMEOUT: Timeout}

_lib = None

EXPORTED = [
    "rmh_create", "rmh_destroy", "rmh_last_error", "rmh_info", "rmh_disable", "rmh_is_disabled",
    "rmh_inject_failures", "rmh_residual_batch", "rmh_residual_plan_create", "rmh_residual_plan_run",
    "rmh_residual_plan_run_range", "rmh_mc_batch", "rmh_mc_plan_create", "rmh_mc_plan_run", "rmh_deblock",
    "rmh_deblock_plan_create", "rmh_deblock_plan_run", "rmh_deblock_to_plan_create", "rmh_sao", "rmh_sao_plan_create", "rmh_sao_plan_run", "rmh_plan_destroy", "rmh_plan_bytes", "rmh_graph_create",
    "rmh_graph_launch", "rmh_graph_destroy", "rmh_transform_matrix", "rmh_intra_subset_decode",
    "rmh_residual_recon_plan_create", "rmh_rd_encode_luma", "rmh_cnn_param_count", "rmh_cnn_create",
    "rmh_cnn_destroy", "rmh_cnn_get_params", "rmh_cnn_set_params", "rmh_cnn_forward", "rmh_cnn_predict",
    "rmh_cnn_gradients", "rmh_cnn_train_step", "rmh_cnn_train_step_soft", "rmh_residual_plan_autotune",
    "rmh_mc_bi_batch", "rmh_mc_bi_plan_create", "rmh_deblock_noalloc", "rmh_intra_predict",
    "rmh_edge_map", "rmh_edge_map_plan_create", "rmh_edge_map_plan_run", "rmh_lmcs_tables",
    "rmh_lmcs_plan_create", "rmh_lmcs_plan_run", "rmh_alf_filter_bank", "rmh_alf_plan_create", "rmh_alf_plan_run",
    "rmh_forward_plan_create", "rmh_forward_plan_run", "rmh_forward_plan_evaluate",
    "rmh_gpm_plan_create", "rmh_gpm_plan_run", "rmh_mc_plan_autotune",
    "rmh_rd_encode_luma_stream", "rmh_cabac_residual_stream", "rmh_luma_stream_decode",
]


def load_library(path=LIB_PATH):
    """Loads librmh.so; raises Unavailable when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise Unavailable("librmh.so is not built (run rustmpeg-next_amd/build.py)")
    L = ctypes.CDLL(path)
    vp, sz, st = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    sig = {
        "rmh_create": ([ctypes.POINTER(abi.Config), ctypes.POINTER(vp)], st),
        "rmh_destroy": ([vp], None),
        "rmh_last_error": ([vp], ctypes.c_char_p),
        "rmh_info": ([vp, ctypes.POINTER(abi.DeviceInfo)], st),
        "rmh_disable": ([vp], None),
        "rmh_is_disabled": ([vp], ctypes.c_int),
        "rmh_inject_failures": ([vp, ctypes.c_uint32], None),
        "rmh_residual_batch": ([vp, vp, sz, vp, sz, vp], st),
        "rmh_residual_plan_create": ([vp, vp, sz, sz, ctypes.POINTER(vp)], st),
        "rmh_residual_plan_run": ([vp, vp, vp, vp], st),
        "rmh_residual_recon_plan_create": ([vp, vp, sz, sz, vp, vp, sz, ctypes.POINTER(vp)], st),
        "rmh_residual_plan_run_range": ([vp, sz, sz, vp, vp, vp], st),
        "rmh_residual_plan_autotune": ([vp, vp, vp, vp, ctypes.c_uint32, vp], st),
        "rmh_mc_batch": ([vp, vp, sz, vp, sz], st),
        "rmh_mc_plan_create": ([vp, vp, sz, vp, sz, ctypes.POINTER(vp)], st),
        "rmh_mc_plan_run": ([vp, vp], st),
        "rmh_mc_bi_batch": ([vp, vp, sz, vp, sz], st),
        "rmh_mc_bi_plan_create": ([vp, vp, sz, vp, sz, ctypes.POINTER(vp)], st),
        "rmh_deblock": ([vp, vp, sz], st),
        "rmh_deblock_plan_create": ([vp, vp, sz, ctypes.POINTER(vp)], st),
        "rmh_deblock_plan_run": ([vp, vp], st),
        "rmh_deblock_to_plan_create": ([vp, vp, sz, vp, ctypes.POINTER(vp)], st),
        "rmh_deblock_noalloc": ([vp, vp, sz, vp], st),
        "rmh_edge_map": ([vp, vp, sz], st),
        "rmh_edge_map_plan_create": ([vp, vp, sz, ctypes.POINTER(vp)], st),
        "rmh_edge_map_plan_run": ([vp, vp], st),
        "rmh_forward_plan_create": ([vp, vp, sz, sz, ctypes.POINTER(vp)], st),
        "rmh_forward_plan_run": ([vp, vp, vp, vp], st),
        "rmh_forward_plan_evaluate": ([vp, vp, vp, vp, vp, vp, vp], st),
        "rmh_gpm_plan_create": ([vp, vp, sz, vp, sz, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(vp)], st),
        "rmh_gpm_plan_run": ([vp, vp], st),
        "rmh_mc_plan_autotune": ([vp, vp, ctypes.POINTER(ctypes.c_uint32)], st),
        "rmh_lmcs_tables": ([vp, ctypes.c_uint32, vp, vp, vp, vp], st),
        "rmh_lmcs_plan_create": ([vp, vp, sz, ctypes.POINTER(vp)], st),
        "rmh_lmcs_plan_run": ([vp, vp], st),
        "rmh_alf_filter_bank": ([vp, ctypes.c_uint32, vp], st),
        "rmh_alf_plan_create": ([vp, vp, sz, ctypes.POINTER(vp)], st),
        "rmh_alf_plan_run": ([vp, vp], st),
        "rmh_sao": ([vp, vp, sz], st),
        "rmh_sao_plan_create": ([vp, vp, sz, ctypes.POINTER(vp)], st),
        "rmh_sao_plan_run": ([vp, vp], st),
        "rmh_plan_destroy": ([vp], None),
        "rmh_plan_bytes": ([vp], ctypes.c_uint64),
        "rmh_graph_create": ([vp, vp, sz, ctypes.POINTER(vp)], st),
        "rmh_graph_launch": ([vp, vp], st),
        "rmh_graph_destroy": ([vp], None),
        "rmh_transform_matrix": ([ctypes.c_int32, ctypes.c_int32, vp], st),
        "rmh_intra_predict": ([vp] + [ctypes.c_uint32] * 5 + [ctypes.c_int, ctypes.c_int, vp], st),
        "rmh_rd_encode_luma": ([vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, vp, ctypes.c_float, ctypes.c_float, vp, vp, vp,
                                ctypes.POINTER(abi.RdStats)], st),
        "rmh_rd_encode_luma_stream": ([vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint32, vp, ctypes.c_float, ctypes.c_float, vp, vp, vp,
                                       ctypes.POINTER(abi.RdStats), vp, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)], st),
        "rmh_cabac_residual_stream": ([vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint32, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)], st),
        "rmh_cnn_param_count": ([], sz),
        "rmh_cnn_create": ([vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(vp)], st),
        "rmh_cnn_destroy": ([vp], None),
        "rmh_cnn_get_params": ([vp, vp], st),
        "rmh_cnn_set_params": ([vp, vp], st),
        "rmh_cnn_forward": ([vp, vp, vp, ctypes.c_uint32, vp, vp], st),
        "rmh_cnn_predict": ([vp, vp, vp, ctypes.c_uint32, vp], st),
        "rmh_cnn_gradients": ([vp, vp, vp, vp, ctypes.c_uint32, vp, vp, vp], st),
        "rmh_cnn_train_step": ([vp, vp, vp, vp, ctypes.c_uint32, ctypes.c_float, vp, vp], st),
        "rmh_cnn_train_step_soft": ([vp, vp, vp, vp, ctypes.c_uint32, ctypes.c_float, vp, vp], st),
        "rmh_intra_subset_decode": ([vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, sz, vp, sz, vp, vp, vp], st),
        "rmh_luma_stream_decode": ([vp, vp, sz, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp, vp], st),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _probs_arg(probs, nb):
    if probs is None:
        return None
    pp = np.ascontiguousarray(probs, dtype=np.float32).reshape(-1)
    if pp.size != nb * 6:
        raise InvalidInput("probs must hold 6 values per 32x32 node")
    return pp


def rd_encode_luma_stream(src, qp, mode=abi.RD_FULL, probs=None, p_prune=0.05, p_fallback=0.6, want_costs=False):
    """The RD search with CABAC as the rate (rmh_rd_encode_luma_stream; host
    code). Returns (recon, splits[nodes_y, nodes_x], costs or None, stats,
    stream bytes); stats["bits"] is the stream's length in bits."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    h, w = src.shape
    nb = (h // 32) * (w // 32)
    recon = np.zeros_like(src)
    splits = np.zeros(nb, np.uint8)
    costs = np.zeros(nb * 6, np.float32) if want_costs else None
    pp = _probs_arg(probs, nb)
    st = abi.RdStats()
    cap = max(4096, w * h * 2)
    buf = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t(0)
    s = load_library().rmh_rd_encode_luma_stream(
        src.ctypes.data, w, h, w, qp, mode, pp.ctypes.data if pp is not None else None, p_prune, p_fallback,
        recon.ctypes.data, splits.ctypes.data, costs.ctypes.data if costs is not None else None, ctypes.byref(st),
        buf.ctypes.data, cap, ctypes.byref(n))
    if s != abi.RMH_OK:
        raise _ERRORS.get(s, Execution)("rmh_rd_encode_luma_stream rejected the input")
    stats = {"bits": st.bits, "sse": int(st.sse), "blocks": st.blocks, "candidates": st.candidates,
             "fallbacks": st.fallbacks}
    return (recon, splits.reshape(h // 32, w // 32), (costs.reshape(nb, 6) if costs is not None else None), stats,
            buf[:n.value].tobytes())


def cabac_residual_stream(levels, qp=26, luma=True, dep_quant=False, sign_hiding=False):
    """residual_coding of a stack of TUs levels[n, h, w] into one CABAC
    substream (rmh_cabac_residual_stream; host code). Returns the bytes."""
    lv = np.ascontiguousarray(levels, dtype=np.int32)
    if lv.ndim == 2:
        lv = lv[None]
    n_blocks, h, w = lv.shape
    flags = (abi.CABAC_LUMA if luma else 0) | (abi.CABAC_DEP_QUANT if dep_quant else 0) | \
        (abi.CABAC_SIGN_HIDING if sign_hiding else 0)
    n = ctypes.c_size_t(0)
    cap = 64 + lv.size * 8
    buf = np.zeros(cap, np.uint8)
    s = load_library().rmh_cabac_residual_stream(lv.ctypes.data, w, h, n_blocks, flags, qp, buf.ctypes.data, cap,
                                                 ctypes.byref(n))
    if s != abi.RMH_OK:
        raise _ERRORS.get(s, Execution)("rmh_cabac_residual_stream rejected the input")
    return buf[:n.value].tobytes()


def rd_encode_luma(src, qp, mode=abi.RD_FULL, probs=None, p_prune=0.05, p_fallback=0.6, want_costs=False):
    """RD search over the six splits of every 32x32 node of an 8-bit luma
    picture (rmh_rd_encode_luma; host code, no device needed). Returns
    (recon, splits[nodes_y, nodes_x], costs or None, stats dict)."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    h, w = src.shape
    nb = (h // 32) * (w // 32)
    recon = np.zeros_like(src)
    splits = np.zeros(nb, np.uint8)
    costs = np.zeros(nb * 6, np.float32) if want_costs else None
    pp = None
    if probs is not None:
        pp = np.ascontiguousarray(probs, dtype=np.float32).reshape(-1)
        if pp.size != nb * 6:
            raise InvalidInput("probs must hold 6 values per 32x32 node")
    st = abi.RdStats()
    s = load_library().rmh_rd_encode_luma(src.ctypes.data, w, h, w, qp, mode, pp.ctypes.data if pp is not None else None,
                                          p_prune, p_fallback, recon.ctypes.data, splits.ctypes.data,
                                          costs.ctypes.data if costs is not None else None, ctypes.byref(st))
    if s != abi.RMH_OK:
        raise _ERRORS.get(s, Execution)("rmh_rd_encode_luma rejected the input")
    stats = {"bits": st.bits, "sse": int(st.sse), "blocks": st.blocks, "candidates": st.candidates,
             "fallbacks": st.fallbacks}
    return recon, splits.reshape(h // 32, w // 32), (costs.reshape(nb, 6) if costs is not None else None), stats


def lmcs_tables(aps, bit_depth):
    """LmcsReshaper16::from_aps (rmh_lmcs_tables) for an abi.LmcsAps: (fwd_lut,
    inv_lut, chroma_adj_lut, reshape_pivot)."""
    n = 1 << bit_depth
    fwd, inv = np.zeros(n, np.uint16), np.zeros(n, np.uint16)
    cadj, piv = np.zeros(16, np.int32), np.zeros(17, np.int32)
    st = load_library().rmh_lmcs_tables(ctypes.addressof(aps), bit_depth, fwd.ctypes.data, inv.ctypes.data,
                                        cadj.ctypes.data, piv.ctypes.data)
    if st != abi.RMH_OK:
        raise _ERRORS.get(st, AcceleratorError)("rmh_lmcs_tables", st)
    return fwd, inv, cadj, piv


def alf_filter_bank(aps, bit_depth):
    """The luma filter bank (rmh_alf_filter_bank) of an abi.AlfAps or None:
    int16 (17 banks, 25 classes, 4 transpositions, 12 coefficients + 12 clips)."""
    out = np.zeros((17, 25, 4, 24), np.int16)
    st = load_library().rmh_alf_filter_bank(ctypes.addressof(aps) if aps is not None else None, bit_depth,
                                            out.ctypes.data)
    if st != abi.RMH_OK:
        raise _ERRORS.get(st, AcceleratorError)("rmh_alf_filter_bank", st)
    return out


def intra_predict(plane, bx, by, n, is_luma, mode):
    """The host PLANAR / DC predictor of the intra-subset decode (rmh_intra_predict)
    for the n x n block at (bx, by) of a C-contiguous uint8 plane."""
    plane = np.ascontiguousarray(plane, np.uint8)
    out = np.zeros(n * n, np.int32)
    st = load_library().rmh_intra_predict(plane.ctypes.data, plane.shape[1], plane.shape[0], bx, by, n, int(is_luma),
                                          mode, out.ctypes.data)
    if st != abi.RMH_OK:
        raise _ERRORS.get(st, AcceleratorError)("rmh_intra_predict", st)
    return out.reshape(n, n)


def transform_matrix(tr_type, n):
    """The integer matrix the kernels use (rmh_transform_matrix)."""
    out = np.zeros(n * n, np.int32)
    if load_library().rmh_transform_matrix(tr_type, n, out.ctypes.data) != abi.RMH_OK:
        raise InvalidInput("unsupported transform (%d, %d)" % (tr_type, n))
    return out.reshape(n, n)


class Plan:
    """A validated, device-resident work description (rmh_plan)."""

    def __init__(self, acc, handle, kind):
        self.acc, self.handle, self.kind = acc, handle, kind

    @property
    def bytes(self):
        return int(load_library().rmh_plan_bytes(self.handle))

    def run(self, d_coeff=0, d_residual=0, stream=0):
        L = load_library()
        if self.kind == "residual":
            self.acc._check(L.rmh_residual_plan_run(self.handle, d_coeff, d_residual, stream))
        elif self.kind == "mc":
            self.acc._check(L.rmh_mc_plan_run(self.handle, stream))
        elif self.kind == "sao":
            self.acc._check(L.rmh_sao_plan_run(self.handle, stream))
        elif self.kind == "edge":
            self.acc._check(L.rmh_edge_map_plan_run(self.handle, stream))
        elif self.kind == "forward":
            self.acc._check(L.rmh_forward_plan_run(self.handle, d_coeff, d_residual, stream))
        elif self.kind == "lmcs":
            self.acc._check(L.rmh_lmcs_plan_run(self.handle, stream))
        elif self.kind == "alf":
            self.acc._check(L.rmh_alf_plan_run(self.handle, stream))
        elif self.kind == "gpm":
            self.acc._check(L.rmh_gpm_plan_run(self.handle, stream))
        else:
            self.acc._check(L.rmh_deblock_plan_run(self.handle, stream))

    def evaluate(self, d_residual, d_pred, d_levels, d_recon, d_sse, stream=0):
        """Forward plans: levels, reconstructed residual and per-TU SSE (rmh_forward_plan_evaluate)."""
        self.acc._check(load_library().rmh_forward_plan_evaluate(self.handle, d_residual, d_pred or None, d_levels,
                                                                 d_recon, d_sse, stream))

    def autotune(self, d_coeff, d_residual, stream=0, reps=5):
        """Chooses the tile split per TU size class by timing (rmh_residual_plan_autotune);
        returns {(w, h): TUs-per-tile divisor}."""
        out = np.zeros(49, np.uint8)
        self.acc._check(load_library().rmh_residual_plan_autotune(self.handle, d_coeff, d_residual, stream, reps,
                                                                  out.ctypes.data))
        return {(1 << (i // 7), 1 << (i % 7)): 1 << int(v) for i, v in enumerate(out) if i // 7 >= 2 and i % 7 >= 2}

    def autotune_mc(self, stream=0):
        """Chooses the work items per workgroup of an inter-prediction plan by
        timing (rmh_mc_plan_autotune); returns the chosen count."""
        w = ctypes.c_uint32(0)
        self.acc._check(load_library().rmh_mc_plan_autotune(self.handle, stream, ctypes.byref(w)))
        return int(w.value)

    def run_range(self, first, count, d_coeff, d_residual, stream=0):
        self.acc._check(load_library().rmh_residual_plan_run_range(self.handle, first, count, d_coeff, d_residual,
                                                                   stream))

    def close(self):
        if self.handle:
            load_library().rmh_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A frame step captured as one hipGraph (rmh_graph)."""

    def __init__(self, acc, handle, keep):
        self.acc, self.handle, self._keep = acc, handle, keep

    def launch(self, stream=0):
        self.acc._check(load_library().rmh_graph_launch(self.handle, stream))

    def close(self):
        if self.handle:
            load_library().rmh_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Accelerator:
    """One device context (rmh_ctx); the counterpart of the Vulkan Accelerator
    façade (crates/rustmpeg-vulkan/src/accelerator.rs)."""

    def __init__(self, device=0, timeout_ms=0):
        L = load_library()
        h = ctypes.c_void_p()
        cfg = abi.Config(device, timeout_ms, 0)
        s = L.rmh_create(ctypes.byref(cfg), ctypes.byref(h))
        if s != abi.RMH_OK:
            raise _ERRORS.get(s, Execution)("rmh_create failed (%s)" % abi.STATUS_NAMES.get(s, s), s)
        self.handle = h.value

    # -- context -----------------------------------------------------------
    def _check(self, status):
        if status != abi.RMH_OK:
            msg = load_library().rmh_last_error(self.handle)
            raise _ERRORS.get(status, Execution)((msg or b"").decode(errors="replace"), status)

    def info(self):
        i = abi.DeviceInfo()
        self._check(load_library().rmh_info(self.handle, ctypes.byref(i)))
        return dict(name=i.name.decode(), arch=i.arch.decode(), compute_units=i.compute_units,
                    max_clock_khz=i.max_clock_khz, global_mem_bytes=i.global_mem_bytes, abi_version=i.abi_version)

    def disable(self):
        load_library().rmh_disable(self.handle)

    @property
    def disabled(self):
        return bool(load_library().rmh_is_disabled(self.handle))

    def inject_failures(self, count):
        load_library().rmh_inject_failures(self.handle, count)

    def close(self):
        if getattr(self, "handle", None):
            load_library().rmh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- host-memory stage calls -------------------------------------------
    def residuals(self, tus, coeffs, out=None):
        tus = np.ascontiguousarray(tus, dtype=TU_DTYPE)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.int16)
        if out is None:
            out = np.zeros_like(coeffs)
        assert out.dtype == np.int16 and out.flags["C_CONTIGUOUS"] and out.size == coeffs.size
        self._check(load_library().rmh_residual_batch(self.handle, tus.ctypes.data, len(tus), coeffs.ctypes.data,
                                                      coeffs.size, out.ctypes.data))
        return out

    def predict(self, frames, pus):
        """frames: list of abi.McFrame (host planes); pus: PU_DTYPE array."""
        pus = np.ascontiguousarray(pus, dtype=PU_DTYPE)
        arr = (abi.McFrame * len(frames))(*frames)
        self._check(load_library().rmh_mc_batch(self.handle, ctypes.addressof(arr), len(frames), pus.ctypes.data,
                                                len(pus)))

    def predict_bi(self, frames, pus):
        """Bi-prediction: frames: list of abi.McBiFrame (host planes); pus: BI_PU_DTYPE array."""
        pus = np.ascontiguousarray(pus, dtype=abi.BI_PU_DTYPE)
        arr = (abi.McBiFrame * len(frames))(*frames)
        self._check(load_library().rmh_mc_bi_batch(self.handle, ctypes.addressof(arr), len(frames),
                                                   pus.ctypes.data, len(pus)))

    def deblock(self, frames):
        """frames: list of abi.DbkFrame (host planes, filtered in place)."""
        arr = (abi.DbkFrame * len(frames))(*frames)
        self._check(load_library().rmh_deblock(self.handle, ctypes.addressof(arr), len(frames)))

    def sao(self, frames):
        """frames: list of abi.SaoFrame (host planes and params); writes each
        frame's dst planes (rmh_sao)."""
        arr = (abi.SaoFrame * len(frames))(*frames)
        self._check(load_library().rmh_sao(self.handle, ctypes.addressof(arr), len(frames)))

    def intra_subset_decode(self, width, height, modes, tus, coeffs):
        """Two-pass decode of an intra-subset picture (rmh_intra_subset_decode):
        the residual batch on the device, prediction and add on the host."""
        modes = np.ascontiguousarray(modes, dtype=np.uint8)
        tus = np.ascontiguousarray(tus, dtype=TU_DTYPE)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.int16)
        y = np.zeros((height, width), np.uint8)
        u = np.zeros((height // 2, width // 2), np.uint8)
        v = np.zeros((height // 2, width // 2), np.uint8)
        self._check(load_library().rmh_intra_subset_decode(self.handle, width, height, modes.ctypes.data,
                                                           tus.ctypes.data, len(tus), coeffs.ctypes.data,
                                                           coeffs.size, y.ctypes.data, u.ctypes.data, v.ctypes.data))
        return y, u, v

    def luma_stream_decode(self, data, width, height, qp):
        """Decodes an rmh_rd_encode_luma_stream substream (rmh_luma_stream_decode):
        host parse, the residual batch on the device, prediction and add on the
        host. Returns (recon[height, width], splits[height/32, width/32])."""
        buf = np.frombuffer(bytes(data), np.uint8)
        rec = np.zeros((height, width), np.uint8)
        sp = np.zeros((height // 32) * (width // 32), np.uint8)
        self._check(load_library().rmh_luma_stream_decode(self.handle, buf.ctypes.data if buf.size else None,
                                                          buf.size, width, height, qp, rec.ctypes.data,
                                                          sp.ctypes.data))
        return rec, sp.reshape(height // 32, width // 32)

    # -- device-resident plans ---------------------------------------------
    def residual_plan(self, tus, n_coeff):
        tus = np.ascontiguousarray(tus, dtype=TU_DTYPE)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_residual_plan_create(self.handle, tus.ctypes.data, len(tus), n_coeff,
                                                            ctypes.byref(h)))
        return Plan(self, h.value, "residual")

    def recon_plan(self, tus, n_coeff, place, frames):
        """Residual plan that adds each TU's residual into a picture plane
        (place: PLACE_DTYPE array, one row per TU; frames: list of abi.ReconFrame
        with device planes). Run it with plan.run(d_coeff, 0)."""
        tus = np.ascontiguousarray(tus, dtype=TU_DTYPE)
        place = np.ascontiguousarray(place, dtype=abi.PLACE_DTYPE)
        if len(place) != len(tus):
            raise InvalidInput("one placement per TU")
        arr = (abi.ReconFrame * len(frames))(*frames)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_residual_recon_plan_create(self.handle, tus.ctypes.data, len(tus), n_coeff,
                                                                  place.ctypes.data, ctypes.addressof(arr),
                                                                  len(frames), ctypes.byref(h)))
        return Plan(self, h.value, "residual")

    def mc_plan(self, frames, pus):
        pus = np.ascontiguousarray(pus, dtype=PU_DTYPE)
        arr = (abi.McFrame * len(frames))(*frames)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_mc_plan_create(self.handle, ctypes.addressof(arr), len(frames),
                                                      pus.ctypes.data, len(pus), ctypes.byref(h)))
        return Plan(self, h.value, "mc")

    def mc_bi_plan(self, frames, pus):
        """Bi-prediction plan over abi.McBiFrame entries with device planes; run with plan.run()."""
        pus = np.ascontiguousarray(pus, dtype=abi.BI_PU_DTYPE)
        arr = (abi.McBiFrame * len(frames))(*frames)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_mc_bi_plan_create(self.handle, ctypes.addressof(arr), len(frames),
                                                         pus.ctypes.data, len(pus), ctypes.byref(h)))
        return Plan(self, h.value, "mc")

    def gpm_plan(self, frames, pus, luma_scale=abi.GPM_LUMA_SCALE, chroma_scale=abi.GPM_CHROMA_SCALE):
        """GPM blend plan over abi.GpmFrame entries with device planes (rmh_gpm_plan_create)."""
        pus = np.ascontiguousarray(pus, dtype=abi.GPM_PU_DTYPE)
        arr = (abi.GpmFrame * len(frames))(*frames)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_gpm_plan_create(self.handle, ctypes.addressof(arr), len(frames),
                                                       pus.ctypes.data, len(pus), luma_scale, chroma_scale,
                                                       ctypes.byref(h)))
        return Plan(self, h.value, "gpm")

    def deblock_plan(self, frames):
        arr = (abi.DbkFrame * len(frames))(*frames)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_deblock_plan_create(self.handle, ctypes.addressof(arr), len(frames),
                                                           ctypes.byref(h)))
        return Plan(self, h.value, "deblock")

    def deblock_to_plan(self, frames, dsts):
        """Out-of-place deblocking plan: frames (abi.DbkFrame, device planes) are
        read only; dsts lists 3 abi.Plane per frame that receive the output."""
        arr = (abi.DbkFrame * len(frames))(*frames)
        if len(dsts) != 3 * len(frames):
            raise InvalidInput("three destination planes per frame")
        darr = (abi.Plane * len(dsts))(*dsts)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_deblock_to_plan_create(self.handle, ctypes.addressof(arr), len(frames),
                                                              ctypes.addressof(darr), ctypes.byref(h)))
        return Plan(self, h.value, "deblock")

    def deblock_noalloc(self, frames, stream=0):
        """In-place deblocking of device planes (abi.DbkFrame entries) on `stream`
        without any device allocation (rmh_deblock_noalloc): the library's
        fixed per-GPU scratch holds the tables and plane copies."""
        arr = (abi.DbkFrame * len(frames))(*frames)
        self._check(load_library().rmh_deblock_noalloc(self.handle, ctypes.addressof(arr), len(frames), stream))

    def edge_map(self, cells, vis_w, vis_h, ctb_size):
        """Edge bytes (uint8[grid_h, grid_w, 4]) of one frame from its cell grid
        (abi.EDGE_CELL_DTYPE[grid_h, grid_w]), by rmh_edge_map."""
        cells = np.ascontiguousarray(cells, dtype=abi.EDGE_CELL_DTYPE)
        gh, gw = cells.shape
        edges = np.zeros((gh, gw, 4), np.uint8)
        f = abi.EdgeFrame(cells.ctypes.data, edges.ctypes.data, gw, gh, vis_w, vis_h, ctb_size, 0)
        self._check(load_library().rmh_edge_map(self.handle, ctypes.byref(f), 1))
        return edges

    def edge_map_plan(self, frames):
        """Edge-map plan over abi.EdgeFrame entries with device cells and edges."""
        arr = (abi.EdgeFrame * len(frames))(*frames)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_edge_map_plan_create(self.handle, ctypes.addressof(arr), len(frames),
                                                            ctypes.byref(h)))
        return Plan(self, h.value, "edge")

    def sao_plan(self, frames):
        """SAO plan over abi.SaoFrame entries with device planes and params; run with plan.run()."""
        arr = (abi.SaoFrame * len(frames))(*frames)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_sao_plan_create(self.handle, ctypes.addressof(arr), len(frames),
                                                       ctypes.byref(h)))
        return Plan(self, h.value, "sao")

    def forward_plan(self, tus, n_coeff):
        """Encoder forward plan (rmh_forward_plan_create) over TU_DTYPE descriptors;
        plan.run(d_residual, d_levels) quantises, plan.evaluate(...) also reconstructs and scores."""
        tus = np.ascontiguousarray(tus, dtype=abi.TU_DTYPE)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_forward_plan_create(self.handle, tus.ctypes.data, len(tus), n_coeff,
                                                           ctypes.byref(h)))
        return Plan(self, h.value, "forward")

    def lmcs_plan(self, frames):
        """LMCS luma-mapping plan over abi.LmcsFrame entries with device planes; run with plan.run()."""
        arr = (abi.LmcsFrame * len(frames))(*frames)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_lmcs_plan_create(self.handle, ctypes.addressof(arr), len(frames),
                                                        ctypes.byref(h)))
        return Plan(self, h.value, "lmcs")

    def alf_plan(self, frames):
        """ALF / CCALF plan over abi.AlfFrame entries (device planes and CTB switches,
        host APS records); run with plan.run()."""
        arr = (abi.AlfFrame * len(frames))(*frames)
        h = ctypes.c_void_p()
        self._check(load_library().rmh_alf_plan_create(self.handle, ctypes.addressof(arr), len(frames),
                                                       ctypes.byref(h)))
        return Plan(self, h.value, "alf")

    def graph(self, nodes):
        """nodes: list of (plan, d_coeff, d_residual) (device addresses; 0 for non-residual plans)."""
        arr = (abi.GraphNode * len(nodes))(*[abi.GraphNode(p.handle, c, r) for p, c, r in nodes])
        h = ctypes.c_void_p()
        self._check(load_library().rmh_graph_create(self.handle, ctypes.addressof(arr), len(nodes), ctypes.byref(h)))
        return Graph(self, h.value, [p for p, _, _ in nodes])


class Cnn:
    """The CU-split predictor (rmh_cnn): BF16 CNN on MFMA, forward and training."""

    def __init__(self, acc, max_batch=256, seed=1):
        self.acc = acc
        h = ctypes.c_void_p()
        acc._check(load_library().rmh_cnn_create(acc.handle, max_batch, seed, ctypes.byref(h)))
        self.handle, self.max_batch = h.value, max_batch
        self.n_params = int(load_library().rmh_cnn_param_count())

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        self.acc._check(load_library().rmh_cnn_get_params(self.handle, out.ctypes.data))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise InvalidInput("expected %d parameters" % self.n_params)
        self.acc._check(load_library().rmh_cnn_set_params(self.handle, p.ctypes.data))

    def predict(self, blocks, qp):
        """Host arrays: blocks uint8 (n, 32, 32), qp (n,) normalised -> probs (n, 6)."""
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
        qp = np.ascontiguousarray(qp, dtype=np.float32)
        n = len(blocks)
        out = np.zeros((n, 6), np.float32)
        for a in range(0, n, self.max_batch):
            b = min(n, a + self.max_batch)
            self.acc._check(load_library().rmh_cnn_predict(self.handle, blocks[a:b].ctypes.data, qp[a:b].ctypes.data,
                                                           b - a, out[a:b].ctypes.data))
        return out

    def forward(self, d_blocks, d_qp, n, d_probs, stream=0):
        self.acc._check(load_library().rmh_cnn_forward(self.handle, d_blocks, d_qp, n, d_probs, stream))

    def gradients(self, d_blocks, d_qp, d_labels, n, d_grads, d_loss, stream=0):
        self.acc._check(load_library().rmh_cnn_gradients(self.handle, d_blocks, d_qp, d_labels, n, d_grads, d_loss,
                                                         stream))

    def train_step(self, d_blocks, d_qp, d_labels, n, lr, d_loss, stream=0):
        self.acc._check(load_library().rmh_cnn_train_step(self.handle, d_blocks, d_qp, d_labels, n, lr, d_loss,
                                                          stream))

    def train_step_soft(self, d_blocks, d_qp, d_targets, n, lr, d_loss, stream=0):
        self.acc._check(load_library().rmh_cnn_train_step_soft(self.handle, d_blocks, d_qp, d_targets, n, lr, d_loss,
                                                               stream))

    def close(self):
        if self.handle:
            load_library().rmh_cnn_destroy(self.handle)
            self.handle = None


class VvcResidualBatch:
    """Provider with the shape of the VvcResidualBatch trait."""

    def __init__(self, acc):
        self.acc = acc

    def residuals(self, tus, coeffs, out=None):
        return self.acc.residuals(tus, coeffs, out)


class VvcInterPredBatch:
    """Provider with the shape of the VvcInterPredBatch trait."""

    def __init__(self, acc):
        self.acc = acc

    def predict(self, frames, pus):
        return self.acc.predict(frames, pus)


class VvcDeblockFilter:
    """Provider with the shape of the VvcDeblockFilter trait (planes in place;
    untouched on error, crates/rustmpeg-video/src/av2/deblock.rs:30-35)."""

    def __init__(self, acc):
        self.acc = acc

    def deblock(self, frames):
        return self.acc.deblock(frames)
class VvcSaoFilter:
    """Provider of the SAO stage that the decoder's apply_sao
    (crates/rustmpeg-video/src/vvc/decode_general.rs:5189-5222) would call:
    dst planes written only on success."""

    def __init__(self, acc):
        self.acc = acc

    def sao(self, frames):
        return self.acc.sao(frames)


from . import learned  # noqa: E402
from . import scheduler  # noqa: E402

"""Pins the oracle's LMCS tables and the whole ALF / CCALF picture filter to
the reference's own code: each reference function is read as text and
evaluated in Python (tests/rust_text.py), then compared bit-exactly with the
oracle. Skipped where the reference tree is absent.

  LmcsReshaper16::from_aps             vvc/filter.rs:2628-2705     orc_lmcs_tables
  chroma_scale_for_luma,
  scale_chroma_residual                vvc/filter.rs:2724-2750     orc_lmcs_scale_chroma
  derive_alf_class_4x4_vb_samples/_16  vvc/filter.rs:1558-1691     orc_alf_classify
  apply_alf with alf_luma_block,
  alf_chroma_block and ccalf_block     vvc/decode_general.rs:5225-5660  orc_alf_frame

The ALF frames mix every CTB switch (off, the 16 fixed sets, the APS; chroma
alternatives past the last; CCALF idc past the filter count), APSs with and
without luma clipping, 8/10/12-bit samples, CTB sizes 16 and 32 (so the
virtual boundary rows fall inside the picture) and visible sizes that end
inside a CTB.
"""
import ctypes
import re
import sys
import types
import os

import numpy as np
import pytest

import rust_text
from conftest import reference_text

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
from gen_lfnst_data import parse_array  # noqa: E402


class Plane16:
    """Plane16 (vvc/predict.rs:2866-2945) as the evaluated reference code uses it."""

    def __init__(self, data, bit_depth):
        self.data = np.array(data, dtype=np.int64)
        self.height, self.width = self.data.shape
        self.bit_depth = bit_depth

    def get(self, x, y):
        return int(self.data[y, x])

    def set(self, x, y, v):
        self.data[y, x] = v

    def max_sample(self):
        return (1 << self.bit_depth) - 1

    def clone(self):
        return Plane16(self.data.copy(), self.bit_depth)


def _table(text, name, shape):
    return np.array(parse_array(text, name)).reshape(shape).tolist()


@pytest.fixture(scope="module")
def ref():
    ftext = reference_text("crates/rustmpeg-video/src/vvc/filter.rs")
    dtext = reference_text("crates/rustmpeg-video/src/vvc/decode_general.rs")
    prec = int(re.search(r"const LMCS_FP_PREC: i32 = (\d+);", ftext).group(1))
    consts = {
        "LMCS_FP_PREC": prec,
        "ALF_CLIP_VALS_8BIT": _table(ftext, "ALF_CLIP_VALS_8BIT", (4,)),
        "ALF_CLIP_VALS_10BIT": _table(ftext, "ALF_CLIP_VALS_10BIT", (4,)),
        "ALF_CLIP_VALS_12BIT": _table(ftext, "ALF_CLIP_VALS_12BIT", (4,)),
        "ALF_TRANSPOSE_PERM": _table(ftext, "ALF_TRANSPOSE_PERM", (4, 12)),
        "ALF_FIXED_FILTER_COEFF": _table(ftext, "ALF_FIXED_FILTER_COEFF", (64, 13)),
        "ALF_CLASS_TO_FILTER_MAPPING": _table(ftext, "ALF_CLASS_TO_FILTER_MAPPING", (16, 25)),
    }
    r16 = "impl LmcsReshaper16"
    r = {n: rust_text.compile_fn(ftext, n, extra=consts, after=r16)
         for n in ("from_aps", "chroma_scale_for_luma", "scale_chroma_residual")}
    samples = rust_text.compile_fn(ftext, "derive_alf_class_4x4_vb_samples")
    r["classify"] = rust_text.compile_fn(ftext, "derive_alf_class_4x4_vb_16",
                                         extra={"derive_alf_class_4x4_vb_samples": samples})
    blocks = dict(consts, derive_alf_class_4x4_vb_16=r["classify"])
    for n in ("alf_luma_block", "alf_chroma_block", "ccalf_block"):
        blocks[n] = rust_text.compile_fn(dtext, n, extra=blocks)
    r["apply_alf"] = rust_text.compile_fn(dtext, "apply_alf", extra=blocks)
    return r


def _reshaper(ref, aps_ns, bd):
    """The evaluated from_aps result with its methods bound (self.chroma_scale_for_luma)."""
    r = ref["from_aps"](aps_ns, bd)
    r.chroma_scale_for_luma = lambda avg: ref["chroma_scale_for_luma"](r, avg)
    return r


def _lmcs_cases(rng):
    yield 0, 15, [0] * 16, 0  # identity (filter.rs:2946-2958)
    yield 1, 14, [0, -3, -2, -1, 0, 1, 2, 2, 1, 0, -1, -2, -1, 0, 3, 0], -2  # filter.rs:2855-2877
    for _ in range(6):
        lo = int(rng.integers(0, 8))
        hi = int(rng.integers(lo, 16))
        yield lo, hi, rng.integers(-12, 13, 16).tolist(), int(rng.integers(-7, 8))
    yield 3, 3, [40] + [0] * 15, 7       # one active bin
    yield 0, 15, [-16] * 8 + [16] * 8, -7  # bins of zero codewords (inv_scale 0)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_lmcs_tables_equal_reference_text(po, ref, bd):
    from importlib import import_module
    A = import_module("rustmpeg-next_amd.abi")
    rng = np.random.default_rng(bd)
    for lo, hi, delta, crs in _lmcs_cases(rng):
        init = 1 << (bd - 4)
        delta = [max(d, -init) for d in delta]
        ns = types.SimpleNamespace(min_bin_idx=lo, max_bin_idx=hi, bin_cw_delta=delta, chroma_res_scaling_offset=crs)
        r = _reshaper(ref, ns, bd)
        fwd, inv, cadj, piv = po.lmcs_tables(A.lmcs_aps(lo, hi, delta, crs), bd)
        assert fwd.tolist() == list(r.fwd_lut), (lo, hi, delta)
        assert inv.tolist() == list(r.inv_lut), (lo, hi, delta)
        assert cadj.tolist() == list(r.chroma_adj_lut)
        assert piv.tolist() == list(r.reshape_pivot)
        for avg in (0, 1, init * 3 + 5, (1 << bd) // 2, (1 << bd) - 1):
            res = rng.integers(-(1 << bd) - 40, (1 << bd) + 40, 24).tolist()
            want = list(res)
            ref["scale_chroma_residual"](r, want, avg)
            got, scale = po.lmcs_scale_chroma(A.lmcs_aps(lo, hi, delta, crs), bd, avg, res)
            assert got.tolist() == want
            assert scale == r.chroma_scale_for_luma(avg)
        if (lo, hi, crs) == (0, 15, 0) and not any(delta):
            assert fwd.tolist() == list(range(1 << bd)) == inv.tolist()


def _content(rng, w, h, bd):
    """Smooth gradients, edges in both directions, texture and noise, so that
    classification reaches every class and transposition."""
    maxv = (1 << bd) - 1
    y, x = np.mgrid[0:h, 0:w]
    img = (x * rng.uniform(-3, 3) + y * rng.uniform(-3, 3)) * (1 << (bd - 8))
    img += np.where((x + 2 * y) % 23 < 11, 1, 0) * rng.integers(0, 60) * (1 << (bd - 8))
    img += np.sin(x / rng.uniform(1, 4)) * np.cos(y / rng.uniform(1, 4)) * 40 * (1 << (bd - 8))
    img += rng.integers(-6, 7, (h, w)) * (1 << (bd - 8))
    img += maxv // 2
    return np.clip(img, 0, maxv).astype(np.int64)


def _vvc_aps(rng, luma_clip, n_luma, n_chroma, n_cc):
    """A VvcAlfAps-shaped namespace with random filters."""
    return types.SimpleNamespace(
        luma_clip=luma_clip,
        luma_filters=[rng.integers(-40, 41, 12).tolist() for _ in range(n_luma)],
        luma_clip_idx=[rng.integers(0, 4, 12).tolist() for _ in range(n_luma)],
        filter_coeff_delta_idx=rng.integers(0, n_luma + 2, 25).tolist(),
        chroma_clip=True,
        chroma_filters=[rng.integers(-40, 41, 6).tolist() for _ in range(n_chroma)],
        chroma_clip_idx=[rng.integers(0, 4, 6).tolist() for _ in range(max(n_chroma - 1, 0))],
        cc_cb_filters=[rng.integers(-64, 65, 7).tolist() for _ in range(n_cc)],
        cc_cr_filters=[rng.integers(-64, 65, 7).tolist() for _ in range(n_cc)],
    )


def _abi_aps(A, a):
    if a is None:
        return None
    return A.alf_aps(a.luma_filters, a.luma_clip_idx, a.luma_clip,
                     [min(v, 255) for v in a.filter_coeff_delta_idx], a.chroma_filters, a.chroma_clip_idx,
                     a.cc_cb_filters, a.cc_cr_filters)


ALF_CASES = [  # (bd, ctu, vis_w, vis_h, seed)
    (10, 32, 80, 72, 1),
    (8, 16, 56, 44, 2),
    (12, 32, 72, 40, 3),
    (10, 16, 40, 36, 4),
]


@pytest.mark.parametrize("bd,ctu,vis_w,vis_h,seed", ALF_CASES)
def test_alf_frame_equals_reference_text(po, ref, bd, ctu, vis_w, vis_h, seed):
    from importlib import import_module
    A = import_module("rustmpeg-next_amd.abi")
    rng = np.random.default_rng(seed)
    cols, rows = -(-vis_w // ctu), -(-vis_h // ctu)
    ctb_cols = cols + 1  # a pitch wider than the picture
    pw, ph = vis_w + 8, vis_h + 4  # planes larger than the visible region
    planes = [_content(rng, pw, ph, bd), _content(rng, -(-pw // 2), -(-ph // 2), bd),
              _content(rng, -(-pw // 2), -(-ph // 2), bd)]
    aps_luma = _vvc_aps(rng, seed % 2 == 1, int(rng.integers(1, 26)), 0, 0)
    aps_chroma = _vvc_aps(rng, True, 1, int(rng.integers(1, 9)), 0)
    aps_cc = _vvc_aps(rng, False, 1, 0, int(rng.integers(1, 5)))
    alf_luma, alf_cb, alf_cr, cc_cb, cc_cr = [], [], [], [], []
    ctbs = np.zeros(ctb_cols * rows, A.ALF_CTB_DTYPE)
    for i in range(ctb_cols * rows):
        k = int(rng.integers(0, 4))
        mode = None if k == 0 else ("Aps", 0) if k == 1 else ("Fixed", int(rng.integers(0, 16)))
        cb = None if rng.random() < 0.3 else int(rng.integers(0, len(aps_chroma.chroma_filters) + 2))
        cr = None if rng.random() < 0.3 else int(rng.integers(0, len(aps_chroma.chroma_filters) + 2))
        ccb, ccr = int(rng.integers(0, len(aps_cc.cc_cb_filters) + 2)), int(rng.integers(0, 3))
        alf_luma.append(mode)
        alf_cb.append(cb)
        alf_cr.append(cr)
        cc_cb.append(ccb)
        cc_cr.append(ccr)
        ctbs[i]["luma"] = 0 if mode is None else A.ALF_LUMA_APS if mode[0] == "Aps" else mode[1] + 1
        ctbs[i]["cb"] = 0 if cb is None else cb + 1
        ctbs[i]["cr"] = 0 if cr is None else cr + 1
        ctbs[i]["cc_cb"], ctbs[i]["cc_cr"] = ccb, ccr
    apss = [None] * 8
    apss[2], apss[5], apss[7] = aps_luma, aps_chroma, aps_cc
    g = types.SimpleNamespace(
        y=Plane16(planes[0], bd), u=Plane16(planes[1], bd), v=Plane16(planes[2], bd), ctu=ctu, vis_w=vis_w,
        vis_h=vis_h, ctu_cols=ctb_cols, alf_luma=alf_luma, alf_cb=alf_cb, alf_cr=alf_cr, ccalf_cb=cc_cb,
        ccalf_cr=cc_cr, bit_depth=bd, aps=types.SimpleNamespace(alf=apss),
        sh=types.SimpleNamespace(alf_aps_ids_luma=[2], alf_aps_id_chroma=5, cc_alf_cb_aps_id=7,
                                 cc_alf_cr_aps_id=None if seed == 2 else 7))
    ref["apply_alf"](g)
    la, ca, cca = _abi_aps(A, aps_luma), _abi_aps(A, aps_chroma), _abi_aps(A, aps_cc)
    out = po.alf_frame([p.astype(np.uint16) for p in planes], bd, ctbs, ctu, ctb_cols, rows, vis_w, vis_h,
                       luma_aps=la, chroma_aps=ca, cc_cb_aps=cca, cc_cr_aps=None if seed == 2 else cca)
    for c, want in enumerate((g.y, g.u, g.v)):
        bad = np.argwhere(out[c].astype(np.int64) != want.data)
        assert bad.size == 0, (c, bad[:5].tolist())
    changed = [(out[c].astype(np.int64) != planes[c]).mean() for c in range(3)]
    assert min(changed) > 0.05, changed  # the filters did act on every plane


@pytest.mark.parametrize("bd", [8, 10])
def test_alf_classification_equals_reference_text(po, ref, bd):
    rng = np.random.default_rng(7 + bd)
    w, h, ctu = 48, 72, 32
    img = _content(rng, w, h, bd)
    plane = Plane16(img, bd)
    seen = set()
    for by in range(0, h, 4):
        for bx in range(0, w, 4):
            for vb_h, vb_pos in ((ctu, ctu - 4), (0, 0)):
                want = ref["classify"](plane, bx, by, w - 3, h - 5, vb_h, vb_pos)
                got = po.alf_classify(img.astype(np.uint16), bd, bx, by, w - 3, h - 5, vb_h, vb_pos)
                assert got == tuple(want), (bx, by, vb_h)
                seen.add(got)
    assert len({c for c, _ in seen}) >= 8 and len({t for _, t in seen}) == 4

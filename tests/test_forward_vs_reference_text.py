"""Pins the oracle's encoder-side forward path to the reference's own code:
each reference function is read as text and evaluated in Python
(tests/rust_text.py), then compared bit-exactly with the oracle on random
blocks of every size, transform pair, zero-out, bit depth and QP. Skipped
where the reference tree is absent.

  forward_transform_rect_bd  vvc/dct.rs:865-991          orc_forward_transform
  quantize_rect_bd           vvc/dct.rs:1113-1147        orc_quant_rect
  quantize_dep_quant_bd      vvc/dct.rs:1252-1301        orc_quant_dep_quant
  forward_lfnst_rect         vvc/lfnst.rs:131-204        orc_forward_lfnst
  quantize_ts_with_mode      vvc/transform.rs:415-427    orc_quant_ts_mode

The evaluated functions read the transform matrices through the oracle's
matrix_for_any (pinned entry by entry to dct.rs / tr_tables.rs in
test_tables_vs_reference.py) and every other table from the reference text.
"""
import re

import numpy as np
import pytest

import rust_text
from conftest import reference_text

from test_alf_vs_reference_text import parse_array

SIZES = [(w, h) for w in (1, 2, 4, 8, 16, 32, 64) for h in (1, 2, 4, 8, 16, 32, 64) if w * h >= 2]


@pytest.fixture(scope="module")
def ref(po):
    dct = reference_text("crates/rustmpeg-video/src/vvc/dct.rs")
    lf = reference_text("crates/rustmpeg-video/src/vvc/lfnst.rs")
    tr = reference_text("crates/rustmpeg-video/src/vvc/transform.rs")
    tabs = reference_text("crates/rustmpeg-video/src/vvc/tr_tables.rs")
    const = lambda text, name: int(re.search(r"const %s: i32 = (-?\d+);" % name, text).group(1))

    def matrix_for_any(tr_type, n):
        m = po.transform_matrix(tr_type, n)
        return None if m is None else m.reshape(-1).tolist()

    env = {
        "TrType_Dct2": 0, "TrType_Dst7": 1, "TrType_Dct8": 2,
        "MAX_LOG2_TR_DYNAMIC_RANGE": const(dct, "MAX_LOG2_TR_DYNAMIC_RANGE"),
        "TRANSFORM_MATRIX_SHIFT": const(dct, "TRANSFORM_MATRIX_SHIFT"),
        "QUANT_SCALES": parse_array(dct, "QUANT_SCALES"),
        "QUANT_SCALES_SQRT2": parse_array(dct, "QUANT_SCALES_SQRT2"),
        "matrix_for_any": matrix_for_any,
        "log2": lambda n: n.bit_length() - 1,
        "clip_transform": lambda v: max(-32768, min(32767, v)),
        "is_supported_rect": lambda w, h: 1 <= w <= 64 and 1 <= h <= 64 and w * h >= 2 and not (w & (w - 1))
        and not (h & (h - 1)),
    }
    env["transform_shift_rect_bd"] = rust_text.compile_fn(dct, "transform_shift_rect_bd", extra=env)
    l8 = np.array(parse_array(tabs, "LFNST_8X8")).reshape(4, 2, 16, 48).tolist()
    l4 = np.array(parse_array(tabs, "LFNST_4X4")).reshape(4, 2, 16, 16).tolist()
    scan = re.search(r"const TOP_LEFT_DIAG_SCAN_8X8: \[\(usize, usize\); 64\] = \[(.*?)\];", lf, re.S).group(1)
    lenv = {"LFNST_8X8": l8, "LFNST_4X4": l4,
            "TOP_LEFT_DIAG_SCAN_8X8": [tuple(map(int, p)) for p in re.findall(r"\((\d+),\s*(\d+)\)", scan)]}
    lut = parse_array(tabs, "LFNST_LUT")
    lenv["lfnst_set"] = lambda mode: lut[mode]  # lfnst_set (lfnst.rs:102-107): LFNST_LUT[mode]
    for n in ("parameters_rect", "transposes"):
        lenv[n] = rust_text.compile_fn(lf, n, extra=lenv)
    return {
        "fwd": rust_text.compile_fn(dct, "forward_transform_rect_bd", extra=env),
        "quant": rust_text.compile_fn(dct, "quantize_rect_bd", extra=env),
        "dq": rust_text.compile_fn(dct, "quantize_dep_quant_bd", extra=env),
        "lfnst": rust_text.compile_fn(lf, "forward_lfnst_rect", extra=lenv),
        "ts": rust_text.compile_fn(tr, "quantize_ts_with_mode"),
    }


def _call(fn, *args):
    try:
        return fn(*args)
    except rust_text._RustErr:
        return None


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_forward_transform_equals_reference_text(po, ref, bd):
    rng = np.random.default_rng(bd)
    for w, h in SIZES:
        for _ in range(2):
            th = int(rng.integers(0, 3)) if 4 <= w <= 32 else 0
            tv = int(rng.integers(0, 3)) if 4 <= h <= 32 else 0
            sw = int(rng.integers(0, max(w // 2, 1) + 1)) if rng.random() < 0.3 else 0
            sh = int(rng.integers(0, max(h // 2, 1) + 1)) if rng.random() < 0.3 else 0
            res = rng.integers(-(1 << bd) + 1, 1 << bd, w * h)
            if rng.random() < 0.3:
                res = np.where(rng.random(w * h) < 0.5, (1 << bd) - 1, -(1 << bd) + 1)  # full swing
            want = _call(ref["fwd"], w, h, res.tolist(), th, tv, sw, sh, bd)
            got = po.forward_transform(w, h, res, th, tv, sw, sh, bd)
            if want is None:
                assert got is None, (w, h, th, tv, sw, sh)
            else:
                assert got is not None and got.reshape(-1).tolist() == list(want), (w, h, th, tv, sw, sh, bd)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_quantisers_equal_reference_text(po, ref, bd):
    rng = np.random.default_rng(20 + bd)
    for w, h in SIZES:
        qp = int(rng.integers(0, 52 + 6 * (bd - 8)))
        c = rng.integers(-4000, 4001, w * h)
        c[rng.random(w * h) < 0.4] = 0
        c[rng.random(w * h) < 0.05] = 32767
        for intra in (True, False):
            want = ref["quant"](w, h, c.tolist(), qp, bd, intra)
            assert po.quant_rect(w, h, c, qp, bd, intra).reshape(-1).tolist() == list(want), (w, h, qp, intra)
        scan = po.coeff_scan(w, h).tolist()
        want = ref["dq"](w, h, c.tolist(), scan, qp, bd)
        assert po.quant_dep_quant(w, h, c, qp, bd).reshape(-1).tolist() == list(want), (w, h, qp)
    for qp in range(0, 64):
        for r in (-4095, -300, -7, -1, 0, 1, 5, 255, 4095):
            for intra in (True, False):
                assert po.quant_ts_mode(r, qp, intra) == ref["ts"](r, qp, intra), (r, qp, intra)


def test_forward_lfnst_equals_reference_text(po, ref):
    rng = np.random.default_rng(5)
    for w, h in [(4, 4), (8, 8), (4, 8), (8, 4), (16, 16), (4, 16), (32, 8), (64, 64), (16, 4)]:
        for idx in (1, 2):
            for mode in rng.integers(0, 95, 6).tolist() + [0, 1, 34, 35, 66, 67, 80, 81, 94]:
                c = rng.integers(-2000, 2001, w * h)
                want = list(c.tolist())
                ref["lfnst"](w, h, want, idx, mode)
                got = po.forward_lfnst(w, h, c, idx, mode)
                assert got.reshape(-1).tolist() == want, (w, h, idx, mode)

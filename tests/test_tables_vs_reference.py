"""Compares every normative table the oracle carries or generates with the
reference's own source text (read as text; skipped where the reference tree
is absent). Tables: vvc/dct.rs:58-208 (DCT-2), vvc/tr_tables.rs:14-482
(DST-7, DCT-8, LFNST kernels, LFNST set table), vvc/inter.rs:57-80 and
vvc/predict.rs:86-119 (interpolation filters), vvc/filter.rs:184-198 (tC,
beta), vvc/lfnst.rs:16-81 (top-left 8x8 diagonal scan)."""
import re

import numpy as np
import pytest

from conftest import reference_text

VVC = "crates/rustmpeg-video/src/vvc/"


def parse_table(text, name):
    m = re.search(r"const %s:[^=]*=\s*\[" % re.escape(name), text)
    assert m, name
    i = m.end() - 1
    depth, j = 0, i
    while True:
        c = text[j]
        if c == "[":
            depth += 1
        elif c == "]":
            depth -= 1
            if depth == 0:
                break
        j += 1
    return np.array([int(v) for v in re.findall(r"-?\d+", text[i:j + 1])])


@pytest.fixture(scope="module")
def src():
    return {f: reference_text(VVC + f) for f in ["dct.rs", "tr_tables.rs", "inter.rs", "predict.rs", "filter.rs", "lfnst.rs"]}


def test_dct2_matrices(po, src):
    for n in [2, 4, 8, 16, 32, 64]:
        assert (parse_table(src["dct.rs"], "DCT2_%d" % n) == po.transform_matrix(0, n).reshape(-1)).all(), n


def test_dst7_dct8_matrices(po, src):
    for t, name in [(1, "DST7"), (2, "DCT8")]:
        for n in [4, 8, 16, 32]:
            ref = parse_table(src["tr_tables.rs"], "%s_%d" % (name, n))
            assert (ref == po.transform_matrix(t, n).reshape(-1)).all(), (name, n)


def test_lfnst_tables(po, src):
    l8 = parse_table(src["tr_tables.rs"], "LFNST_8X8").reshape(4, 2, 16, 48)
    l4 = parse_table(src["tr_tables.rs"], "LFNST_4X4").reshape(4, 2, 16, 16)
    lib = po.lib()
    for s in range(4):
        for k in range(2):
            got8 = np.array([[lib.orc_lfnst_coeff(s, k, 1, i, j) for j in range(48)] for i in range(16)])
            got4 = np.array([[lib.orc_lfnst_coeff(s, k, 0, i, j) for j in range(16)] for i in range(16)])
            assert (got8 == l8[s, k]).all() and (got4 == l4[s, k]).all()
    lut = parse_table(src["tr_tables.rs"], "LFNST_LUT")
    assert [lib.orc_lfnst_set(m) for m in range(95)] == lut.tolist()


def test_top_left_scan(po, src):
    body = src["lfnst.rs"]
    m = re.search(r"TOP_LEFT_DIAG_SCAN_8X8[^=]*=\s*\[(.*?)\];", body, re.S)
    pairs = [(int(a), int(b)) for a, b in re.findall(r"\((\d+),\s*(\d+)\)", m.group(1))]
    assert len(pairs) == 64
    got = [(r % 8, r // 8) for r in po.coeff_scan(8, 8)]
    assert got == pairs


def test_interpolation_filters(po, src):
    from test_oracle_mc import LUMA8, _impulse_taps
    ref = parse_table(src["inter.rs"], "LUMA_FILTER_8TAP").reshape(16, 8)
    assert (ref == LUMA8).all()
    for f in range(16):
        assert (_impulse_taps(po, f) == ref[f]).all()
    alt = parse_table(src["inter.rs"], "LUMA_ALT_HPEL_FILTER_8TAP")
    p = np.zeros((16, 32), np.uint16)
    p[8, 16] = 1
    out = po.interp_luma(p, 8, 0, 8, 32, 1, 8, 0, bi=True, alt_hpel=True)[0] + 8192
    assert (out[[19, 18, 17, 16, 15, 14, 13, 12]] == alt).all()
    ch = parse_table(src["predict.rs"], "CHROMA_FILTER_4TAP").reshape(32, 4)
    for f in range(32):
        p = np.zeros((8, 32), np.uint16)
        p[4, 16] = 1
        o = po.interp_chroma(p, 8, 0, 4, 32, 1, f, 0, bi=True)[0] + 8192
        assert (o[[17, 16, 15, 14]] == ch[f]).all(), f


def test_deblocking_tables(po, src):
    tc = parse_table(src["filter.rs"], "DBF_TC_TABLE")
    beta = parse_table(src["filter.rs"], "DBF_BETA_TABLE")
    assert len(tc) == 66 and len(beta) == 64
    for idx in range(66):
        t, b = po.tc_beta(idx, 1, 0, 0, 10)
        assert t == tc[idx]
        if idx < 64:
            assert b == beta[idx] << 2


def _tr_shift(w, h, bd):
    """transform_shift_rect_bd (vvc/dct.rs:858-860) with the -1 of odd areas."""
    lsum = int(np.log2(w)) + int(np.log2(h))
    return 15 - bd - (lsum >> 1) - (lsum & 1)


def test_dequant_scales(po, src):
    """INV_QUANT_SCALES / INV_QUANT_SCALES_SQRT2 (vvc/dct.rs:37-40) read out of
    the oracle's dequantize_rect_bd: a level of 1 at a QP whose right shift is
    0 comes out as the scale itself, for every qp % 6, every bit depth and
    every size class (the odd-area classes use the sqrt(2) table)."""
    inv = parse_table(src["dct.rs"], "INV_QUANT_SCALES")
    inv2 = parse_table(src["dct.rs"], "INV_QUANT_SCALES_SQRT2")
    for bd in range(8, 13):
        for lw in range(1, 7):
            for lh in range(1, 7):
                w, h = 1 << lw, 1 << lh
                if w * h < 4:
                    continue
                # right_shift = 6 - (tr_shift + qp / 6) = 0
                per = 6 - _tr_shift(w, h, bd)
                for rem in range(6):
                    qp = 6 * per + rem
                    if qp < 4 or qp > 63 + 6 * (bd - 8):
                        continue
                    lv = np.zeros((h, w), np.int32)
                    lv[0, 0] = 1
                    lv[h - 1, w - 1] = -1
                    out = po.dequant_rect(w, h, lv, qp, bd)
                    want = (inv2 if (lw + lh) & 1 else inv)[rem]
                    assert out[0, 0] == want and out[h - 1, w - 1] == -want, (w, h, bd, qp)


def test_quant_scales(po, src):
    """QUANT_SCALES / QUANT_SCALES_SQRT2 (vvc/dct.rs:43-46) read out of the
    oracle's quantize_rect_bd: a coefficient of 2^iqbits quantises to the
    scale itself (the rounding offset is below 2^iqbits), with iqbits =
    14 + qp / 6 + tr_shift (vvc/dct.rs:1131-1134)."""
    q = parse_table(src["dct.rs"], "QUANT_SCALES")
    q2 = parse_table(src["dct.rs"], "QUANT_SCALES_SQRT2")
    for bd in (8, 10):
        for lw in range(2, 6):
            for lh in range(2, 6):
                w, h = 1 << lw, 1 << lh
                for qp in range(4, 30):
                    iq = 14 + qp // 6 + _tr_shift(w, h, bd)
                    c = np.zeros((h, w), np.int32)
                    c[0, 0] = 1 << iq
                    c[0, 1] = -(1 << iq)
                    for intra in (True, False):
                        out = po.quant_rect(w, h, c, qp, bd, intra)
                        want = (q2 if (lw + lh) & 1 else q)[qp % 6]
                        assert out[0, 0] == want and out[0, 1] == -want, (w, h, bd, qp, intra)

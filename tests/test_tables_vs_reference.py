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

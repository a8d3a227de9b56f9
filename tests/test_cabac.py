"""CABAC: the library's encoder side (rmh_cabac_residual_stream,
rmh_rd_encode_luma_stream; host code) read back by the oracle's decoder side
(orc_cabac.c), which restates the reference's BinDecoder and
decode_residual_with_options independently of the encoder.

Pins beyond the round trip:
- the context initialisation tables and context-set offsets equal the
  reference text (vvc/cabac.rs:101-260, 307-398);
- the length of a stream of random signs is bounded below by the entropy of
  those signs and above by that entropy plus the cost of bins that adaptation
  makes near-certain: a coder that priced bypass bins wrongly or did not adapt
  its contexts fails one side;
- a stream of the RD search decodes, in the oracle, to the search's own
  reconstruction and splits.
"""
import re

import numpy as np
import pytest

from conftest import load_package, reference_text

rmh = load_package()

CTX_NAMES = ["SPLIT_FLAG", "SPLIT_QT_FLAG", "SPLIT_HV_FLAG", "SPLIT_12_FLAG", "INTRA_LUMA_MPM_FLAG",
             "INTRA_LUMA_PLANAR_FLAG", "QT_CBF_Y", "SIG_COEFF_GROUP_LUMA", "SIG_COEFF_GROUP_CHROMA", "SIG_FLAG_LUMA",
             "SIG_FLAG_CHROMA", "SIG_FLAG_LUMA_1", "SIG_FLAG_CHROMA_1", "SIG_FLAG_LUMA_2", "SIG_FLAG_CHROMA_2",
             "PAR_FLAG_LUMA", "PAR_FLAG_CHROMA", "GTX2_LUMA", "GTX2_CHROMA", "GTX1_LUMA", "GTX1_CHROMA",
             "LAST_X_LUMA", "LAST_X_CHROMA", "LAST_Y_LUMA", "LAST_Y_CHROMA"]


def _header_tables(path, prefix):
    text = open(path).read()
    body = text[text.index("%s_CTX_INIT[3]" % prefix):]
    vals = [int(v) for v in re.findall(r"\{(\d+), (\d+)\}", body) for v in v]
    return np.array(vals).reshape(3, -1, 2)


def test_context_tables_and_offsets_equal_reference_text(po):
    import os
    text = reference_text("crates/rustmpeg-video/src/vvc/cabac.rs")
    ref = []
    for name in ("CTX_INIT_I", "CTX_INIT_P", "CTX_INIT_B"):
        m = re.search(r"pub const %s: \[\(u8, u8\); NUM_CONTEXTS\] = \[(.*?)\];" % name, text, re.S)
        ref.append([(int(a), int(b)) for a, b in re.findall(r"\((\d+),\s*(\d+)\)", m.group(1))])
    ref = np.array(ref)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    np.testing.assert_array_equal(_header_tables(os.path.join(root, "oracle", "cabac_data.h"), "ORC"), ref)
    np.testing.assert_array_equal(
        _header_tables(os.path.join(root, "rustmpeg-next_amd", "csrc", "cabac_data.h"), "RMH"), ref)
    ctx_mod = text[text.index("pub mod ctx {"):]
    product = open(os.path.join(root, "rustmpeg-next_amd", "csrc", "cabac.h")).read()
    for i, name in enumerate(CTX_NAMES):
        want = int(re.search(r"pub const %s: usize = (\d+);" % name, ctx_mod).group(1))
        assert po.cabac_ctx_offset(i) == want, name
        assert int(re.search(r"\b%s = (\d+)," % name, product).group(1)) == want, name
    assert int(re.search(r"NUM_CONTEXTS = (\d+),", product).group(1)) == ref.shape[1]


def _levels(rng, n, w, h, density, big=False):
    lv = np.zeros((n, h, w), np.int32)
    cw, ch = min(w, 32), min(h, 32)
    for b in range(n):
        mask = rng.random((ch, cw)) < density
        mag = rng.geometric(0.35, (ch, cw))
        if big:
            mag = np.where(rng.random((ch, cw)) < 0.1, rng.integers(1, 32768, (ch, cw)), mag)
        sign = np.where(rng.random((ch, cw)) < 0.5, -1, 1)
        blk = (mask * mag * sign).astype(np.int32)
        if not blk.any():
            blk[rng.integers(ch), rng.integers(cw)] = rng.choice([-1, 1]) * int(rng.integers(1, 9))
        lv[b, :ch, :cw] = blk
    return lv


SIZES = [(w, h) for w in (1, 2, 4, 8, 16, 32, 64) for h in (1, 2, 4, 8, 16, 32, 64) if w * h >= 4]


@pytest.mark.parametrize("w,h", SIZES)
def test_residual_round_trip_every_size(po, w, h):
    rng = np.random.default_rng(w * 131 + h)
    for luma in (True, False):
        for density in (0.05, 0.4, 1.0):
            lv = _levels(rng, 6, w, h, density, big=density == 0.4)
            data = rmh.cabac_residual_stream(lv, qp=32, luma=luma)
            got = po.cabac_decode_residuals(data, w, h, len(lv), qp=32, luma=luma)
            assert got is not None, (luma, density)
            np.testing.assert_array_equal(got, lv)


@pytest.mark.parametrize("qp", [0, 22, 37, 51, 63])
def test_residual_round_trip_dependent_quantisation(po, qp):
    rng = np.random.default_rng(qp)
    for w, h in [(4, 4), (8, 8), (16, 8), (32, 32), (64, 16), (4, 32)]:
        lv = _levels(rng, 8, w, h, 0.3, big=True)
        for luma in (True, False):
            data = rmh.cabac_residual_stream(lv, qp=qp, luma=luma, dep_quant=True)
            np.testing.assert_array_equal(po.cabac_decode_residuals(data, w, h, len(lv), qp=qp, luma=luma,
                                                                    dep_quant=True), lv)


def _hide_signs(po, lv):
    """Sign data hiding: in each 4x4 group whose first and last non-zero scan
    positions are more than 3 apart, the sign of the first one is the parity
    of the group's absolute sum (decode_subblock, vvc/residual.rs:858-886)."""
    out = lv.copy()
    n, h, w = lv.shape
    scan = po.coeff_scan(w, h)
    for b in range(n):
        flat = out[b].reshape(-1)
        for g in range(0, len(scan), 16):
            pos = scan[g:g + 16]
            nz = [i for i in range(16) if flat[pos[i]] != 0]
            if nz and nz[-1] - nz[0] > 3:
                s = int(np.abs(flat[pos]).sum())
                flat[pos[nz[0]]] = -abs(flat[pos[nz[0]]]) if s & 1 else abs(flat[pos[nz[0]]])
    return out


def test_residual_round_trip_sign_hiding(po):
    rng = np.random.default_rng(7)
    for w, h in [(4, 4), (8, 4), (16, 16), (32, 8), (32, 32)]:
        lv = _hide_signs(po, _levels(rng, 10, w, h, 0.5))
        data = rmh.cabac_residual_stream(lv, qp=27, sign_hiding=True)
        np.testing.assert_array_equal(po.cabac_decode_residuals(data, w, h, len(lv), qp=27, sign_hiding=True), lv)
        # with the hidden sign left free the stream is shorter by one bypass bin per hiding group
        plain = rmh.cabac_residual_stream(lv, qp=27)
        assert len(data) <= len(plain)


def test_remainder_escape_codes_round_trip(po):
    """Levels whose Rice remainder takes the Exp-Golomb escape, up to the
    int16 limit (the reference's remainder round trip covers 0..40967)."""
    vals = [4, 5, 6, 13, 14, 24, 25, 88, 2004, 32767, -32768 + 1]
    lv = np.zeros((len(vals), 8, 8), np.int32)
    for i, v in enumerate(vals):
        lv[i, 0, 0] = v
        lv[i, 1, 2] = -v // 3 or 1
    data = rmh.cabac_residual_stream(lv, qp=22)
    np.testing.assert_array_equal(po.cabac_decode_residuals(data, 8, 8, len(lv), qp=22), lv)


def test_stream_length_is_bounded_by_the_entropy_of_its_signs(po):
    """N 4x4 TUs each holding one level +-1 at DC: per TU the syntax is two
    last-position bins, one greater-than-1 bin (always the same value, so the
    adaptive contexts make them near-certain), one bypass sign and the end
    flag. Random signs carry exactly N bits, so the stream holds at least N
    bits; adaptation must keep the rest to a small fraction of a bit per TU."""
    n = 4000
    rng = np.random.default_rng(3)
    lv = np.zeros((n, 4, 4), np.int32)
    lv[:, 0, 0] = np.where(rng.random(n) < 0.5, -1, 1)
    data = rmh.cabac_residual_stream(lv, qp=30)
    bits = 8 * len(data)
    assert n <= bits < 1.25 * n, bits
    np.testing.assert_array_equal(po.cabac_decode_residuals(data, 4, 4, n, qp=30), lv)


def test_residual_stream_rejects_bad_input():
    with pytest.raises(rmh.InvalidInput):
        rmh.cabac_residual_stream(np.zeros((1, 4, 4), np.int32))  # no non-zero level
    with pytest.raises(rmh.InvalidInput):
        rmh.cabac_residual_stream(np.ones((1, 4, 6), np.int32))  # not a power of two
    lv = np.zeros((1, 64, 64), np.int32)
    lv[0, 40, 3] = 1  # outside the coded 32x32
    with pytest.raises(rmh.InvalidInput):
        rmh.cabac_residual_stream(lv)


def _content(seed, w=160, h=96):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    img = 60 + 0.8 * xx + 0.5 * yy + rng.normal(0, 6, (h, w))
    img[20:60, 30:80] += 70
    img[50:90, 100:150] = 40 + 60 * ((xx[50:90, 100:150] // 4 + yy[50:90, 100:150] // 4) % 2)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("qp", [22, 27, 32, 37])
def test_rd_stream_decodes_to_the_search_reconstruction(po, qp):
    src = _content(qp)
    rec, splits, costs, st, data = rmh.rd_encode_luma_stream(src, qp, want_costs=True)
    assert st["bits"] == 8 * len(data)
    out = po.decode_luma_stream(data, src.shape[1], src.shape[0], qp)
    assert out is not None
    np.testing.assert_array_equal(out[0], rec)
    np.testing.assert_array_equal(out[1], splits)
    # the chosen split of every node is its cheapest candidate
    np.testing.assert_array_equal(splits.reshape(-1), np.argmin(costs, axis=1))
    assert len(set(splits.reshape(-1).tolist())) > 1  # the content makes the search split somewhere


def test_rd_stream_pruned_and_heuristic_modes_decode(po):
    src = _content(5)
    rng = np.random.default_rng(5)
    nb = (src.shape[0] // 32) * (src.shape[1] // 32)
    probs = rng.dirichlet(np.ones(6) * 0.3, nb).astype(np.float32)
    for mode, p in ((rmh.abi.RD_PRUNED, probs), (rmh.abi.RD_HEURISTIC, None)):
        rec, splits, _, st, data = rmh.rd_encode_luma_stream(src, 32, mode=mode, probs=p)
        out = po.decode_luma_stream(data, src.shape[1], src.shape[0], 32)
        assert out is not None, mode
        np.testing.assert_array_equal(out[0], rec)
        np.testing.assert_array_equal(out[1], splits)


def test_rd_stream_searches_a_second_level_under_quad_splits(po):
    """The children of a quad split choose their own splits: the stream holds
    CUs that only a split of a 16x16 child makes (8x8, 16x4, 4x16 ...), and
    the decoder rebuilds them."""
    src = _content(9)
    rec, splits, costs, st, data = rmh.rd_encode_luma_stream(src, 22, want_costs=True)
    out = po.decode_luma_stream(data, src.shape[1], src.shape[0], 22)
    np.testing.assert_array_equal(out[0], rec)
    np.testing.assert_array_equal(out[1], splits)
    cu = out[2]
    second_level = ((cu[..., 0] < 16) & (cu[..., 1] < 32)) | ((cu[..., 1] < 16) & (cu[..., 0] < 32))
    assert second_level.any()
    # every CU sits inside a node whose split is quad exactly when it is second-level
    qt_nodes = np.repeat(np.repeat(splits == 1, 8, axis=0), 8, axis=1)
    assert not (second_level & ~qt_nodes).any()

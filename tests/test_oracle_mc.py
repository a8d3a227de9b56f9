"""Pins the inter-prediction oracle (oracle/orc_mc.c).

Known answers: crates/rustmpeg-video/src/vvc/inter.rs:2842-2879 (tap sums,
constant plane at every fraction, BCW average of flat planes). The 10-bit
path has no reference caller or test (inter.rs:2010-2134 has a defective
uni-prediction 2-D shift); it is pinned by (a) equality with the 8-bit
function at bd = 8 (same code, head_room = 14 - bd), (b) the constant-plane
invariant at bd = 10, (c) agreement with the reference's 16-bit function as
written on every branch that function gets right, and (d) an independent
numpy separable-FIR computation over an edge-padded plane.
"""
import numpy as np
import pytest

LUMA8 = np.array([
    [0, 0, 0, 64, 0, 0, 0, 0], [0, 1, -3, 63, 4, -2, 1, 0], [-1, 2, -5, 62, 8, -3, 1, 0],
    [-1, 3, -8, 60, 13, -4, 1, 0], [-1, 4, -10, 58, 17, -5, 1, 0], [-1, 4, -11, 52, 26, -8, 3, -1],
    [-1, 3, -9, 47, 31, -10, 4, -1], [-1, 4, -11, 45, 34, -10, 4, -1], [-1, 4, -11, 40, 40, -11, 4, -1],
    [-1, 4, -10, 34, 45, -11, 4, -1], [-1, 4, -10, 31, 47, -9, 3, -1], [-1, 3, -8, 26, 52, -11, 4, -1],
    [0, 1, -5, 17, 58, -10, 4, -1], [0, 1, -4, 13, 60, -8, 3, -1], [0, 1, -3, 8, 62, -5, 2, -1],
    [0, 1, -2, 4, 63, -3, 1, 0]])


def _impulse_taps(po, frac, bd=8):
    """Reads the oracle's horizontal taps back through bi-prediction of an impulse."""
    p = np.zeros((16, 32), np.uint16)
    p[8, 16] = 1
    out = po.interp_luma(p, bd, 0, 8, 32, 1, frac, 0, bi=True)
    return (out[0] + 8192)[[19, 18, 17, 16, 15, 14, 13, 12]]


def test_luma_taps_sum_to_64_and_match_table(po):
    # vvc/inter.rs:2842-2853
    for f in range(16):
        taps = _impulse_taps(po, f)
        assert taps.sum() == 64
        assert (taps == LUMA8[f]).all()


def test_constant_plane_at_all_256_fractions_8bit(po):
    # vvc/inter.rs:2856-2868
    p = np.full((32, 32), 115, np.uint16)
    for fy in range(16):
        for fx in range(16):
            assert (po.interp_luma(p, 8, 8, 8, 8, 8, fx, fy) == 115).all()


def test_constant_plane_at_all_fractions_10bit(po):
    p = np.full((32, 32), 512, np.uint16)
    c = np.full((16, 16), 700, np.uint16)
    for fy in range(16):
        for fx in range(16):
            assert (po.interp_luma(p, 10, 8, 8, 8, 8, fx, fy) == 512).all()
    for fy in range(32):
        for fx in range(32):
            assert (po.interp_chroma(c, 10, 4, 4, 4, 4, fx, fy) == 700).all()


def test_reference_16bit_uni_2d_saturates_as_written(po):
    """The defect the generalised oracle avoids: the reference's 16-bit uni 2-D
    branch maps a constant 512 (10-bit) and 100 (8-bit) to the maximum."""
    p10 = np.full((32, 32), 512, np.uint16)
    assert (po.interp_luma_16_as_written(p10, 10, 8, 8, 8, 8, 5, 7) == 1023).all()
    p8 = np.full((32, 32), 100, np.uint16)
    assert (po.interp_luma_16_as_written(p8, 8, 8, 8, 8, 8, 5, 7) == 255).all()


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_generalised_agrees_with_reference_16bit_on_correct_branches(po, bd):
    rng = np.random.default_rng(bd)
    p = rng.integers(0, 1 << bd, (40, 48)).astype(np.uint16)
    for mvx, mvy in [(0, 0), (16 * 3, -16 * 2), (5, 0), (0, 9), (-7, 0), (0, -13), (37, 0), (0, 8)]:
        for bi in (False, True):
            a = po.interp_luma(p, bd, 12, 10, 16, 8, mvx, mvy, bi=bi)
            b = po.interp_luma_16_as_written(p, bd, 12, 10, 16, 8, mvx, mvy, bi=bi)
            assert (a == b).all(), (mvx, mvy, bi)
    # bi-prediction 2-D also agrees
    for mvx, mvy in [(5, 3), (-9, 14), (33, -41)]:
        a = po.interp_luma(p, bd, 12, 10, 16, 8, mvx, mvy, bi=True)
        b = po.interp_luma_16_as_written(p, bd, 12, 10, 16, 8, mvx, mvy, bi=True)
        assert (a == b).all()


def _np_fir_luma(plane, bd, bx, by, w, h, mvx, mvy):
    """Independent separable FIR over an edge-padded plane (int64), uni-pred."""
    pad = 80
    P = np.pad(plane.astype(np.int64), pad, mode="edge")
    ix, iy = bx + (mvx >> 4), by + (mvy >> 4)
    fx, fy = mvx & 15, mvy & 15
    hr = 14 - bd
    win = P[pad + iy - 3: pad + iy + h + 4, pad + ix - 3: pad + ix + w + 4]
    if fx == 0 and fy == 0:
        return win[3:3 + h, 3:3 + w]
    if fy == 0:
        s = sum(LUMA8[fx][k] * win[3:3 + h, k:k + w] for k in range(8))
        return np.clip((s + 32) >> 6, 0, (1 << bd) - 1)
    if fx == 0:
        s = sum(LUMA8[fy][k] * win[k:k + h, 3:3 + w] for k in range(8))
        return np.clip((s + 32) >> 6, 0, (1 << bd) - 1)
    hs = sum(LUMA8[fx][k] * win[:, k:k + w] for k in range(8))          # (h+7) x w
    tmp = (hs >> (6 - hr)) - 8192
    vs = sum(LUMA8[fy][k] * tmp[k:k + h] for k in range(8))
    return np.clip((vs + 8192 * 64 + (1 << (5 + hr))) >> (6 + hr), 0, (1 << bd) - 1)


@pytest.mark.parametrize("bd", [8, 10])
def test_luma_matches_numpy_fir_including_picture_edges(po, bd):
    rng = np.random.default_rng(100 + bd)
    p = rng.integers(0, 1 << bd, (64, 80)).astype(np.uint16)
    for _ in range(200):
        w, h = 1 << int(rng.integers(3, 6)), 1 << int(rng.integers(3, 6))
        bx, by = int(rng.integers(0, 80 - 8)), int(rng.integers(0, 64 - 8))
        mvx, mvy = int(rng.integers(-40 * 16, 40 * 16)), int(rng.integers(-40 * 16, 40 * 16))
        got = po.interp_luma(p, bd, bx, by, w, h, mvx, mvy)
        assert (got == _np_fir_luma(p, bd, bx, by, w, h, mvx, mvy)).all(), (w, h, bx, by, mvx, mvy)


def test_half_pel_of_even_ramp_is_midpoint(po):
    x = np.arange(64, dtype=np.uint16) * 2
    p = np.tile(x, (16, 1))
    out = po.interp_luma(p, 8, 8, 4, 16, 4, 8, 0)
    assert (out == (np.arange(8, 24) * 2 + 1)[None, :]).all()
    c = po.interp_chroma(p, 8, 8, 4, 16, 4, 16, 0)
    assert (c == (np.arange(8, 24) * 2 + 1)[None, :]).all()


def test_chroma_quarter_sample_taps(po):
    """Chroma frac 16 (half sample) of an impulse reads back CHROMA_FILTER_4TAP[16]."""
    p = np.zeros((8, 32), np.uint16)
    p[4, 16] = 1
    out = po.interp_chroma(p, 8, 0, 4, 32, 1, 16, 0, bi=True)[0] + 8192
    assert out[14:18].tolist()[::-1] == [-4, 36, 36, -4]


def test_bcw_equal_weights_average_flat_planes(po):
    # vvc/inter.rs:2871-2879
    p0 = np.full((16, 16), 80, np.uint16)
    p1 = np.full((16, 16), 160, np.uint16)
    y0 = po.interp_luma(p0, 8, 4, 4, 8, 8, 0, 0, bi=True)
    y1 = po.interp_luma(p1, 8, 4, 4, 8, 8, 0, 0, bi=True)
    assert (po.weighted_average_bi(y0, y1, 2, 8) == 120).all()


# ---- bi-prediction (orc_mc_bi_batch): the bi branch of predict_inter_pu
# without DMVR / BDOF, vvc/inter.rs:1753-1801 -------------------------------

BCW = [-2, 3, 4, 5, 10]  # BCW_WEIGHTS, vvc/inter.rs:128


def _bi_setup(refs0, refs1, bd, fill=7):
    from frames import mc_bi_frame_struct
    dsts = [np.full_like(p, fill) for p in refs0]
    f, keep = mc_bi_frame_struct(refs0, refs1, dsts, bd)
    return f, keep, dsts


def _uni(po, refs, bd, pus_bi, fill=7):
    """The uni-prediction oracle over the L0 half of bi PUs."""
    from frames import mc_frame_struct
    import workloads as W
    dsts = [np.full_like(p, fill) for p in refs]
    f, keep = mc_frame_struct(refs, dsts, bd)
    pus = np.zeros(len(pus_bi), W.PU_DTYPE)
    for k in ("x", "y", "log2_w", "log2_h", "frame"):
        pus[k] = pus_bi[k]
    pus["mv_x"], pus["mv_y"] = pus_bi["mv0_x"], pus_bi["mv0_y"]
    assert po.mc_batch(f, pus) == 0
    return dsts


@pytest.mark.parametrize("bd", [8, 10])
def test_bi_with_identical_lists_equals_uni_at_every_fraction_and_weight(po, bd):
    # p0 = p1 = p: (w0 p + w1 p + off) >> (hr + 3) with w0 + w1 = 8 equals the
    # uni-prediction rounding (nested floors), for every BCW index
    import workloads as W
    d = W.bi_frame(5, width=96, height=64, bit_depth=bd, mv_range=6)
    refs = d["ref0"] if bd > 8 else [p.astype(np.uint8) for p in d["ref0"]]
    pus = d["pus"].copy()
    n = len(pus)
    pus["mv0_x"] = (np.arange(n) * 7) % 16 + 16 * ((np.arange(n) % 5) - 2)
    pus["mv0_y"] = (np.arange(n) * 3) % 16 - 16
    pus["mv1_x"], pus["mv1_y"] = pus["mv0_x"], pus["mv0_y"]
    want = _uni(po, refs, bd, pus)
    for idx in range(8):
        pus["bcw_idx"] = idx
        f, keep, got = _bi_setup(refs, refs, bd)
        assert po.mc_bi_batch(f, pus) == 0
        for c in range(3):
            assert (got[c] == want[c]).all(), (idx, c)


@pytest.mark.parametrize("bd", [8, 10])
def test_bi_flat_planes_closed_form_for_every_bcw_index(po, bd):
    # every interpolation of a flat plane a gives a * 2^hr - 8192, so the
    # output is (w0 a + w1 b + 4) >> 3 clipped (BCW index clamped to 4)
    import workloads as W
    rng = np.random.default_rng(11)
    maxv = (1 << bd) - 1
    d = W.bi_frame(6, width=64, height=48, bit_depth=bd)
    dt = np.uint16 if bd > 8 else np.uint8
    for trial in range(4):
        a = rng.integers(0, maxv + 1, 3)
        b = rng.integers(0, maxv + 1, 3)
        refs0 = [np.full(p.shape, a[c], dt) for c, p in enumerate(d["ref0"])]
        refs1 = [np.full(p.shape, b[c], dt) for c, p in enumerate(d["ref0"])]
        pus = d["pus"].copy()
        pus["bcw_idx"] = (np.arange(len(pus)) + trial) % 8
        f, keep, got = _bi_setup(refs0, refs1, bd)
        assert po.mc_bi_batch(f, pus) == 0
        for i, pu in enumerate(pus):
            w1 = BCW[min(int(pu["bcw_idx"]), 4)]
            w0 = 8 - w1
            x, y = int(pu["x"]), int(pu["y"])
            w, h = 1 << (int(pu["log2_w"]) & 15), 1 << int(pu["log2_h"])
            for c in range(3):
                s = 0 if c == 0 else 1
                want = min(max((w0 * int(a[c]) + w1 * int(b[c]) + 4) >> 3, 0), maxv)
                blk = got[c][y >> s:(y >> s) + (h >> s), x >> s:(x >> s) + (w >> s)]
                assert (blk == want).all(), (trial, i, c)


def test_bi_swapping_lists_with_mirrored_weights_is_the_same(po):
    # BCW_WEIGHTS is symmetric about 4: index k on (L0, L1) equals index 4 - k
    # on (L1, L0)
    import workloads as W
    d = W.bi_frame(7, width=128, height=96, bit_depth=10, mv_range=20)
    pus = d["pus"].copy()
    pus["bcw_idx"] = np.arange(len(pus)) % 5
    f, keep, got = _bi_setup(d["ref0"], d["ref1"], 10)
    assert po.mc_bi_batch(f, pus) == 0
    sw = pus.copy()
    sw["mv0_x"], sw["mv1_x"] = pus["mv1_x"], pus["mv0_x"]
    sw["mv0_y"], sw["mv1_y"] = pus["mv1_y"], pus["mv0_y"]
    sw["bcw_idx"] = 4 - pus["bcw_idx"]
    f2, keep2, got2 = _bi_setup(d["ref1"], d["ref0"], 10)
    assert po.mc_bi_batch(f2, sw) == 0
    for c in range(3):
        assert (got[c] == got2[c]).all()
    # and distinct lists do differ from either uni prediction somewhere
    assert any((got[c] != _uni(po, d["ref0"], 10, pus)[c]).any() for c in range(3))


def test_bi_batch_known_answer_flat_80_160(po):
    # vvc/inter.rs:2871-2879 through the PU driver: equal weights -> 120
    import workloads as W
    pus = np.zeros(1, W.BI_PU_DTYPE)
    pus["x"], pus["y"], pus["log2_w"], pus["log2_h"], pus["bcw_idx"] = 4, 4, 3, 3, 2
    refs0 = [np.full((16, 16), 80, np.uint8), np.full((8, 8), 80, np.uint8), np.full((8, 8), 80, np.uint8)]
    refs1 = [np.full((16, 16), 160, np.uint8), np.full((8, 8), 160, np.uint8), np.full((8, 8), 160, np.uint8)]
    f, keep, got = _bi_setup(refs0, refs1, 8, fill=0)
    assert po.mc_bi_batch(f, pus) == 0
    assert (got[0][4:12, 4:12] == 120).all() and (got[1][2:6, 2:6] == 120).all()
    assert got[0][0, 0] == 0 and got[0][12, 12] == 0

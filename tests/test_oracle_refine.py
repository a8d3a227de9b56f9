"""Pins the oracle's decoder-side bi-prediction refinements (orc_bdof_luma,
orc_dmvr_refine, and the refined bi branch of orc_mc_bi_batch) to the
reference: apply_bdof_luma (crates/rustmpeg-video/src/vvc/inter.rs:1204-1339)
and refine_dmvr_subblock with dmvr_subpel_1d (inter.rs:1341-1459) read as
text and evaluated in Python (tests/rust_text.py), on intermediates from the
pinned interpolators and on random planes and MVs; the reference's own known
answer (identical predictions are kept, inter.rs:2882-2887); and, for DMVR,
what the search must find: a picture pair displaced by a symmetric integer
offset is aligned at that offset (to within the sub-sample step). The text comparisons are skipped
where the reference tree is absent; the known answers are not."""
import re
import types

import numpy as np
import pytest

import rust_text
import workloads as W
from conftest import reference_text


class Mv(types.SimpleNamespace):
    """The Mv methods refine_dmvr_subblock calls (inter.rs:150-219):
    component-wise add / sub and the clip to the 18-bit storage range."""

    def add(self, o):
        return Mv(hor=self.hor + o.hor, ver=self.ver + o.ver)

    def sub(self, o):
        return Mv(hor=self.hor - o.hor, ver=self.ver - o.ver)

    def clip_to_storage(self):
        c = lambda v: max(-(1 << 17), min((1 << 17) - 1, v))
        return Mv(hor=c(self.hor), ver=c(self.ver))


@pytest.fixture(scope="module")
def ref():
    text = reference_text("crates/rustmpeg-video/src/vvc/inter.rs")
    prec = int(re.search(r"pub const IF_INTERNAL_PREC: i32 = (\d+);", text).group(1))
    frac = int(re.search(r"pub const MV_FRACTIONAL_BITS_INTERNAL: i32 = (\d+);", text).group(1))
    table = re.search(r"pub const BILINEAR_FILTER_PREC4: \[\[i32; 2\]; 16\] = \[(.*?)\];", text, re.S).group(1)
    bil = [[int(a), int(b)] for a, b in re.findall(r"\[(-?\d+),\s*(-?\d+)\]", table)]
    assert len(bil) == 16
    env = {"IF_INTERNAL_PREC": prec, "IF_INTERNAL_OFFS": 1 << (prec - 1), "MV_FRACTIONAL_BITS_INTERNAL": frac,
           "BILINEAR_FILTER_PREC4": bil,
           "Mv_from_int_pel": lambda dx, dy: Mv(hor=dx << frac, ver=dy << frac),
           "sads_init": lambda: [[(2 ** 63 - 1) // 4] * 5 for _ in range(5)],
           "sample_clamped": lambda p, x, y: int(p[min(max(y, 0), p.shape[0] - 1), min(max(x, 0), p.shape[1] - 1)])}
    env["right_shift_msb"] = rust_text.compile_fn(text, "right_shift_msb", extra=env)
    env["dmvr_subpel_1d"] = rust_text.compile_fn(text, "dmvr_subpel_1d", extra=env)
    dmvr_text = text.replace("[[i64::MAX / 4; 5]; 5]", "sads_init()").replace("Mv::from_int_pel", "Mv_from_int_pel")
    return {"bdof": rust_text.compile_fn(text, "apply_bdof_luma", extra=env),
            "dmvr": rust_text.compile_fn(dmvr_text, "refine_dmvr_subblock", extra=env)}


def _intermediates(po, rng, w, h, mv=None):
    """Two 14-bit bi intermediates of an 8-bit textured picture, as
    interpolate_luma_block(.., bi = true) gives them."""
    plane = W.textured_plane(rng, 96, 96, 8)
    if mv is None:
        mv = [(int(rng.integers(-200, 200)), int(rng.integers(-200, 200))) for _ in range(2)]
    return [po.interp_luma(plane, 8, 32, 32, w, h, m[0], m[1], bi=True) for m in mv]


def test_bdof_keeps_identical_flat_predictions(po):
    """bdof_preserves_identical_bi_predictions (inter.rs:2882-2887): a flat 99
    plane, both lists the same, stays 99."""
    plane = np.full((16, 16), 99, np.uint16)
    y0 = po.interp_luma(plane, 8, 4, 4, 8, 8, 0, 0, bi=True)
    assert (po.bdof_luma(y0, y0, 8, 8) == 99).all()


@pytest.mark.parametrize("w,h", [(4, 4), (8, 8), (16, 8), (8, 16), (16, 16), (32, 8)])
def test_bdof_equals_reference_text(po, ref, w, h):
    rng = np.random.default_rng(w * 100 + h)
    for k in range(6):
        if k < 4:
            p0, p1 = _intermediates(po, rng, w, h)
        else:  # the full 14-bit range, gradients and differences at their extremes
            p0 = rng.integers(-8192, 8192 + 8 * 255, (h, w))
            p1 = rng.integers(-8192, 8192 + 8 * 255, (h, w))
        want = ref["bdof"](p0.reshape(-1).tolist(), p1.reshape(-1).tolist(), w, h)
        got = po.bdof_luma(p0, p1, w, h)
        assert got.reshape(-1).tolist() == list(want), (w, h, k)


def test_bdof_differs_from_the_average_on_moving_content(po):
    """BDOF is not the plain average: on two predictions of a moving ramp it
    changes samples (a restatement that ignored the optical flow would not)."""
    rng = np.random.default_rng(3)
    p0, p1 = _intermediates(po, rng, 16, 16, mv=[(-24, 8), (24, -8)])
    bd = po.bdof_luma(p0, p1, 16, 16)
    avg = po.weighted_average_bi(p0, p1, 2, 8).reshape(16, 16)
    assert (bd != avg).any()
    assert np.abs(bd.astype(int) - avg.astype(int)).max() <= 16


@pytest.mark.parametrize("dx,dy", [(8, 16), (16, 8), (16, 16), (32, 8), (8, 32)])
def test_dmvr_equals_reference_text(po, ref, dx, dy):
    rng = np.random.default_rng(dx * 7 + dy)
    for k in range(5):
        r0 = W.textured_plane(rng, 80, 72, 8)
        r1 = np.clip(r0.astype(int) + rng.integers(-6, 7, r0.shape), 0, 255).astype(np.uint16)
        if k % 2:
            r1 = np.roll(r1, (int(rng.integers(-2, 3)), int(rng.integers(-2, 3))), axis=(0, 1))
        bx, by = int(rng.integers(-4, 60)), int(rng.integers(-4, 50))
        m0 = (int(rng.integers(-60, 60)), int(rng.integers(-60, 60)))
        m1 = (int(rng.integers(-60, 60)), int(rng.integers(-60, 60)))
        a, b = ref["dmvr"](r0, r1, bx, by, dx, dy, Mv(hor=m0[0], ver=m0[1]), Mv(hor=m1[0], ver=m1[1]))
        got = po.dmvr_refine(r0, r1, bx, by, dx, dy, m0, m1)
        assert got == ((a.hor, a.ver), (b.hor, b.ver)), (dx, dy, k)


def test_dmvr_finds_a_symmetric_integer_displacement(po):
    """L0 = picture moved by +(2 ox, 2 oy) relative to L1 around the block:
    the bilateral search must settle on the integer offset (ox, oy) (cost 0
    there, every other offset costlier on textured content), so L0 + o and
    L1 - o point at the same content. Inside the 3x3 core the parabolic step
    then moves it by less than half a sample along each axis; on the border of
    the 5x5 grid it stays integer. The pair stays symmetric."""
    rng = np.random.default_rng(11)
    for ox, oy in [(1, 0), (0, -1), (2, 1), (-2, -2), (1, -1)]:
        base = W.textured_plane(rng, 120, 120, 8).astype(np.int64)
        base = np.clip(base + rng.integers(-20, 21, base.shape), 0, 255)
        # r0(x, y) = base(x - ox, y - oy) and r1(x, y) = base(x + ox, y + oy):
        # L0 at +o and L1 at -o read the same base samples
        r0 = np.roll(base, (oy, ox), axis=(0, 1)).astype(np.uint16)
        r1 = np.roll(base, (-oy, -ox), axis=(0, 1)).astype(np.uint16)
        m0, m1 = po.dmvr_refine(r0, r1, 40, 40, 16, 16, (0, 0), (0, 0))
        assert m1 == (-m0[0], -m0[1]), (ox, oy, m0, m1)
        if abs(ox) == 2 or abs(oy) == 2:
            assert m0 == (16 * ox, 16 * oy), (ox, oy, m0)
        else:
            assert abs(m0[0] - 16 * ox) < 8 and abs(m0[1] - 16 * oy) < 8, (ox, oy, m0)


def test_dmvr_keeps_the_pair_when_the_centre_is_good(po):
    """Identical references: the centre cost is 0 < dx * dy, so the MVs stay."""
    rng = np.random.default_rng(5)
    r = W.textured_plane(rng, 64, 64, 8)
    assert po.dmvr_refine(r, r, 8, 8, 16, 16, (5, -3), (5, -3)) == ((5, -3), (5, -3))


def test_refined_bi_batch_composes_the_reference_steps(po, ref):
    """orc_mc_bi_batch with the refinement flags: the bi branch of
    predict_inter_pu (inter.rs:1763-1800) composed from the reference-text
    DMVR and BDOF, the pinned interpolators and weighted average; the
    conditions (default BCW weights; DMVR on blocks of at least 8x8 and 128
    samples) decide per PU."""
    import ctypes
    from frames import mc_bi_frame_struct
    rng = np.random.default_rng(21)
    d = W.bi_frame(3, width=96, height=64, bit_depth=8, mv_range=12, ctu=32)
    pus = d["pus"]
    pus["refine"] = rng.integers(0, 4, len(pus))
    pus["bcw_idx"] = np.where(rng.random(len(pus)) < 0.7, 2, pus["bcw_idx"])
    out = [np.zeros_like(p) for p in d["ref0"]]
    f, keep = mc_bi_frame_struct(d["ref0"], d["ref1"], out, 8)
    arr = (type(f) * 1)(f)
    assert po.lib().orc_mc_bi_batch(ctypes.addressof(arr), 1, pus.ctypes.data, len(pus)) == 0
    r0, r1 = d["ref0"], d["ref1"]
    for pu in pus:
        w, h = 1 << (int(pu["log2_w"]) & 15), 1 << int(pu["log2_h"])
        x, y, alt = int(pu["x"]), int(pu["y"]), bool(pu["log2_w"] & 0x10)
        m0, m1 = (int(pu["mv0_x"]), int(pu["mv0_y"])), (int(pu["mv1_x"]), int(pu["mv1_y"]))
        default = int(pu["bcw_idx"]) == 2
        if (pu["refine"] & W.BI_DMVR) and default and w >= 8 and h >= 8 and w * h >= 128:
            a, b = ref["dmvr"](r0[0], r1[0], x, y, w, h, Mv(hor=m0[0], ver=m0[1]), Mv(hor=m1[0], ver=m1[1]))
            m0, m1 = (a.hor, a.ver), (b.hor, b.ver)
        y0 = po.interp_luma(r0[0], 8, x, y, w, h, m0[0], m0[1], bi=True, alt_hpel=alt)
        y1 = po.interp_luma(r1[0], 8, x, y, w, h, m1[0], m1[1], bi=True, alt_hpel=alt)
        if (pu["refine"] & W.BI_BDOF) and default:
            want = np.array(ref["bdof"](y0.reshape(-1).tolist(), y1.reshape(-1).tolist(), w, h)).reshape(h, w)
        else:
            want = po.weighted_average_bi(y0, y1, int(pu["bcw_idx"]), 8).reshape(h, w)
        assert (out[0][y:y + h, x:x + w] == want[:64 - y, :96 - x]).all(), pu
        for c in (1, 2):
            u0 = po.interp_chroma(r0[c], 8, x // 2, y // 2, w // 2, h // 2, m0[0], m0[1], bi=True)
            u1 = po.interp_chroma(r1[c], 8, x // 2, y // 2, w // 2, h // 2, m1[0], m1[1], bi=True)
            wc = po.weighted_average_bi(u0, u1, int(pu["bcw_idx"]), 8).reshape(h // 2, w // 2)
            assert (out[c][y // 2:y // 2 + h // 2, x // 2:x // 2 + w // 2] == wc).all(), (pu, c)

"""Pins the deblocking oracle (oracle/orc_deblock.c).

Known answers: crates/rustmpeg-video/src/vvc/filter.rs:2921-2944 (the
bit-depth form at 8 bits equals the 8-sample short-filter form) and
filter.rs:2962-2969 (a step edge is smoothed); the filter-length tables of
decode_general.rs:5718-5739. The rest are properties the filters must have:
flat and linear content is left alone by the weak filters, the filters are
mirror-symmetric in P and Q up to one unit of rounding, every change stays within
the tC clipping bounds,
and the edge-map driver reproduces the reference's lattice plane drivers.
"""
import numpy as np
import pytest

from workloads import chroma_lens, edge_byte, luma_lens


def test_luma_length_table():
    # decode_general.rs:5718-5728
    assert luma_lens(4, 32) == (1, 1)
    assert luma_lens(32, 4) == (1, 1)
    assert luma_lens(8, 8) == (3, 3)
    assert luma_lens(32, 8) == (7, 3)
    assert luma_lens(8, 64) == (3, 7)
    assert luma_lens(16, 16) == (3, 3)


def test_chroma_length_table():
    # decode_general.rs:5730-5739
    assert chroma_lens(8, 8, False) == (3, 3)
    assert chroma_lens(16, 8, False) == (3, 3)
    assert chroma_lens(8, 8, True) == (1, 3)
    assert chroma_lens(4, 16, False) == (1, 1)
    assert chroma_lens(16, 4, True) == (1, 1)


def test_tc_beta_bit_depth_scaling(po):
    # vvc/filter.rs:653-666: tc = (T + 2^(9-bd)) >> (10-bd) below 10 bits, T << (bd-10) above
    for qp in range(0, 64):
        tc10, b10 = po.tc_beta(qp, 1, 0, 0, 10)
        tc8, b8 = po.tc_beta(qp, 1, 0, 0, 8)
        tc12, b12 = po.tc_beta(qp, 1, 0, 0, 12)
        assert tc8 == (tc10 + 2) >> 2
        assert tc12 == tc10 << 2
        assert b10 == b8 << 2 and b12 == b8 << 4


def test_bd_form_equals_narrow_8bit_form(po):
    # vvc/filter.rs:2921-2944: with lengths 3/3 only p3..q3 matter
    narrow = np.array([[40, 55, 70, 82, 96, 110, 125, 140], [41, 56, 71, 83, 97, 111, 126, 141],
                       [42, 57, 72, 84, 98, 112, 127, 142], [43, 58, 73, 85, 99, 113, 128, 143]])
    for pad in (0, 255, 17):
        wide = np.full((4, 16), pad)
        wide[:, 4:12] = narrow
        a = po.deblock_luma_edge(wide, 32, 32, 2, 0, 0, 8, 3, 3)
        b = po.deblock_luma_edge(np.pad(narrow, ((0, 0), (4, 4))), 32, 32, 2, 0, 0, 8, 3, 3)
        assert (a[:, 4:12] == b[:, 4:12]).all()
        assert (a[:, :4] == pad).all() and (a[:, 12:] == pad).all()


def test_step_edge_is_smoothed(po):
    # vvc/filter.rs:2962-2969
    line = [0] * 4 + [100] * 4 + [108] * 4 + [0] * 4
    out = po.deblock_luma_edge([line] * 4, 30, 30, 2, 0, 0, 8, 3, 3)
    assert out[0, 7] > 100 and out[0, 8] < 108


@pytest.mark.parametrize("bd", [8, 10])
def test_flat_lines_are_unchanged(po, bd):
    for v in [0, 37, (1 << bd) - 1]:
        for bs in (1, 2):
            for lp, lq in [(1, 1), (3, 3), (7, 7), (7, 3), (3, 7), (5, 5), (5, 3)]:
                for qp in (20, 37, 51, 63):
                    l = np.full((4, 16), v)
                    assert (po.deblock_luma_edge(l, qp, qp, bs, 0, 0, bd, lp, lq) == v).all()
                    c = np.full((2, 8), v)
                    assert (po.deblock_chroma_edge(c, qp, bs, 0, 0, min(lp, 3), min(lq, 3), bd) == v).all()


def test_weak_filter_leaves_linear_ramps(po):
    """delta = (9(q0-p0) - 3(q1-p1) + 8) >> 4 vanishes on a ramp (9s - 9s)."""
    for slope in (1, 2, 3):
        line = (np.arange(16) * slope + 100)
        out = po.deblock_luma_edge([line] * 4, 45, 45, 2, 0, 0, 10, 1, 1)
        assert (out == line).all()


def _mirror(lines):
    return np.asarray(lines)[:, ::-1]


@pytest.mark.parametrize("bd", [8, 10])
def test_luma_filter_is_mirror_symmetric(po, bd):
    rng = np.random.default_rng(5 + bd)
    maxv = (1 << bd) - 1
    for _ in range(2000):
        base = rng.integers(0, maxv + 1)
        step = rng.integers(-maxv // 16, maxv // 16 + 1)
        noise = rng.integers(-4 << (bd - 8), (4 << (bd - 8)) + 1, (4, 16))
        lines = np.clip(base + np.where(np.arange(16) >= 8, step, 0) + noise, 0, maxv)
        lp, lq = [(1, 1), (3, 3), (7, 7), (7, 3), (3, 7), (5, 5), (5, 7), (3, 5)][rng.integers(0, 8)]
        bs, qp = int(rng.integers(1, 3)), int(rng.integers(20, 64))
        a = po.deblock_luma_edge(lines, qp, qp, bs, 0, 0, bd, lp, lq)
        b = po.deblock_luma_edge(_mirror(lines), qp, qp, bs, 0, 0, bd, lq, lp)
        # the weak filters round (x + 8) >> 4 toward -inf, so a mirror may differ by one
        assert np.abs(a - _mirror(b)).max() <= 1
        tc, _ = po.tc_beta(qp, bs, 0, 0, bd)
        assert np.abs(a - lines).max() <= 3 * tc


@pytest.mark.parametrize("bd", [8, 10])
def test_chroma_filter_is_mirror_symmetric_for_equal_lengths(po, bd):
    rng = np.random.default_rng(9 + bd)
    maxv = (1 << bd) - 1
    for _ in range(2000):
        base = rng.integers(0, maxv + 1)
        step = rng.integers(-maxv // 16, maxv // 16 + 1)
        lines = np.clip(base + np.where(np.arange(8) >= 4, step, 0) + rng.integers(-3, 4, (2, 8)), 0, maxv)
        l = [1, 3][rng.integers(0, 2)]
        bs, qp = int(rng.integers(1, 3)), int(rng.integers(20, 64))
        a = po.deblock_chroma_edge(lines, qp, bs, 0, 0, l, l, bd)
        b = po.deblock_chroma_edge(_mirror(lines), qp, bs, 0, 0, l, l, bd)
        assert np.abs(a - _mirror(b)).max() <= 1
        tc, _ = po.tc_beta(qp, bs, 0, 0, bd)
        assert np.abs(a - lines).max() <= tc


def test_chroma_bs1_short_sides_is_a_no_op(po):
    # deblock_chroma_edge returns early for lengths 1/1 unless bS == 2 (vvc/filter.rs:947-951)
    lines = np.array([[10, 20, 30, 40, 90, 100, 110, 120]] * 2)
    assert (po.deblock_chroma_edge(lines, 40, 1, 0, 0, 1, 1, 8) == lines).all()
    assert not (po.deblock_chroma_edge(lines, 40, 2, 0, 0, 1, 1, 8) == lines).all()


def _lattice_edges(w, h, step, ctb, bs):
    gw, gh = (w + 3) // 4, (h + 3) // 4
    e = np.zeros((gh, gw, 4), np.uint8)
    ll = 1 if step <= 4 else (7 if step >= 32 else 3)
    for j in range(gh):
        for i in range(gw):
            x, y = 4 * i, 4 * j
            if x % step == 0 and x >= 8 and x + 8 <= w and y + 4 <= h:
                e[j, i, 0] = edge_byte(bs, ll, ll)
            if y % step == 0 and y >= 8 and y + 8 <= h and x + 4 <= w:
                e[j, i, 1] = edge_byte(bs, ll, ll, ctb=(y % ctb == 0))
    return e


def test_edge_map_driver_equals_lattice_plane_driver(po, rmh_types=None):
    """apply_deblock-style driver with a uniform edge map == deblock_luma_plane_16 / deblock_chroma_plane_16."""
    import ctypes
    from workloads import textured_plane
    rng = np.random.default_rng(21)
    for bd, step, ctb in [(10, 8, 32), (8, 16, 64), (10, 32, 64)]:
        w, h = 128, 96
        Y = textured_plane(rng, w, h, bd)
        Y[:, 64:] = np.clip(Y[:, 64:].astype(np.int64) + 40, 0, (1 << bd) - 1)
        U = textured_plane(rng, w // 2, h // 2, bd)
        V = textured_plane(rng, w // 2, h // 2, bd)
        for bs in (1, 2):
            e = _lattice_edges(w, h, step, ctb, bs)
            cstep = max(step // 2, 8)
            for j in range(e.shape[0]):
                for i in range(e.shape[1]):
                    cx, cy = 2 * i, 2 * j
                    if cx % cstep == 0 and cx >= 4 and cx + 4 <= w // 2 and cy + 2 <= h // 2:
                        e[j, i, 2] = edge_byte(bs, 3, 3)
                    if cy % cstep == 0 and cy >= 4 and cy + 4 <= h // 2 and cx + 2 <= w // 2:
                        lp = 1 if cy % max(ctb // 2, cstep) == 0 else 3
                        e[j, i, 3] = edge_byte(bs, lp, 3)
            planes = [Y.copy(), U.copy(), V.copy()]
            _run_frame_driver(po, planes, e, bd, [37, 35, 35])
            ey = po.deblock_luma_plane(Y, bd, step, ctb, 37, bs)
            eu = po.deblock_chroma_plane(U, bd, cstep, ctb // 2, 35, bs)
            ev = po.deblock_chroma_plane(V, bd, cstep, ctb // 2, 35, bs)
            assert (planes[0] == ey).all()
            assert (planes[1] == eu).all() and (planes[2] == ev).all()


def _run_frame_driver(po, planes, edges, bd, qps):
    from frames import dbk_frame_struct
    f, keep = dbk_frame_struct(planes, edges, bd, qps)
    po.lib().orc_deblock_frames(ctypes_addr(f), 1)
    del keep


def ctypes_addr(s):
    import ctypes
    return ctypes.addressof(s)

"""The SAO oracle (oracle/orc_sao.c) against the reference's own SAO tests
(crates/rustmpeg-video/src/vvc/filter.rs:2971-2988, 3023-3081) and what the
definition fixes: edge categories and directions, the visible-boundary rule,
the four bands, and a vectorised restatement of the picture driver."""
import numpy as np
import pytest

import pyoracle as po

A = po._abi()


def _lcg_plane(w, h, seed=12345):
    """The reference test's generator: st = st * 1103515245 + 12345 (mod 2^32),
    sample = (st >> 16) & 0xff."""
    st = seed
    out = np.zeros(w * h, np.uint8)
    for i in range(w * h):
        st = (st * 1103515245 + 12345) & 0xFFFFFFFF
        out[i] = (st >> 16) & 0xFF
    return out.reshape(h, w)


BAND, EDGE = A.SAO_BAND, A.SAO_EDGE


def test_reference_band_offset_known_answer():
    """filter.rs:2971-2988: a constant 64 plane (band 64 >> 3 = 8), band_pos 8,
    offsets [5, -2, 0, 0] -> every sample 69, in both forms."""
    src = np.full((8, 8), 64, np.uint8)
    for eight in (False, True):
        out = po.sao_region(src, 8, 0, 0, 8, 8, 8, 8, (BAND, 8, [5, -2, 0, 0]), eight_bit_form=eight)
        assert (out == 69).all()


def test_16bit_form_equals_8bit_form_reference_cases():
    """filter.rs:3023-3081: the reference's six cases over its LCG plane, region
    (2, 2) 10x10 in a 12x12 visible picture."""
    plane = _lcg_plane(16, 16)
    cases = [(BAND, 0, [1, -1, 2, -2]), (BAND, 30, [3, 4, -3, -4])] + [(EDGE, k, [1, 2, -1, -2]) for k in range(4)]
    for c in cases:
        a = po.sao_region(plane, 8, 2, 2, 10, 10, 12, 12, c, eight_bit_form=True)
        b = po.sao_region(plane, 8, 2, 2, 10, 10, 12, 12, c)
        assert (a == b).all(), c


def test_16bit_form_equals_8bit_form_random():
    rng = np.random.default_rng(3)
    for _ in range(200):
        w, h = int(rng.integers(4, 40)), int(rng.integers(4, 40))
        plane = rng.integers(0, 256, (h, w)).astype(np.uint8)
        t = int(rng.integers(1, 3))
        c = (t, int(rng.integers(0, 32 if t == BAND else 4)), [int(v) for v in rng.integers(-40, 41, 4)])
        bx, by = int(rng.integers(0, w)), int(rng.integers(0, h))
        rw, rh = int(rng.integers(1, w + 4)), int(rng.integers(1, h + 4))
        bw, bh = int(rng.integers(1, w + 4)), int(rng.integers(1, h + 4))
        a = po.sao_region(plane, 8, bx, by, rw, rh, bw, bh, c, eight_bit_form=True)
        b = po.sao_region(plane, 8, bx, by, rw, rh, bw, bh, c)
        assert (a == b).all(), (c, bx, by, rw, rh, bw, bh)


DIRS = {0: ((-1, 0), (1, 0)), 1: ((0, -1), (0, 1)), 2: ((-1, -1), (1, 1)), 3: ((1, -1), (-1, 1))}


@pytest.mark.parametrize("bd", [8, 10])
def test_edge_categories_and_directions(bd):
    """A 3x3 neighbourhood where only the two neighbours along one direction
    differ from the centre: only the class of that direction changes the
    centre, by the offset of the category that the signs give."""
    dt = np.uint8 if bd == 8 else np.uint16
    offs = [7, 3, -3, -7]
    # (sign of c - a, sign of c - b) -> offset index; (0, 0) -> unchanged
    cat = {(-1, -1): 0, (-1, 0): 1, (0, -1): 1, (1, 0): 2, (0, 1): 2, (1, 1): 3}
    for d, ((ax, ay), (bx, by)) in DIRS.items():
        for (sa, sb), k in cat.items():
            p = np.full((3, 3), 100, dt)
            p[1 + ay, 1 + ax] = 100 - 5 * sa
            p[1 + by, 1 + bx] = 100 - 5 * sb
            for cls in range(4):
                out = po.sao_region(p, bd, 1, 1, 1, 1, 3, 3, (EDGE, cls, offs))
                want = 100 + offs[k] if cls == d else 100
                assert int(out[1, 1]) == want, (d, sa, sb, cls)


def test_edge_offset_leaves_constant_plane_and_boundary():
    rng = np.random.default_rng(5)
    flat = np.full((12, 12), 300, np.uint16)
    for cls in range(4):
        assert (po.sao_region(flat, 10, 0, 0, 12, 12, 12, 12, (EDGE, cls, [9, 9, -9, -9])) == 300).all()
    plane = rng.integers(0, 1024, (12, 12)).astype(np.uint16)
    big = [31, 31, -31, -31]
    for cls in range(4):
        out = po.sao_region(plane, 10, 0, 0, 12, 12, 9, 10, (EDGE, cls, big))
        # samples whose neighbour pair leaves the 9x10 visible picture stay
        (ax, ay), (bx, by) = DIRS[cls]
        for y in range(12):
            for x in range(12):
                inside = all(0 <= x + dx < 9 and 0 <= y + dy < 10 for dx, dy in ((ax, ay), (bx, by)))
                if not inside:
                    assert out[y, x] == plane[y, x], (cls, x, y)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_band_offset_touches_only_its_four_bands(bd):
    rng = np.random.default_rng(bd)
    dt = np.uint8 if bd == 8 else np.uint16
    plane = rng.integers(0, 1 << bd, (24, 24)).astype(dt)
    for band_pos in (0, 13, 29, 30, 31):
        offs = [int(v) for v in rng.integers(-60, 61, 4)]
        out = po.sao_region(plane, bd, 0, 0, 24, 24, 24, 24, (BAND, band_pos, offs)).astype(np.int64)
        band = plane.astype(np.int64) >> (bd - 5)
        k = (band - band_pos) % 32
        want = plane.astype(np.int64).copy()
        sel = k < 4
        want[sel] = np.clip(want[sel] + np.array(offs)[k[sel]], 0, (1 << bd) - 1)
        assert (out == want).all()


def _np_sao_picture(planes, bd, params, ctb, ctb_cols, vis_w, vis_h):
    """Vectorised restatement of apply_sao: per component, a per-sample CTB
    parameter lookup, band and edge offsets from shifted copies of the plane."""
    outs = []
    for c, p in enumerate(planes):
        p = p.astype(np.int64)
        H, W = p.shape
        s = 0 if c == 0 else 1
        vw, vh = (vis_w + s) >> s, (vis_h + s) >> s
        cs = ctb >> s
        ys, xs = np.mgrid[0:H, 0:W]
        vis = (xs < vw) & (ys < vh)
        idx = ((np.minimum(ys // cs, params.shape[0] - 1)) * ctb_cols + np.minimum(xs // cs, ctb_cols - 1))
        prm = params.reshape(-1, 3)[idx.ravel(), c].reshape(H, W)
        typ, boc, offs = prm["type"], prm["band_or_class"].astype(np.int64), prm["offsets"].astype(np.int64)
        out = p.copy()
        k = ((p >> (bd - 5)) - boc) % 32
        bsel = vis & (typ == BAND) & (k < 4)
        kk = np.minimum(k, 3)
        bval = np.take_along_axis(offs, kk[..., None], axis=2)[..., 0]
        out[bsel] = np.clip(p[bsel] + bval[bsel], 0, (1 << bd) - 1)
        pw, ph = min(vw, W), min(vh, H)
        for cls, ((ax, ay), (bx, by)) in DIRS.items():
            sel = vis & (typ == EDGE) & ((boc & 3) == cls)
            ok = (xs + ax >= 0) & (xs + ax < pw) & (ys + ay >= 0) & (ys + ay < ph)
            ok &= (xs + bx >= 0) & (xs + bx < pw) & (ys + by >= 0) & (ys + by < ph)
            sel &= ok
            if not sel.any():
                continue
            a = p[np.clip(ys + ay, 0, H - 1), np.clip(xs + ax, 0, W - 1)]
            b = p[np.clip(ys + by, 0, H - 1), np.clip(xs + bx, 0, W - 1)]
            e = 2 + np.sign(p - a) + np.sign(p - b)
            lut = np.stack([offs[..., 0], offs[..., 1], np.zeros_like(p), offs[..., 2], offs[..., 3]], axis=-1)
            val = np.take_along_axis(lut, e[..., None], axis=2)[..., 0]
            out[sel] = np.clip(p[sel] + val[sel], 0, (1 << bd) - 1)
        outs.append(out)
    return outs


def random_sao_params(rng, ctb_rows, ctb_cols, bd):
    prm = np.zeros((ctb_rows, ctb_cols, 3), A.SAO_PARAM_DTYPE)
    prm["type"] = rng.integers(0, 3, prm.shape)
    band = rng.integers(0, 32, prm.shape)
    cls = rng.integers(0, 4, prm.shape)
    prm["band_or_class"] = np.where(prm["type"] == BAND, band, cls)
    lim = (1 << (min(bd, 10) - 5)) - 1
    o = rng.integers(0, lim + 1, prm.shape + (4,))
    sign = np.where(prm["type"][..., None] == EDGE, np.array([1, 1, -1, -1]), rng.choice([-1, 1], prm.shape + (4,)))
    prm["offsets"] = o * sign
    return prm


@pytest.mark.parametrize("bd,W,H,ctb,vis", [(8, 96, 64, 32, (90, 60)), (10, 128, 72, 64, (128, 72)),
                                            (10, 136, 80, 32, (129, 77)), (12, 64, 64, 16, (64, 64))])
def test_picture_driver_matches_vectorised_form(bd, W, H, ctb, vis):
    rng = np.random.default_rng(bd * 1000 + W)
    dt = np.uint8 if bd == 8 else np.uint16
    planes = [rng.integers(0, 1 << bd, (H, W)).astype(dt),
              rng.integers(0, 1 << bd, (H // 2, W // 2)).astype(dt),
              rng.integers(0, 1 << bd, (H // 2, W // 2)).astype(dt)]
    rows, cols = -(-vis[1] // ctb), -(-vis[0] // ctb)
    prm = random_sao_params(rng, rows, cols, bd)
    got = po.sao_frame(planes, bd, prm, ctb, cols, rows, vis[0], vis[1])
    want = _np_sao_picture(planes, bd, prm, ctb, cols, vis[0], vis[1])
    for g, w_ in zip(got, want):
        assert (g.astype(np.int64) == w_).all()

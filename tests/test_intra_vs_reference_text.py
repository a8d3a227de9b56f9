"""Pins the PLANAR / DC intra prediction of the oracle (oracle/orc_intra.c)
and of the library's intra-subset decoder (rmh_intra_predict, host code) to
the reference: the known answers of crates/rustmpeg-video/src/vvc/predict.rs's
own tests (4247-4315, 4557-4565), and the reference's gather_references,
RefSamples::filtered, predict_planar, predict_dc and apply_pdpc read as text
from predict.rs (1084-1139, 1233-1273, 1688-1754) and evaluated in Python
(tests/rust_text.py), composed as predict_block composes them (2619-2646).
The text comparisons are skipped where the reference tree is absent; the
known answers are not."""
import numpy as np
import pytest

import rust_text
from conftest import load_package, reference_text

PLANAR, DC = 0, 1


def oracle_predict(po, plane, bx, by, w, h, luma, mode):
    import ctypes
    p = np.ascontiguousarray(plane, np.uint8)
    out = np.zeros(w * h, np.int32)
    po.lib().orc_intra_predict(p.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), p.shape[1], p.shape[0], bx, by, w, h,
                               int(luma), mode, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out.reshape(h, w)


def oracle_choose(po, rec, src, bx, by, n):
    import ctypes
    U8P = ctypes.POINTER(ctypes.c_uint8)
    rec = np.ascontiguousarray(rec, np.uint8)
    src = np.ascontiguousarray(src, np.uint8)
    return po.lib().orc_intra_choose_mode(rec.ctypes.data_as(U8P), src.ctypes.data_as(U8P), rec.shape[1],
                                          rec.shape[0], bx, by, n)


class _Plane:
    """The Plane accessors gather_references uses (vvc/predict.rs:987-1000)."""

    def __init__(self, a):
        self.a = a
        self.height, self.width = a.shape

    def get(self, x, y):
        return int(self.a[y, x])


@pytest.fixture(scope="module")
def ref():
    text = reference_text("crates/rustmpeg-video/src/vvc/predict.rs")
    fs = {}
    for name in ["log2_exact", "gather_references", "filtered", "predict_planar", "predict_dc", "apply_pdpc"]:
        fs[name] = rust_text.compile_fn(text, name, extra=fs)

    def predict_block(plane, bx, by, w, h, luma, mode):
        refs = fs["gather_references"](_Plane(plane), bx, by, w, h)
        # uses_reference_filter (predict.rs:1287-1292) for PLANAR / DC; its
        # rule is pinned by the reference's own known answers below
        used = fs["filtered"](refs) if (mode == PLANAR and luma and w * h > 32) else refs
        pred = (fs["predict_planar"] if mode == PLANAR else fs["predict_dc"])(used, w, h)
        fs["apply_pdpc"](pred, used, w, h)
        return np.array(pred, np.int64).reshape(h, w)

    return predict_block


def _planes(rng, n):
    out = []
    for k in range(n):
        w, h = int(rng.choice([40, 64, 72, 96])), int(rng.choice([36, 64, 80]))
        kind = k % 3
        if kind == 0:
            a = rng.integers(0, 256, (h, w))
        elif kind == 1:
            yy, xx = np.mgrid[0:h, 0:w]
            a = (xx * rng.integers(1, 5) + yy * rng.integers(-3, 4) + rng.integers(0, 60)) % 256
        else:
            a = np.full((h, w), int(rng.integers(0, 256)))
            a[rng.integers(0, h):, :] = rng.integers(0, 256)
        out.append(a.astype(np.uint8))
    return out


def test_oracle_planar_dc_equal_reference_text(po, ref):
    rng = np.random.default_rng(11)
    n = 0
    for plane in _planes(rng, 30):
        ph, pw = plane.shape
        for w in (4, 8, 16, 32):
            for h in (4, 8, 16, 32):
                if w > pw or h > ph:
                    continue
                for bx, by in [(0, 0), (pw - w, 0), (0, ph - h), (pw - w, ph - h),
                               (int(rng.integers(0, pw - w + 1)), int(rng.integers(0, ph - h + 1)))]:
                    for luma in (True, False):
                        for mode in (PLANAR, DC):
                            want = ref(plane, bx, by, w, h, luma, mode)
                            got = oracle_predict(po, plane, bx, by, w, h, luma, mode)
                            assert (got == want).all(), (w, h, bx, by, luma, mode)
                            n += 1
    assert n > 3000


def test_library_predictor_equals_oracle_and_reference_text(po, ref):
    rmh = load_package()
    rng = np.random.default_rng(12)
    for plane in _planes(rng, 20):
        ph, pw = plane.shape
        for n in (4, 8, 16, 32):
            if n > pw or n > ph:
                continue
            for bx, by in [(0, 0), (pw - n, ph - n), (int(rng.integers(0, pw - n + 1)), int(rng.integers(0, ph - n + 1)))]:
                for luma in (True, False):
                    for mode in (PLANAR, DC):
                        got = rmh.intra_predict(plane, bx, by, n, luma, mode)
                        assert (got == oracle_predict(po, plane, bx, by, n, n, luma, mode)).all(), (n, bx, by)
                        assert (got == ref(plane, bx, by, n, n, luma, mode)).all(), (n, bx, by)


# ---- the reference's own known answers (vvc/predict.rs:4247-4315, 4557-4565)

def _both(po, plane, bx, by, n, luma, mode):
    rmh = load_package()
    a = oracle_predict(po, plane, bx, by, n, n, luma, mode)
    b = rmh.intra_predict(plane, bx, by, n, luma, mode)
    assert (a == b).all()
    return a


def test_planar_and_dc_of_flat_reconstruction_are_flat(po):
    assert (_both(po, np.full((64, 64), 77, np.uint8), 32, 32, 32, True, PLANAR) == 77).all()
    assert (_both(po, np.full((64, 64), 42, np.uint8), 32, 32, 32, True, DC) == 42).all()


def test_top_left_corner_block_uses_neutral_references(po):
    assert (_both(po, np.zeros((32, 32), np.uint8), 0, 0, 32, True, PLANAR) == 128).all()


def test_reference_padding_repeats_last_available_sample(po, ref):
    """gather_references at (0, 16) of a plane of 9s: the missing left column
    is padded from above, so every reference, and every prediction, is 9."""
    plane = np.full((32, 32), 9, np.uint8)
    for mode in (PLANAR, DC):
        assert (_both(po, plane, 0, 16, 16, True, mode) == 9).all()


def test_choose_intra_mode_prefers_dc_on_half_dark_half_bright(po):
    recon = np.zeros((64, 64), np.uint8)
    recon[31, :] = 200
    src = np.zeros((64, 64), np.uint8)
    src[32:, 32:] = 100
    assert oracle_choose(po, recon, src, 32, 32, 32) == DC


def test_choose_intra_mode_prefers_planar_on_a_tie(po):
    recon = np.zeros((32, 32), np.uint8)
    src = np.full((32, 32), 128, np.uint8)
    assert oracle_choose(po, recon, src, 0, 0, 32) == PLANAR


def test_reference_filter_only_for_luma_planar_above_32_samples(po, ref):
    """uses_reference_filter(true, Planar, 4, 8) and (4, 4) are false, (8, 8)
    true, and chroma never filters (predict.rs:4557-4565): seen in the
    prediction, which equals the unfiltered or filtered composition."""
    import types
    text = reference_text("crates/rustmpeg-video/src/vvc/predict.rs")
    fs = {}
    for name in ["log2_exact", "gather_references", "filtered", "predict_planar", "apply_pdpc"]:
        fs[name] = rust_text.compile_fn(text, name, extra=fs)

    def planar(plane, bx, by, w, h, filt):
        refs = fs["gather_references"](_Plane(plane), bx, by, w, h)
        used = fs["filtered"](refs) if filt else refs
        pred = fs["predict_planar"](used, w, h)
        fs["apply_pdpc"](pred, used, w, h)
        return np.array(pred).reshape(h, w)

    rng = np.random.default_rng(13)
    plane = rng.integers(0, 256, (64, 64)).astype(np.uint8)
    for (w, h, luma, filt) in [(4, 8, True, False), (4, 4, True, False), (8, 8, True, True),
                               (32, 32, False, False)]:
        got = oracle_predict(po, plane, 16, 16, w, h, luma, PLANAR)
        assert (got == planar(plane, 16, 16, w, h, filt)).all(), (w, h, luma)
        assert not (got == planar(plane, 16, 16, w, h, not filt)).all(), (w, h, luma)

"""Oracle of the residual add (orc_recon_batch) and the P-frame workload."""
import ctypes

import numpy as np
import pytest

import pipeline as PL
import workloads as W
from conftest import load_package

abi = load_package().abi


def _recon_numpy(po, tus, coeff, place, planes):
    """pred + residual clipped to [0, 2^bd - 1] with residuals from the
    (separately pinned) residual oracle, placed by numpy slicing."""
    res = po.residual_batch(tus, coeff)
    out = [[p.astype(np.int64) for p in fr] for fr in planes]
    for t, pl in zip(tus, place):
        w, h = 1 << int(t["log2_w"]), 1 << int(t["log2_h"])
        r = res[int(t["offset"]):int(t["offset"]) + w * h].reshape(h, w)
        tgt = out[int(pl["frame"])][int(pl["plane"])]
        x, y = int(pl["x"]), int(pl["y"])
        tgt[y:y + h, x:x + w] += r
    return [[np.clip(p, 0, (1 << int(t0)) - 1) for p, t0 in zip(fr, bds)]
            for fr, bds in zip(out, [[tus["bit_depth"][0]] * 3] * len(out))]


def _run_oracle(po, tus, coeff, place, planes, bd):
    frames = (abi.ReconFrame * len(planes))(*[abi.ReconFrame((abi.Plane * 3)(*[abi.plane_from_array(p, bd)
                                                                                for p in fr])) for fr in planes])
    return po.lib().orc_recon_batch(tus.ctypes.data, len(tus), coeff.ctypes.data_as(po.I16P), coeff.size,
                                    place.ctypes.data, ctypes.addressof(frames), len(planes))


@pytest.mark.parametrize("bd", [8, 10])
def test_recon_oracle_matches_numpy_add(po, bd):
    f = W.pframe(21, width=256, height=128, bit_depth=bd)
    rng = np.random.default_rng(5)
    maxv = (1 << bd) - 1
    dt = PL.sample_dtype(bd)
    # predictions with many samples at the range ends so both clips fire
    planes = [[np.ascontiguousarray(rng.choice([0, 1, maxv - 1, maxv, maxv // 2], size=p.shape).astype(dt))
               for p in f["ref"]]]
    want = _recon_numpy(po, f["tus"], f["coeff"], f["place"], planes)
    assert _run_oracle(po, f["tus"], f["coeff"], f["place"], planes, bd) == 0
    for c in range(3):
        assert (planes[0][c] == want[0][c]).all(), c
    lo = sum(int((p == 0).sum()) for p in planes[0])
    hi = sum(int((p == maxv).sum()) for p in planes[0])
    assert lo > 1000 and hi > 1000


def test_recon_oracle_zero_coefficients_keep_prediction(po):
    f = W.pframe(22, width=128, height=64, bit_depth=10)
    planes = [[p.copy() for p in f["ref"]]]
    before = [p.copy() for p in planes[0]]
    assert _run_oracle(po, f["tus"], np.zeros_like(f["coeff"]), f["place"], planes, 10) == 0
    for c in range(3):
        assert (planes[0][c] == before[c]).all()


def test_recon_oracle_rejects_block_outside_plane(po):
    f = W.pframe(23, width=128, height=64, bit_depth=10)
    place = f["place"].copy()
    place["x"][0] = 126
    planes = [[p.copy() for p in f["ref"]]]
    assert _run_oracle(po, f["tus"], f["coeff"], place, planes, 10) == 1


@pytest.mark.parametrize("size", [(1920, 1080), (856, 480), (200, 120)])
def test_pframe_tus_tile_the_coded_cus(size):
    """Each non-skipped CU is covered exactly once by its luma TUs and its
    chroma TUs cover the half-size area in each chroma plane; skipped CUs
    have no TUs; nothing overlaps."""
    w, h = size
    f = W.pframe(31, width=w, height=h, bit_depth=10)
    cover = [np.zeros((h, w), np.int32), np.zeros(((h + 1) // 2, (w + 1) // 2), np.int32),
             np.zeros(((h + 1) // 2, (w + 1) // 2), np.int32)]
    for t, pl in zip(f["tus"], f["place"]):
        bw, bh = 1 << int(t["log2_w"]), 1 << int(t["log2_h"])
        assert 2 <= t["log2_w"] <= 6 and 2 <= t["log2_h"] <= 6
        cover[int(pl["plane"])][int(pl["y"]):int(pl["y"]) + bh, int(pl["x"]):int(pl["x"]) + bw] += 1
    assert max(int(c.max()) for c in cover) == 1
    coded = 0
    for (x, y, cw, ch) in f["cus"]:
        s = cover[0][y:y + ch, x:x + cw]
        assert s.min() == s.max()
        coded += int(s.min())
        for c in (1, 2):
            sc = cover[c][y // 2:(y + ch) // 2, x // 2:(x + cw) // 2]
            assert sc.min() == sc.max() == s.min()
    assert 0.6 * len(f["cus"]) < coded < len(f["cus"])


def test_joint_cbcr_oracle_equals_reference_text(po):
    """Joint Cb-Cr: the mode selection and residual derivation of
    reconstruct_joint_cbcr (vvc/decode_general.rs:3908-3981), read as text
    and evaluated in Python (rust_text), applied to the coded residual and
    arbitrary values for the other component (which the reference computes
    from that component's own coefficients and then overwrites), equal the
    oracle's recon of both planes over a mid-grey prediction."""
    import types

    import rust_text
    from conftest import reference_text
    text = reference_text("crates/rustmpeg-video/src/vvc/decode_general.rs")
    _, body = rust_text.fn_source(text, "reconstruct_joint_cbcr")
    src = "\n".join(rust_text.block(body[body.index("let mode ="):body.index("if g.lmcs_on")], ""))
    tus, coeff, place, blocks = W.joint_cbcr_frame(3, bit_depth=10)
    res = po.residual_batch(tus, coeff)
    rng = np.random.default_rng(9)
    planes = [[np.full((64, 128), 512, np.uint16), np.full((64, 128), 512, np.uint16),
               np.full((64, 128), 512, np.uint16)]]
    sel = [i for i, b in enumerate(blocks) if b[5]]
    assert len(sel) > 20 and {blocks[i][5] for i in sel} == {1, 2, 3, 5, 6, 7}
    assert _run_oracle(po, tus[sel], coeff, place[sel], planes, 10) == 0
    for i in sel:
        plane, x, y, w, h, joint = blocks[i]
        r = res[int(tus[i]["offset"]):int(tus[i]["offset"]) + w * h].astype(int).tolist()
        other = rng.integers(-300, 300, w * h).tolist()
        ns = {"res_cb": list(r) if plane == 1 else other, "res_cr": list(r) if plane == 2 else list(other),
              "cw": w, "ch": h, "cb": types.SimpleNamespace(joint=joint & 3),
              "g": types.SimpleNamespace(ph=types.SimpleNamespace(joint_cbcr_sign=bool(joint & 4))),
              "_range": range}
        exec(src, ns)
        for c, key in ((1, "res_cb"), (2, "res_cr")):
            got = planes[0][c][y:y + h, x:x + w].astype(int).reshape(-1)
            assert (got == np.clip(512 + np.array(ns[key]), 0, 1023)).all(), (i, joint, c)

"""Pins the oracle's geometric-partitioning blend (orc_gpm_params,
orc_gpm_blend, orc_gpm_batch) to the reference: derive_gpm_params_table and
predict_gpm_block (crates/rustmpeg-video/src/vvc/inter.rs:1641-1730), with
GEO_DIS and GEO_ANGLE_TO_MASK parsed from the same file, read as text and
evaluated in Python (tests/rust_text.py), over every split direction, block
sizes 8..64 and the position scales the decoder uses; the reference's own
known answers (64 directions, the first (0, 1), flat inputs kept,
inter.rs:2889-2906); and the mask's meaning: the two sides of the split take
prediction 0 and prediction 1 unblended away from the line."""
import re

import numpy as np
import pytest

import rust_text
from conftest import reference_text, load_package


def _table(text, name):
    body = re.search(r"pub const %s: \[i(?:8|32); 32\] = \[(.*?)\];" % name, text, re.S).group(1)
    vals = [int(v) for v in re.findall(r"-?\d+", body)]
    assert len(vals) == 32
    return vals


class _Param:
    def __init__(self, angle_idx, distance_idx):
        self.angle_idx, self.distance_idx = angle_idx, distance_idx


@pytest.fixture(scope="module")
def ref():
    text = reference_text("crates/rustmpeg-video/src/vvc/inter.rs")
    text = text.replace("let mut table = [GeoParam {\n        angle_idx: 0,\n        distance_idx: 0,\n    }; 64];",
                        "let mut table = table_init();")
    text = re.sub(r"table\[mode_idx\] = GeoParam \{\s*angle_idx,\s*distance_idx,\s*\};",
                  "table[mode_idx] = GeoParam_new(angle_idx, distance_idx);", text)
    env = {"GEO_DIS": _table(text, "GEO_DIS"), "GEO_ANGLE_TO_MASK": _table(text, "GEO_ANGLE_TO_MASK"),
           "table_init": lambda: [_Param(0, 0) for _ in range(64)], "GeoParam_new": _Param}
    env["derive_gpm_params_table"] = rust_text.compile_fn(text, "derive_gpm_params_table", extra=env)
    return {"table": env["derive_gpm_params_table"], "blend": rust_text.compile_fn(text, "predict_gpm_block", extra=env)}


def test_params_table_equals_reference_text(po, ref):
    want = [(p.angle_idx, p.distance_idx) for p in ref["table"]()]
    assert po.gpm_params() == want


def test_params_table_known_answers(po):
    t = po.gpm_params()
    assert len(t) == 64 and t[0] == (0, 1)
    assert len(set(t)) == 64


@pytest.mark.parametrize("w,h", [(8, 8), (16, 8), (8, 32), (32, 16), (64, 64), (4, 4)])
@pytest.mark.parametrize("scale", [(0, 0), (1, 1), (2, 2)])
def test_blend_equals_reference_text(po, ref, w, h, scale):
    rng = np.random.default_rng(w * 131 + h * 7 + scale[0])
    for d in range(64):
        p0 = rng.integers(0, 256, w * h)
        p1 = rng.integers(0, 256, w * h)
        want = ref["blend"](p0.tolist(), p1.tolist(), w, h, d, True, scale[0], scale[1])
        got = po.gpm_blend(p0, p1, w, h, d, scale[0], scale[1])
        assert got.reshape(-1).tolist() == list(want), (w, h, d, scale)


def test_flat_inputs_are_kept(po):
    """gpm_derives_all_64_valid_partition_modes: 60 and 60 blend to 60."""
    for d in range(64):
        assert (po.gpm_blend(np.full(64, 60), np.full(64, 60), 8, 8, d, 0, 0) == 60).all()


def test_the_two_sides_take_one_prediction_each(po):
    """Away from the split line the weight saturates at 8 or 0: blending 0
    and 200 gives only 0 and 200 there, both occur, and the blended band is
    narrow (a mask that never saturates would pass the flat test)."""
    for d in range(64):
        out = po.gpm_blend(np.full(32 * 32, 200), np.zeros(32 * 32, np.int64), 32, 32, d, 0, 0)
        vals, counts = np.unique(out, return_counts=True)
        assert 0 in vals and 200 in vals, d
        assert counts[(vals != 0) & (vals != 200)].sum() < 0.3 * out.size, d


def test_batch_composes_the_blend_per_pu(po, ref):
    from frames import abi
    rng = np.random.default_rng(4)
    W, H = 72, 40  # PUs past the right and bottom edges
    planes = lambda: [rng.integers(0, 256, (H, W)).astype(np.uint8),
                      rng.integers(0, 256, (H // 2, W // 2)).astype(np.uint8),
                      rng.integers(0, 256, (H // 2, W // 2)).astype(np.uint8)]
    p0, p1 = planes(), planes()
    dst = [np.zeros_like(p) for p in p0]
    f = abi.GpmFrame((abi.Plane * 3)(*[abi.plane_from_array(p, 8) for p in p0]),
                     (abi.Plane * 3)(*[abi.plane_from_array(p, 8) for p in p1]),
                     (abi.Plane * 3)(*[abi.plane_from_array(p, 8) for p in dst]))
    pus = np.zeros(6, abi.GPM_PU_DTYPE)
    for k, (x, y, lw, lh) in enumerate([(0, 0, 4, 4), (16, 0, 5, 4), (48, 0, 5, 5), (0, 16, 3, 4), (8, 16, 3, 3),
                                        (8, 24, 5, 4)]):
        pus[k] = (x, y, lw, lh, 0, (k * 11) % 64, (0, 0, 0))
    assert po.gpm_batch([f], pus, 1, 2) == 0
    for pu in pus:
        w, h = 1 << int(pu["log2_w"]), 1 << int(pu["log2_h"])
        for c, (sx, sc) in enumerate([(0, 1), (1, 2), (1, 2)]):
            x0, y0, bw, bh = int(pu["x"]) >> sx, int(pu["y"]) >> sx, w >> sx, h >> sx
            a = np.zeros((bh, bw), int)
            b = np.zeros((bh, bw), int)
            vis = p0[c][y0:y0 + bh, x0:x0 + bw]
            a[:vis.shape[0], :vis.shape[1]] = vis
            b[:vis.shape[0], :vis.shape[1]] = p1[c][y0:y0 + bh, x0:x0 + bw]
            want = np.array(ref["blend"](a.reshape(-1).tolist(), b.reshape(-1).tolist(), bw, bh,
                                         int(pu["split_dir"]), c == 0, sc, sc)).reshape(bh, bw)
            got = dst[c][y0:y0 + bh, x0:x0 + bw]
            assert (got == want[:got.shape[0], :got.shape[1]]).all(), (pu, c)

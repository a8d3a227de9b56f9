"""The coding-tree split RD search (rmh_rd_encode_luma, host code)."""
import numpy as np
import pytest

from conftest import load_package

rmh = load_package()


def _content(seed, w=128, h=96):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    img = 60 + 0.8 * xx + 0.5 * yy + rng.normal(0, 6, (h, w))
    img[20:60, 30:80] += 70  # a rectangle with sharp edges
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def _planar_dc(top, left, w, h, mode):
    """PLANAR / DC over a rectangle from top[0..w], left[0..h] (H.266 8.4.5.2.11-12)."""
    lw, lh = int(np.log2(w)), int(np.log2(h))
    if mode == 0:
        y, x = np.mgrid[0:h, 0:w]
        pv = ((h - 1 - y) * top[x] + (y + 1) * left[h]) << lw
        ph = ((w - 1 - x) * left[y] + (x + 1) * top[w]) << lh
        return (pv + ph + w * h) >> (lw + lh + 1)
    if w == h:
        dc = (int(top[:w].sum() + left[:h].sum()) + w) >> (lw + 1)
    elif w > h:
        dc = (int(top[:w].sum()) + (w >> 1)) >> lw
    else:
        dc = (int(left[:h].sum()) + (h >> 1)) >> lh
    return np.full((h, w), dc, np.int64)


def test_nosplit_picture_matches_oracle_transform_path(po):
    """With only NoSplit allowed every node is one 32x32 leaf: rebuild the
    recon independently (numpy PLANAR/DC on the reconstructed neighbours,
    the oracle's forward_transform_rect_bd, quantize_rect_bd, dequant and
    inverse) and compare sample for sample."""
    src = _content(1)
    qp = 27
    h, w = src.shape
    probs = np.tile(np.array([1, 0, 0, 0, 0, 0], np.float32), (h // 32) * (w // 32))
    rec, splits, _, st = rmh.rd_encode_luma(src, qp, rmh.abi.RD_PRUNED, probs)
    assert (splits == 0).all() and st["candidates"] == st["blocks"]
    want = np.zeros((h, w), np.int64)
    for by in range(0, h, 32):
        for bx in range(0, w, 32):
            # reference line: left column (bottom-up from y0+32), corner, top row to x0+32
            def avail(x, y):
                return 0 <= x < w and 0 <= y < h and (y < by or (y < by + 32 and x < bx))
            pts = [(bx - 1, by + 32 - i) for i in range(33)] + [(bx - 1, by - 1)] + [(bx + i, by - 1) for i in range(33)]
            vals = [int(want[y, x]) if avail(x, y) else None for (x, y) in pts]
            first = next((v for v in vals if v is not None), None)
            if first is None:
                vals = [128] * len(vals)
            else:
                prev = first
                for i, v in enumerate(vals):
                    if v is None:
                        vals[i] = prev
                    else:
                        prev = v
            left = np.array([vals[32 - j] for j in range(33)], np.int64)
            top = np.array(vals[34:], np.int64)
            blk = src[by:by + 32, bx:bx + 32].astype(np.int64)
            preds = [_planar_dc(top, left, 32, 32, m) for m in (0, 1)]
            sads = [int(np.abs(blk - p).sum()) for p in preds]
            pred = preds[1] if sads[1] < sads[0] else preds[0]
            coef = po.forward_transform(32, 32, (blk - pred).reshape(-1), 0, 0, bd=8)
            lv = po.quant_rect(32, 32, coef, qp, 8, is_intra=True)
            res = po.inverse_transform(32, 32, po.dequant_rect(32, 32, lv, qp, 8), 0, 0, bd=8).reshape(32, 32)
            want[by:by + 32, bx:bx + 32] = np.clip(pred + res, 0, 255)
    assert (rec == want).all()


def test_full_search_picks_the_cheapest_split_and_pruning_is_consistent():
    src = _content(2)
    for qp in (22, 37):
        rec, splits, costs, st = rmh.rd_encode_luma(src, qp, want_costs=True)
        assert st["candidates"] == 6 * st["blocks"]
        assert (costs >= 0).all()
        assert (splits.reshape(-1) == costs.argmin(axis=1)).all()
        assert len(set(splits.reshape(-1).tolist())) > 1
        # one-hot probabilities on the full search's choice reproduce it exactly
        onehot = np.eye(6, dtype=np.float32)[splits.reshape(-1)]
        rec2, splits2, costs2, st2 = rmh.rd_encode_luma(src, qp, rmh.abi.RD_PRUNED, onehot, want_costs=True)
        assert st2["candidates"] == st2["blocks"] and st2["fallbacks"] == 0
        assert (splits2 == splits).all() and (rec2 == rec).all()
        assert st2["bits"] == st["bits"] and st2["sse"] == st["sse"]
        # flat probabilities fall back to the full search
        flat = np.full(onehot.shape, 1 / 6, np.float32)
        rec3, splits3, _, st3 = rmh.rd_encode_luma(src, qp, rmh.abi.RD_PRUNED, flat)
        assert st3["fallbacks"] == st3["blocks"] and (rec3 == rec).all()


def test_heuristic_mode_codes_one_split_per_node():
    src = _content(3)
    _, splits, costs, st = rmh.rd_encode_luma(src, 32, rmh.abi.RD_HEURISTIC, want_costs=True)
    assert st["candidates"] == st["blocks"]
    assert ((costs >= 0).sum(axis=1) == 1).all()
    assert set(splits.reshape(-1).tolist()) <= {0, 1, 2, 3}


def test_rejects_bad_sizes():
    with pytest.raises(rmh.InvalidInput):
        rmh.rd_encode_luma(np.zeros((40, 64), np.uint8), 30)
    with pytest.raises(rmh.InvalidInput):
        rmh.rd_encode_luma(np.zeros((32, 64), np.uint8), 30, rmh.abi.RD_PRUNED, np.zeros(5, np.float32))

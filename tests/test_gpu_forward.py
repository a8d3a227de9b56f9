"""The encoder forward path on the GPU (rmh_forward_plan_*) against the
oracle's restatement (orc_forward_batch, pinned to the reference text in
test_forward_vs_reference_text.py), bit for bit: every transform size from
2 to 64x64 including one-dimensional blocks, every transform pair a size
allows, extra zero-out, LFNST over all modes, flat (intra and inter rounding),
dependent and transform-skip quantisation, bit depths 8 to 12, residuals up
to the int16 limits; then the candidate evaluation (levels -> decoder-side
residual -> per-TU SSE of the clipped reconstruction) against the oracle's
residual batch and a numpy SSE."""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu

SIZES = [(w, h) for w in (1, 2, 4, 8, 16, 32, 64) for h in (1, 2, 4, 8, 16, 32, 64) if w * h >= 2]


def _batch(rng, abi, reps=3, extreme=True):
    tus, blocks = [], []
    off = 0
    for w, h in SIZES:
        for rep in range(reps):
            bd = int(rng.integers(8, 13))
            t = np.zeros(1, abi.TU_DTYPE)[0]
            t["offset"], t["log2_w"], t["log2_h"] = off, w.bit_length() - 1, h.bit_length() - 1
            t["bit_depth"] = bd
            t["qp"] = int(rng.integers(0, 52 + 6 * (bd - 8)))
            kind = rng.integers(0, 5)
            if kind == 0:
                t["flags"] = abi.TU_TRANSFORM_SKIP
            else:
                t["tr_hor"] = int(rng.integers(0, 3)) if 4 <= w <= 32 and rng.random() < 0.5 else 0
                t["tr_ver"] = int(rng.integers(0, 3)) if 4 <= h <= 32 and rng.random() < 0.5 else 0
                if w >= 4 and h >= 4 and rng.random() < 0.35:
                    t["lfnst_idx"], t["lfnst_mode"] = int(rng.integers(1, 3)), int(rng.integers(0, 95))
                elif rng.random() < 0.2:
                    t["skip_w"] = int(rng.integers(0, max(w // 2, 1))) if w > 1 else 0
                    t["skip_h"] = int(rng.integers(0, max(h // 2, 1))) if h > 1 else 0
                if kind == 1:
                    t["flags"] = abi.TU_DEP_QUANT
            if rng.random() < 0.4:
                t["flags"] |= abi.TU_QUANT_INTER
            maxv = (1 << bd) - 1
            r = rng.integers(-maxv, maxv + 1, w * h)
            if rep == 1:
                r = (r * rng.random(w * h) ** 3).astype(np.int64)  # mostly small, as residuals are
            if extreme and rep == 2 and rng.random() < 0.5:
                r = np.where(rng.random(w * h) < 0.5, 32767, -32768)
            tus.append(t)
            blocks.append(r.astype(np.int16))
            off += w * h
    return np.array(tus, abi.TU_DTYPE), np.concatenate(blocks)


def test_levels_equal_oracle(acc, po):
    print("forward batch levels", flush=True)
    pkg = load_package()
    hip, abi = pkg.hip, pkg.abi
    rng = np.random.default_rng(1)
    tus, res = _batch(rng, abi)
    want = po.forward_batch(tus, res)
    plan = acc.forward_plan(tus, res.size)
    # the level buffer starts filled with a non-zero pattern: every position
    # of every TU must be written, including the zeroed parts of 64-sample sides
    d_res = hip.DeviceBuffer.from_array(res)
    d_lv = hip.DeviceBuffer.from_array(np.full(res.size, 0x5A5A, np.int16))
    s = hip.Stream()
    plan.run(d_res.ptr, d_lv.ptr, s.handle)
    s.synchronize()
    got = d_lv.download(np.int16)
    bad = np.nonzero(got != want)[0]
    if bad.size:
        k = int(np.searchsorted(tus["offset"], bad[0], side="right")) - 1
        pytest.fail("TU %d %s: first mismatch at %d (%d vs %d)" % (k, tus[k], bad[0], got[bad[0]], want[bad[0]]))
    assert (want != 0).mean() > 0.05


def test_candidate_evaluation(acc, po):
    """encode_luma's scoring: the same 32x32 residual under DCT-2, the four MTS
    pairs and LFNST 1/2, each candidate's SSE of the clipped reconstruction."""
    print("forward candidates", flush=True)
    pkg = load_package()
    hip, abi = pkg.hip, pkg.abi
    rng = np.random.default_rng(2)
    tus, res_list, pred_list = [], [], []
    off = 0
    for blk in range(40):
        bd = [8, 10][blk % 2]
        maxv = (1 << bd) - 1
        n = 32 if blk % 3 else 16
        pred = rng.integers(0, maxv + 1, n * n)
        src = np.clip(pred + rng.normal(0, 0.05 * maxv, n * n).astype(np.int64), 0, maxv)
        qp = int(rng.integers(22, 38)) + 6 * (bd - 8)
        cands = [(0, 0, 0), (1, 1, 0), (2, 1, 0), (1, 2, 0), (2, 2, 0), (0, 0, 1), (0, 0, 2)]
        for th, tv, li in cands:
            t = np.zeros(1, abi.TU_DTYPE)[0]
            t["offset"], t["log2_w"], t["log2_h"], t["bit_depth"], t["qp"] = off, n.bit_length() - 1, \
                n.bit_length() - 1, bd, qp
            t["tr_hor"], t["tr_ver"], t["lfnst_idx"], t["lfnst_mode"] = th, tv, li, int(rng.integers(0, 67))
            if blk % 4 == 3:
                t["flags"] = abi.TU_DEP_QUANT
            tus.append(t)
            res_list.append((src - pred).astype(np.int16))
            pred_list.append(pred.astype(np.int16))
            off += n * n
    tus = np.array(tus, abi.TU_DTYPE)
    res, pred = np.concatenate(res_list), np.concatenate(pred_list)
    plan = acc.forward_plan(tus, res.size)
    d_res, d_pred = hip.DeviceBuffer.from_array(res), hip.DeviceBuffer.from_array(pred)
    d_lv, d_rec = hip.DeviceBuffer(res.nbytes), hip.DeviceBuffer(res.nbytes)
    d_sse = hip.DeviceBuffer(8 * len(tus))
    s = hip.Stream()
    plan.evaluate(d_res.ptr, d_pred.ptr, d_lv.ptr, d_rec.ptr, d_sse.ptr, s.handle)
    s.synchronize()
    levels, recon, sse = d_lv.download(np.int16), d_rec.download(np.int16), d_sse.download(np.uint64)
    assert (levels == po.forward_batch(tus, res)).all()
    want_rec = po.residual_batch(tus, levels)
    assert (recon == want_rec).all()
    for k, t in enumerate(tus):
        n = 1 << (int(t["log2_w"]) + int(t["log2_h"]))
        o = int(t["offset"])
        p, r, rr = pred[o:o + n].astype(np.int64), res[o:o + n].astype(np.int64), want_rec[o:o + n].astype(np.int64)
        want = int((((p + r) - np.clip(p + rr, 0, (1 << int(t["bit_depth"])) - 1)) ** 2).sum())
        assert int(sse[k]) == want, (k, int(sse[k]), want)
    # without a prediction: the residual-domain SSE
    plan.evaluate(d_res.ptr, 0, d_lv.ptr, d_rec.ptr, d_sse.ptr, s.handle)
    s.synchronize()
    sse = d_sse.download(np.uint64)
    for k, t in enumerate(tus[:10]):
        n = 1 << (int(t["log2_w"]) + int(t["log2_h"]))
        o = int(t["offset"])
        assert int(sse[k]) == int(((res[o:o + n].astype(np.int64) - want_rec[o:o + n]) ** 2).sum())


def test_rejects_joint_and_bad_flags(acc):
    abi = load_package().abi
    t = np.zeros(1, abi.TU_DTYPE)
    t["log2_w"], t["log2_h"], t["bit_depth"], t["qp"] = 3, 3, 8, 30
    t["joint"] = 1
    with pytest.raises(load_package().InvalidInput):
        acc.forward_plan(t, 64)
    t["joint"], t["flags"] = 0, 8
    with pytest.raises(load_package().InvalidInput):
        acc.forward_plan(t, 64)

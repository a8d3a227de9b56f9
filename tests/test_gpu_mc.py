"""GPU inter prediction (librmh) against the oracle, bit-exact."""
import numpy as np
import pytest

import workloads as W
from frames import mc_frame_struct

pytestmark = pytest.mark.gpu


def _run(acc, po, d, bps=2):
    bd = d["bit_depth"]
    dt = np.uint16 if bps == 2 else np.uint8
    refs = [np.ascontiguousarray(p.astype(dt)) for p in d["ref"]]
    got = [np.full_like(p, 7) for p in refs]
    want = [np.full_like(p, 7) for p in refs]
    f1, k1 = mc_frame_struct(refs, got, bd)
    f2, k2 = mc_frame_struct(refs, want, bd)
    acc.predict([f1], d["pus"])
    import ctypes
    arr = (type(f2) * 1)(f2)
    assert po.lib().orc_mc_batch(ctypes.addressof(arr), 1, d["pus"].ctypes.data, len(d["pus"])) == 0
    for c in range(3):
        bad = np.argwhere(got[c] != want[c])
        assert bad.size == 0, "plane %d first mismatch at %s: got %d want %d" % (
            c, bad[0], got[c][tuple(bad[0])], want[c][tuple(bad[0])])


def test_1080p10_p_frame_bit_exact(acc, po):
    print("mc 1080p10 seed 0", flush=True)
    _run(acc, po, W.inter_frame(0))


def test_8bit_planes_bit_exact(acc, po):
    print("mc 8-bit 480x272", flush=True)
    _run(acc, po, W.inter_frame(1, width=480, height=272, bit_depth=8), bps=1)


def test_motion_far_outside_the_picture(acc, po):
    print("mc far MVs", flush=True)
    d = W.inter_frame(2, width=256, height=128, bit_depth=10, mv_range=400)
    _run(acc, po, d)


def test_odd_picture_size_and_every_fraction(acc, po):
    print("mc every fraction", flush=True)
    d = W.inter_frame(3, width=200, height=136, bit_depth=10, mv_range=8)
    pus = d["pus"]
    pus["mv_x"] = np.arange(len(pus)) % 16 - 8 * 16 + (np.arange(len(pus)) * 7) % 16
    pus["mv_y"] = (np.arange(len(pus)) * 5) % 32 - 16
    _run(acc, po, d)


# ---- bi-prediction with BCW (rmh_mc_bi_batch / rmh_mc_bi_plan_create) ----

def _run_bi(acc, po, d, bps=2):
    from frames import mc_bi_frame_struct
    bd = d["bit_depth"]
    dt = np.uint16 if bps == 2 else np.uint8
    r0 = [np.ascontiguousarray(p.astype(dt)) for p in d["ref0"]]
    r1 = [np.ascontiguousarray(p.astype(dt)) for p in d["ref1"]]
    got = [np.full_like(p, 7) for p in r0]
    want = [np.full_like(p, 7) for p in r0]
    f1, k1 = mc_bi_frame_struct(r0, r1, got, bd)
    f2, k2 = mc_bi_frame_struct(r0, r1, want, bd)
    acc.predict_bi([f1], d["pus"])
    assert po.mc_bi_batch(f2, d["pus"]) == 0
    for c in range(3):
        bad = np.argwhere(got[c] != want[c])
        assert bad.size == 0, "plane %d first mismatch at %s: got %d want %d" % (
            c, bad[0], got[c][tuple(bad[0])], want[c][tuple(bad[0])])
    return want


def test_bi_1080p10_bit_exact(acc, po):
    print("mc bi 1080p10 seed 0", flush=True)
    _run_bi(acc, po, W.bi_frame(0))


def test_bi_8bit_bit_exact(acc, po):
    print("mc bi 8-bit 480x272", flush=True)
    _run_bi(acc, po, W.bi_frame(1, width=480, height=272, bit_depth=8), bps=1)


def test_bi_far_motion_and_every_fraction(acc, po):
    print("mc bi far MVs, every fraction, 12-bit", flush=True)
    d = W.bi_frame(2, width=200, height=136, bit_depth=12, mv_range=300)
    pus = d["pus"]
    n = len(pus)
    pus["mv0_x"] = np.arange(n) % 16 + 16 * ((np.arange(n) * 37) % 600 - 300)
    pus["mv0_y"] = (np.arange(n) * 5) % 16 - 16 * 3
    pus["mv1_x"] = (np.arange(n) * 3) % 16 - 16 * 2
    pus["mv1_y"] = (np.arange(n) * 11) % 16 + 16 * ((np.arange(n) * 29) % 500 - 250)
    pus["bcw_idx"] = np.arange(n) % 8
    _run_bi(acc, po, d)


def test_bi_plan_on_device_planes(acc, po, rmh):
    # the plan path (device planes, rmh_mc_plan_run) against the host batch
    print("mc bi plan 416x240", flush=True)
    hip = rmh.hip
    d = W.bi_frame(3, width=416, height=240, bit_depth=10)
    want = _run_bi(acc, po, d)
    bufs, planes = [], {}
    for key in ("ref0", "ref1", "dst"):
        src = d["ref0"] if key == "dst" else d[key]
        ps = []
        for p in src:
            b = hip.DeviceBuffer(p.nbytes)
            bufs.append(b)
            if key != "dst":
                b.upload(np.ascontiguousarray(p))
            else:
                b.upload(np.full_like(p, 7))
            ps.append(rmh.abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], 2, 10))
        planes[key] = (rmh.abi.Plane * 3)(*ps)
    frame = rmh.abi.McBiFrame(planes["ref0"], planes["ref1"], planes["dst"])
    plan = acc.mc_bi_plan([frame], d["pus"])
    plan.run()
    hip.device_synchronize()
    for c in range(3):
        got = bufs[6 + c].download(np.uint16).reshape(want[c].shape)
        assert (got == want[c]).all(), c
    plan.close()
    for b in bufs:
        b.close()


def test_bi_invalid_descriptors_are_rejected(acc, rmh):
    from frames import mc_bi_frame_struct
    d = W.bi_frame(4, width=64, height=64, bit_depth=10)
    dst = [np.zeros_like(p) for p in d["ref0"]]
    f, keep = mc_bi_frame_struct(d["ref0"], d["ref1"], dst, 10)
    pus = d["pus"].copy()
    pus["reserved"][0] = (1, 0)
    with pytest.raises(rmh.InvalidInput):
        acc.predict_bi([f], pus)
    pus = d["pus"].copy()
    pus["refine"][0] = W.BI_BDOF  # the refinements exist for 8-bit frames only
    with pytest.raises(rmh.InvalidInput):
        acc.predict_bi([f], pus)
    pus = d["pus"].copy()
    pus["refine"][0] = 4
    with pytest.raises(rmh.InvalidInput):
        acc.predict_bi([f], pus)
    pus = d["pus"].copy()
    pus["mv1_x"][0] = 1 << 17
    with pytest.raises(rmh.InvalidInput):
        acc.predict_bi([f], pus)
    assert all((p == 0).all() for p in dst)

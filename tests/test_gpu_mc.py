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

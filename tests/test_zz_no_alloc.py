"""rmh_deblock_noalloc (in-place deblocking that allocates nothing: tables
and plane copies in the library's fixed per-GPU scratch) against the
allocating entry point it varies, rmh_deblock_plan_create + rmh_deblock_plan_run,
bit for bit, on two calls in a row (the second with a smaller batch, so stale
scratch contents from the first must not leak into it)."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def _device_frames(rmh, seeds, width, height, bd=10):
    """Two device copies of each frame's planes (one per entry point) with the
    frame's edge map; returns (frames_a, frames_b, keepalive, host planes)."""
    hip, abi = rmh.hip, rmh.abi
    fa, fb, keep, host = [], [], [], []
    for seed in seeds:
        d = W.inter_frame(seed, width=width, height=height, bit_depth=bd)
        edges = hip.DeviceBuffer.from_array(np.ascontiguousarray(d["edges"]))
        keep.append(edges)
        sets = []
        for _ in range(2):
            planes = []
            for p in d["ref"]:
                b = hip.DeviceBuffer.from_array(p)
                keep.append(b)
                planes.append(abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], 2, bd))
            sets.append((planes, [keep[-3], keep[-2], keep[-1]]))
        gw, gh = d["edges"].shape[1], d["edges"].shape[0]
        fa.append((abi.dbk_frame(sets[0][0], edges.ptr, gw, gh, d["qp"]), sets[0][1]))
        fb.append((abi.dbk_frame(sets[1][0], edges.ptr, gw, gh, d["qp"]), sets[1][1]))
        host.append(d["ref"])
    return fa, fb, keep, host


@pytest.mark.parametrize("rounds", [[(2, 1920, 1080), (1, 640, 368)]])
def test_noalloc_deblock_equals_plan_deblock(acc, rmh, rounds):
    hip = rmh.hip
    s = hip.Stream()
    for call, (n, w, h) in enumerate(rounds):
        print("no-alloc deblock call %d: %d frame(s) %dx%d 10-bit" % (call, n, w, h), flush=True)
        fa, fb, keep, host = _device_frames(rmh, range(40 + 10 * call, 40 + 10 * call + n), w, h)
        acc.deblock_noalloc([f for f, _ in fa], s.handle)
        plan = acc.deblock_plan([f for f, _ in fb])
        plan.run(stream=s.handle)
        s.synchronize()
        plan.close()
        for k in range(n):
            changed = 0
            for c in range(3):
                shape = host[k][c].shape
                got = fa[k][1][c].download(np.uint16).reshape(shape)
                want = fb[k][1][c].download(np.uint16).reshape(shape)
                assert (got == want).all(), (call, k, c)
                changed += int((got != host[k][c]).sum())
            assert changed > 0, "the filter changed nothing"


def test_noalloc_deblock_rejects_batches_beyond_the_scratch(acc, rmh):
    """Six 1080p 10-bit frames (37 MB of plane copies) need more than the
    fixed scratch holds: the call fails before any device work and leaves the
    planes untouched."""
    fa, _, keep, host = _device_frames(rmh, range(70, 76), 1920, 1080)
    with pytest.raises(rmh.LimitExceeded):
        acc.deblock_noalloc([f for f, _ in fa], 0)
    for k in range(6):
        for c in range(3):
            got = fa[k][1][c].download(np.uint16).reshape(host[k][c].shape)
            assert (got == host[k][c]).all()

"""SAO on the GPU (rmh_sao, rmh_sao_plan_*) against the oracle's apply_sao
restatement, bit for bit: full 1080p10 pictures, 8-bit pictures with
unaligned rows and a cropped visible region, mixed batches, plans on device
buffers and graph replays, edge and band offsets at the clip limits, and the
error contract (dst untouched on error)."""
import numpy as np
import pytest

import workloads as W
from conftest import load_package
from test_oracle_sao import random_sao_params

pytestmark = pytest.mark.gpu


def _picture(rng, w, h, bd, textured=True):
    dt = np.uint8 if bd == 8 else np.uint16
    if textured:
        planes = [W.textured_plane(rng, w, h, bd), W.textured_plane(rng, (w + 1) // 2, (h + 1) // 2, bd),
                  W.textured_plane(rng, (w + 1) // 2, (h + 1) // 2, bd)]
    else:
        planes = [rng.integers(0, 1 << bd, (h, w)), rng.integers(0, 1 << bd, ((h + 1) // 2, (w + 1) // 2)),
                  rng.integers(0, 1 << bd, ((h + 1) // 2, (w + 1) // 2))]
    return [np.ascontiguousarray(np.asarray(p).astype(dt)) for p in planes]


def _case(rng, w, h, bd, ctb, vis=None, textured=True):
    vis = vis or (w, h)
    planes = _picture(rng, w, h, bd, textured)
    rows, cols = -(-vis[1] // ctb), -(-vis[0] // ctb)
    prm = np.ascontiguousarray(random_sao_params(rng, rows, cols, bd))
    return planes, prm, ctb, cols, rows, vis


def _host_run(acc, cases, bd_of):
    abi = load_package().abi
    frames, outs, keep = [], [], []
    for (planes, prm, ctb, cols, rows, vis), bd in zip(cases, bd_of):
        out = [np.full_like(p, 7) for p in planes]
        frames.append(abi.sao_frame([abi.plane_from_array(p, bd) for p in planes],
                                    [abi.plane_from_array(o, bd) for o in out], prm.ctypes.data, ctb, cols, rows,
                                    vis[0], vis[1]))
        outs.append(out)
        keep.append((planes, prm))
    acc.sao(frames)
    return outs


def _check(po, case, bd, got):
    planes, prm, ctb, cols, rows, vis = case
    want = po.sao_frame(planes, bd, prm, ctb, cols, rows, vis[0], vis[1])
    for c in range(3):
        bad = np.argwhere(got[c] != want[c])
        assert bad.size == 0, "plane %d: first mismatch at %s of %d" % (c, bad[0], len(bad))


def test_1080p10_bit_exact(acc, po):
    print("sao 1080p10", flush=True)
    rng = np.random.default_rng(11)
    case = _case(rng, 1920, 1080, 10, 64)
    got = _host_run(acc, [case], [10])[0]
    _check(po, case, 10, got)
    assert any((got[c] != case[0][c]).any() for c in range(3))


@pytest.mark.parametrize("w,h,bd,ctb,vis", [(856, 480, 8, 32, (854, 480)), (854, 480, 8, 64, (854, 477)),
                                            (333, 201, 10, 16, (331, 199)), (130, 66, 12, 128, (130, 66)),
                                            (64, 64, 8, 16, (64, 64))])
def test_sizes_cropping_and_unaligned_rows(acc, po, w, h, bd, ctb, vis):
    print("sao %dx%d bd %d ctb %d vis %s" % (w, h, bd, ctb, vis), flush=True)
    rng = np.random.default_rng(w * 7 + h)
    case = _case(rng, w, h, bd, ctb, vis, textured=False)
    _check(po, case, bd, _host_run(acc, [case], [bd])[0])


def test_mixed_batch_and_clip_limits(acc, po):
    """8-bit and 10-bit pictures in one call; offsets at +-(2^bd) so both
    clip ends are met."""
    print("sao mixed batch", flush=True)
    rng = np.random.default_rng(5)
    cases, bds = [], []
    for bd, (w, h) in [(8, (256, 128)), (10, (320, 184)), (8, (200, 120)), (12, (96, 72))]:
        case = _case(rng, w, h, bd, 32, textured=False)
        big = case[1]["type"] != 0
        case[1]["offsets"][big] = rng.choice([-(1 << bd), (1 << bd)], (int(big.sum()), 4))
        cases.append(case)
        bds.append(bd)
    outs = _host_run(acc, cases, bds)
    for case, bd, got in zip(cases, bds, outs):
        _check(po, case, bd, got)


def test_plan_and_graph_replays(acc, po):
    print("sao plan and graph", flush=True)
    pkg = load_package()
    hip, abi = pkg.hip, pkg.abi
    rng = np.random.default_rng(9)
    cases = [_case(rng, 640, 360, 10, 64), _case(rng, 856, 480, 8, 32, (854, 480))]
    bds = [10, 8]
    frames, bufs, outs = [], [], []
    for (planes, prm, ctb, cols, rows, vis), bd in zip(cases, bds):
        srcs, dsts = [], []
        for p in planes:
            b = hip.DeviceBuffer.from_array(p)
            o = hip.DeviceBuffer(p.nbytes)
            bufs += [b, o]
            srcs.append(abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], p.dtype.itemsize, bd))
            dsts.append(abi.plane_from_address(o.ptr, p.shape[1], p.shape[0], p.shape[1], p.dtype.itemsize, bd))
            outs.append((o, p))
        dp = hip.DeviceBuffer.from_array(prm)
        bufs.append(dp)
        frames.append(abi.sao_frame(srcs, dsts, dp.ptr, ctb, cols, rows, vis[0], vis[1]))
    plan = acc.sao_plan(frames)
    assert plan.bytes == sum(2 * p.nbytes for c in cases for p in c[0]) + sum(c[1].nbytes for c in cases)
    s = hip.Stream()
    graph = acc.graph([(plan, 0, 0)])

    def check():
        k = 0
        for case, bd in zip(cases, bds):
            got = [outs[k + c][0].download(outs[k + c][1].dtype).reshape(outs[k + c][1].shape) for c in range(3)]
            _check(po, case, bd, got)
            k += 3

    plan.run(stream=s.handle)
    s.synchronize()
    check()
    for rep in range(3):
        for o, _ in outs:
            o.zero(s.handle)
        s.synchronize()
        graph.launch(s.handle)
        s.synchronize()
        check()
    graph.close()
    plan.close()


def test_errors_leave_dst_untouched(acc, rmh):
    print("sao errors", flush=True)
    abi = load_package().abi
    rng = np.random.default_rng(2)
    planes, prm, ctb, cols, rows, vis = _case(rng, 128, 64, 10, 64)
    out = [np.full_like(p, 3) for p in planes]

    def frame(dsts, params):
        return abi.sao_frame([abi.plane_from_array(p, 10) for p in planes], [abi.plane_from_array(o, 10) for o in dsts],
                             params.ctypes.data, ctb, cols, rows, vis[0], vis[1])

    with pytest.raises(rmh.InvalidInput):
        acc.sao([frame(planes, prm)])  # dst overlaps src
    bad = prm.copy()
    bad["type"][0, 0, 0] = 3
    with pytest.raises(rmh.InvalidInput):
        acc.sao([frame(out, bad)])
    bad = prm.copy()
    bad["offsets"][0, 0, 1] = 1 << 20
    with pytest.raises(rmh.InvalidInput):
        acc.sao([frame(out, bad)])
    assert all((o == 3).all() for o in out)
    acc.sao([])  # zero work is RMH_OK

"""GPU geometric-partitioning blend (rmh_gpm_plan_*) against the oracle
(orc_gpm_batch, pinned to the reference text in test_oracle_gpm.py),
bit-exact: a 1080p picture tiled with GPM PUs of 8..64 at every split
direction, a picture whose PUs run past its right and bottom edges, the
decoder's position scales (1, 2) and the spec's (0, 1), and graph replays."""
import numpy as np
import pytest

import workloads as W
from conftest import load_package

pytestmark = pytest.mark.gpu


def _case(seed, width, height, tiled_w=None, tiled_h=None):
    """Planes of width x height; GPM PUs tiling tiled_w x tiled_h (larger
    than the picture: PUs that run past its edges)."""
    rng = np.random.default_rng([12, seed])
    cw, ch = (width + 1) // 2, (height + 1) // 2
    planes = lambda: [rng.integers(0, 256, (height, width)).astype(np.uint8),
                      rng.integers(0, 256, (ch, cw)).astype(np.uint8),
                      rng.integers(0, 256, (ch, cw)).astype(np.uint8)]
    cus = W._tile(rng, tiled_w or width, tiled_h or height, 64, 64, 8,
                  lambda w, h: {64: 0.7, 32: 0.5, 16: 0.4}.get(max(w, h), 0.0))
    cus = [c for c in cus if c[0] < width and c[1] < height]
    abi = load_package().abi
    pus = np.zeros(len(cus), abi.GPM_PU_DTYPE)
    pus["x"] = [c[0] for c in cus]
    pus["y"] = [c[1] for c in cus]
    pus["log2_w"] = [int(c[2]).bit_length() - 1 for c in cus]
    pus["log2_h"] = [int(c[3]).bit_length() - 1 for c in cus]
    pus["split_dir"] = np.arange(len(cus)) % 64
    return planes(), planes(), pus


def _run(acc, po, seed, width, height, scales, graph_reps=0, tiled=(None, None)):
    rmh = load_package()
    hip, abi = rmh.hip, rmh.abi
    p0, p1, pus = _case(seed, width, height, *tiled)
    want = [np.full_like(p, 9) for p in p0]
    fh = abi.GpmFrame((abi.Plane * 3)(*[abi.plane_from_array(p, 8) for p in p0]),
                      (abi.Plane * 3)(*[abi.plane_from_array(p, 8) for p in p1]),
                      (abi.Plane * 3)(*[abi.plane_from_array(p, 8) for p in want]))
    assert po.gpm_batch([fh], pus, *scales) == 0
    bufs = []

    def dev(p, fill=None):
        b = hip.DeviceBuffer.from_array(p if fill is None else np.full_like(p, fill))
        bufs.append(b)
        return abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], 1, 8)

    d0 = [dev(p) for p in p0]
    d1 = [dev(p) for p in p1]
    dd = [dev(p, 9) for p in p0]
    frame = abi.GpmFrame((abi.Plane * 3)(*d0), (abi.Plane * 3)(*d1), (abi.Plane * 3)(*dd))
    plan = acc.gpm_plan([frame], pus, *scales)
    plan.run()
    hip.device_synchronize()
    out = lambda: [bufs[6 + c].download(np.uint8).reshape(want[c].shape) for c in range(3)]
    for c, g in enumerate(out()):
        bad = np.argwhere(g != want[c])
        assert bad.size == 0, "plane %d first mismatch at %s" % (c, bad[0])
    if graph_reps:
        graph = acc.graph([(plan, 0, 0)])
        s = hip.Stream()
        for _ in range(graph_reps):
            for c in range(3):
                bufs[6 + c].upload(np.full_like(want[c], 9))
            graph.launch(s.handle)
            s.synchronize()
            assert all((g == w).all() for g, w in zip(out(), want))
        graph.close()
    plan.close()
    for b in bufs:
        b.close()


def test_gpm_1080p_bit_exact(acc, po):
    print("gpm 1080p", flush=True)
    _run(acc, po, 0, 1920, 1080, (1, 2))


def test_gpm_edges_and_spec_scales(acc, po):
    print("gpm 196x132 scales (0, 1) and (1, 2)", flush=True)
    _run(acc, po, 1, 196, 132, (0, 1), tiled=(200, 136))
    _run(acc, po, 2, 196, 132, (1, 2), graph_reps=3, tiled=(200, 136))


def test_gpm_rejects_bad_descriptors(acc):
    rmh = load_package()
    abi = rmh.abi
    p0, p1, pus = _case(3, 64, 64)
    hip = rmh.hip
    bufs = [hip.DeviceBuffer.from_array(p) for p in p0 + p1 + p0]
    planes = [abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], 1, 8) for b, p in zip(bufs, p0 + p1 + p0)]
    frame = abi.GpmFrame((abi.Plane * 3)(*planes[0:3]), (abi.Plane * 3)(*planes[3:6]), (abi.Plane * 3)(*planes[6:9]))
    for field, value in (("split_dir", 64), ("log2_w", 2), ("frame", 1)):
        bad = pus.copy()
        bad[field][0] = value
        with pytest.raises(rmh.InvalidInput):
            acc.gpm_plan([frame], bad)
    with pytest.raises(rmh.InvalidInput):
        acc.gpm_plan([frame], pus, 4, 2)
    # dst aliasing a prediction plane
    alias = abi.GpmFrame((abi.Plane * 3)(*planes[0:3]), (abi.Plane * 3)(*planes[3:6]), (abi.Plane * 3)(*planes[0:3]))
    with pytest.raises(rmh.InvalidInput):
        acc.gpm_plan([alias], pus)
    for b in bufs:
        b.close()

"""The device edge-map builder (rmh_edge_map, rmh_edge_map_plan_*) against
the oracle (orc_edge_map, pinned to the reference's rules), bit for bit,
over coded pictures with visible sizes that are not multiples of 4 and CTB
sizes 32..128; and a frame deblocked with the device-built map equals the
oracle's deblocking with the oracle's map."""
import ctypes

import numpy as np
import pytest

import workloads as W
from frames import dbk_frame_struct
from test_oracle_edges import oracle_edges

pytestmark = pytest.mark.gpu

CASES = [(11, 640, 368, 128), (12, 250, 134, 64), (13, 1920, 1080, 128), (14, 96, 64, 32)]


@pytest.mark.parametrize("seed,w,h,ctb", CASES)
def test_edge_map_host_entry_matches_oracle(acc, po, seed, w, h, ctb):
    cells = W.edge_cells(seed, width=w, height=h, ctu=ctb)
    assert (acc.edge_map(cells, w, h, ctb) == oracle_edges(po, cells, w, h, ctb)).all()


def test_edge_map_plan_over_many_frames(acc, po, rmh):
    hip, abi = rmh.hip, rmh.abi
    frames, keep, want = [], [], []
    for seed, w, h, ctb in CASES:
        cells = np.ascontiguousarray(W.edge_cells(seed, width=w, height=h, ctu=ctb))
        dc = hip.DeviceBuffer.from_array(cells)
        de = hip.DeviceBuffer(cells.size * 4)
        keep += [dc, de]
        frames.append(abi.EdgeFrame(dc.ptr, de.ptr, cells.shape[1], cells.shape[0], w, h, ctb, 0))
        want.append(oracle_edges(po, cells, w, h, ctb))
    plan = acc.edge_map_plan(frames)
    s = hip.Stream()
    plan.run(stream=s.handle)
    s.synchronize()
    for k, wnt in enumerate(want):
        got = keep[2 * k + 1].download(np.uint8).reshape(wnt.shape)
        assert (got == wnt).all(), k
    plan.close()


def test_deblocking_with_the_device_built_map(acc, po):
    w, h = 640, 368
    cells = W.edge_cells(21, width=w, height=h, ctu=128)
    edges = acc.edge_map(cells, w, h, 128)
    d = W.inter_frame(21, width=w, height=h)
    got = [p.copy() for p in d["ref"]]
    want = [p.copy() for p in d["ref"]]
    f, keep = dbk_frame_struct(got, edges, 10, d["qp"])
    acc.deblock([f])
    g, keep2 = dbk_frame_struct(want, oracle_edges(po, cells, w, h, 128), 10, d["qp"])
    arr = (type(g) * 1)(g)
    po.lib().orc_deblock_frames(ctypes.addressof(arr), 1)
    assert all((a == b).all() for a, b in zip(got, want))
    assert any((a != r).any() for a, r in zip(got, d["ref"]))

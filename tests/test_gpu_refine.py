"""GPU bi-prediction with DMVR and BDOF (rmh_bi_pu_desc.refine) against the
oracle (orc_mc_bi_batch, whose refinements are pinned to the reference text in
test_oracle_refine.py), bit-exact: whole 8-bit frames with every combination
of flags and BCW indices (so each condition of predict_inter_pu decides per
PU), pictures whose PUs run past the right and bottom edges, MVs far outside
the picture at every fraction, the plan path on device planes, and a captured
graph replayed several times (DMVR rewrites its MV slots from the signalled
pairs on every run)."""
import numpy as np
import pytest

import workloads as W
from frames import mc_bi_frame_struct

pytestmark = pytest.mark.gpu


def _frame(seed, width, height, mv_range=64, p_default=0.75):
    d = W.bi_frame(seed, width=width, height=height, bit_depth=8, mv_range=mv_range)
    rng = np.random.default_rng([9, seed])
    pus = d["pus"]
    pus["refine"] = rng.integers(0, 4, len(pus))
    pus["bcw_idx"] = np.where(rng.random(len(pus)) < p_default, 2, pus["bcw_idx"])
    return d


def _oracle(po, d):
    r0 = [np.ascontiguousarray(p.astype(np.uint8)) for p in d["ref0"]]
    r1 = [np.ascontiguousarray(p.astype(np.uint8)) for p in d["ref1"]]
    want = [np.full_like(p, 7) for p in r0]
    f, keep = mc_bi_frame_struct(r0, r1, want, 8)
    assert po.mc_bi_batch(f, d["pus"]) == 0
    return r0, r1, want


def _check(got, want):
    for c in range(3):
        bad = np.argwhere(got[c] != want[c])
        assert bad.size == 0, "plane %d first mismatch at %s: got %d want %d" % (
            c, bad[0], got[c][tuple(bad[0])], want[c][tuple(bad[0])])


def _run_host(acc, po, d):
    r0, r1, want = _oracle(po, d)
    got = [np.full_like(p, 7) for p in r0]
    f, keep = mc_bi_frame_struct(r0, r1, got, 8)
    acc.predict_bi([f], d["pus"])
    _check(got, want)
    return want


def test_refined_1080p_frame_bit_exact(acc, po):
    print("refine 1080p 8-bit seed 0", flush=True)
    d = _frame(0, 1920, 1080)
    pus = d["pus"]
    assert ((pus["refine"] & W.BI_DMVR) > 0).sum() > 100 and ((pus["refine"] & W.BI_BDOF) > 0).sum() > 100
    _run_host(acc, po, d)


def test_refined_edges_far_motion_and_every_fraction(acc, po):
    print("refine 200x136 far MVs", flush=True)
    d = _frame(1, 200, 136, mv_range=300, p_default=1.0)
    pus = d["pus"]
    n = len(pus)
    pus["mv0_x"] = np.arange(n) % 16 + 16 * ((np.arange(n) * 37) % 600 - 300)
    pus["mv0_y"] = (np.arange(n) * 5) % 16 - 16 * 3
    pus["mv1_x"] = (np.arange(n) * 3) % 16 - 16 * 2
    pus["mv1_y"] = (np.arange(n) * 11) % 16 + 16 * ((np.arange(n) * 29) % 500 - 250)
    pus["refine"] = 3
    _run_host(acc, po, d)


def test_refined_small_motion_changes_predictions(acc, po):
    """On content where the two references differ by a small motion, DMVR
    moves pairs and BDOF changes samples: the refined frame differs from the
    same PUs without refinement (the GPU must agree with the oracle on both)."""
    print("refine effect 256x128", flush=True)
    d = _frame(2, 256, 128, mv_range=2, p_default=1.0)
    d["pus"]["refine"] = 3
    refined = _run_host(acc, po, d)
    plain = dict(d, pus=d["pus"].copy())
    plain["pus"]["refine"] = 0
    unrefined = _run_host(acc, po, plain)
    assert (refined[0] != unrefined[0]).mean() > 0.05


def test_refined_plan_and_graph_replays(acc, po, rmh):
    print("refine plan + graph 416x240", flush=True)
    hip = rmh.hip
    d = _frame(3, 416, 240)
    r0, r1, want = _oracle(po, d)
    bufs, planes = [], {}
    for key, src in (("ref0", r0), ("ref1", r1), ("dst", r0)):
        ps = []
        for p in src:
            b = hip.DeviceBuffer(p.nbytes)
            bufs.append(b)
            b.upload(np.ascontiguousarray(p) if key != "dst" else np.full_like(p, 7))
            ps.append(rmh.abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], 1, 8))
        planes[key] = (rmh.abi.Plane * 3)(*ps)
    frame = rmh.abi.McBiFrame(planes["ref0"], planes["ref1"], planes["dst"])
    plan = acc.mc_bi_plan([frame], d["pus"])
    dst = lambda: [bufs[6 + c].download(np.uint8).reshape(want[c].shape) for c in range(3)]
    plan.run()
    hip.device_synchronize()
    _check(dst(), want)
    graph = acc.graph([(plan, 0, 0)])
    s = hip.Stream()
    for rep in range(3):
        for c in range(3):
            bufs[6 + c].upload(np.full_like(want[c], 7))
        graph.launch(s.handle)
        s.synchronize()
        _check(dst(), want)
    graph.close()
    plan.close()
    for b in bufs:
        b.close()

"""Device-resident plans, ranged (unbatched) runs and frame-step hipGraphs
against the host-memory entry points and the oracle."""
import ctypes

import numpy as np
import pytest

import workloads as W
from conftest import load_package

pytestmark = pytest.mark.gpu


def _dev_plane(hip, arr, bd, abi):
    b = hip.DeviceBuffer.from_array(arr)
    return b, abi.plane_from_address(b.ptr, arr.shape[1], arr.shape[0], arr.shape[1], arr.dtype.itemsize, bd)


def test_residual_plan_and_ranges_match_oracle(acc, po):
    hip = load_package().hip
    tus, coeff, _ = W.residual_frame(7, width=640, height=368)
    ref = po.residual_batch(tus, coeff)
    plan = acc.residual_plan(tus, coeff.size)
    assert plan.bytes == 4 * coeff.size + 16 * len(tus)
    dc = hip.DeviceBuffer.from_array(coeff)
    dr = hip.DeviceBuffer(coeff.nbytes)
    s = hip.Stream()
    plan.run(dc.ptr, dr.ptr, s.handle)
    s.synchronize()
    assert (dr.download(np.int16) == ref).all()
    dr.upload(np.zeros_like(coeff))
    n = len(tus)
    for a in range(0, n, 997):
        plan.run_range(a, min(997, n - a), dc.ptr, dr.ptr, s.handle)
    s.synchronize()
    assert (dr.download(np.int16) == ref).all()
    plan.close()


def test_frame_step_graph_replays(acc, po):
    pkg = load_package()
    hip, abi = pkg.hip, pkg.abi
    bd = 10
    tus, coeff, _ = W.residual_frame(8, width=320, height=192)
    d = W.inter_frame(8, width=320, height=192)
    ref_res = po.residual_batch(tus, coeff)
    # oracle: MC then deblock on the prediction
    from frames import dbk_frame_struct, mc_frame_struct
    want = [np.zeros_like(p) for p in d["ref"]]
    f, k = mc_frame_struct(d["ref"], want, bd)
    arr = (type(f) * 1)(f)
    po.lib().orc_mc_batch(ctypes.addressof(arr), 1, d["pus"].ctypes.data, len(d["pus"]))
    g, k2 = dbk_frame_struct(want, d["edges"], bd, d["qp"])
    arr2 = (type(g) * 1)(g)
    po.lib().orc_deblock_frames(ctypes.addressof(arr2), 1)

    bufs, refs, dsts = [], [], []
    for p in d["ref"]:
        b, pl = _dev_plane(hip, p, bd, abi)
        b2, pl2 = _dev_plane(hip, np.zeros_like(p), bd, abi)
        bufs.append((b, b2))
        refs.append(pl)
        dsts.append(pl2)
    edges = hip.DeviceBuffer.from_array(np.ascontiguousarray(d["edges"]))
    mc = acc.mc_plan([abi.mc_frame(refs, dsts)], d["pus"])
    db = acc.deblock_plan([abi.dbk_frame(dsts, edges.ptr, d["edges"].shape[1], d["edges"].shape[0], d["qp"])])
    rp = acc.residual_plan(tus, coeff.size)
    dc = hip.DeviceBuffer.from_array(coeff)
    dr = hip.DeviceBuffer(coeff.nbytes)
    graph = acc.graph([(rp, dc.ptr, dr.ptr), (mc, 0, 0), (db, 0, 0)])
    s = hip.Stream()
    for rep in range(4):
        dr.upload(np.zeros_like(coeff))
        for _, b2 in bufs:
            b2.upload(np.zeros(b2.nbytes // 2, np.uint16))
        graph.launch(s.handle)
        s.synchronize()
        assert (dr.download(np.int16) == ref_res).all(), rep
        for c in range(3):
            got = bufs[c][1].download(np.uint16).reshape(want[c].shape)
            assert (got == want[c]).all(), (rep, c)
        # eager work on the same stream between replays
        plan2 = acc.residual_plan(tus[:10], coeff.size)
        plan2.run(dc.ptr, dr.ptr, s.handle)
        s.synchronize()
        plan2.close()
    graph.close()


def test_plans_run_on_torch_tensors_when_torch_loads_first():
    """A caller that imports torch before librmh runs librmh on torch's HIP
    runtime; plans then take torch device pointers and streams directly."""
    import os
    import subprocess
    import sys
    code = r"""
import sys, numpy as np
sys.path.insert(0, %r); sys.path.insert(0, %r)
import torch
assert torch.cuda.is_available()
from conftest import load_package
import pyoracle as po, workloads as W
rmh = load_package()
acc = rmh.Accelerator(0)
tus, coeff, _ = W.residual_frame(11, width=256, height=128)
plan = acc.residual_plan(tus, coeff.size)
dc = torch.from_numpy(coeff).cuda(); dr = torch.zeros_like(dc)
plan.run(dc.data_ptr(), dr.data_ptr(), torch.cuda.current_stream().cuda_stream)
torch.cuda.synchronize()
assert (dr.cpu().numpy() == po.residual_batch(tus, coeff)).all()
print("torch interop ok")
""" % (os.path.join(os.path.dirname(__file__)), os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "torch interop ok" in out.stdout


def test_disable_latch(po):
    rmh = load_package()
    a = rmh.Accelerator(0)
    tus, coeff, _ = W.residual_frame(9, width=64, height=64)
    a.residuals(tus, coeff)
    a.disable()
    assert a.disabled
    with pytest.raises(rmh.Unavailable):
        a.residuals(tus, coeff)
    a.close()


def test_autotuned_residual_plan_is_bit_exact(acc, po):
    hip = load_package().hip
    tus, coeff, _ = W.residual_frame(9, width=640, height=368)
    ref = po.residual_batch(tus, coeff)
    plan = acc.residual_plan(tus, coeff.size)
    dc = hip.DeviceBuffer.from_array(coeff)
    dr = hip.DeviceBuffer(coeff.nbytes)
    s = hip.Stream()
    splits = plan.autotune(dc.ptr, dr.ptr, s.handle, reps=2)
    assert set(splits.values()) <= {1, 2, 4, 8}
    dr.upload(np.zeros_like(coeff))
    plan.run(dc.ptr, dr.ptr, s.handle)
    s.synchronize()
    assert (dr.download(np.int16) == ref).all()
    plan.close()


def test_autotune_refused_while_a_graph_holds_the_plan(acc, po, rmh):
    """A graph replays the tile list and grid it captured, so the autotuner
    must not rewrite them while the graph lives; after the graph is destroyed
    it may, and a recaptured graph is bit-exact."""
    hip = rmh.hip
    tus, coeff, _ = W.residual_frame(12, width=320, height=192)
    ref = po.residual_batch(tus, coeff)
    plan = acc.residual_plan(tus, coeff.size)
    dc = hip.DeviceBuffer.from_array(coeff)
    dr = hip.DeviceBuffer(coeff.nbytes)
    s = hip.Stream()
    graph = acc.graph([(plan, dc.ptr, dr.ptr)])
    with pytest.raises(rmh.InvalidInput):
        plan.autotune(dc.ptr, dr.ptr, s.handle, reps=1)
    graph.close()
    plan.autotune(dc.ptr, dr.ptr, s.handle, reps=1)
    graph = acc.graph([(plan, dc.ptr, dr.ptr)])
    for _ in range(2):
        dr.upload(np.zeros_like(coeff))
        graph.launch(s.handle)
        s.synchronize()
        assert (dr.download(np.int16) == ref).all()
    graph.close()
    plan.close()


def test_cross_frame_aliasing_is_rejected(acc, rmh):
    """Written planes that overlap another frame's planes are refused at plan
    creation (before any device work): deblocking in place, out of place and
    SAO."""
    abi, hip = rmh.abi, rmh.hip
    w, h, bd = 64, 32, 10
    buf = hip.DeviceBuffer(8 * w * h * 2)
    edges = hip.DeviceBuffer(16 * 8 * 4)

    def planes(base):
        y = abi.plane_from_address(base, w, h, w, 2, bd)
        u = abi.plane_from_address(base + w * h * 2, w // 2, h // 2, w // 2, 2, bd)
        v = abi.plane_from_address(base + w * h * 2 + (w // 2) * (h // 2) * 2, w // 2, h // 2, w // 2, 2, bd)
        return [y, u, v]

    frame_bytes = w * h * 2 * 3 // 2
    a, b = planes(buf.ptr), planes(buf.ptr + frame_bytes)
    qp = (32, 32, 32)
    fa = abi.dbk_frame(a, edges.ptr, w // 4, h // 4, qp)
    # in place: two frames over the same planes
    with pytest.raises(rmh.InvalidInput):
        acc.deblock_plan([fa, abi.dbk_frame(a, edges.ptr, w // 4, h // 4, qp)])
    # out of place: frame 1's destination is frame 0's source
    fb = abi.dbk_frame(b, edges.ptr, w // 4, h // 4, qp)
    c = planes(buf.ptr + 2 * frame_bytes)
    with pytest.raises(rmh.InvalidInput):
        acc.deblock_to_plan([fa, fb], c + a)
    # out of place: two destinations overlap
    with pytest.raises(rmh.InvalidInput):
        acc.deblock_to_plan([fa, fb], c + c)
    ok = acc.deblock_to_plan([fa, fb], c + planes(buf.ptr + 3 * frame_bytes))
    ok.close()


def test_mc_autotune_keeps_the_prediction_bit_exact(acc, po, rmh):
    """The MC autotuner (rmh_mc_plan_autotune) picks 1, 2 or 4 work items
    per workgroup; the prediction is the oracle's whatever it picks, for uni
    and bi plans, and the tuner is refused while a graph holds the plan."""
    print("mc autotune", flush=True)
    hip, abi = rmh.hip, rmh.abi
    bd = 10
    d = W.inter_frame(9, width=320, height=192)
    from frames import mc_frame_struct
    want = [np.zeros_like(p) for p in d["ref"]]
    f, k = mc_frame_struct(d["ref"], want, bd)
    arr = (type(f) * 1)(f)
    po.lib().orc_mc_batch(ctypes.addressof(arr), 1, d["pus"].ctypes.data, len(d["pus"]))
    bufs, refs, dsts = [], [], []
    for p in d["ref"]:
        b, pl = _dev_plane(hip, p, bd, abi)
        b2, pl2 = _dev_plane(hip, np.zeros_like(p), bd, abi)
        bufs.append((b, b2))
        refs.append(pl)
        dsts.append(pl2)
    mc = acc.mc_plan([abi.mc_frame(refs, dsts)], d["pus"])
    s = hip.Stream()
    waves = mc.autotune_mc(s.handle)
    assert waves in (1, 2, 4)
    for _, b2 in bufs:
        b2.upload(np.zeros(b2.nbytes // 2, np.uint16))
    mc.run(stream=s.handle)
    s.synchronize()
    for c in range(3):
        assert (bufs[c][1].download(np.uint16).reshape(want[c].shape) == want[c]).all(), (waves, c)
    graph = acc.graph([(mc, 0, 0)])
    with pytest.raises(rmh.InvalidInput):
        mc.autotune_mc(s.handle)
    graph.close()
    mc.close()
    # bi
    from frames import mc_bi_frame_struct
    bi = W.bi_frame(9, width=320, height=192)
    want = [np.zeros_like(p) for p in bi["ref0"]]
    fb, kb = mc_bi_frame_struct(bi["ref0"], bi["ref1"], want, bd)
    assert po.mc_bi_batch(fb, bi["pus"]) == 0
    planes, keep = {}, []
    for key in ("ref0", "ref1", "dst"):
        ps = []
        for p in bi["ref0" if key == "dst" else key]:
            b, pl = _dev_plane(hip, np.zeros_like(p) if key == "dst" else p, bd, abi)
            keep.append(b)
            ps.append(pl)
        planes[key] = (abi.Plane * 3)(*ps)
    plan = acc.mc_bi_plan([abi.McBiFrame(planes["ref0"], planes["ref1"], planes["dst"])], bi["pus"])
    assert plan.autotune_mc(s.handle) in (1, 2, 4)
    plan.run(stream=s.handle)
    s.synchronize()
    for c in range(3):
        assert (keep[6 + c].download(np.uint16).reshape(want[c].shape) == want[c]).all(), c
    plan.close()

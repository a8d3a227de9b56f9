"""P-frame reconstruction on the GPU: MC into the picture, the residual added
in place by a recon plan, then deblocking; bit-exact against the oracle."""
import numpy as np
import pytest

import pipeline as PL
import workloads as W
from conftest import load_package

pytestmark = pytest.mark.gpu


def _check(got, want):
    for fi, (g, w) in enumerate(zip(got, want)):
        for c in range(3):
            if not (g[c] == w[c]).all():
                bad = np.argwhere(g[c] != w[c])
                raise AssertionError("frame %d plane %d: %d samples differ, first at %s (got %d, want %d)" % (
                    fi, c, len(bad), bad[0], g[c][tuple(bad[0])], w[c][tuple(bad[0])]))


def test_pframe_1080p10_pipeline_matches_oracle(acc, po):
    rmh = load_package()
    frames = [W.pframe(41, width=1920, height=1080, bit_depth=10)]
    want = PL.oracle(po, rmh.abi, frames)
    b = PL.DeviceBatch(acc, rmh, frames)
    s = rmh.hip.Stream()
    b.run(s.handle)
    s.synchronize()
    _check(b.download(), want)
    b.close()


def test_pframe_built_in_scratch_deblocked_out_of_place(acc, po):
    """MC and the residual add into scratch planes, deblocking out of place
    into the picture (8-bit and 10-bit frames in one batch); the graph is
    replayed over cleared pictures and every replay checked."""
    print("pipeline out of place", flush=True)
    rmh = load_package()
    frames = [W.pframe(43, width=640, height=360, bit_depth=10), W.pframe(44, width=856, height=480, bit_depth=8)]
    want = PL.oracle(po, rmh.abi, frames)
    b = PL.DeviceBatch(acc, rmh, frames, out_of_place=True)
    s = rmh.hip.Stream()
    b.run(s.handle)
    s.synchronize()
    _check(b.download(), want)
    g = b.graph()
    for rep in range(3):
        for fr in b.dst:
            for buf, _, _ in fr:
                buf.zero(s.handle)
        s.synchronize()
        g.launch(s.handle)
        s.synchronize()
        _check(b.download(), want)
    g.close()
    b.close()


def test_mixed_streams_one_batch_graph_replays(acc, po):
    """Frames of three streams (8-bit 856x480 twice, 10-bit 640x368) in one
    batch, captured as one frame-step graph and replayed."""
    rmh = load_package()
    frames = [W.pframe(51, width=856, height=480, bit_depth=8),
              W.pframe(52, width=856, height=480, bit_depth=8),
              W.pframe(53, width=640, height=368, bit_depth=10)]
    want = PL.oracle(po, rmh.abi, frames)
    b = PL.DeviceBatch(acc, rmh, frames)
    g = b.graph()
    s = rmh.hip.Stream()
    for rep in range(3):
        for fr in b.dst:
            for buf, _, _ in fr:
                buf.upload(np.zeros(buf.nbytes, np.uint8))
        g.launch(s.handle)
        s.synchronize()
        _check(b.download(), want)
    g.close()
    b.close()


def test_in_loop_chain_with_sao_graph_replays(acc, po):
    """MC -> residual add -> out-of-place deblocking -> SAO back into the
    picture, over a 1080p10 frame and an 8-bit 856x480 frame, as one
    frame-step graph replayed three times."""
    print("pipeline with SAO", flush=True)
    rmh = load_package()
    frames = [W.pframe(61, width=1920, height=1080, bit_depth=10), W.pframe(62, width=856, height=480, bit_depth=8)]
    want = PL.oracle(po, rmh.abi, frames, with_sao=True)
    b = PL.DeviceBatch(acc, rmh, frames, with_sao=True)
    g = b.graph()
    s = rmh.hip.Stream()
    for rep in range(3):
        for fr in b.dst:
            for buf, _, _ in fr:
                buf.upload(np.zeros(buf.nbytes, np.uint8))
        g.launch(s.handle)
        s.synchronize()
        _check(b.download(), want)
    g.close()
    b.close()


def test_out_of_place_deblock_equals_in_place(acc, po):
    """rmh_deblock_to_plan_create: the source planes are untouched and the
    destination (with its own, wider pitch) equals the in-place result."""
    print("out-of-place deblock", flush=True)
    rmh = load_package()
    hip, abi = rmh.hip, rmh.abi
    from frames import dbk_frame_struct
    d = W.inter_frame(3, width=640, height=360)
    planes = [np.ascontiguousarray(p) for p in d["ref"]]
    want = [p.copy() for p in planes]
    f, keep = dbk_frame_struct(want, d["edges"], 10, d["qp"])
    acc.deblock([f])
    srcs, dsts, bufs = [], [], []
    for p in planes:
        sb = hip.DeviceBuffer.from_array(p)
        pitch = p.shape[1] + 24
        ob = hip.DeviceBuffer(pitch * p.shape[0] * 2)
        ob.upload(np.full(pitch * p.shape[0], 0x7777, np.uint16))
        bufs += [sb, ob]
        srcs.append(abi.plane_from_address(sb.ptr, p.shape[1], p.shape[0], p.shape[1], 2, 10))
        dsts.append(abi.plane_from_address(ob.ptr, p.shape[1], p.shape[0], pitch, 2, 10))
    e = np.ascontiguousarray(d["edges"])
    eb = hip.DeviceBuffer.from_array(e)
    plan = acc.deblock_to_plan([abi.dbk_frame(srcs, eb.ptr, e.shape[1], e.shape[0], d["qp"])], dsts)
    s = hip.Stream()
    plan.run(stream=s.handle)
    s.synchronize()
    for c, p in enumerate(planes):
        pitch = p.shape[1] + 24
        got = bufs[2 * c + 1].download(np.uint16).reshape(p.shape[0], pitch)
        assert (got[:, :p.shape[1]] == want[c]).all()
        assert (got[:, p.shape[1]:] == 0x7777).all()
        assert (bufs[2 * c].download(np.uint16).reshape(p.shape) == p).all()
    with pytest.raises(rmh.InvalidInput):
        acc.deblock_to_plan([abi.dbk_frame(srcs, eb.ptr, e.shape[1], e.shape[0], d["qp"])], srcs)
    plan.close()


@pytest.mark.parametrize("bd,pad", [(10, 0), (10, 1), (8, 0), (8, 1)])
def test_recon_plan_equals_residual_plan_plus_add(acc, po, bd, pad):
    """The fused add against the residual plan's int16 output added on the host,
    on planes whose row pitch is the width (pairs of samples added with one
    access) or the width + 1 (odd pitch: sample by sample)."""
    print("recon plan bd %d pad %d" % (bd, pad), flush=True)
    rmh = load_package()
    hip, abi = rmh.hip, rmh.abi
    f = W.pframe(61, width=512, height=256, bit_depth=bd)
    tus, coeff, place = f["tus"], f["coeff"], f["place"]
    dt = np.uint8 if bd == 8 else np.uint16
    pred = [np.ascontiguousarray(p.astype(dt)) for p in f["ref"]]
    rp = acc.residual_plan(tus, coeff.size)
    dc = hip.DeviceBuffer.from_array(coeff)
    dr = hip.DeviceBuffer(coeff.nbytes)
    rp.run(dc.ptr, dr.ptr, 0)
    hip.device_synchronize()
    res = dr.download(np.int16)
    want = [p.astype(np.int64) for p in pred]
    for t, pl in zip(tus, place):
        w, h = 1 << int(t["log2_w"]), 1 << int(t["log2_h"])
        x, y = int(pl["x"]), int(pl["y"])
        want[int(pl["plane"])][y:y + h, x:x + w] += res[int(t["offset"]):int(t["offset"]) + w * h].reshape(h, w)
    want = [np.clip(p, 0, (1 << bd) - 1) for p in want]
    # planes with pitch width + pad, the padding column set to a sentinel
    padded = []
    for p in pred:
        q = np.full((p.shape[0], p.shape[1] + pad), 77, dt)
        q[:, :p.shape[1]] = p
        padded.append(q)
    bufs = [hip.DeviceBuffer.from_array(q) for q in padded]
    planes = [abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1] + pad, p.itemsize, bd)
              for b, p in zip(bufs, pred)]
    plan = acc.recon_plan(tus, coeff.size, place, [abi.ReconFrame((abi.Plane * 3)(*planes))])
    plan.run(dc.ptr, 0, 0)
    hip.device_synchronize()
    for c in range(3):
        got = bufs[c].download(dt).reshape(padded[c].shape)
        assert (got[:, :pred[c].shape[1]] == want[c]).all(), c
        assert (got[:, pred[c].shape[1]:] == 77).all(), c
    plan.close()
    rp.close()


def test_recon_plan_rejects_bad_placement(acc):
    rmh = load_package()
    hip, abi = rmh.hip, rmh.abi
    f = W.pframe(62, width=128, height=64, bit_depth=10)
    bufs = [hip.DeviceBuffer(p.nbytes) for p in f["ref"]]
    planes = [abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], 2, 10) for b, p in zip(bufs, f["ref"])]
    frames = [abi.ReconFrame((abi.Plane * 3)(*planes))]
    cases = []
    p = f["place"].copy()
    p["x"][0] = 127  # block leaves the plane
    cases.append((f["tus"], p))
    p = f["place"].copy()
    p[1] = p[0]  # two TUs on one block
    t = f["tus"].copy()
    t[1]["log2_w"], t[1]["log2_h"] = t[0]["log2_w"], t[0]["log2_h"]
    t[1]["offset"] = t[0]["offset"]
    cases.append((t, p))
    p = f["place"].copy()
    p["frame"][0] = 1  # no such frame
    cases.append((f["tus"], p))
    p = f["place"].copy()
    p["reserved"][0] = 1
    cases.append((f["tus"], p))
    t = f["tus"].copy()
    t["bit_depth"] = 8  # planes are 10-bit
    t["qp"] = 30
    cases.append((t, f["place"]))
    for i, (tt, pp) in enumerate(cases):
        with pytest.raises(rmh.InvalidInput):
            acc.recon_plan(tt, f["coeff"].size, pp, frames)
    acc.recon_plan(f["tus"], f["coeff"].size, f["place"], frames).close()

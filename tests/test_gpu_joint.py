"""Joint Cb-Cr (ICT) residuals in recon plans on the GPU against the oracle
(reconstruct_joint_cbcr, vvc/decode_general.rs:3908-3981), bit-exact: every
mode and sign, with transform skip, over predictions that reach both clip
ends; and the validation that keeps joint TUs out of residual-only plans and
rejects a second TU over the derived block."""
import numpy as np
import pytest

import pipeline as PL
import workloads as W
from test_oracle_recon import _run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bd", [8, 10])
def test_joint_cbcr_recon_plan_matches_oracle(acc, po, rmh, bd):
    hip, abi = rmh.hip, rmh.abi
    tus, coeff, place, blocks = W.joint_cbcr_frame(11 + bd, width=320, height=192, bit_depth=bd)
    assert {b[5] for b in blocks} >= {1, 2, 3, 5, 6, 7}
    rng = np.random.default_rng(bd)
    maxv = (1 << bd) - 1
    dt = PL.sample_dtype(bd)
    pred = [np.ascontiguousarray(rng.choice([0, 3, maxv // 2, maxv - 2, maxv], size=s).astype(dt))
            for s in [(192, 320), (96, 160), (96, 160)]]
    want = [[p.copy() for p in pred]]
    assert _run_oracle(po, tus, coeff, place, want, bd) == 0
    bufs = [hip.DeviceBuffer.from_array(p) for p in pred]
    planes = [abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], p.itemsize, bd)
              for b, p in zip(bufs, pred)]
    plan = acc.recon_plan(tus, coeff.size, place, [abi.ReconFrame((abi.Plane * 3)(*planes))])
    dc = hip.DeviceBuffer.from_array(coeff)
    s = hip.Stream()
    plan.run(dc.ptr, 0, s.handle)
    s.synchronize()
    for c in range(3):
        got = bufs[c].download(dt).reshape(pred[c].shape)
        assert (got == want[0][c]).all(), c
    plan.close()


def test_joint_cbcr_validation(acc, rmh):
    abi, hip = rmh.abi, rmh.hip
    tus, coeff, place, blocks = W.joint_cbcr_frame(5, width=64, height=64, bit_depth=10)
    with pytest.raises(rmh.InvalidInput):
        acc.residual_plan(tus, coeff.size)  # joint TUs need a recon plan
    bufs = [hip.DeviceBuffer(64 * 64 * 2), hip.DeviceBuffer(32 * 32 * 2), hip.DeviceBuffer(32 * 32 * 2)]
    planes = [abi.plane_from_address(bufs[0].ptr, 64, 64, 64, 2, 10)] + \
             [abi.plane_from_address(b.ptr, 32, 32, 32, 2, 10) for b in bufs[1:]]
    frames = [abi.ReconFrame((abi.Plane * 3)(*planes))]
    i = next(k for k, b in enumerate(blocks) if b[5])
    # a second TU over the derived block of a joint TU
    extra = tus[i:i + 1].copy()
    extra["joint"] = 0
    pl = place[i:i + 1].copy()
    pl["plane"] = 3 - int(place[i]["plane"])
    with pytest.raises(rmh.InvalidInput):
        acc.recon_plan(np.concatenate([tus, extra]), coeff.size, np.concatenate([place, pl]), frames)
    # a joint TU placed on the component it does not code
    bad = place.copy()
    bad[i]["plane"] = 3 - int(place[i]["plane"])
    with pytest.raises(rmh.InvalidInput):
        acc.recon_plan(tus, coeff.size, bad, frames)
    ok = acc.recon_plan(tus, coeff.size, place, frames)
    ok.close()

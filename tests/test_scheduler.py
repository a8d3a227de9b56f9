"""The placement scheduler's host logic (rustmpeg-next_amd/scheduler.py)
against the reference's Scheduler (crates/rustmpeg-vulkan/src/scheduler.rs):
fit_line's two-point line with its floors and fallback (535-557), place's
ten percent margin and its CPU answers when the GPU is off, uncalibrated or
the frame is empty (481-499), and synthetic calibration workloads that the
oracle accepts. No device is needed; test_gpu_scheduler.py calibrates on the
GPU."""
import numpy as np
import pytest

from conftest import load_package

S = load_package().scheduler


def test_fit_line_through_two_points():
    assert S.fit_line(100.0, 1100.0, 10.0, 110.0) == (10.0, 0.0)
    assert S.fit_line(300.0, 1300.0, 10.0, 110.0) == (10.0, 200.0)


def test_fit_line_floors_and_fallback():
    # a noisy pair (large faster than small): slope floored, fixed part from the small point
    slope, fixed = S.fit_line(500.0, 400.0, 10.0, 110.0)
    assert slope == S.MIN_NS_PER_PIXEL and fixed == pytest.approx(500.0 - 10.0 * S.MIN_NS_PER_PIXEL)
    # no span: the large frame's rate, no fixed part
    assert S.fit_line(5.0, 2000.0, 100.0, 100.0) == (20.0, 0.0)
    assert S.fit_line(5.0, 0.0, 100.0, 100.0) == (S.MIN_NS_PER_PIXEL, 0.0)
    # a fixed part never goes negative
    assert S.fit_line(0.0, 1000.0, 10.0, 110.0)[1] == 0.0


def _calibrated(cpu_rate, gpu_rate, gpu_fixed=0.0):
    s = S.Scheduler(None, {})
    s.gpu_enabled = True
    s.calibrated = True
    s.costs = {st: S.StageCost(0.0, cpu_rate, gpu_fixed, gpu_rate) for st in S.Stage}
    s.measurements = {st: {"cpu_ns": [0, 0], "gpu_ns": [0, 0]} for st in S.Stage}
    return s


def test_place_needs_the_ten_percent_margin():
    big = S.FrameSize(1920, 1080)
    assert _calibrated(1.0, 0.5).place(S.Stage.VVC_RESIDUAL, big) is S.Placement.GPU
    # gpu * 1.1 must be strictly below the CPU estimate
    assert _calibrated(1.1, 1.0).place(S.Stage.VVC_RESIDUAL, big) is S.Placement.CPU
    assert _calibrated(1.0, 0.95).place(S.Stage.VVC_DEBLOCK, big) is S.Placement.CPU
    # a fixed GPU cost keeps small frames on the CPU and lets large ones go
    s = _calibrated(1.0, 0.1, gpu_fixed=100000.0)
    assert s.place(S.Stage.VVC_INTER_PRED, S.FrameSize(128, 128)) is S.Placement.CPU
    assert s.place(S.Stage.VVC_INTER_PRED, big) is S.Placement.GPU
    assert s.estimate(S.Stage.VVC_INTER_PRED, big) == pytest.approx((1920 * 1080, 100000.0 + 0.1 * 1920 * 1080))


def test_place_is_cpu_when_disabled_uncalibrated_or_empty():
    s = _calibrated(1.0, 0.1)
    assert s.place(S.Stage.VVC_RESIDUAL, S.FrameSize(0, 1080)) is S.Placement.CPU
    s.disable_gpu()
    assert s.place(S.Stage.VVC_RESIDUAL, S.FrameSize(1920, 1080)) is S.Placement.CPU
    assert s.estimate(S.Stage.VVC_RESIDUAL, S.FrameSize(1920, 1080)) is not None  # the estimate stays
    u = S.Scheduler(None, {})
    assert u.place(S.Stage.VVC_RESIDUAL, S.FrameSize(1920, 1080)) is S.Placement.CPU
    assert u.estimate(S.Stage.VVC_RESIDUAL, S.FrameSize(1920, 1080)) is None
    with pytest.raises(load_package().Unavailable):
        u.calibrate()


def test_calibration_workloads_are_valid_for_the_oracle(po):
    import ctypes
    from frames import dbk_frame_struct, mc_frame_struct
    size = S.FrameSize(128, 128)
    tus, coeff = S.residual_workload(size)
    assert len(tus) == 64 + 2 * 64 and coeff.size == 128 * 128 + 2 * 64 * 64
    assert (coeff != 0).mean() > 0.1
    out = po.residual_batch(tus, coeff)
    assert (out != 0).any()
    ref, pus = S.inter_workload(size)
    dst = [np.zeros_like(p) for p in ref]
    f, keep = mc_frame_struct(ref, dst, 10)
    arr = (type(f) * 1)(f)
    assert po.lib().orc_mc_batch(ctypes.addressof(arr), 1, pus.ctypes.data, len(pus)) == 0
    assert all(d.any() for d in dst)
    planes, edges, qp = S.deblock_workload(size)
    work = [p.copy() for p in planes]
    f, keep = dbk_frame_struct(work, edges, 10, (qp, qp, qp))
    arr = (type(f) * 1)(f)
    po.lib().orc_deblock_frames(ctypes.addressof(arr), 1)
    assert any((a != b).any() for a, b in zip(work, planes))

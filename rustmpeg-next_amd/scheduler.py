"""Per-stage CPU/GPU placement with a calibrated cost model, for the VVC
reconstruction stages.

Mirrors Scheduler of crates/rustmpeg-vulkan/src/scheduler.rs with the stages
Stage::{VvcResidual, VvcInterPred, VvcDeblock} in place of its colour and
metrics stages:
  - calibrate (scheduler.rs:310-453): both paths run on synthetic frames of
    two sizes (CALIBRATION_SIZES, 128x128 and 1920x1080), after eight warm-up
    runs of the GPU path at the large size; every GPU result is compared with
    the CPU result, and a mismatch disables the GPU for good and raises
    Execution, so a disagreeing kernel never changes an output; each path is
    timed as the best of TIMED_RUNS runs after one untimed run (best_of_three,
    scheduler.rs:559-579);
  - the cost of each path is a line through the two points, its slope at least
    MIN_NS_PER_PIXEL and its fixed part at least 0 (fit_line, 535-557);
  - place (481-499): the GPU when it is enabled and calibrated and its
    predicted time times GPU_MARGIN (1.1) is below the CPU's.

The GPU path is this library's host entry points (rmh_residual_batch,
rmh_mc_batch, rmh_deblock): copies in, kernels, copies out, as the reference
times its Vulkan kernels with their uploads and read-backs. The CPU path is
the caller's own implementation of each stage (in rustmpeg the decoder's
inverse_residual, interpolate_* and deblocking functions): this library has
no CPU path of its own, so `cpu` maps each stage to a callable that takes the
stage's workload and returns its output.

Workloads (per frame size, 10-bit 4:2:0): every TU of the frame (16x16 luma
and 8x8 chroma, DCT-II / DST-VII / DCT-VIII, sparse levels); one 16x16 PU per
luma block with an MV within +-16 samples at any fraction; deblocking on the
8x8 grid with bS 1 or 2 and lengths 3. The outputs compared are the residual
array, the three prediction planes and the three deblocked planes.
"""
import enum
import math
import time
from dataclasses import dataclass

import numpy as np

from . import abi

GPU_MARGIN = 1.1
MIN_NS_PER_PIXEL = 1e-6
TIMED_RUNS = 5


class Stage(enum.Enum):
    VVC_RESIDUAL = "vvc_residual"
    VVC_INTER_PRED = "vvc_inter_pred"
    VVC_DEBLOCK = "vvc_deblock"

    @property
    def label(self):
        return self.value


class Placement(enum.Enum):
    CPU = "cpu"
    GPU = "gpu"


@dataclass(frozen=True)
class FrameSize:
    width: int
    height: int

    def pixels(self):
        return max(0, self.width) * max(0, self.height)

    def is_empty(self):
        return self.width == 0 or self.height == 0


CALIBRATION_SIZES = (FrameSize(128, 128), FrameSize(1920, 1080))


@dataclass
class StageCost:
    cpu_fixed_ns: float
    cpu_ns_per_pixel: float
    gpu_fixed_ns: float
    gpu_ns_per_pixel: float


def fit_line(small_ns, large_ns, small_px, large_px):
    """(ns per pixel, fixed ns) of the line through two measurements; a
    non-finite slope falls back to the large frame's rate, and neither part
    may go below its floor."""
    span = large_px - small_px
    slope = (large_ns - small_ns) / span if span > 0 else math.nan
    if not math.isfinite(slope):
        rate = max(large_ns / large_px, MIN_NS_PER_PIXEL) if large_px > 0 else MIN_NS_PER_PIXEL
        return rate, 0.0
    slope = max(slope, MIN_NS_PER_PIXEL)
    return slope, max(small_ns - slope * small_px, 0.0)


def stage_cost(cpu_small, cpu_large, gpu_small, gpu_large, small_px, large_px):
    cs, cf = fit_line(cpu_small, cpu_large, small_px, large_px)
    gs, gf = fit_line(gpu_small, gpu_large, small_px, large_px)
    return StageCost(cpu_fixed_ns=cf, cpu_ns_per_pixel=cs, gpu_fixed_ns=gf, gpu_ns_per_pixel=gs)


def best_of(run, runs=TIMED_RUNS):
    """Fastest wall time of `runs` calls in ns, after one untimed call."""
    run()
    best = math.inf
    for _ in range(runs):
        t0 = time.perf_counter_ns()
        run()
        best = min(best, time.perf_counter_ns() - t0)
    return float(best)


# ------------------------------------------------------------ workloads

def _rng(seed):
    return np.random.default_rng([11, seed])


def _planes(size, bd, seed):
    """A 4:2:0 picture of smooth texture plus noise, uint16."""
    w, h = size.width, size.height
    maxv = (1 << bd) - 1
    out = []
    for c, (pw, ph) in enumerate([(w, h), ((w + 1) // 2, (h + 1) // 2), ((w + 1) // 2, (h + 1) // 2)]):
        yy, xx = np.mgrid[0:ph, 0:pw].astype(np.float64)
        v = 0.5 * maxv + 0.2 * maxv * np.sin(0.07 * xx + 0.05 * yy + seed + c) + 0.1 * maxv * np.cos(0.011 * xx * (1 + c))
        noise = _rng(seed * 7 + c).integers(-16, 17, (16, 16))
        v = v + np.tile(noise, (ph // 16 + 1, pw // 16 + 1))[:ph, :pw]
        out.append(np.clip(np.rint(v), 0, maxv).astype(np.uint16))
    return out


def residual_workload(size, bd=10, seed=1):
    """(tus, coeff): 16x16 luma TUs and 8x8 chroma TUs covering the frame."""
    blocks = []
    for plane, (pw, ph, n) in enumerate([(size.width, size.height, 16),
                                         ((size.width + 1) // 2, (size.height + 1) // 2, 8),
                                         ((size.width + 1) // 2, (size.height + 1) // 2, 8)]):
        for y in range(0, ph - n + 1, n):
            for x in range(0, pw - n + 1, n):
                blocks.append(n)
    n_tu = len(blocks)
    tus = np.zeros(n_tu, abi.TU_DTYPE)
    sizes = np.array(blocks) ** 2
    tus["offset"] = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    lg = np.log2(blocks).astype(np.uint8)
    tus["log2_w"], tus["log2_h"] = lg, lg
    rng = _rng(seed)
    tus["qp"] = rng.integers(22, 38, n_tu) + 6 * (bd - 8)
    tus["bit_depth"] = bd
    tus["tr_hor"], tus["tr_ver"] = rng.integers(0, 3, n_tu), rng.integers(0, 3, n_tu)
    total = int(sizes.sum())
    vals = rng.integers(-20, 21, total)
    keep = rng.random(total) < 0.15
    return tus, np.where(keep, vals, 0).astype(np.int16)


def inter_workload(size, bd=10, seed=2):
    """(reference planes, PUs): one 16x16 PU per luma block."""
    ref = _planes(size, bd, seed)
    pus = []
    for y in range(0, size.height - 15, 16):
        for x in range(0, size.width - 15, 16):
            pus.append((x, y))
    p = np.zeros(len(pus), abi.PU_DTYPE)
    p["x"], p["y"] = [q[0] for q in pus], [q[1] for q in pus]
    p["log2_w"], p["log2_h"] = 4, 4
    rng = _rng(seed)
    p["mv_x"], p["mv_y"] = rng.integers(-256, 257, len(pus)), rng.integers(-256, 257, len(pus))
    return ref, p


def deblock_workload(size, bd=10, seed=3):
    """(planes, edges, qp): bS 1 or 2 with lengths 3 on the 8x8 grid."""
    planes = _planes(size, bd, seed)
    gw, gh = (size.width + 3) // 4, (size.height + 3) // 4
    edges = np.zeros((gh, gw, 4), np.uint8)
    bs = _rng(seed).integers(1, 3, (gh, gw)).astype(np.uint8)
    lens = (1 << 2) | (1 << 4)  # lenP = lenQ = 3
    gx, gy = np.meshgrid(np.arange(gw), np.arange(gh))
    edges[..., 0] = np.where((gx % 2 == 0) & (gx > 0), bs | lens, 0)
    edges[..., 1] = np.where((gy % 2 == 0) & (gy > 0), bs | lens, 0)
    edges[..., 2] = np.where((gx % 4 == 0) & (gx > 0), 2, 0)
    edges[..., 3] = np.where((gy % 4 == 0) & (gy > 0), 2, 0)
    return planes, edges, 32


def gpu_run(acc, stage, work, bd=10):
    """The stage on the accelerator's host entry points; returns its output."""
    if stage is Stage.VVC_RESIDUAL:
        tus, coeff = work
        return acc.residuals(tus, coeff)
    if stage is Stage.VVC_INTER_PRED:
        ref, pus = work
        out = [np.zeros_like(p) for p in ref]
        f = abi.mc_frame([abi.plane_from_array(p, bd) for p in ref], [abi.plane_from_array(p, bd) for p in out])
        acc.predict([f], pus)
        return out
    planes, edges, qp = work
    out = [p.copy() for p in planes]
    f = abi.dbk_frame([abi.plane_from_array(p, bd) for p in out], edges.ctypes.data, edges.shape[1], edges.shape[0],
                      (qp, qp, qp))
    acc.deblock([f])
    return out


WORKLOADS = {Stage.VVC_RESIDUAL: residual_workload, Stage.VVC_INTER_PRED: inter_workload,
             Stage.VVC_DEBLOCK: deblock_workload}


def _equal(a, b):
    if isinstance(a, (list, tuple)):
        return len(a) == len(b) and all(_equal(x, y) for x, y in zip(a, b))
    return np.array_equal(np.asarray(a), np.asarray(b))


class Scheduler:
    """Places each stage on the CPU or the GPU from a calibrated cost model.

    acc: an Accelerator or None (no GPU); cpu: {Stage: callable(workload) ->
    output}, the caller's CPU implementation of each stage."""

    def __init__(self, acc, cpu):
        self.acc = acc
        self.cpu = dict(cpu)
        self.gpu_enabled = acc is not None
        self.calibrated = False
        self.costs = None
        self.measurements = None

    def disable_gpu(self):
        """Permanently: calibration never re-enables it."""
        self.gpu_enabled = False

    def calibrate(self, stages=tuple(Stage)):
        from . import Execution, Unavailable
        if self.acc is None:
            raise Unavailable("scheduler has no GPU")
        works = {st: [WORKLOADS[st](sz) for sz in CALIBRATION_SIZES] for st in stages}
        for _ in range(8):  # raise the clocks before any measurement
            for st in stages:
                gpu_run(self.acc, st, works[st][1])
        costs, meas = {}, {}
        small_px, large_px = CALIBRATION_SIZES[0].pixels(), CALIBRATION_SIZES[1].pixels()
        for st in stages:
            cpu_ns, gpu_ns = [], []
            for work in works[st]:
                if not _equal(gpu_run(self.acc, st, work), self.cpu[st](work)):
                    self.disable_gpu()
                    raise Execution("%s kernel disagrees with the CPU path" % st.label)
                cpu_ns.append(best_of(lambda: self.cpu[st](work)))
                gpu_ns.append(best_of(lambda: gpu_run(self.acc, st, work)))
            costs[st] = stage_cost(cpu_ns[0], cpu_ns[1], gpu_ns[0], gpu_ns[1], small_px, large_px)
            meas[st] = {"cpu_ns": cpu_ns, "gpu_ns": gpu_ns}
        self.costs, self.measurements = costs, meas
        self.calibrated = True

    def estimate(self, stage, size):
        """(cpu ns, gpu ns) predicted for one frame of `size`, or None."""
        if not self.calibrated or size.pixels() == 0 or stage not in self.costs:
            return None
        c, px = self.costs[stage], float(size.pixels())
        return c.cpu_fixed_ns + c.cpu_ns_per_pixel * px, c.gpu_fixed_ns + c.gpu_ns_per_pixel * px

    def place(self, stage, size):
        if not self.gpu_enabled or not self.calibrated or size.is_empty():
            return Placement.CPU
        est = self.estimate(stage, size)
        if est is None:
            return Placement.CPU
        cpu_ns, gpu_ns = est
        if not math.isfinite(cpu_ns) or cpu_ns <= 0 or not math.isfinite(gpu_ns) or gpu_ns < 0:
            return Placement.CPU
        return Placement.GPU if gpu_ns * GPU_MARGIN < cpu_ns else Placement.CPU

    def report(self, sizes=CALIBRATION_SIZES):
        """Both paths' measurements and fitted costs, and the placement at each
        size, per stage (what the reference logs after calibrating)."""
        out = {}
        for st, c in (self.costs or {}).items():
            out[st.label] = {"measured_ns": self.measurements[st], "cost": c.__dict__,
                             "placement": {"%dx%d" % (s.width, s.height): self.place(st, s).value for s in sizes}}
        return out

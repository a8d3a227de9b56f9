#!/usr/bin/env python3
"""Benchmark: bit-exact batched VVC residual reconstruction on one MI355X.

Workload (one step = one pass of the residual stage over one batch): every TU
of F synthetic 1920x1080 10-bit 4:2:0 frames (recursive tiling 4x4..64x64
favouring 8x8/16x16, DCT-II/DST-VII/DCT-VIII drawn per TU, int16 Laplacian
coefficients with ~85% zeros toward high frequencies, QP 22..37), dequantised
and inverse transformed to int16 residuals by one launch of the batched kernel.
Inputs (descriptors + coefficients) are resident in HBM before timing starts.

Reported beside the batched rate: the same kernel run unbatched (one launch
per frame), the roofline of the kernel against MI355X HBM (8 TB/s), and the
CPU oracle timed on a bounded sample on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--frames F]
    python bench.py --copy-calibration      (copy kernel of known size, for rocprofv3 --pmc scaling)
    python bench.py --cnn-kernels           (the CU-split network's forward pass and training step alone)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
METRIC = "bit-exact 1080p10 VVC recon Mpix/s vs unbatched; encode fps learned vs full RDO"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(REPO, "profiles", "residual_tile_kernel_1080p10x32_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_calibration():
    """A plain device copy of known size (torch's copy kernel), profiled with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to scale those counters."""
    import torch
    n = 256 << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n, "copies": 5}), flush=True)


def read_traffic():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, None
    vals = {}
    for line in open(TRAFFIC_FILE):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"]), vals.get("frames")
    except (KeyError, ValueError):
        return None, None


def load_package():
    import importlib.util
    name = "rustmpeg_next_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(REPO, "rustmpeg-next_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"), submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def host_info():
    """CPU model and the compiler of the oracle build, for the CPU baseline."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        import subprocess
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True, timeout=20).stdout.splitlines()[0]
    except Exception:
        cc = None
    return model, cc


def host_threads():
    """Threads for the all-cores leg: this process's CPU share (the affinity
    mask), not the machine's core count."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def median_rate(fn, work_units, runs=7):
    """Units per second of fn(), as the median of `runs` timed calls after one
    warm-up call (the method of crates/rustmpeg-vulkan/examples/benchmark.rs:66-81)."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return work_units / ts[len(ts) // 2], ts[len(ts) // 2]


def cpu_baseline(tus, coeff, luma_px, n_frames):
    """The CPU oracle (oracle/liboracle.so, scalar C) over the whole residual
    workload: one host thread, and every host thread of this process (OpenMP
    over TUs); median of 7 after a warm-up. `value` is the all-threads rate."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as po
    L = po.lib()
    n = host_threads()
    L.orc_set_threads(1)
    r1, t1 = median_rate(lambda: po.residual_batch(tus, coeff), luma_px)
    L.orc_set_threads(n)
    rn, tn = median_rate(lambda: po.residual_batch(tus, coeff), luma_px)
    L.orc_set_threads(1)
    model, cc = host_info()
    return {"value": round(rn / 1e6, 2), "unit": "Mpix/s", "cores": n, "kind": "port",
            "sample": "the whole workload (%d frames of 1080p10 TUs), median of 7 runs after a warm-up; scalar C "
                      "oracle (oracle/orc_residual.c), OpenMP over TUs on %d threads" % (n_frames, n),
            "ms_per_step": round(tn * 1e3, 2),
            "single_thread": {"value": round(r1 / 1e6, 3), "unit": "Mpix/s", "cores": 1, "ms_per_step": round(t1 * 1e3, 1)},
            "cpu_model": model, "compiler": cc}


def placement_report(rmh, acc):
    """The placement scheduler (rustmpeg-next_amd/scheduler.py, after
    Scheduler::calibrate / place of the reference's Vulkan backend) calibrated
    on this box with the CPU oracle on every host thread as the CPU path: both
    paths' measured times at 128x128 and 1920x1080, the fitted costs and the
    placement of each stage at both sizes. The GPU path is the host-memory
    entry points, PCIe copies included."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import ctypes
    import pyoracle as po
    from frames import dbk_frame_struct, mc_frame_struct
    S = rmh.scheduler

    def inter(work):
        ref, pus = work
        dst = [np.zeros_like(p) for p in ref]
        f, keep = mc_frame_struct(ref, dst, 10)
        arr = (type(f) * 1)(f)
        po.lib().orc_mc_batch(ctypes.addressof(arr), 1, pus.ctypes.data, len(pus))
        return dst

    def deblock(work):
        planes, edges, qp = work
        out = [p.copy() for p in planes]
        f, keep = dbk_frame_struct(out, edges, 10, (qp, qp, qp))
        arr = (type(f) * 1)(f)
        po.lib().orc_deblock_frames(ctypes.addressof(arr), 1)
        return out

    po.lib().orc_set_threads(host_threads())
    try:
        sch = S.Scheduler(acc, {S.Stage.VVC_RESIDUAL: lambda w: po.residual_batch(*w),
                                S.Stage.VVC_INTER_PRED: inter, S.Stage.VVC_DEBLOCK: deblock})
        sch.calibrate()
    finally:
        po.lib().orc_set_threads(1)
    rep = sch.report()
    for k, v in rep.items():
        log("placement %s: cpu %s ns, gpu %s ns -> %s" % (k, [round(x) for x in v["measured_ns"]["cpu_ns"]],
                                                          [round(x) for x in v["measured_ns"]["gpu_ns"]],
                                                          v["placement"]))
    return {"cpu_path": "oracle on %d host threads" % host_threads(),
            "gpu_path": "host-memory entry points (PCIe copies included)", "stages": rep}


def cpu_pipeline_baseline(PL, po, rmh, frames):
    """The CPU oracle over the P-frame pipeline (MC -> residual add ->
    deblock) of `frames`: one thread and every thread of this process (OpenMP
    over PUs, TUs and frames), median of 7 after a warm-up."""
    L = po.lib()
    px = sum(int(f["width"]) * int(f["height"]) for f in frames)
    n = host_threads()
    L.orc_set_threads(1)
    r1, t1 = median_rate(lambda: PL.oracle(po, rmh.abi, frames), px)
    L.orc_set_threads(n)
    rn, tn = median_rate(lambda: PL.oracle(po, rmh.abi, frames), px)
    L.orc_set_threads(1)
    return {"value": round(rn / 1e6, 2), "unit": "Mpix/s", "cores": n, "kind": "port",
            "sample": "the same %d P-frames, median of 7 runs after a warm-up; oracle MC, recon and deblock "
                      "(OpenMP over PUs, TUs and frames)" % len(frames),
            "ms_per_step": round(tn * 1e3, 2),
            "single_thread": {"value": round(r1 / 1e6, 3), "unit": "Mpix/s", "cores": 1,
                              "ms_per_step": round(t1 * 1e3, 1)}}


def timed(hip, stream, fn, steps, warmup):
    """Average ms of fn(stream) over `steps` calls after `warmup` calls, by HIP
    events on `stream`."""
    for _ in range(warmup):
        fn(stream.handle)
    hip.device_synchronize()
    ev0, ev1 = hip.Event(), hip.Event()
    ev0.record(stream.handle)
    for _ in range(steps):
        fn(stream.handle)
    ev1.record(stream.handle)
    ev1.synchronize()
    hip.device_synchronize()
    return ev0.elapsed_ms(ev1) / steps


def stage_roofline(plan, ms):
    gbs = plan.bytes / (ms / 1e3) / 1e9
    return {"ms": round(ms, 4), "algorithmic_bytes": plan.bytes, "achieved_gbs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def check_batch(PL, po, rmh, batch, frames, which):
    """Bit-exact gate: the frames at indices `which` against the oracle."""
    got = batch.download()
    for i in which:
        want = PL.oracle(po, rmh.abi, [frames[i]])[0]
        for c in range(3):
            if not (got[i][c] == want[c]).all():
                log("bench: frame %d plane %d differs from the oracle" % (i, c))
                return False
    return True


def bench_pipeline(rmh, acc, W, PL, po, n_frames, steps, warmup, cpu=True, autotune=True):
    """P-frame reconstruction, MC -> residual add -> deblock, over n_frames
    1080p10 frames: one frame-step graph over the batch against the same
    plans built per frame (unbatched)."""
    hip = rmh.hip
    t0 = time.perf_counter()
    frames = [W.pframe(2000 + i) for i in range(n_frames)]
    log("pipeline: generated %d P-frames in %.1f s" % (n_frames, time.perf_counter() - t0))
    batch = PL.DeviceBatch(acc, rmh, frames)
    stream = hip.Stream()
    tuned = None
    if autotune:
        # shapes chosen before the graph captures the plans
        tuned = {"mc_waves": batch.mc.autotune_mc(stream.handle),
                 "recon_split": {"%dx%d" % k: v for k, v in batch.recon.autotune(batch.d_coeff.ptr, 0,
                                                                                   stream.handle).items() if v > 1}}
        stream.synchronize()
        log("pipeline: autotuned %s" % tuned)
    graph = batch.graph()
    graph.launch(stream.handle)
    stream.synchronize()
    if not check_batch(PL, po, rmh, batch, frames, (0, n_frames - 1)):
        return None
    ms = timed(hip, stream, graph.launch, steps, warmup)
    px = batch.luma_pixels()
    stages = {"mc": stage_roofline(batch.mc, timed(hip, stream, lambda s: batch.mc.run(stream=s), steps, 1)),
              "recon": stage_roofline(batch.recon, timed(hip, stream, lambda s: batch.recon.run(batch.d_coeff.ptr, 0, s),
                                                          steps, 1)),
              "deblock": stage_roofline(batch.deblock, timed(hip, stream, lambda s: batch.deblock.run(stream=s),
                                                             steps, 1))}
    singles = [PL.DeviceBatch(acc, rmh, [f]) for f in frames]

    def unbatched(s):
        for b in singles:
            b.run(s)
    ums = timed(hip, stream, unbatched, steps, warmup)
    for b in singles:
        b.close()
    graph.close()
    batch.close()
    out = {"workload": "%d P-frames 1920x1080 10-bit 4:2:0: MC (8/4-tap, MVs +-64 px) -> residual add "
                       "(transform trees of the coded CUs) -> deblock, one frame-step graph" % n_frames,
           "value": round(px / (ms / 1e3) / 1e6, 1), "unit": "Mpix/s", "ms_per_step": round(ms, 4),
           "unbatched": {"value": round(px / (ums / 1e3) / 1e6, 1), "ms_per_step": round(ums, 4),
                         "launches_per_step": 4 * n_frames},
           "stages": stages, "autotune": tuned}
    out["vs_unbatched"] = round(out["value"] / out["unbatched"]["value"], 3)
    log("pipeline: %.3f ms/step %.1f Mpix/s, unbatched %.3f ms (x%.2f)" % (ms, out["value"], ums, out["vs_unbatched"]))
    if cpu:
        out["cpu_baseline"] = cpu_pipeline_baseline(PL, po, rmh, frames)
        log("pipeline cpu oracle: %s Mpix/s on %d threads, %s on one" % (
            out["cpu_baseline"]["value"], out["cpu_baseline"]["cores"], out["cpu_baseline"]["single_thread"]["value"]))

    # the same chain built in scratch planes and deblocked out of place into
    # the picture (no shadow copy of the pre-deblocking samples)
    ob = PL.DeviceBatch(acc, rmh, frames, out_of_place=True)
    ograph = ob.graph()
    ograph.launch(stream.handle)
    stream.synchronize()
    want = PL.oracle(po, rmh.abi, [frames[0]])[0]
    got = ob.download()[0]
    if not all((got[c] == want[c]).all() for c in range(3)):
        log("bench: out-of-place pipeline differs from the oracle")
        out["out_of_place"] = None
    else:
        oms = timed(hip, stream, ograph.launch, steps, warmup)
        out["out_of_place"] = {
            "workload": "the same frames: MC and residual add into scratch planes, deblock out of place into "
                        "the picture, one frame-step graph",
            "value": round(px / (oms / 1e3) / 1e6, 1), "unit": "Mpix/s", "ms_per_step": round(oms, 4)}
        log("pipeline out of place: %.3f ms/step" % oms)
    ograph.close()
    ob.close()

    # the in-loop chain with SAO: deblocking out of place into scratch planes
    # (no shadow copy), SAO back into the picture, one graph
    sb = PL.DeviceBatch(acc, rmh, frames, with_sao=True)
    sgraph = sb.graph()
    sgraph.launch(stream.handle)
    stream.synchronize()
    want = PL.oracle(po, rmh.abi, [frames[0]], with_sao=True)[0]
    got = sb.download()[0]
    if not all((got[c] == want[c]).all() for c in range(3)):
        log("bench: pipeline with SAO differs from the oracle")
        out["with_sao"] = None
    else:
        sms = timed(hip, stream, sgraph.launch, steps, warmup)
        out["with_sao"] = {
            "workload": "the same frames: MC -> residual add -> deblock (out of place) -> SAO (CTB 64, random "
                        "band/edge parameters) back into the picture, one frame-step graph",
            "value": round(px / (sms / 1e3) / 1e6, 1), "unit": "Mpix/s", "ms_per_step": round(sms, 4),
            "stages": {"deblock_out_of_place": stage_roofline(sb.deblock, timed(hip, stream,
                                                                                 lambda s: sb.deblock.run(stream=s),
                                                                                 steps, 1)),
                       "sao": stage_roofline(sb.sao, timed(hip, stream, lambda s: sb.sao.run(stream=s), steps, 1))}}
        log("pipeline with SAO: %.3f ms/step" % sms)
    sgraph.close()
    sb.close()
    return out


def bench_streams(rmh, acc, PL, po, steps, reps, autotune=True):
    """The batched multi-stream configuration: 16 concurrent streams
    (4x 854x480 8-bit coded as 856x480, 8x 1080p10, 4x 2160p10) x `steps`
    chained frame steps with the residual tool mix; each step is one graph
    over every stream, and its output is the next step's reference. Against
    the same steps with one launch per stage per stream."""
    hip = rmh.hip
    t0 = time.perf_counter()
    refs, frames = PL.generate_streams(PL.STREAMS, steps, tools=True, seed=3000, workers=min(16, host_threads()))
    log("streams: generated %d frames in %.1f s" % (steps * len(PL.STREAMS), time.perf_counter() - t0))
    chain = PL.ChainedStreams(acc, rmh, refs, frames, autotune=autotune)
    if autotune:
        log("streams: autotuned %s" % chain.tuned)
    stream = hip.Stream()
    chain.run(stream.handle)
    stream.synchronize()
    # bit-exact gate: three streams (one of each size) after the last step
    got = chain.output(steps - 1)
    for k in (0, 4, 12):
        want = PL.oracle_chain(po, rmh.abi, refs[k], [frames[t][k] for t in range(steps)])
        if not all((got[k][c] == want[c]).all() for c in range(3)):
            log("bench: stream %d differs from the oracle after %d chained steps" % (k, steps))
            chain.close()
            return None
    ms = timed(hip, stream, chain.run, reps, 1) / steps
    px = chain.luma_pixels_per_step
    step_bytes = sum(chain.bytes.values()) / steps
    # per-stage time: each stage's plans of every step run alone
    stage_ms = {}
    for k, name in enumerate(("mc", "recon", "deblock")):
        def run_stage(st, k=k):
            for plans in chain.plans:
                if k == 1:
                    plans[1].run(plans[3].ptr, 0, st)
                else:
                    plans[k].run(stream=st)
        stage_ms[name] = timed(hip, stream, run_stage, reps, 1) / steps
    chain.build_unbatched(rmh, frames)
    ums = timed(hip, stream, chain.run_unbatched, reps, 1) / steps
    n_tu = sum(len(f["tus"]) for fr in frames for f in fr)
    n_tool = sum(int(((f["tus"]["lfnst_idx"] > 0) | (f["tus"]["flags"] > 0)).sum()) for fr in frames for f in fr)
    achieved = step_bytes / (ms / 1e3) / 1e9
    out = {"workload": "16 streams (4x 856x480 8-bit, 8x 1920x1080 10-bit, 4x 3840x2160 10-bit, 4:2:0) x %d chained "
                       "frame steps (each step's output is the next one's reference; distinct PUs, TUs and edges per "
                       "step), MC -> residual add -> deblock, residual tool mix (MTS, LFNST, transform skip, dependent "
                       "quantisation on %.0f %% of %d TUs), one graph per frame step" % (steps, 100.0 * n_tool / n_tu,
                                                                                        n_tu),
           "value": round(px / (ms / 1e3) / 1e6, 1), "unit": "Mpix/s", "ms_per_frame_step": round(ms, 4),
           "luma_px_per_frame_step": px,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "algorithmic_bytes_per_frame_step": int(step_bytes)},
           "stages": {name: {"ms": round(t, 4), "algorithmic_bytes": int(chain.bytes[name] / steps),
                             "achieved_gbs": round(chain.bytes[name] / steps / (t / 1e3) / 1e9, 1),
                             "frac": round(chain.bytes[name] / steps / (t / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
                      for name, t in stage_ms.items()},
           "unbatched": {"value": round(px / (ums / 1e3) / 1e6, 1), "ms_per_frame_step": round(ums, 4),
                         "launches_per_frame_step": 4 * len(PL.STREAMS)}}
    out["vs_unbatched"] = round(out["value"] / out["unbatched"]["value"], 3)
    if chain.tuned:
        out["autotune"] = {"mc_waves_per_step": [t["mc_waves"] for t in chain.tuned],
                           "recon_split_step0": chain.tuned[0]["recon_split"]}
    log("streams: %.3f ms/frame step %.1f Mpix/s (%.1f GB/s), unbatched x%.2f" % (ms, out["value"], achieved,
                                                                                  out["vs_unbatched"]))
    chain.close()

    # built in scratch planes, deblocked out of place into the pictures
    ob = PL.ChainedStreams(acc, rmh, refs, frames, out_of_place=True)
    ob.run(stream.handle)
    stream.synchronize()
    got = ob.output(steps - 1)
    want = PL.oracle_chain(po, rmh.abi, refs[12], [frames[t][12] for t in range(steps)])
    if all((got[12][c] == want[c]).all() for c in range(3)):
        oms = timed(hip, stream, ob.run, reps, 1) / steps
        out["out_of_place"] = {"value": round(px / (oms / 1e3) / 1e6, 1), "unit": "Mpix/s",
                               "ms_per_frame_step": round(oms, 4),
                               "achieved_gbs": round(step_bytes / (oms / 1e3) / 1e9, 1)}
        log("streams out of place: %.3f ms/frame step" % oms)
    else:
        log("bench: out-of-place stream 12 differs from the oracle")
        out["out_of_place"] = None
    ob.close()
    return out


def bench_bi(rmh, acc, W, po, n_frames, steps, warmup, refine=False):
    """Bi-prediction with BCW weights over n_frames 1080p10 frames (two
    references each, one bi PU per CU): one launch of the bi plan. refine:
    8-bit frames with DMVR and BDOF on every PU (default BCW weights, so both
    apply wherever the PU size allows)."""
    hip, abi = rmh.hip, rmh.abi
    t0 = time.perf_counter()
    bd = 8 if refine else 10
    dt = np.uint8 if refine else np.uint16
    frames = [W.bi_frame(4000 + i, bit_depth=bd) for i in range(n_frames)]
    for d in frames:
        for k in ("ref0", "ref1"):
            d[k] = [np.ascontiguousarray(p.astype(dt)) for p in d[k]]
        if refine:
            d["pus"]["refine"] = W.BI_DMVR | W.BI_BDOF
            d["pus"]["bcw_idx"] = 2
    log("bi: generated %d frames in %.1f s" % (n_frames, time.perf_counter() - t0))
    bufs, structs = [], []
    for d in frames:
        planes = {}
        for key in ("ref0", "ref1", "dst"):
            ps = []
            for p in d["ref0" if key == "dst" else key]:
                b = hip.DeviceBuffer.from_array(np.zeros_like(p) if key == "dst" else p)
                bufs.append(b)
                ps.append(abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], p.itemsize, d["bit_depth"]))
            planes[key] = (abi.Plane * 3)(*ps)
        structs.append(abi.McBiFrame(planes["ref0"], planes["ref1"], planes["dst"]))
    pus = []
    for i, d in enumerate(frames):
        p = d["pus"].copy()
        p["frame"] = i
        pus.append(p)
    pus = np.concatenate(pus)
    plan = acc.mc_bi_plan(structs, pus)
    stream = hip.Stream()
    plan.run(stream=stream.handle)
    stream.synchronize()
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from frames import mc_bi_frame_struct
    for fi in (0, n_frames - 1):
        d = frames[fi]
        want = [np.zeros_like(p) for p in d["ref0"]]
        f, keep = mc_bi_frame_struct(d["ref0"], d["ref1"], want, d["bit_depth"])
        po.mc_bi_batch(f, d["pus"])
        for c in range(3):
            got = bufs[9 * fi + 6 + c].download(dt).reshape(want[c].shape)
            if not (got == want[c]).all():
                log("bench: bi-prediction frame %d plane %d differs from the oracle" % (fi, c))
                return None
    ms = timed(hip, stream, lambda s: plan.run(stream=s), steps, warmup)
    px = n_frames * 1920 * 1080
    if refine:
        work = ("%d frames 1920x1080 8-bit 4:2:0, one bi PU per CU (8x8..128x128), independent L0/L1 MVs +-64 px at "
                "any fraction, DMVR (every PU of at least 8x16) and BDOF on every PU, one plan run (DMVR launch, "
                "then the refined prediction)" % n_frames)
        kern = "dmvr_kernel + mc_bi_refined_kernel"
    else:
        work = ("%d frames 1920x1080 10-bit 4:2:0, one bi PU per CU (8x8..128x128), independent L0/L1 MVs "
                "+-64 px at any fraction, BCW weights, one launch" % n_frames)
        kern = "mc_bi_kernel"
    out = {"workload": work, "value": round(px / (ms / 1e3) / 1e6, 1), "unit": "Mpix/s", "ms_per_step": round(ms, 4),
           "roofline": dict(stage_roofline(plan, ms), kernel=kern)}
    log("bi%s: %.4f ms/step %.1f Mpix/s (%.1f GB/s)" % (" refined" if refine else "", ms, out["value"],
                                                        out["roofline"]["achieved_gbs"]))
    plan.close()
    for b in bufs:
        b.close()
    return out


RESIDUAL_CLASSES = [(4, 4), (8, 8), (16, 8), (8, 16), (16, 16), (32, 16), (32, 32), (64, 64)]


def residual_classes(steps, warmup):
    """The tiled residual launch on TU batches of one size class each (about
    26 M samples, the workload's coefficient statistics, default tiling):
    us per launch by HIP events. Shows which classes cost the most per sample."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import workloads as W
    rmh = load_package()
    hip = rmh.hip
    acc = rmh.Accelerator(0)
    out = {}
    for (w, h) in RESIDUAL_CLASSES:
        n = (25 << 20) // (w * h)
        tus, coeff = W._tu_batch(np.random.default_rng([7, w, h]), [(0, 0, 0, w, h)] * n, 10, (22, 37))
        plan = acc.residual_plan(tus, coeff.size)
        dc, dr = hip.DeviceBuffer.from_array(coeff), hip.DeviceBuffer(coeff.nbytes)
        stream = hip.Stream()
        ms = timed(hip, stream, lambda s: plan.run(dc.ptr, dr.ptr, s), steps, warmup)
        out["%dx%d" % (w, h)] = {"tus": n, "samples": int(coeff.size), "us": round(ms * 1e3, 1)}
        log("%dx%d: %.1f us" % (w, h, ms * 1e3))
        plan.close()
        dc.close()
        dr.close()
    print(json.dumps({"residual_classes": out}), flush=True)
    acc.close()


def bench_sao(rmh, acc, W, po, n_frames, steps, warmup):
    """SAO (band and edge offsets per CTB, random per component) over
    n_frames deblocked 1080p10 pictures: one launch of the SAO plan, output to
    separate planes."""
    hip, abi = rmh.hip, rmh.abi
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from test_oracle_sao import random_sao_params
    rng = np.random.default_rng(77)
    w, h, bd, ctb = 1920, 1080, 10, 64
    rows, cols = -(-h // ctb), -(-w // ctb)
    pics, prms, bufs, structs, outs = [], [], [], [], []
    for i in range(n_frames):
        planes = [np.ascontiguousarray(W.textured_plane(rng, pw, ph, bd).astype(np.uint16))
                  for pw, ph in ((w, h), (w // 2, h // 2), (w // 2, h // 2))]
        prm = np.ascontiguousarray(random_sao_params(rng, rows, cols, bd))
        srcs, dsts, fouts = [], [], []
        for p in planes:
            b, o = hip.DeviceBuffer.from_array(p), hip.DeviceBuffer(p.nbytes)
            bufs += [b, o]
            srcs.append(abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], 2, bd))
            dsts.append(abi.plane_from_address(o.ptr, p.shape[1], p.shape[0], p.shape[1], 2, bd))
            fouts.append(o)
        dp = hip.DeviceBuffer.from_array(prm)
        bufs.append(dp)
        structs.append(abi.sao_frame(srcs, dsts, dp.ptr, ctb, cols, rows, w, h))
        pics.append(planes)
        prms.append(prm)
        outs.append(fouts)
    plan = acc.sao_plan(structs)
    stream = hip.Stream()
    plan.run(stream=stream.handle)
    stream.synchronize()
    for fi in (0, n_frames - 1):
        want = po.sao_frame(pics[fi], bd, prms[fi], ctb, cols, rows, w, h)
        for c in range(3):
            got = outs[fi][c].download(np.uint16).reshape(want[c].shape)
            if not (got == want[c]).all():
                log("bench: SAO frame %d plane %d differs from the oracle" % (fi, c))
                return None
    ms = timed(hip, stream, lambda s: plan.run(stream=s), steps, warmup)
    px = n_frames * w * h
    out = {"workload": "%d deblocked 1920x1080 10-bit 4:2:0 pictures, CTB 64, per-CTB band or edge offset (random "
                       "class, band and offsets per component), output to separate planes, one launch" % n_frames,
           "value": round(px / (ms / 1e3) / 1e6, 1), "unit": "Mpix/s", "ms_per_step": round(ms, 4),
           "roofline": dict(stage_roofline(plan, ms), kernel="sao_ctb_kernel<uint16_t>")}
    log("sao: %.4f ms/step %.1f Mpix/s (%.1f GB/s)" % (ms, out["value"], out["roofline"]["achieved_gbs"]))
    plan.close()
    for b in bufs:
        b.close()
    return out


def bench_tool_mix(rmh, acc, W, po, n_frames, steps, warmup):
    """The residual stage with LFNST, transform skip and dependent
    quantisation in the mix: TUs with those tools run on the generic
    one-TU-per-workgroup kernel, the rest on the tiles."""
    hip = rmh.hip
    frames = [W.residual_frame(5000 + i, with_lfnst=True, with_ts=True, with_dq=True)[:2] for i in range(n_frames)]
    tus, coeff = W.concat_frames(frames)
    plan = acc.residual_plan(tus, coeff.size)
    d_coeff = hip.DeviceBuffer.from_array(coeff)
    d_res = hip.DeviceBuffer(coeff.nbytes)
    stream = hip.Stream()
    plan.run(d_coeff.ptr, d_res.ptr, stream.handle)
    stream.synchronize()
    got = d_res.download(np.int16)
    t, c = frames[0]
    if not (got[:c.size] == po.residual_batch(t, c)).all():
        log("bench: tool-mix residuals differ from the oracle on frame 0")
        return None
    ms = timed(hip, stream, lambda s: plan.run(d_coeff.ptr, d_res.ptr, s), steps, warmup)
    generic = (tus["lfnst_idx"] > 0) | ((tus["flags"] & (W.TS_FLAG | W.DQ_FLAG)) != 0)
    px = n_frames * 1920 * 1080
    out = {"workload": "%d frames of 1920x1080 10-bit TUs with LFNST (30 %%), transform skip (10 %%) and dependent "
                       "quantisation (20 %%) drawn per TU, one plan run" % n_frames,
           "value": round(px / (ms / 1e3) / 1e6, 1), "unit": "Mpix/s", "ms_per_step": round(ms, 4),
           "tus": int(len(tus)), "tus_on_generic_kernel": int(generic.sum()),
           "roofline": dict(stage_roofline(plan, ms), kernel="residual_tile_kernel<false> + residual_kernel")}
    log("tool mix: %.4f ms/step %.1f Mpix/s, %d of %d TUs on the generic kernel" % (
        ms, out["value"], out["tus_on_generic_kernel"], out["tus"]))
    plan.close()
    d_coeff.close()
    d_res.close()
    return out


CNN_MACS_PER_BLOCK = 32 * 32 * 32 * 9 + 2 * 32 * 32 * 32 * 288 + 33 * 64 + 64 * 6


def cnn_kernel_rates(rmh, net, steps):
    """Forward pass and training step of the CU-split network on one batch of
    256 blocks of synthetic content, by HIP events: (forward ms, step ms,
    forward FLOPs). The network's parameters are restored afterwards."""
    hip = rmh.hip
    L = rmh.learned
    n = 256
    frames = L.synthetic_sequence(2, 1)
    blocks = L.node_blocks(L.pad_to_nodes(frames[0]))[:n]
    db = hip.DeviceBuffer.from_array(np.ascontiguousarray(blocks))
    dq = hip.DeviceBuffer.from_array(np.full(n, L.qp_feature(32), np.float32))
    dl = hip.DeviceBuffer.from_array(np.zeros(n, np.uint8))
    dp = hip.DeviceBuffer(n * 24)
    dloss = hip.DeviceBuffer(n * 4)
    stream = hip.Stream()
    fwd_ms = timed(hip, stream, lambda s: net.forward(db.ptr, dq.ptr, n, dp.ptr, s), steps, 3)
    saved = net.params()
    step_ms = timed(hip, stream, lambda s: net.train_step(db.ptr, dq.ptr, dl.ptr, n, 1e-6, dloss.ptr, s), steps, 3)
    net.set_params(saved)
    for b in (db, dq, dl, dp, dloss):
        b.close()
    return fwd_ms, step_ms, 2.0 * CNN_MACS_PER_BLOCK * n


def bench_learned(rmh, acc, steps):
    """The learned CU-split predictor: training on full-RD pairs, calibration,
    full / pruned / heuristic encodes of held-out content, and the network's
    forward and training-step rates on MFMA."""
    hip = rmh.hip
    L = rmh.learned
    net, rep = L.run_experiment(acc, log=log)
    t = rep["test"]
    log("learned: pruned %.2f fps vs full %.2f fps (x%.2f), BD-rate %.3f %%, heuristic BD-rate %.3f %%" % (
        t["pruned"]["fps"], t["full"]["fps"], t["speedup_pruned_vs_full"], t["pruned"]["bd_rate_vs_full"],
        t["heuristic"]["bd_rate_vs_full"]))
    fwd_ms, step_ms, flops = cnn_kernel_rates(rmh, net, steps)
    tflops = flops / (fwd_ms / 1e3) / 1e12
    net.close()
    def strip(d):
        return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items() if k != "splits"}
    out = {"workload": "CU-split CNN (3 conv3x3 x 32ch + pool + MLP 64 -> 6, BF16 on MFMA, FP32 accumulate) trained "
                       "on full-RD pairs of synthetic 1280x720 8-bit content at QP 22/27/32/37; encodes of held-out "
                       "sequences",
           "encode_fps_full": round(t["full"]["fps"], 3), "encode_fps_pruned": round(t["pruned"]["fps"], 3),
           "encode_fps_heuristic": round(t["heuristic"]["fps"], 3),
           "speedup_pruned_vs_full": round(t["speedup_pruned_vs_full"], 3),
           "bd_rate_pruned_vs_full_pct": round(t["pruned"]["bd_rate_vs_full"], 3),
           "bd_rate_heuristic_vs_full_pct": round(t["heuristic"]["bd_rate_vs_full"], 3),
           "pruned": strip(t["pruned"]), "heuristic": strip(t["heuristic"]),
           "untrained_same_scale": {k: round(v, 3) for k, v in rep["untrained"].items()},
           "training": {k: rep[k] for k in ("train_pairs", "collect_seconds", "train_seconds", "epochs", "loss_first",
                                            "loss_last", "logit_scale", "validation")},
           "rd_points": t["points"],
           "forward_ms_256": round(fwd_ms, 4), "train_step_ms_256": round(step_ms, 4),
           "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                        "frac": round(tflops / 2500.0, 5), "traffic": None,
                        "kernel": "forward pass (conv_fwd_kernel x3 + pool + MLP), 256 blocks",
                        "algorithmic_flops_per_launch": flops}}
    log("learned: forward %.4f ms / 256 blocks = %.1f TFLOP/s; train step %.4f ms" % (fwd_ms, tflops, step_ms))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-unbatched", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--pipeline-frames", type=int, default=8)
    ap.add_argument("--no-pipeline", action="store_true")
    ap.add_argument("--no-streams", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-bi", action="store_true")
    ap.add_argument("--cnn-kernels", action="store_true")
    ap.add_argument("--no-tools", action="store_true")
    ap.add_argument("--no-sao", action="store_true")
    ap.add_argument("--residual-classes", action="store_true")
    args = ap.parse_args()

    if args.residual_classes:
        residual_classes(args.steps, args.warmup)
        return 0

    if args.copy_calibration:
        copy_calibration()
        return 0
    if args.cnn_kernels:
        # the CU-split network's kernels alone (for kernel traces): forward and
        # training step on one 256-block batch, untrained weights
        rmh = load_package()
        acc = rmh.Accelerator(0)
        net = rmh.Cnn(acc, max_batch=256, seed=1)
        fwd_ms, step_ms, flops = cnn_kernel_rates(rmh, net, args.steps)
        print(json.dumps({"forward_ms_256": round(fwd_ms, 4), "train_step_ms_256": round(step_ms, 4),
                          "forward_tflops": round(flops / (fwd_ms / 1e3) / 1e12, 2)}), flush=True)
        net.close()
        acc.close()
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the residual stage shards by "
              "stream with no exchange). Run with --gpus 1.", file=sys.stderr)
        return 2

    sys.path.insert(0, os.path.join(REPO, "tests"))
    import workloads as W

    rmh = load_package()
    hip = rmh.hip
    t0 = time.perf_counter()
    frames = []
    for f in range(args.frames):
        tus, coeff, _ = W.residual_frame(1000 + f)
        frames.append((tus, coeff))
    tus, coeff = W.concat_frames(frames)
    ranges, first = [], 0
    for t, _ in frames:
        ranges.append((first, len(t)))
        first += len(t)
    log("generated %d frames (%d TUs, %d samples) in %.1f s" % (args.frames, len(tus), coeff.size,
                                                                 time.perf_counter() - t0))

    acc = rmh.Accelerator(0)
    info = acc.info()
    plan = acc.residual_plan(tus, coeff.size)
    d_coeff = hip.DeviceBuffer.from_array(coeff)
    d_res = hip.DeviceBuffer(coeff.nbytes)
    stream = hip.Stream()
    s = stream.handle

    # correctness gate: the first and last frame of the batch against the oracle
    plan.run(d_coeff.ptr, d_res.ptr, s)
    stream.synchronize()
    got = d_res.download(np.int16)
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as po
    for fi in (0, args.frames - 1):
        t, c = frames[fi]
        base = int(tus[ranges[fi][0]]["offset"])
        want = po.residual_batch(t, c)
        if not (got[base:base + c.size] == want).all():
            log("bench: GPU residuals differ from the oracle on frame %d" % fi)
            return 1
    log("bit-exact check passed (frames 0 and %d)" % (args.frames - 1))

    tuned = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuned = plan.autotune(d_coeff.ptr, d_res.ptr, s)
        stream.synchronize()
        plan.run(d_coeff.ptr, d_res.ptr, s)
        stream.synchronize()
        got = d_res.download(np.int16)
        for fi in (0, args.frames - 1):
            t, c = frames[fi]
            base = int(tus[ranges[fi][0]]["offset"])
            if not (got[base:base + c.size] == po.residual_batch(t, c)).all():
                log("bench: autotuned residuals differ from the oracle on frame %d" % fi)
                return 1
        log("autotuned tile splits in %.2f s: %s" % (time.perf_counter() - t0,
                                                      {"%dx%d" % k: v for k, v in tuned.items() if v > 1}))

    ev0, ev1 = hip.Event(), hip.Event()
    for _ in range(args.warmup):
        plan.run(d_coeff.ptr, d_res.ptr, s)
    hip.device_synchronize()
    ev0.record(s)
    for _ in range(args.steps):
        plan.run(d_coeff.ptr, d_res.ptr, s)
    ev1.record(s)
    ev1.synchronize()
    hip.device_synchronize()
    ms = ev0.elapsed_ms(ev1) / args.steps
    luma_px = args.frames * 1920 * 1080
    value = luma_px / (ms / 1e3) / 1e6
    log("batched: %.4f ms/step, %.1f Mpix/s" % (ms, value))

    unb = None
    if not args.no_unbatched:
        # the same kernels, one plan (one launch) per frame over that frame's buffers
        fplans, base = [], 0
        for t, c in frames:
            fplans.append((acc.residual_plan(t, c.size), base))
            base += c.nbytes
        for _ in range(max(1, args.warmup)):
            for fp, b in fplans:
                fp.run(d_coeff.ptr + b, d_res.ptr + b, s)
        hip.device_synchronize()
        ev0.record(s)
        for _ in range(args.steps):
            for fp, b in fplans:
                fp.run(d_coeff.ptr + b, d_res.ptr + b, s)
        ev1.record(s)
        ev1.synchronize()
        hip.device_synchronize()
        ums = ev0.elapsed_ms(ev1) / args.steps
        unb = {"value": round(luma_px / (ums / 1e3) / 1e6, 1), "ms_per_step": round(ums, 4),
               "launches_per_step": len(ranges)}
        log("unbatched: %.4f ms/step, %.1f Mpix/s" % (ums, unb["value"]))

    achieved = plan.bytes / (ms / 1e3) / 1e9
    traffic, t_frames = read_traffic()
    if t_frames is None or int(t_frames) != args.frames:
        traffic = None  # the committed counter run covers another batch size
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "residual_tile_kernel<false>", "algorithmic_bytes_per_launch": plan.bytes}

    sys.path.insert(0, os.path.join(REPO, "tests"))
    import pipeline as PL
    pipe = None if args.no_pipeline else bench_pipeline(rmh, acc, W, PL, po, args.pipeline_frames, args.steps,
                                                         args.warmup, cpu=not args.no_cpu,
                                                         autotune=not args.no_autotune)
    streams = None if args.no_streams else bench_streams(rmh, acc, PL, po, 8, max(2, args.steps // 5),
                                                         autotune=not args.no_autotune)

    bi = None if args.no_bi else bench_bi(rmh, acc, W, po, args.pipeline_frames, args.steps, args.warmup)
    if bi is not None:
        bi["refined"] = bench_bi(rmh, acc, W, po, args.pipeline_frames, args.steps, args.warmup, refine=True)
    tools = None if args.no_tools else bench_tool_mix(rmh, acc, W, po, 8, args.steps, args.warmup)
    sao = None if args.no_sao else bench_sao(rmh, acc, W, po, args.pipeline_frames, args.steps, args.warmup)
    learned = None if args.no_learned else bench_learned(rmh, acc, args.steps)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(tus, coeff, luma_px, args.frames)
        cpu["cores_on_host"] = os.cpu_count()
        cpu["placement"] = placement_report(rmh, acc)
        log("cpu oracle: %s Mpix/s on %d threads, %s on one" % (cpu["value"], cpu["cores"],
                                                              cpu["single_thread"]["value"]))

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int16", "data": "synthetic",
        "config": {"workload": "batched dequant + inverse DCT-II/DST-VII/DCT-VIII over every TU of %d "
                               "synthetic 1920x1080 10-bit 4:2:0 frames per step" % args.frames,
                   "frames_per_step": args.frames, "tus_per_step": int(len(tus)), "samples_per_step": int(coeff.size),
                   "parallelism": "single GPU"},
        "unbatched": unb,
        "vs_unbatched": round(value / unb["value"], 3) if unb else None,
        "autotune": None if tuned is None else {"tus_per_tile_divisor": {"%dx%d" % k: v for k, v in tuned.items()
                                                                          if v > 1}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pframe_pipeline": pipe,
        "multi_stream": streams,
        "bi_prediction": bi,
        "residual_tool_mix": tools,
        "sao": sao,
        "learned": learned,
        "device": info,
    }
    print(json.dumps(out), flush=True)
    plan.close()
    acc.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

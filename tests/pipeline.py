"""P-frame reconstruction pipeline (MC -> residual add -> deblock, optionally
-> SAO) over many frames, for the parity tests and the benchmark.

A batch is a list of P-frames (dicts from workloads.pframe), possibly of
different sizes and bit depths, as the frames of several concurrent streams.
`DeviceBatch` holds their planes in device memory and the three batched plans
(one MC plan, one recon plan, one deblock plan over every frame); `oracle`
runs the CPU oracle over host copies of the same frames. Both produce the
reconstructed, deblocked [Y, Cb, Cr] of every frame.
"""
import ctypes

import numpy as np

import workloads as W


def sample_dtype(bit_depth):
    return np.uint8 if bit_depth == 8 else np.uint16


def ref_planes(f):
    dt = sample_dtype(f["bit_depth"])
    return [np.ascontiguousarray(p.astype(dt)) for p in f["ref"]]


def concat_tus(frames):
    """Every frame's TUs in one descriptor list; placements point at frame i."""
    tus, coeff = W.concat_frames([(f["tus"], f["coeff"]) for f in frames])
    place = np.concatenate([f["place"] for f in frames]).copy()
    place["frame"] = np.concatenate([np.full(len(f["place"]), i, np.uint16) for i, f in enumerate(frames)])
    pus = np.concatenate([f["pus"] for f in frames]).copy()
    pus["frame"] = np.concatenate([np.full(len(f["pus"]), i, np.uint16) for i, f in enumerate(frames)])
    return tus, coeff, place, pus


SAO_CTB = 64


def frame_sao(f, i):
    """Seeded per-CTB SAO parameters of frame i (CTB 64): (params, cols, rows)."""
    from test_oracle_sao import random_sao_params
    rows, cols = -(-int(f["height"]) // SAO_CTB), -(-int(f["width"]) // SAO_CTB)
    rng = np.random.default_rng([9, i, int(f["width"])])
    return np.ascontiguousarray(random_sao_params(rng, rows, cols, int(f["bit_depth"]))), cols, rows


def oracle(po, abi, frames, with_sao=False):
    """The CPU oracle: orc_mc_batch into fresh planes, orc_recon_batch, then
    orc_deblock_frames (and orc_sao_frame). Returns [[Y, Cb, Cr]] per frame."""
    tus, coeff, place, pus = concat_tus(frames)
    refs = [ref_planes(f) for f in frames]
    outs = [[np.zeros_like(p) for p in r] for r in refs]
    mcf = (abi.McFrame * len(frames))(*[abi.mc_frame([abi.plane_from_array(p, f["bit_depth"]) for p in r],
                                                     [abi.plane_from_array(p, f["bit_depth"]) for p in o])
                                        for f, r, o in zip(frames, refs, outs)])
    assert po.lib().orc_mc_batch(ctypes.addressof(mcf), len(frames), pus.ctypes.data, len(pus)) == 0
    rf = (abi.ReconFrame * len(frames))(*[abi.ReconFrame((abi.Plane * 3)(*[abi.plane_from_array(p, f["bit_depth"])
                                                                           for p in o]))
                                          for f, o in zip(frames, outs)])
    bad = po.lib().orc_recon_batch(tus.ctypes.data, len(tus), coeff.ctypes.data_as(po.I16P), coeff.size,
                                   place.ctypes.data, ctypes.addressof(rf), len(frames))
    assert bad == 0, bad
    edges = [np.ascontiguousarray(f["edges"]) for f in frames]
    dbf = (abi.DbkFrame * len(frames))(*[abi.dbk_frame([abi.plane_from_array(p, f["bit_depth"]) for p in o],
                                                       e.ctypes.data, e.shape[1], e.shape[0], f["qp"])
                                         for f, o, e in zip(frames, outs, edges)])
    assert po.lib().orc_deblock_frames(ctypes.addressof(dbf), len(frames)) == 0
    if with_sao:
        for i, f in enumerate(frames):
            prm, cols, rows = frame_sao(f, i)
            outs[i] = po.sao_frame(outs[i], f["bit_depth"], prm, SAO_CTB, cols, rows, int(f["width"]),
                                   int(f["height"]))
    return outs


class DeviceBatch:
    """Device planes and batched plans for a list of P-frames."""

    def __init__(self, acc, rmh, frames, with_sao=False, out_of_place=False):
        """with_sao: deblocking writes to scratch planes (out of place) and SAO
        writes the final samples back into the frame planes. out_of_place
        (without SAO): MC and the residual add build the picture in scratch
        planes and deblocking writes it out of place into the frame planes,
        so no shadow copy of the pre-deblocking picture is made."""
        if with_sao and out_of_place:
            raise ValueError("with_sao already deblocks out of place")
        self.hip, abi = rmh.hip, rmh.abi
        hip = self.hip
        self.frames = frames
        self.tus, self.coeff, self.place, self.pus = concat_tus(frames)
        self.keep = []
        self.dst = []  # per frame: [(DeviceBuffer, shape, dtype)] x 3
        mc_frames, rc_frames, db_frames, tmp_planes, sao_frames = [], [], [], [], []
        for i, f in enumerate(frames):
            bd = f["bit_depth"]
            refs, dsts, dbufs = [], [], []
            for p in ref_planes(f):
                rb = hip.DeviceBuffer.from_array(p)
                db = hip.DeviceBuffer(p.nbytes)
                db.zero()
                self.keep += [rb, db]
                refs.append(abi.plane_from_address(rb.ptr, p.shape[1], p.shape[0], p.shape[1], p.itemsize, bd))
                dsts.append(abi.plane_from_address(db.ptr, p.shape[1], p.shape[0], p.shape[1], p.itemsize, bd))
                dbufs.append((db, p.shape, p.dtype))
            self.dst.append(dbufs)
            e = np.ascontiguousarray(f["edges"])
            eb = hip.DeviceBuffer.from_array(e)
            self.keep.append(eb)
            work = dsts
            if out_of_place:
                work = []
                for (db, shape, dt), d in zip(dbufs, dsts):
                    wb = hip.DeviceBuffer(int(np.prod(shape)) * np.dtype(dt).itemsize)
                    self.keep.append(wb)
                    work.append(abi.plane_from_address(wb.ptr, d.width, d.height, d.stride, d.bytes_per_sample, bd))
                tmp_planes += dsts
            mc_frames.append(abi.mc_frame(refs, work))
            rc_frames.append(abi.ReconFrame((abi.Plane * 3)(*work)))
            db_frames.append(abi.dbk_frame(work, eb.ptr, e.shape[1], e.shape[0], f["qp"]))
            if with_sao:
                tmps = []
                for (db, shape, dt), d in zip(dbufs, dsts):
                    tb = hip.DeviceBuffer(int(np.prod(shape)) * np.dtype(dt).itemsize)
                    self.keep.append(tb)
                    tmps.append(abi.plane_from_address(tb.ptr, d.width, d.height, d.stride, d.bytes_per_sample, bd))
                tmp_planes += tmps
                prm, cols, rows = frame_sao(f, i)
                pb = hip.DeviceBuffer.from_array(prm)
                self.keep.append(pb)
                sao_frames.append(abi.sao_frame(tmps, dsts, pb.ptr, SAO_CTB, cols, rows, int(f["width"]),
                                                int(f["height"])))
        self.d_coeff = hip.DeviceBuffer.from_array(self.coeff)
        hip.device_synchronize()
        self.mc = acc.mc_plan(mc_frames, self.pus)
        self.recon = acc.recon_plan(self.tus, self.coeff.size, self.place, rc_frames)
        if with_sao:
            self.deblock = acc.deblock_to_plan(db_frames, tmp_planes)
            self.sao = acc.sao_plan(sao_frames)
        elif out_of_place:
            self.deblock = acc.deblock_to_plan(db_frames, tmp_planes)
            self.sao = None
        else:
            self.deblock = acc.deblock_plan(db_frames)
            self.sao = None
        self.acc = acc

    def run(self, stream):
        """One frame step: the batched plans on one stream."""
        self.mc.run(stream=stream)
        self.recon.run(self.d_coeff.ptr, 0, stream)
        self.deblock.run(stream=stream)
        if self.sao:
            self.sao.run(stream=stream)

    def graph(self):
        nodes = [(self.mc, 0, 0), (self.recon, self.d_coeff.ptr, 0), (self.deblock, 0, 0)]
        if self.sao:
            nodes.append((self.sao, 0, 0))
        return self.acc.graph(nodes)

    def download(self):
        return [[b.download(dt).reshape(shape) for b, shape, dt in fr] for fr in self.dst]

    def luma_pixels(self):
        return sum(int(f["width"]) * int(f["height"]) for f in self.frames)

    def close(self):
        for p in (self.mc, self.recon, self.deblock, self.sao):
            if p:
                p.close()


# ---------------------------------------------------------------- streams

STREAMS = [(856, 480, 8)] * 4 + [(1920, 1080, 10)] * 8 + [(3840, 2160, 10)] * 4


def generate_streams(streams, steps, tools=True, seed=0, workers=8):
    """Host data of chained streams: (refs, frames) with refs[s] the first
    reference picture of stream s and frames[t][s] the P-frame of step t
    (workloads.pframe_step; distinct PUs, TUs and edges at every step).
    Generated on a thread pool (numpy releases the GIL for the bulk work)."""
    import concurrent.futures
    jobs = [(t, s) for t in range(steps) for s in range(len(streams))]

    def one(job):
        t, s = job
        w, h, bd = streams[s]
        return W.pframe_step(seed + 1000 * s + t, width=w, height=h, bit_depth=bd, tools=tools)

    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        refs = list(ex.map(lambda s: W.stream_reference(seed + s, *streams[s]), range(len(streams))))
        flat = list(ex.map(one, jobs))
    frames = [[flat[t * len(streams) + s] for s in range(len(streams))] for t in range(steps)]
    for s, (w, h, bd) in enumerate(streams):
        dt = sample_dtype(bd)
        refs[s] = [np.ascontiguousarray(p.astype(dt)) for p in refs[s]]
    return refs, frames


def oracle_chain(po, abi, ref, steps_frames, on_step=None):
    """The CPU oracle over one stream's chained steps: step t predicts from
    step t-1's output (step 0 from ref). Returns the last step's planes;
    on_step(t, planes) sees every step's output."""
    cur = ref
    for t, f in enumerate(steps_frames):
        g = dict(f)
        g["ref"] = cur
        cur = oracle(po, abi, [g])[0]
        if on_step:
            on_step(t, cur)
    return cur


class ChainedStreams:
    """Chained frame steps of several streams on the device: at step t every
    stream predicts from the picture its step t-1 produced (step 0 from its
    first reference picture), adds step t's residuals and deblocks. Two
    device pictures per stream alternate as reference and output; each step
    is one MC plan, one recon plan and one deblock plan over every stream,
    captured as one graph. With out_of_place, MC and the residual add build
    the picture in a scratch set and deblocking writes it out of place into
    the output picture."""

    def __init__(self, acc, rmh, refs, frames, out_of_place=False, autotune=False):
        """autotune: each step's MC plan picks its work items per workgroup and
        its recon plan its tile splits (rmh_mc_plan_autotune,
        rmh_residual_plan_autotune) before its graph is captured; the choices
        are in self.tuned."""
        hip, abi = rmh.hip, rmh.abi
        self.tuned = []
        self.hip, self.acc = hip, acc
        self.n_streams, self.n_steps = len(refs), len(frames)
        self.keep, self.plans, self.graphs = [], [], []
        self.shapes = [[p.shape for p in r] for r in refs]
        self.dtypes = [r[0].dtype for r in refs]
        self.bds = [frames[0][s]["bit_depth"] for s in range(self.n_streams)]
        # pics[k][s] = (buffers, planes) of picture set k of stream s
        self.pics = [[], []]
        self.scratch = []
        for s, r in enumerate(refs):
            for k in range(2):
                bufs = [hip.DeviceBuffer.from_array(p) if k == 0 else hip.DeviceBuffer(p.nbytes) for p in r]
                self.keep += bufs
                self.pics[k].append((bufs, [abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1],
                                                                    p.itemsize, self.bds[s])
                                            for b, p in zip(bufs, r)]))
            if out_of_place:
                bufs = [hip.DeviceBuffer(p.nbytes) for p in r]
                self.keep += bufs
                self.scratch.append([abi.plane_from_address(b.ptr, p.shape[1], p.shape[0], p.shape[1], p.itemsize,
                                                            self.bds[s]) for b, p in zip(bufs, r)])
        self.bytes = {"mc": 0, "recon": 0, "deblock": 0}
        for t, fr in enumerate(frames):
            src, dst = self.pics[t % 2], self.pics[(t + 1) % 2]
            tus, coeff, place, pus = concat_tus(fr)
            dc = hip.DeviceBuffer.from_array(coeff)
            edges = [hip.DeviceBuffer.from_array(np.ascontiguousarray(f["edges"])) for f in fr]
            self.keep += [dc] + edges
            work = [self.scratch[s] if out_of_place else dst[s][1] for s in range(self.n_streams)]
            mc = acc.mc_plan([abi.mc_frame(src[s][1], work[s]) for s in range(self.n_streams)], pus)
            rc = acc.recon_plan(tus, coeff.size, place, [abi.ReconFrame((abi.Plane * 3)(*work[s]))
                                                         for s in range(self.n_streams)])
            dbf = [abi.dbk_frame(work[s], edges[s].ptr, f["edges"].shape[1], f["edges"].shape[0], f["qp"])
                   for s, f in enumerate(fr)]
            if out_of_place:
                db = acc.deblock_to_plan(dbf, [p for s in range(self.n_streams) for p in dst[s][1]])
            else:
                db = acc.deblock_plan(dbf)
            self.plans.append((mc, rc, db, dc))
            self.bytes["mc"] += mc.bytes
            self.bytes["recon"] += rc.bytes
            self.bytes["deblock"] += db.bytes
            if autotune:
                self.tuned.append({"mc_waves": mc.autotune_mc(),
                                   "recon_split": {"%dx%d" % k: v for k, v in rc.autotune(dc.ptr, 0).items() if v > 1}})
            self.graphs.append(acc.graph([(mc, 0, 0), (rc, dc.ptr, 0), (db, 0, 0)]))
        self.luma_pixels_per_step = sum(int(f["width"]) * int(f["height"]) for f in frames[0])
        if autotune:
            # the tuning runs wrote into the picture sets, the first reference included
            hip.device_synchronize()
            self.reset(refs)

    def run(self, stream, steps=None):
        """Launches the frame-step graphs 0 .. steps-1 in order on stream."""
        for g in self.graphs[:self.n_steps if steps is None else steps]:
            g.launch(stream)

    def run_unbatched(self, stream):
        """The same steps with one launch per stage per stream: plans over a
        single frame, built on first use."""
        if not hasattr(self, "_single"):
            raise RuntimeError("call build_unbatched first")
        for step in self._single:
            for mc, rc, db, dc in step:
                mc.run(stream=stream)
                rc.run(dc.ptr, 0, stream)
                db.run(stream=stream)

    def build_unbatched(self, rmh, frames):
        """Per-stream plans of every step over the same device pictures."""
        abi, hip = rmh.abi, self.hip
        self._single = []
        for t, fr in enumerate(frames):
            src, dst = self.pics[t % 2], self.pics[(t + 1) % 2]
            row = []
            for s, f in enumerate(fr):
                pus = f["pus"].copy()
                pus["frame"] = 0
                place = f["place"].copy()
                place["frame"] = 0
                dc = hip.DeviceBuffer.from_array(f["coeff"])
                e = hip.DeviceBuffer.from_array(np.ascontiguousarray(f["edges"]))
                self.keep += [dc, e]
                mc = self.acc.mc_plan([abi.mc_frame(src[s][1], dst[s][1])], pus)
                rc = self.acc.recon_plan(f["tus"], f["coeff"].size, place, [abi.ReconFrame((abi.Plane * 3)(*dst[s][1]))])
                db = self.acc.deblock_plan([abi.dbk_frame(dst[s][1], e.ptr, f["edges"].shape[1], f["edges"].shape[0],
                                                          f["qp"])])
                row.append((mc, rc, db, dc))
            self._single.append(row)

    def output(self, t):
        """[[Y, Cb, Cr]] per stream after step t (valid until step t + 2 runs)."""
        k = (t + 1) % 2
        return [[b.download(self.dtypes[s]).reshape(self.shapes[s][c]) for c, b in enumerate(self.pics[k][s][0])]
                for s in range(self.n_streams)]

    def reset(self, refs):
        """Restores every stream's first reference picture (set 0)."""
        for s, r in enumerate(refs):
            for b, p in zip(self.pics[0][s][0], r):
                b.upload(p)

    def close(self):
        for g in self.graphs:
            g.close()
        for mc, rc, db, _ in self.plans:
            mc.close()
            rc.close()
            db.close()
        for row in getattr(self, "_single", []):
            for mc, rc, db, _ in row:
                mc.close()
                rc.close()
                db.close()
        for b in self.keep:
            b.close()

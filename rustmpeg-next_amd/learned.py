"""Learned CU-split pruning: data, training, pruned encode and BD-rate.

The pieces, all on librmh:
  - synthetic 8-bit luma content: seeded mixes of gradients, band-limited
    noise and moving rectangles;
  - (block, best split) pairs from the full RD search (RMH_RD_FULL) at
    QP 22 / 27 / 32 / 37, its rate the bits of the CABAC substream it writes
    (rmh_rd_encode_luma_stream; the static bit model of rmh_rd_encode_luma
    when RATE is "model");
  - training of the BF16 CNN on the GPU (rmh_cnn_train_step);
  - the encode with the network pruning each node's splits (p < p_prune
    dropped, full search when max p < p_fallback) against the full search and
    the reference heuristic (choose_qt_mtt_partition_rdo,
    crates/rustmpeg-video/src/vvc/slice.rs:4127-4167);
  - BD-rate (bd_rate, scripts/bench/bench.py:777-813).
"""
import concurrent.futures as cf
import math
import queue
import time

import numpy as np

from . import abi, hip as _hip

QPS = (22, 27, 32, 37)
NODE = 32
RATE = "cabac"  # "cabac": coded bits of the luma substream; "model": the static bit model


# ------------------------------------------------------------------ content

def _band_noise(rng, h, w, lo, hi):
    """White noise band-passed to spatial frequencies in [lo, hi) cycles/sample."""
    f = np.fft.rfft2(rng.normal(0, 1, (h, w)))
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    r = np.sqrt(fx * fx + fy * fy)
    f[(r < lo) | (r >= hi)] = 0
    n = np.fft.irfft2(f, s=(h, w))
    return n / (n.std() + 1e-9)


def synthetic_sequence(seed, n_frames, width=1280, height=720):
    """n_frames of 8-bit luma (height, width): a gradient, band-limited noise
    whose strength varies over the picture, and rectangles moving between
    frames. Deterministic in seed."""
    rng = np.random.default_rng([5, seed])
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    ang = rng.uniform(0, 2 * np.pi)
    grad = 128 + rng.uniform(20, 90) * ((np.cos(ang) * xx + np.sin(ang) * yy) / max(width, height) - 0.5)
    noise_fine = _band_noise(rng, height, width, 0.15, 0.5)
    noise_mid = _band_noise(rng, height, width, 0.02, 0.15)
    mask = _band_noise(rng, height, width, 0.0005, 0.01)
    mask = np.clip(0.5 + 0.6 * mask, 0, 1)
    base = grad + mask * (rng.uniform(4, 16) * noise_fine + rng.uniform(6, 24) * noise_mid)
    rects = []
    for _ in range(int(rng.integers(6, 14))):
        rw, rh = rng.integers(24, 260, 2)
        rects.append(dict(x=rng.uniform(0, width), y=rng.uniform(0, height), w=int(rw), h=int(rh),
                          vx=rng.uniform(-12, 12), vy=rng.uniform(-8, 8), v=rng.uniform(-80, 80),
                          tex=rng.uniform(0, 20)))
    frames = []
    for t in range(n_frames):
        img = base.copy()
        for r in rects:
            x0, y0 = int(r["x"] + r["vx"] * t) % width, int(r["y"] + r["vy"] * t) % height
            x1, y1 = min(width, x0 + r["w"]), min(height, y0 + r["h"])
            img[y0:y1, x0:x1] += r["v"] + r["tex"] * noise_fine[y0:y1, x0:x1]
        frames.append(np.clip(np.rint(img), 0, 255).astype(np.uint8))
    return frames


def pad_to_nodes(img):
    """Edge-replicates the picture to multiples of 32 (720 -> 736)."""
    h, w = img.shape
    H, W = -(-h // NODE) * NODE, -(-w // NODE) * NODE
    return np.pad(img, ((0, H - h), (0, W - w)), mode="edge")


def node_blocks(img):
    """(nodes, 32, 32) in raster order of a padded picture."""
    h, w = img.shape
    return np.ascontiguousarray(img.reshape(h // NODE, NODE, w // NODE, NODE).transpose(0, 2, 1, 3).reshape(-1, NODE, NODE))


def qp_feature(qp):
    return np.float32(qp / 51.0)


# ---------------------------------------------------------------- RD search

def _rd(args):
    from . import rd_encode_luma, rd_encode_luma_stream
    img, qp, mode, probs, p_prune, p_fallback = args[:6]
    want_costs = len(args) > 6 and args[6]
    t0 = time.perf_counter()
    if RATE == "cabac":
        rec, splits, costs, st, _ = rd_encode_luma_stream(img, qp, mode, probs, p_prune, p_fallback, want_costs)
    else:
        rec, splits, costs, st = rd_encode_luma(img, qp, mode, probs, p_prune, p_fallback, want_costs)
    st["seconds"] = time.perf_counter() - t0
    return rec, splits, st, costs


def collect_pairs(frames, qps=QPS, workers=8, with_costs=False):
    """Full RD search of every frame at every QP: (blocks, qp features, labels)
    and, with_costs, the RD cost of every split of every node (nodes, 6)."""
    jobs = [(pad_to_nodes(f), qp, abi.RD_FULL, None, 0.05, 0.6, with_costs) for f in frames for qp in qps]
    with cf.ThreadPoolExecutor(workers) as ex:
        res = list(ex.map(_rd, jobs))
    blocks, qf, labels, costs = [], [], [], []
    for (img, qp, *_), (_, splits, _, c) in zip(jobs, res):
        b = node_blocks(img)
        blocks.append(b)
        qf.append(np.full(len(b), qp_feature(qp), np.float32))
        labels.append(splits.reshape(-1))
        if with_costs:
            costs.append(c)
    out = (np.concatenate(blocks), np.concatenate(qf), np.concatenate(labels).astype(np.uint8))
    return out + (np.concatenate(costs),) if with_costs else out


def soft_targets(costs, tau_rel=0.005):
    """Target distribution over the six splits of each node:
    softmax(-(J - J_min) / (tau_rel * J_min)); a split whose cost is 3 * tau_rel
    above the best gets about 5 % of the best one's weight."""
    c = np.asarray(costs, np.float64)
    jmin = c.min(axis=1, keepdims=True)
    z = -(c - jmin) / (tau_rel * np.maximum(jmin, 1.0))
    e = np.exp(z - z.max(axis=1, keepdims=True))
    return (e / e.sum(axis=1, keepdims=True)).astype(np.float32)


# ----------------------------------------------------------------- training

def train(net, blocks, qf, labels, epochs=8, batch=256, lr=2e-3, seed=0, log=None):
    """Adam on the GPU over shuffled batches with a cosine learning-rate decay.
    labels: split indices (n,) or target distributions (n, 6). Returns the mean
    loss per epoch."""
    rng = np.random.default_rng(seed)
    n = len(blocks)
    soft = np.asarray(labels).ndim == 2
    batch = min(batch, net.max_batch)
    db = _hip.DeviceBuffer(batch * NODE * NODE)
    dq = _hip.DeviceBuffer(batch * 4)
    dl = _hip.DeviceBuffer(batch * (24 if soft else 1))
    dloss = _hip.DeviceBuffer(batch * 4)
    s = _hip.Stream()
    history = []
    steps = max(1, epochs * (n // batch))
    k = 0
    for ep in range(epochs):
        perm = rng.permutation(n)
        tot, cnt = 0.0, 0
        for a in range(0, n - batch + 1, batch):
            idx = perm[a:a + batch]
            db.upload(np.ascontiguousarray(blocks[idx]))
            dq.upload(np.ascontiguousarray(qf[idx]))
            dl.upload(np.ascontiguousarray(labels[idx]))
            step_lr = lr * 0.5 * (1.0 + math.cos(math.pi * k / steps))
            if soft:
                net.train_step_soft(db.ptr, dq.ptr, dl.ptr, batch, max(step_lr, 1e-6), dloss.ptr, s.handle)
            else:
                net.train_step(db.ptr, dq.ptr, dl.ptr, batch, max(step_lr, 1e-6), dloss.ptr, s.handle)
            k += 1
            s.synchronize()
            tot += float(dloss.download(np.float32).mean())
            cnt += 1
        history.append(tot / max(cnt, 1))
        if log:
            log("epoch %d loss %.4f" % (ep, history[-1]))
    for b in (db, dq, dl, dloss):
        b.close()
    return history


def set_logit_scale(net, scale):
    """Multiplies the last layer (W5, b5) by `scale`: softmax(scale * z). A scale
    above 1 sharpens the split probabilities the pruning thresholds act on."""
    p = net.params()
    o = p.size - (6 * 64 + 6)
    p[o:] *= np.float32(scale)
    net.set_params(p)


# ------------------------------------------------------------------- encode

def psnr(a, b):
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)
    return 99.0 if mse == 0 else 10 * math.log10(255.0 ** 2 / mse)


class NetPool:
    """Inference instances of one network for concurrent pruned jobs: a
    predict call takes a free instance, and instances run on their own
    streams, so predictions of different frames overlap on the device."""

    def __init__(self, nets):
        self.nets = list(nets)
        self.free = queue.Queue()
        for n in self.nets:
            self.free.put(n)

    def predict(self, blocks, qp):
        n = self.free.get()
        try:
            return n.predict(blocks, qp)
        finally:
            self.free.put(n)

    def params(self):
        return self.nets[0].params()

    def set_params(self, p):
        for n in self.nets:
            n.set_params(p)

    def close(self):
        for n in self.nets:
            n.close()


def encode(frames, qps, mode, net=None, p_prune=0.05, p_fallback=0.6, workers=8):
    """Encodes every frame at every QP as one pool of (frame, QP) jobs; returns
    (per-QP totals, wall seconds of the whole encode, network inference
    included). Per QP: bits, PSNR (on the padded picture), candidates per
    node, fallback rate, the splits chosen. In pruned mode each job runs the
    network on its own frame first (net: a network or a NetPool; one network
    serves one job at a time), overlapping other jobs' RD search."""
    imgs = [pad_to_nodes(f) for f in frames]
    if mode == abi.RD_PRUNED and not isinstance(net, NetPool):
        net = NetPool([net])

    def job(args):
        img, qp = args
        probs = None
        if mode == abi.RD_PRUNED:
            b = node_blocks(img)
            probs = net.predict(b, np.full(len(b), qp_feature(qp), np.float32))
        return _rd((img, qp, mode, probs, p_prune, p_fallback))

    jobs = [(img, qp) for qp in qps for img in imgs]
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(workers) as ex:
        res = list(ex.map(job, jobs))
    sec = time.perf_counter() - t0
    out = []
    for k, qp in enumerate(qps):
        r = res[k * len(imgs):(k + 1) * len(imgs)]
        sse = sum(x[2]["sse"] for x in r)
        px = sum(img.size for img in imgs)
        nodes = sum(x[2]["blocks"] for x in r)
        mse = sse / px
        out.append(dict(qp=qp, bits=sum(x[2]["bits"] for x in r),
                        psnr=99.0 if mse == 0 else 10 * math.log10(255.0 ** 2 / mse),
                        candidates_per_node=sum(x[2]["candidates"] for x in r) / nodes,
                        fallback_rate=sum(x[2]["fallbacks"] for x in r) / nodes,
                        splits=np.concatenate([x[1].reshape(-1) for x in r])))
    return out, sec


# ------------------------------------------------------------------ BD-rate

def bd_rate(anchor, test):
    """Percent BD-rate of test against anchor (bd_rate, scripts/bench/bench.py:777-813).
    Samples are (PSNR dB, rate); a cubic (or lower) least-squares fit of
    log10(rate) over PSNR, integrated over the overlapping PSNR interval.
    Positive: the test needs more bits for the same quality. None when the
    intervals do not overlap."""
    if len(anchor) < 3 or len(test) < 3:
        return None
    anchor, test = sorted(anchor), sorted(test)
    low, high = max(anchor[0][0], test[0][0]), min(anchor[-1][0], test[-1][0])
    if high - low < 1e-3:
        return None
    deg = min(3, len(anchor) - 1, len(test) - 1)
    fa = np.polyfit([p[0] for p in anchor], [math.log10(p[1]) for p in anchor], deg)
    ft = np.polyfit([p[0] for p in test], [math.log10(p[1]) for p in test], deg)
    d = np.polyint(ft - fa)
    avg = (np.polyval(d, high) - np.polyval(d, low)) / (high - low)
    return (10 ** avg - 1) * 100


def timed_encode(frames, qps, mode, net=None, p_prune=0.05, p_fallback=0.6, workers=8, repeats=1):
    """encode() `repeats` times (the results are identical) keeping the
    shortest wall time, so one slow run does not set the fps."""
    out, sec = encode(frames, qps, mode, net, p_prune, p_fallback, workers)
    for _ in range(repeats - 1):
        sec = min(sec, encode(frames, qps, mode, net, p_prune, p_fallback, workers)[1])
    return out, sec


def baselines(frames, qps=QPS, workers=8, repeats=1):
    """The full-RD and heuristic encodes evaluate() compares against."""
    return {"full": timed_encode(frames, qps, abi.RD_FULL, workers=workers, repeats=repeats),
            "heuristic": timed_encode(frames, qps, abi.RD_HEURISTIC, workers=workers, repeats=repeats)}


def evaluate(frames, net, qps=QPS, p_prune=0.05, p_fallback=0.6, workers=8, base=None, repeats=1):
    """Full RD, CNN-pruned and heuristic encodes of the same frames at every
    QP: fps, candidates per node, BD-rates against full RD, and how often
    each fast method picks the full search's split. base: baselines() of the
    same frames, to reuse; repeats: wall-time samples per encode (best kept)."""
    base = base or baselines(frames, qps, workers, repeats)
    out, secs = {}, {}
    for k in ("full", "heuristic"):
        out[k], secs[k] = base[k]
    out["pruned"], secs["pruned"] = timed_encode(frames, qps, abi.RD_PRUNED, net, p_prune, p_fallback, workers,
                                                 repeats)
    pts = {k: [(r["psnr"], r["bits"]) for r in v] for k, v in out.items()}
    summ = {}
    for k, v in out.items():
        summ[k] = dict(fps=len(frames) * len(qps) / secs[k],
                       candidates_per_node=float(np.mean([r["candidates_per_node"] for r in v])),
                       bd_rate_vs_full=None if k == "full" else bd_rate(pts["full"], pts[k]),
                       split_agreement=float(np.mean(np.concatenate([r["splits"] for r in v]) ==
                                                     np.concatenate([r["splits"] for r in out["full"]]))))
        if k == "pruned":
            summ[k]["fallback_rate"] = float(np.mean([r["fallback_rate"] for r in v]))
    summ["speedup_pruned_vs_full"] = summ["pruned"]["fps"] / summ["full"]["fps"]
    summ["points"] = {k: [(round(p, 4), round(b, 1)) for p, b in v] for k, v in pts.items()}
    return summ


TRAIN_SEEDS = (1, 3, 4, 5, 6, 7, 8, 9) + tuple(range(17, 73))


def run_experiment(acc, train_seeds=TRAIN_SEEDS, val_seeds=(10, 11, 13, 14), test_seeds=(2, 12, 15, 16),
                   frames_per_seq=6, eval_frames=2, width=1280, height=720, epochs=24, lr=3e-3, batch=256,
                   scales=(2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 10.0, 13.0), bd_target=1.5, workers=16, seed=1,
                   infer_instances=4, test_frames=4, timing_repeats=3, log=None):
    """Collects full-RD pairs on the training sequences, trains the network on
    the GPU, picks the largest logit scale whose pruned encode stays within
    bd_target % BD-rate of the full search on the validation sequences, and
    reports full / pruned / heuristic encodes on the test sequences."""
    from . import Cnn
    say = log or (lambda s: None)
    t0 = time.perf_counter()
    tr = sum([synthetic_sequence(s, frames_per_seq, width, height) for s in train_seeds], [])
    blocks, qf, labels = collect_pairs(tr, workers=workers)
    t_collect = time.perf_counter() - t0
    say("collected %d pairs in %.1f s" % (len(blocks), t_collect))
    net = Cnn(acc, max_batch=batch, seed=seed)
    untrained = net.params()
    t0 = time.perf_counter()
    hist = train(net, blocks, qf, labels, epochs=epochs, batch=batch, lr=lr, seed=seed)
    t_train = time.perf_counter() - t0
    say("trained %d epochs in %.1f s, loss %.3f -> %.3f" % (epochs, t_train, hist[0], hist[-1]))
    val = sum([synthetic_sequence(s, eval_frames, width, height) for s in val_seeds], [])
    trained = net.params()
    # inference runs on a second instance sized for a whole frame's nodes, so
    # each pruned job predicts its frame in one call
    per_frame = len(node_blocks(pad_to_nodes(val[0])))
    infer = NetPool(Cnn(acc, max_batch=per_frame, seed=seed) for _ in range(infer_instances))
    chosen, val_curve = scales[0], []
    val_base = baselines(val, workers=workers)
    for sc in scales:
        infer.set_params(trained)
        set_logit_scale(infer, sc)
        v = evaluate(val, infer, workers=workers, base=val_base)
        val_curve.append((sc, round(v["pruned"]["bd_rate_vs_full"], 3), round(v["pruned"]["candidates_per_node"], 3)))
        say("scale %.2f: validation BD-rate %.3f %%, %.2f candidates per node" % val_curve[-1])
        if v["pruned"]["bd_rate_vs_full"] <= bd_target:
            chosen = sc
    infer.set_params(trained)
    set_logit_scale(infer, chosen)
    test = sum([synthetic_sequence(s, test_frames, width, height) for s in test_seeds], [])
    test_base = baselines(test, workers=workers, repeats=timing_repeats)
    res = evaluate(test, infer, workers=workers, base=test_base, repeats=timing_repeats)
    # the same scale on the untrained network: what training buys
    infer.set_params(untrained)
    set_logit_scale(infer, chosen)
    res_untrained = evaluate(test, infer, workers=workers, base=test_base)
    infer.close()
    net.set_params(trained)
    set_logit_scale(net, chosen)
    report = dict(train_pairs=int(len(blocks)), collect_seconds=round(t_collect, 2), train_seconds=round(t_train, 2),
                  epochs=epochs, loss_first=round(hist[0], 4), loss_last=round(hist[-1], 4), logit_scale=chosen,
                  validation=val_curve, test=res,
                  untrained=dict(bd_rate_vs_full=res_untrained["pruned"]["bd_rate_vs_full"],
                                 candidates_per_node=res_untrained["pruned"]["candidates_per_node"],
                                 speedup_vs_full=res_untrained["speedup_pruned_vs_full"]))
    return net, report

"""Seeded synthetic workloads for the parity tests and the benchmark.

Frame-level generators for the three stages:

  residual_frame   one 4:2:0 frame's worth of TUs (recursive tiling, sizes 4..64,
                   MTS types, Laplacian coefficients with most zeros at high
                   frequencies, QP 22..37)
  inter_frame      a textured 4:2:0 reference, a seeded CU quadtree over 128x128
                   CTUs, one uni-prediction MV per CU (integer part +-64, any
                   fraction), and a deblocking edge map on the 8x8 grid with a
                   random bS per 4-sample segment and filter lengths from the CU
                   sizes (luma_lens / chroma_lens, vvc/decode_general.rs:4916-4935)
  bi_frame         two textured 4:2:0 references and one bi-predicted PU per CU
                   (independent L0/L1 MVs, BCW index, alternative half-sample
                   filter on some PUs)
  pframe           inter_frame plus the residual TUs of the non-skipped CUs
                   (a transform tree per CU) and where each TU lands

The numpy dtypes mirror the C structs of include/rmh.h byte for byte.
"""
import numpy as np

TU_DTYPE = np.dtype([
    ("offset", "<u4"), ("log2_w", "u1"), ("log2_h", "u1"), ("qp", "u1"), ("bit_depth", "u1"),
    ("tr_hor", "u1"), ("tr_ver", "u1"), ("skip_w", "u1"), ("skip_h", "u1"),
    ("lfnst_idx", "u1"), ("lfnst_mode", "u1"), ("flags", "u1"), ("joint", "u1")])
assert TU_DTYPE.itemsize == 16

PU_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("log2_w", "u1"), ("log2_h", "u1"),
                     ("frame", "<u2"), ("mv_x", "<i4"), ("mv_y", "<i4")])
assert PU_DTYPE.itemsize == 16

BI_PU_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("log2_w", "u1"), ("log2_h", "u1"), ("frame", "<u2"),
                        ("mv0_x", "<i4"), ("mv0_y", "<i4"), ("mv1_x", "<i4"), ("mv1_y", "<i4"),
                        ("bcw_idx", "u1"), ("refine", "u1"), ("reserved", "u1", (2,))])
assert BI_PU_DTYPE.itemsize == 28
BI_DMVR, BI_BDOF = 1, 2

EDGE_CELL_DTYPE = np.dtype([("tu_y", "<u4"), ("tu_c", "<u4"), ("tu_wy", "u1"), ("tu_hy", "u1"), ("tu_wc", "u1"),
                            ("tu_hc", "u1"), ("flags", "u1"), ("reserved", "u1", (3,)), ("mv_x", "<i4"),
                            ("mv_y", "<i4"), ("ref_poc", "<i4")])
assert EDGE_CELL_DTYPE.itemsize == 28
CELL_INTRA, CELL_CBF = 1, 2

DCT2, DST7, DCT8 = 0, 1, 2
TS_FLAG, DQ_FLAG = 1, 2


# ----------------------------------------------------------------- tiling

def _tile(rng, W, H, ctu, max_tu, min_tu, p_split):
    """Recursive seeded tiling of a W x H plane into power-of-two blocks."""
    out = []

    def rec(x, y, w, h):
        if x >= W or y >= H:
            return
        fits = x + w <= W and y + h <= H
        big = max(w, h) > max_tu
        can_h = h > min_tu
        can_w = w > min_tu
        if fits and not big:
            if not (can_h or can_w) or rng.random() >= p_split(w, h):
                out.append((x, y, w, h))
                return
        # choose a split: quad, horizontal binary or vertical binary
        if not fits:
            if x + w > W and y + h > H and can_w and can_h:
                kind = 0
            elif x + w > W and can_w:
                kind = 2
            elif can_h:
                kind = 1
            else:
                kind = 2
        else:
            opts = []
            if can_w and can_h and w == h:
                opts += [0, 0]
            if can_h and h >= w // 2 * 1 and h * 4 > w:
                opts.append(1)
            if can_w and w * 4 > h:
                opts.append(2)
            if can_h and h > w:
                opts.append(1)
            if can_w and w > h:
                opts.append(2)
            if not opts:
                opts = [1] if can_h else [2]
            kind = opts[int(rng.integers(0, len(opts)))]
        if kind == 0:
            hw, hh = w // 2, h // 2
            rec(x, y, hw, hh)
            rec(x + hw, y, hw, hh)
            rec(x, y + hh, hw, hh)
            rec(x + hw, y + hh, hw, hh)
        elif kind == 1:
            rec(x, y, w, h // 2)
            rec(x, y + h // 2, w, h // 2)
        else:
            rec(x, y, w // 2, h)
            rec(x + w // 2, y, w // 2, h)

    for cy in range(0, H, ctu):
        for cx in range(0, W, ctu):
            rec(cx, cy, ctu, ctu)
    return out


def _p_split_residual(w, h):
    m = max(w, h)
    return {64: 0.93, 32: 0.72, 16: 0.42, 8: 0.16}.get(m, 0.0)


# ---------------------------------------------------------------- residual

def residual_frame(seed, width=1920, height=1080, bit_depth=10, qp_range=(22, 37), with_lfnst=False,
                   with_ts=False, with_dq=False, zero_frac=0.85):
    """One frame of TUs: returns (tus, coeff int16, blocks) where blocks lists
    (plane, x, y, w, h) of every TU in descriptor order."""
    rng = np.random.default_rng([2, seed])
    blocks = []
    for plane, (W, H, ctu, maxtu) in enumerate([(width, height, 64, 64),
                                                 ((width + 1) // 2, (height + 1) // 2, 32, 32),
                                                 ((width + 1) // 2, (height + 1) // 2, 32, 32)]):
        for (x, y, w, h) in _tile(rng, W, H, ctu, maxtu, 4, _p_split_residual):
            blocks.append((plane, x, y, w, h))
    tus, coeff = _tu_batch(rng, blocks, bit_depth, qp_range, with_lfnst, with_ts, with_dq, zero_frac)
    return tus, coeff, blocks


def _tu_batch(rng, blocks, bit_depth, qp_range, with_lfnst=False, with_ts=False, with_dq=False, zero_frac=0.85):
    """Descriptors and coefficients for blocks [(plane, x, y, w, h)], in order."""
    n = len(blocks)
    tus = np.zeros(n, TU_DTYPE)
    ws = np.array([b[3] for b in blocks])
    hs = np.array([b[4] for b in blocks])
    sizes = ws * hs
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    tus["offset"] = offsets
    tus["log2_w"] = np.log2(ws).astype(np.uint8)
    tus["log2_h"] = np.log2(hs).astype(np.uint8)
    qp = rng.integers(qp_range[0], qp_range[1] + 1, n)
    tus["qp"] = qp + 6 * (bit_depth - 8)
    tus["bit_depth"] = bit_depth
    th = rng.integers(0, 3, n)
    tv = rng.integers(0, 3, n)
    th[ws > 32] = DCT2
    tv[hs > 32] = DCT2
    tus["tr_hor"] = th
    tus["tr_ver"] = tv
    if with_lfnst:
        sel = (rng.random(n) < 0.3) & (ws >= 4) & (hs >= 4)
        tus["lfnst_idx"][sel] = rng.integers(1, 3, int(sel.sum()))
        tus["lfnst_mode"][sel] = rng.integers(0, 95, int(sel.sum()))
        tus["tr_hor"][sel] = DCT2
        tus["tr_ver"][sel] = DCT2
    if with_ts:
        sel = (rng.random(n) < 0.1) & (sizes <= 1024)
        tus["flags"][sel] |= TS_FLAG
    if with_dq:
        sel = (rng.random(n) < 0.2) & ((tus["flags"] & TS_FLAG) == 0)
        tus["flags"][sel] |= DQ_FLAG
    total = int(sizes.sum())
    # per-sample local coordinates
    tu_of = np.repeat(np.arange(n), sizes)
    local = np.arange(total) - np.repeat(offsets, sizes)
    lx = local % np.repeat(ws, sizes)
    ly = local // np.repeat(ws, sizes)
    scale = np.repeat((ws + hs) / 2.0, sizes)
    # P(nonzero) decays with x + y; density tuned so ~zero_frac of samples are 0
    p = np.exp(-(lx + ly) / (0.22 * scale + 1.0))
    keep = rng.random(total) < p * (1.0 - zero_frac) / max(p.mean(), 1e-9)
    mag = np.rint(rng.laplace(0.0, 12.0, total)).astype(np.int64)
    coeff = np.where(keep, mag, 0)
    # rare large values exercise the int16 clip of the dequantiser
    big = rng.random(total) < 2e-4
    coeff = np.where(big & keep, rng.integers(-32768, 32768, total), coeff)
    coeff = np.clip(coeff, -32768, 32767).astype(np.int16)
    del tu_of
    return tus, coeff


def concat_frames(frames):
    """Concatenate [(tus, coeff)] into one batch, shifting offsets."""
    all_tus, all_c, base = [], [], 0
    for tus, coeff in frames:
        t = tus.copy()
        t["offset"] += base
        base += coeff.size
        all_tus.append(t)
        all_c.append(coeff)
    return np.concatenate(all_tus), np.concatenate(all_c)


# ------------------------------------------------------------------- inter

def luma_lens(size_p, size_q):
    """luma_lens (vvc/decode_general.rs:4916-4922)."""
    if size_p <= 4 or size_q <= 4:
        return 1, 1
    return (7 if size_p >= 32 else 3), (7 if size_q >= 32 else 3)


def chroma_lens(size_p, size_q, ctb_boundary):
    """chroma_lens (vvc/decode_general.rs:4928-4935)."""
    if size_p >= 8 and size_q >= 8:
        return (1, 3) if ctb_boundary else (3, 3)
    return 1, 1


def _len_code(n):
    return (n - 1) // 2


def edge_byte(bs, lp, lq, ctb=False):
    return (bs & 3) | (_len_code(lp) << 2) | (_len_code(lq) << 4) | (0x40 if ctb else 0)


def textured_plane(rng, w, h, bit_depth):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    maxv = (1 << bit_depth) - 1
    v = 0.5 * maxv
    for _ in range(4):
        fx, fy = rng.uniform(0.005, 0.12, 2)
        ph = rng.uniform(0, 2 * np.pi)
        v = v + rng.uniform(0.05, 0.18) * maxv * np.sin(2 * np.pi * (fx * xx + fy * yy) + ph)
    v = v + rng.normal(0, 0.02 * maxv, (h, w))
    return np.clip(np.rint(v), 0, maxv).astype(np.uint16)


def _p_split_inter(w, h):
    return {128: 0.9, 64: 0.6, 32: 0.45, 16: 0.3}.get(max(w, h), 0.0)


def inter_frame(seed, width=1920, height=1080, bit_depth=10, mv_range=64, qp=32, ctu=128):
    """Returns dict(ref=[Y,Cb,Cr] uint16, pus=PU array, cus=[(x,y,w,h)], edges=uint8[gh,gw,4],
    qp=[3], ctu=ctu)."""
    rng = np.random.default_rng([3, seed])
    cw, ch = (width + 1) // 2, (height + 1) // 2
    ref = [textured_plane(rng, width, height, bit_depth), textured_plane(rng, cw, ch, bit_depth),
           textured_plane(rng, cw, ch, bit_depth)]
    d = _inter_layout(rng, width, height, bit_depth, mv_range, qp, ctu)
    d["ref"] = ref
    return d


def _inter_layout(rng, width, height, bit_depth, mv_range, qp, ctu):
    """The coded layout of a P-frame drawn from rng: CUs of a seeded quadtree,
    one uni-prediction PU per CU, the edge map and the QPs."""
    cw, ch = (width + 1) // 2, (height + 1) // 2
    cus = _tile(rng, width, height, ctu, ctu, 8, _p_split_inter)
    n = len(cus)
    pus = np.zeros(n, PU_DTYPE)
    pus["x"] = [c[0] for c in cus]
    pus["y"] = [c[1] for c in cus]
    pus["log2_w"] = [int(np.log2(c[2])) for c in cus]
    pus["log2_h"] = [int(np.log2(c[3])) for c in cus]
    pus["mv_x"] = 16 * rng.integers(-mv_range, mv_range + 1, n) + rng.integers(0, 16, n)
    pus["mv_y"] = 16 * rng.integers(-mv_range, mv_range + 1, n) + rng.integers(0, 16, n)
    # CU-size maps on the 4x4 grid
    gw, gh = (width + 3) // 4, (height + 3) // 4
    cu_id = np.full((gh, gw), -1, np.int64)
    cu_w = np.zeros((gh, gw), np.int64)
    cu_h = np.zeros((gh, gw), np.int64)
    for i, (x, y, w, h) in enumerate(cus):
        cu_id[y // 4:(y + h + 3) // 4, x // 4:(x + w + 3) // 4] = i
        cu_w[y // 4:(y + h + 3) // 4, x // 4:(x + w + 3) // 4] = w
        cu_h[y // 4:(y + h + 3) // 4, x // 4:(x + w + 3) // 4] = h
    edges = np.zeros((gh, gw, 4), np.uint8)
    bs_v = rng.integers(0, 3, (gh, gw))
    bs_h = rng.integers(0, 3, (gh, gw))
    cgw, cgh = (cw + 1) // 2, (ch + 1) // 2
    ii = np.arange(gw)[None, :]
    jj = np.arange(gh)[:, None]
    # CU boundaries crossing the left / top side of each 4x4 cell
    diff_v = np.zeros((gh, gw), bool)
    diff_v[:, 1:] = cu_id[:, 1:] != cu_id[:, :-1]
    diff_h = np.zeros((gh, gw), bool)
    diff_h[1:, :] = cu_id[1:, :] != cu_id[:-1, :]
    w_p = np.zeros_like(cu_w)
    w_p[:, 1:] = cu_w[:, :-1]
    h_p = np.zeros_like(cu_h)
    h_p[1:, :] = cu_h[:-1, :]

    def lens_luma(sp, sq):  # luma_lens, elementwise
        small = (sp <= 4) | (sq <= 4)
        return np.where(small, 1, np.where(sp >= 32, 7, 3)), np.where(small, 1, np.where(sq >= 32, 7, 3))

    def byte(bs, lp, lq, ctb):  # edge_byte, elementwise
        return ((bs & 3) | (((lp - 1) // 2) << 2) | (((lq - 1) // 2) << 4) | np.where(ctb, 0x40, 0)).astype(np.uint8)

    lp, lq = lens_luma(w_p, cu_w)
    edges[..., 0] = np.where(((4 * ii) % 8 == 0) & diff_v, byte(bs_v, lp, lq, False), 0)
    lp, lq = lens_luma(h_p, cu_h)
    edges[..., 1] = np.where(((4 * jj) % 8 == 0) & diff_h, byte(bs_h, lp, lq, (4 * jj) % ctu == 0), 0)
    # chroma cell (i, j) covers chroma x 2i..2i+1, y 2j..2j+1 (luma 4i.., 4j..): chroma_lens
    in_c = (ii < cgw) & (jj < cgh)
    big = (w_p // 2 >= 8) & (cu_w // 2 >= 8)
    lc = np.where(big, 3, 1)
    edges[..., 2] = np.where(in_c & ((2 * ii) % 8 == 0) & diff_v, byte(bs_v, lc, lc, False), 0)
    big = (h_p // 2 >= 8) & (cu_h // 2 >= 8)
    at_ctb = (2 * jj) % (ctu // 2) == 0
    lp = np.where(big, np.where(at_ctb, 1, 3), 1)
    lq = np.where(big, 3, 1)
    edges[..., 3] = np.where(in_c & ((2 * jj) % 8 == 0) & diff_h, byte(bs_h, lp, lq, False), 0)
    chroma_qp = qp if qp <= 26 else qp - 1  # default chroma QP mapping, vvc/hls.rs:778-781
    return dict(pus=pus, cus=cus, edges=edges, qp=[qp, chroma_qp, chroma_qp], ctu=ctu,
                bit_depth=bit_depth, width=width, height=height)


def bi_frame(seed, width=1920, height=1080, bit_depth=10, mv_range=64, ctu=128):
    """Returns dict(ref0=[Y,Cb,Cr], ref1=[Y,Cb,Cr] uint16, pus=BI_PU_DTYPE array, bit_depth, width,
    height). BCW indices are drawn from 0..4 with a few 5..7, which the reference clamps to 4."""
    rng = np.random.default_rng([4, seed])
    cw, ch = (width + 1) // 2, (height + 1) // 2
    refs = [[textured_plane(rng, width, height, bit_depth), textured_plane(rng, cw, ch, bit_depth),
             textured_plane(rng, cw, ch, bit_depth)] for _ in range(2)]
    cus = _tile(rng, width, height, ctu, ctu, 8, _p_split_inter)
    n = len(cus)
    pus = np.zeros(n, BI_PU_DTYPE)
    pus["x"] = [c[0] for c in cus]
    pus["y"] = [c[1] for c in cus]
    pus["log2_w"] = [int(np.log2(c[2])) for c in cus]
    pus["log2_w"] |= np.where(rng.random(n) < 0.25, 0x10, 0).astype(np.uint8)
    pus["log2_h"] = [int(np.log2(c[3])) for c in cus]
    for k in ("mv0_x", "mv0_y", "mv1_x", "mv1_y"):
        pus[k] = 16 * rng.integers(-mv_range, mv_range + 1, n) + rng.integers(0, 16, n)
    pus["bcw_idx"] = np.where(rng.random(n) < 0.05, rng.integers(5, 8, n), rng.integers(0, 5, n))
    return dict(ref0=refs[0], ref1=refs[1], pus=pus, bit_depth=bit_depth, width=width, height=height)


# --------------------------------------------------------------- P-frames

PLACE_DTYPE = np.dtype([("frame", "<u2"), ("plane", "u1"), ("reserved", "u1"), ("x", "<u2"), ("y", "<u2")])
assert PLACE_DTYPE.itemsize == 8


def _split_block(rng, x, y, w, h, max_tu, min_tu, p_split, out):
    """Transform tree of one block: forced splits down to max_tu, then seeded
    quad / binary splits while both halves stay >= min_tu."""
    if w > max_tu or h > max_tu:
        if w > max_tu and h > max_tu:
            for (dx, dy) in ((0, 0), (w // 2, 0), (0, h // 2), (w // 2, h // 2)):
                _split_block(rng, x + dx, y + dy, w // 2, h // 2, max_tu, min_tu, p_split, out)
        elif w > max_tu:
            _split_block(rng, x, y, w // 2, h, max_tu, min_tu, p_split, out)
            _split_block(rng, x + w // 2, y, w // 2, h, max_tu, min_tu, p_split, out)
        else:
            _split_block(rng, x, y, w, h // 2, max_tu, min_tu, p_split, out)
            _split_block(rng, x, y + h // 2, w, h // 2, max_tu, min_tu, p_split, out)
        return
    can_w, can_h = w // 2 >= min_tu, h // 2 >= min_tu
    if (can_w or can_h) and rng.random() < p_split:
        kind = int(rng.integers(0, 3))
        if kind == 0 and can_w and can_h:
            for (dx, dy) in ((0, 0), (w // 2, 0), (0, h // 2), (w // 2, h // 2)):
                _split_block(rng, x + dx, y + dy, w // 2, h // 2, max_tu, min_tu, p_split * 0.5, out)
            return
        if (kind == 1 and can_h) or not can_w:
            _split_block(rng, x, y, w, h // 2, max_tu, min_tu, p_split * 0.5, out)
            _split_block(rng, x, y + h // 2, w, h // 2, max_tu, min_tu, p_split * 0.5, out)
            return
        _split_block(rng, x, y, w // 2, h, max_tu, min_tu, p_split * 0.5, out)
        _split_block(rng, x + w // 2, y, w // 2, h, max_tu, min_tu, p_split * 0.5, out)
        return
    out.append((x, y, w, h))


def pframe(seed, width=1920, height=1080, bit_depth=10, mv_range=64, qp=32, ctu=128, p_skip=0.2, frame=0):
    """A P-frame for the MC -> residual add -> deblock pipeline: inter_frame's
    reference, PUs and edge map, plus the residual TUs of every CU that is not
    skipped (luma transform tree inside the CU, max 64; chroma likewise at half
    size, max 32), with their placement. Returns inter_frame's dict plus
    tus, coeff, place (PLACE_DTYPE, frame index `frame`)."""
    d = inter_frame(seed, width, height, bit_depth, mv_range, qp, ctu)
    _add_residuals(np.random.default_rng([4, seed]), d, bit_depth, qp, p_skip, frame)
    return d


def _add_residuals(rng, d, bit_depth, qp, p_skip, frame, tools=False):
    """The residual TUs of every CU of layout d that is not skipped (luma
    transform tree inside the CU, max 64; chroma likewise at half size, max
    32) with their placement; tools adds LFNST, transform skip and dependent
    quantisation at the rates of _tu_batch."""
    blocks = []
    for (x, y, w, h) in d["cus"]:
        if rng.random() < p_skip:
            continue
        lum = []
        _split_block(rng, x, y, w, h, 64, 4, 0.35, lum)
        blocks += [(0, bx, by, bw, bh) for (bx, by, bw, bh) in lum]
        for plane in (1, 2):
            ch = []
            _split_block(rng, x // 2, y // 2, w // 2, h // 2, 32, 4, 0.2, ch)
            blocks += [(plane, bx, by, bw, bh) for (bx, by, bw, bh) in ch]
    tus, coeff = _tu_batch(rng, blocks, bit_depth, (qp - 5, qp + 5), tools, tools, tools)
    place = np.zeros(len(blocks), PLACE_DTYPE)
    place["frame"] = frame
    place["plane"] = [b[0] for b in blocks]
    place["x"] = [b[1] for b in blocks]
    place["y"] = [b[2] for b in blocks]
    d.update(tus=tus, coeff=coeff, place=place, blocks=blocks)


def stream_reference(seed, width, height, bit_depth):
    """The first reference picture of a synthetic stream: seeded textured
    planes [Y, Cb, Cr] (4:2:0) as uint16."""
    rng = np.random.default_rng([6, seed])
    cw, ch = (width + 1) // 2, (height + 1) // 2
    return [textured_plane(rng, width, height, bit_depth), textured_plane(rng, cw, ch, bit_depth),
            textured_plane(rng, cw, ch, bit_depth)]


def pframe_step(seed, width=1920, height=1080, bit_depth=10, mv_range=64, qp=32, ctu=128, p_skip=0.2, tools=True):
    """One frame step of a chained stream: pframe's layout and residuals
    without a reference picture (the previous step's output is the
    reference). tools: the residual tool mix (LFNST, transform skip,
    dependent quantisation)."""
    rng = np.random.default_rng([5, seed])
    d = _inter_layout(rng, width, height, bit_depth, mv_range, qp, ctu)
    _add_residuals(rng, d, bit_depth, qp, p_skip, 0, tools)
    return d


def joint_cbcr_frame(seed, width=256, height=128, bit_depth=10, p_joint=0.5):
    """A frame of recon TUs where chroma blocks are either a separate Cb and
    Cr TU or one joint Cb-Cr TU (mode 1..3, random sign; the coded component
    is Cr for mode 1 and Cb otherwise, as reconstruct_joint_cbcr reads them).
    Returns (tus, coeff, place, blocks) with blocks listing (plane, x, y, w, h,
    joint) per TU; luma TUs are plain. Some TUs use transform skip."""
    rng = np.random.default_rng([7, seed])
    blocks = [(0, x, y, w, h) for (x, y, w, h) in _tile(rng, width, height, 64, 64, 4, _p_split_residual)]
    joints = [0] * len(blocks)
    cw, ch = (width + 1) // 2, (height + 1) // 2
    for (x, y, w, h) in _tile(rng, cw, ch, 32, 32, 4, _p_split_residual):
        if rng.random() < p_joint:
            mode = int(rng.integers(1, 4))
            code = mode | (4 if rng.random() < 0.5 else 0)
            blocks.append((2 if mode == 1 else 1, x, y, w, h))
            joints.append(code)
        else:
            blocks += [(1, x, y, w, h), (2, x, y, w, h)]
            joints += [0, 0]
    tus, coeff = _tu_batch(rng, blocks, bit_depth, (22, 37), with_ts=True)
    tus["joint"] = joints
    place = np.zeros(len(blocks), PLACE_DTYPE)
    place["plane"] = [b[0] for b in blocks]
    place["x"] = [b[1] for b in blocks]
    place["y"] = [b[2] for b in blocks]
    return tus, coeff, place, [b + (j,) for b, j in zip(blocks, joints)]


def edge_cells(seed, width=640, height=368, p_intra=0.15, ctu=128):
    """The cell grid (EDGE_CELL_DTYPE[gh, gw]) of a coded P-picture for the
    edge-map builder: pframe_step's CUs and transform trees (a skipped CU is
    one TU of its size, split at 64 luma / 32 chroma), a coded-block flag per
    luma TU from its levels, intra CUs at random, and per inter CU its MV and
    one of two reference POCs."""
    # the layout is coded at the next multiple of 8; the grid covers the visible size
    d = pframe_step(seed, width=(width + 7) // 8 * 8, height=(height + 7) // 8 * 8, bit_depth=10, tools=False, ctu=ctu)
    rng = np.random.default_rng([8, seed])
    gw, gh = (width + 3) // 4, (height + 3) // 4
    cells = np.zeros((gh, gw), EDGE_CELL_DTYPE)
    cells["tu_y"] = 0xffffffff
    cells["tu_c"] = 0xffffffff
    tu_id = 0
    coded = {(b[1], b[2], b[0]) for b in d["blocks"]}
    luma, chroma = [], []
    for k, (x, y, w, h) in enumerate(d["cus"]):
        intra = rng.random() < p_intra
        poc = int(rng.integers(0, 2))
        pu = d["pus"][k]
        sl = (slice(y // 4, (y + h + 3) // 4), slice(x // 4, (x + w + 3) // 4))
        cells["flags"][sl] = CELL_INTRA if intra else 0
        cells["mv_x"][sl] = 0 if intra else int(pu["mv_x"])
        cells["mv_y"][sl] = 0 if intra else int(pu["mv_y"])
        cells["ref_poc"][sl] = 0 if intra else poc
    for b in d["blocks"]:
        (luma if b[0] == 0 else chroma if b[0] == 1 else []).append(b[1:])
    # skipped CUs: one TU of the CU's size (at most 64 luma, 32 chroma)
    covered = np.zeros(((int(d["height"]) + 3) // 4, (int(d["width"]) + 3) // 4), bool)
    for (x, y, w, h) in luma:
        covered[y // 4:(y + h + 3) // 4, x // 4:(x + w + 3) // 4] = True
    for (x, y, w, h) in d["cus"]:
        if covered[y // 4, x // 4]:
            continue
        for yy in range(y, y + h, 64):
            for xx in range(x, x + w, 64):
                luma.append((xx, yy, min(64, w), min(64, h)))
        for yy in range(y // 2, (y + h) // 2, 32):
            for xx in range(x // 2, (x + w) // 2, 32):
                chroma.append((xx, yy, min(32, w // 2), min(32, h // 2)))
    nz = {}
    for t, b in zip(d["tus"], d["blocks"]):
        if b[0] == 0:
            n = 1 << (int(t["log2_w"]) + int(t["log2_h"]))
            nz[b[1:]] = bool(d["coeff"][int(t["offset"]):int(t["offset"]) + n].any())
    for (x, y, w, h) in luma:
        sl = (slice(y // 4, min(gh, (y + h + 3) // 4)), slice(x // 4, min(gw, (x + w + 3) // 4)))
        cells["tu_y"][sl] = tu_id
        cells["tu_wy"][sl] = w
        cells["tu_hy"][sl] = h
        if nz.get((x, y, w, h), False):
            cells["flags"][sl] |= CELL_CBF
        tu_id += 1
    for (x, y, w, h) in chroma:
        sl = (slice(y // 2, min(gh, (y + h + 1) // 2)), slice(x // 2, min(gw, (x + w + 1) // 2)))
        cells["tu_c"][sl] = tu_id
        cells["tu_wc"][sl] = w
        cells["tu_hc"][sl] = h
        tu_id += 1
    return cells

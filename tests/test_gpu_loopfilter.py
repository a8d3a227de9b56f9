"""LMCS luma mapping and ALF / CCALF on the GPU (rmh_lmcs_plan_*,
rmh_alf_plan_*) against the oracle's restatements of apply_*_plane_16
(vvc/filter.rs:2707-2722) and apply_alf (vvc/decode_general.rs:5225-5660),
bit for bit: 1080p10 pictures, 8-bit pictures with rows that are not 8-byte
aligned, 12-bit pictures with a visible region smaller than the planes, CTB
sizes 16 to 128, mixed batches, in-place mapping and graph replays."""
import ctypes

import numpy as np
import pytest

import workloads as W
from conftest import load_package

pytestmark = pytest.mark.gpu


def _dev_plane(hip, abi, a, bd, pad, keep):
    """a copied to a device plane whose rows are `pad` samples longer than a's."""
    h, w = a.shape
    buf = np.zeros((h, w + pad), a.dtype)
    buf[:, :w] = a
    d = hip.DeviceBuffer.from_array(buf)
    keep.append(d)
    return abi.plane_from_address(d.ptr, w, h, w + pad, a.dtype.itemsize, bd), d, (h, w + pad)


def _download(d, shape, dtype, w):
    return d.download(dtype).reshape(shape)[:, :w]


LMCS_CASES = [  # (w, h, bd, pad, direction, in_place)
    (1920, 1080, 10, 0, 0, True),
    (416, 240, 8, 3, 1, False),
    (333, 77, 12, 5, 0, False),
    (640, 360, 10, 1, 1, True),
]


def test_lmcs_mapping_batch(acc, po):
    print("lmcs batch", flush=True)
    pkg = load_package()
    hip, abi = pkg.hip, pkg.abi
    rng = np.random.default_rng(3)
    frames, checks, keep = [], [], []
    for w, h, bd, pad, direction, in_place in LMCS_CASES:
        lo = int(rng.integers(0, 4))
        hi = int(rng.integers(11, 16))
        init = 1 << (bd - 4)
        delta = [int(max(d, -init)) for d in rng.integers(-init // 2, init // 2 + 1, 16)]
        aps = abi.lmcs_aps(lo, hi, delta, int(rng.integers(-7, 8)))
        dt = np.uint8 if bd == 8 else np.uint16
        img = rng.integers(0, 1 << bd, (h, w)).astype(dt)
        if bd > 8:
            img[0, : w // 4] = rng.integers(1 << bd, 1 << 16, w // 4)  # above 2^bd - 1: the last LUT entry
        src, dsrc, shape = _dev_plane(hip, abi, img, bd, pad, keep)
        if in_place:
            dst, ddst = src, dsrc
        else:
            dst, ddst, _ = _dev_plane(hip, abi, np.zeros_like(img), bd, pad, keep)
        f = abi.LmcsFrame()
        f.src, f.dst, f.aps, f.direction = src, dst, aps, direction
        frames.append(f)
        fwd, inv, _, _ = po.lmcs_tables(aps, bd)
        checks.append((ddst, shape, dt, w, po.lmcs_map(img, fwd if direction else inv, bd)))
    plan = acc.lmcs_plan(frames)
    s = hip.Stream()
    plan.run(stream=s.handle)
    s.synchronize()
    for ddst, shape, dt, w, want in checks:
        got = _download(ddst, shape, dt, w)
        assert (got == want).all(), np.argwhere(got != want)[:5]


def _alf_aps(abi, rng, n_luma=25, n_chroma=8, n_cc=4, luma_clip=True):
    return abi.alf_aps([rng.integers(-128, 129, 12).tolist() for _ in range(n_luma)],
                       [rng.integers(0, 4, 12).tolist() for _ in range(n_luma)], luma_clip,
                       rng.integers(0, n_luma + 3, 25).tolist(),
                       [rng.integers(-128, 129, 6).tolist() for _ in range(n_chroma)],
                       [rng.integers(0, 4, 6).tolist() for _ in range(n_chroma)],
                       [rng.integers(-64, 65, 7).tolist() for _ in range(n_cc)],
                       [rng.integers(-64, 65, 7).tolist() for _ in range(n_cc)])


def _alf_case(abi, rng, w, h, bd, ctb, vis, seed):
    dt = np.uint8 if bd == 8 else np.uint16
    planes = [W.textured_plane(rng, w, h, bd).astype(dt), W.textured_plane(rng, (w + 1) // 2, (h + 1) // 2, bd).astype(dt),
              W.textured_plane(rng, (w + 1) // 2, (h + 1) // 2, bd).astype(dt)]
    cols, rows = -(-vis[0] // ctb) + 1, -(-vis[1] // ctb)
    ctbs = np.zeros(cols * rows, abi.ALF_CTB_DTYPE)
    ctbs["luma"] = rng.integers(0, 19, ctbs.size)   # off, the 16 fixed sets, the APS, and 18 (off)
    ctbs["cb"] = rng.integers(0, 11, ctbs.size)
    ctbs["cr"] = rng.integers(0, 11, ctbs.size)
    ctbs["cc_cb"] = rng.integers(0, 6, ctbs.size)
    ctbs["cc_cr"] = rng.integers(0, 6, ctbs.size)
    apss = [_alf_aps(abi, rng, luma_clip=seed % 2 == 0), _alf_aps(abi, rng, 1, int(rng.integers(1, 9)), 0),
            _alf_aps(abi, rng, 1, 0, int(rng.integers(1, 5)))]
    return planes, ctbs, ctb, cols, rows, vis, apss


ALF_CASES = [  # (w, h, bd, ctb, vis)
    (1920, 1080, 10, 128, (1920, 1080)),
    (416, 240, 8, 64, (416, 240)),
    (333, 201, 12, 16, (328, 200)),
    (856, 480, 8, 32, (854, 478)),
    (640, 368, 10, 64, (640, 360)),
]


def _alf_frames(hip, abi, cases, bds, keep, pads):
    frames, outs = [], []
    for (planes, ctbs, ctb, cols, rows, vis, apss), bd, pad in zip(cases, bds, pads):
        srcs, dsts, out = [], [], []
        for p in planes:
            sp, _, _ = _dev_plane(hip, abi, p, bd, pad, keep)
            dp, dd, shape = _dev_plane(hip, abi, np.zeros_like(p), bd, pad, keep)
            srcs.append(sp)
            dsts.append(dp)
            out.append((dd, shape, p.dtype, p.shape[1]))
        dc = hip.DeviceBuffer.from_array(ctbs)
        keep.append(dc)
        frames.append(abi.alf_frame(srcs, dsts, dc.ptr, ctb, cols, rows, vis[0], vis[1], luma_aps=apss[0],
                                    chroma_aps=apss[1], cc_cb_aps=apss[2], cc_cr_aps=apss[2]))
        outs.append(out)
    return frames, outs


def _alf_check(po, case, bd, got):
    planes, ctbs, ctb, cols, rows, vis, apss = case
    want = po.alf_frame(planes, bd, ctbs, ctb, cols, rows, vis[0], vis[1], luma_aps=apss[0], chroma_aps=apss[1],
                        cc_cb_aps=apss[2], cc_cr_aps=apss[2])
    for c in range(3):
        bad = np.argwhere(got[c] != want[c])
        assert bad.size == 0, "plane %d: first mismatch at %s of %d" % (c, bad[0], len(bad))
    assert all((want[c] != planes[c]).any() for c in range(3))


def test_alf_batch_and_graph_replays(acc, po):
    pkg = load_package()
    hip, abi = pkg.hip, pkg.abi
    rng = np.random.default_rng(21)
    cases, bds = [], []
    for k, (w, h, bd, ctb, vis) in enumerate(ALF_CASES):
        print("alf case %dx%d bd %d ctb %d vis %s" % (w, h, bd, ctb, vis), flush=True)
        cases.append(_alf_case(abi, rng, w, h, bd, ctb, vis, k))
        bds.append(bd)
    keep = []
    frames, outs = _alf_frames(hip, abi, cases, bds, keep, pads=[0, 3, 1, 2, 0])
    plan = acc.alf_plan(frames)
    assert plan.bytes == sum(2 * p.nbytes for c in cases for p in c[0])
    s = hip.Stream()

    def check():
        for case, bd, out in zip(cases, bds, outs):
            _alf_check(po, case, bd, [_download(d, shape, dt, w) for d, shape, dt, w in out])

    plan.run(stream=s.handle)
    s.synchronize()
    check()
    graph = acc.graph([(plan, 0, 0)])
    for rep in range(3):
        for out in outs:
            for d, _, _, _ in out:
                d.zero(s.handle)
        s.synchronize()
        graph.launch(s.handle)
        s.synchronize()
        check()


def test_alf_rejects_bad_frames(acc):
    pkg = load_package()
    hip, abi = pkg.hip, pkg.abi
    rng = np.random.default_rng(2)
    keep = []
    case = _alf_case(abi, rng, 64, 64, 10, 32, (64, 64), 0)
    frames, _ = _alf_frames(hip, abi, [case], [10], keep, pads=[0])
    bad = abi.AlfFrame.from_buffer_copy(frames[0])
    bad.dst[0] = bad.src[0]  # dst overlapping src
    with pytest.raises(load_package().InvalidInput):
        acc.alf_plan([bad])
    bad = abi.AlfFrame.from_buffer_copy(frames[0])
    bad.ctb_size = 8
    with pytest.raises(load_package().InvalidInput):
        acc.alf_plan([bad])
    aps = _alf_aps(abi, rng)
    aps.luma_coeff[0][0] = 129
    bad = abi.AlfFrame.from_buffer_copy(frames[0])
    bad.luma_aps = ctypes.addressof(aps)
    with pytest.raises(load_package().InvalidInput):
        acc.alf_plan([bad])

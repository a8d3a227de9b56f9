"""GPU deblocking (librmh) against the sequential oracle driver, bit-exact.

The oracle filters edges one after another in the reference order
(apply_deblock, vvc/decode_general.rs:4977-5187); the GPU filters every edge of
a direction at once. Equality on random bS and content is the proof that the
edge-parallel schedule is exact for these filter lengths."""
import ctypes

import numpy as np
import pytest

import workloads as W
from frames import dbk_frame_struct

pytestmark = pytest.mark.gpu


def _run(acc, po, planes, edges, bd, qps, beta=(0, 0, 0), tc=(0, 0, 0)):
    got = [p.copy() for p in planes]
    want = [p.copy() for p in planes]
    f1, k1 = dbk_frame_struct(got, edges, bd, qps, beta, tc)
    f2, k2 = dbk_frame_struct(want, edges, bd, qps, beta, tc)
    acc.deblock([f1])
    arr = (type(f2) * 1)(f2)
    po.lib().orc_deblock_frames(ctypes.addressof(arr), 1)
    for c in range(3):
        bad = np.argwhere(got[c] != want[c])
        assert bad.size == 0, "plane %d first mismatch at %s (%d mismatches)" % (c, bad[0], len(bad))
    changed = sum(int((got[c] != planes[c]).sum()) for c in range(3))
    assert changed > 0


def test_1080p10_random_bs_bit_exact(acc, po):
    print("deblock 1080p10 seed 0", flush=True)
    d = W.inter_frame(0)
    _run(acc, po, d["ref"], d["edges"], 10, d["qp"])


@pytest.mark.parametrize("seed,bd,qp,beta,tc", [(1, 8, 37, (0, 0, 0), (0, 0, 0)), (2, 10, 45, (2, -1, 1), (3, 2, -2)),
                                                (3, 12, 30, (-3, 0, 0), (6, 0, 0))])
def test_blocky_content_all_filter_forms(acc, po, seed, bd, qp, beta, tc):
    """Piecewise-flat CU content with small offsets triggers strong, long and
    weak filters and both chroma forms."""
    print("deblock blocky seed %d bd %d" % (seed, bd), flush=True)
    d = W.inter_frame(seed, width=512, height=288, bit_depth=bd)
    rng = np.random.default_rng(seed)
    maxv = (1 << bd) - 1
    planes = []
    for c, p in enumerate(d["ref"]):
        q = np.zeros_like(p)
        s = 1 if c == 0 else 2
        for (x, y, w, h) in d["cus"]:
            q[y // s:(y + h) // s, x // s:(x + w) // s] = rng.integers(maxv // 4, 3 * maxv // 4)
        q = np.clip(q.astype(np.int64) + rng.integers(-2, 3, q.shape) + (d["ref"][c].astype(np.int64) >> 6), 0, maxv)
        planes.append(np.ascontiguousarray(q.astype(np.uint16 if bd > 8 else np.uint8)))
    _run(acc, po, planes, d["edges"], bd, [qp, qp - 1, qp - 1], beta, tc)


def test_4x4_grid_edges(acc, po):
    """Edges on the 4-sample grid with lengths (1, 1), as the general decoder
    produces for 4-wide transform blocks (decode_general.rs:4990-4993)."""
    rng = np.random.default_rng(9)
    w, h, bd = 128, 64, 10
    Y = W.textured_plane(rng, w, h, bd)
    U, V = W.textured_plane(rng, w // 2, h // 2, bd), W.textured_plane(rng, w // 2, h // 2, bd)
    e = np.zeros((h // 4, w // 4, 4), np.uint8)
    for j in range(h // 4):
        for i in range(w // 4):
            if i > 0:
                e[j, i, 0] = W.edge_byte(int(rng.integers(0, 3)), 1, 1)
            if j > 0:
                e[j, i, 1] = W.edge_byte(int(rng.integers(0, 3)), 1, 1)
    _run(acc, po, [Y, U, V], e, bd, [40, 39, 39])


@pytest.mark.parametrize("w,h,bd", [(200, 120, 8), (200, 136, 10), (456, 72, 12), (196, 100, 10)])
def test_partial_tiles_and_unaligned_rows(acc, po, w, h, bd):
    """Widths that end inside a 64-sample tile, and planes whose row pitch is
    not a multiple of the 8- or 4-sample chunks the kernel moves as one access
    (8-bit chroma rows of 100 bytes, 10-bit chroma rows of 200 bytes): chunks
    crossing the right picture edge and unaligned planes go sample by sample.
    Frames are cropped from a larger generated frame, edge map included."""
    print("deblock cropped %dx%d bd %d" % (w, h, bd), flush=True)
    d = W.inter_frame(11 + w + h, width=(w + 63) // 64 * 64, height=(h + 31) // 32 * 32, bit_depth=bd)
    dt = np.uint8 if bd == 8 else np.uint16
    cw, ch = (w + 1) // 2, (h + 1) // 2
    planes = [np.ascontiguousarray(d["ref"][0][:h, :w].astype(dt)),
              np.ascontiguousarray(d["ref"][1][:ch, :cw].astype(dt)),
              np.ascontiguousarray(d["ref"][2][:ch, :cw].astype(dt))]
    edges = np.ascontiguousarray(d["edges"][:(h + 3) // 4, :(w + 3) // 4])
    _run(acc, po, planes, edges, bd, d["qp"])

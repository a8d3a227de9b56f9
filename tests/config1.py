"""The intra-subset workload: 8-bit 4:2:0 synthetic frames coded at one QP.

Source per frame f (seed 1000 + f): Y = clip(round(255 (x + y) / 126 + N(0, 8^2)),
0, 255); U, V = clip(round(128 + N(0, 4^2))). Sizes not a multiple of 32 are
padded by edge replication (pad_i420, crates/rustmpeg-video/src/vvc/mod.rs:891-915)."""
import hashlib

import numpy as np


def source(f, w=64, h=64):
    rng = np.random.default_rng(1000 + f)
    yy, xx = np.mgrid[0:h, 0:w]
    Y = np.clip(np.rint(255 * (xx + yy) / 126 + rng.normal(0, 8, (h, w))), 0, 255).astype(np.uint8)
    U = np.clip(np.rint(128 + rng.normal(0, 4, (h // 2, w // 2))), 0, 255).astype(np.uint8)
    V = np.clip(np.rint(128 + rng.normal(0, 4, (h // 2, w // 2))), 0, 255).astype(np.uint8)
    return Y, U, V


def pad(planes, w32, h32):
    out = []
    for i, p in enumerate(planes):
        tw, th = (w32, h32) if i == 0 else (w32 // 2, h32 // 2)
        sy = np.minimum(np.arange(th), p.shape[0] - 1)
        sx = np.minimum(np.arange(tw), p.shape[1] - 1)
        out.append(np.ascontiguousarray(p[sy][:, sx]))
    return out


def md5(planes):
    h = hashlib.md5()
    for p in planes:
        h.update(np.ascontiguousarray(p).tobytes())
    return h.hexdigest()

"""Writes the golden fixtures of tests/golden/ from the oracle.

The reference (Rust) cannot be built or imported here, so these vectors come
from the oracle (oracle/, itself pinned by the reference's known answers and
the invariants in tests/test_oracle_*.py). Freezing them pins the oracle and
the HIP path against later changes of either: tests/test_golden.py checks
both against these bytes. Inputs are seeded synthetic workloads
(tests/workloads.py) at small sizes.

    python tests/golden/make_fixtures.py      (rewrites the .npz files)
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))

import workloads as W  # noqa: E402
import pyoracle as po  # noqa: E402
from frames import dbk_frame_struct, mc_bi_frame_struct, mc_frame_struct  # noqa: E402


def residual_cases():
    out = {}
    for name, kw in (("residual_tools_10bit", dict(bit_depth=10, with_lfnst=True, with_ts=True, with_dq=True)),
                     ("residual_plain_8bit", dict(bit_depth=8))):
        tus, coeff, _ = W.residual_frame(77, width=128, height=64, **kw)
        out[name] = dict(tus=tus, coeff=coeff, residual=po.residual_batch(tus, coeff))
    return out


def mc_case():
    d = W.inter_frame(78, width=96, height=64, bit_depth=10, mv_range=24)
    refs = [np.ascontiguousarray(p) for p in d["ref"]]
    dst = [np.zeros_like(p) for p in refs]
    f, keep = mc_frame_struct(refs, dst, 10)
    assert po.mc_batch(f, d["pus"]) == 0
    return dict(ref_y=refs[0], ref_u=refs[1], ref_v=refs[2], pus=d["pus"], out_y=dst[0], out_u=dst[1], out_v=dst[2])


def mc_bi_case():
    d = W.bi_frame(79, width=96, height=64, bit_depth=10, mv_range=24)
    r0 = [np.ascontiguousarray(p) for p in d["ref0"]]
    r1 = [np.ascontiguousarray(p) for p in d["ref1"]]
    dst = [np.zeros_like(p) for p in r0]
    f, keep = mc_bi_frame_struct(r0, r1, dst, 10)
    assert po.mc_bi_batch(f, d["pus"]) == 0
    return dict(ref0_y=r0[0], ref0_u=r0[1], ref0_v=r0[2], ref1_y=r1[0], ref1_u=r1[1], ref1_v=r1[2], pus=d["pus"],
                out_y=dst[0], out_u=dst[1], out_v=dst[2])


def deblock_case():
    # piecewise-flat CU content so strong, long and weak filters all occur
    d = W.inter_frame(80, width=128, height=64, bit_depth=10)
    rng = np.random.default_rng(80)
    planes = []
    for c, p in enumerate(d["ref"]):
        q = np.zeros_like(p)
        s = 1 if c == 0 else 2
        for (x, y, w, h) in d["cus"]:
            q[y // s:(y + h) // s, x // s:(x + w) // s] = rng.integers(256, 768)
        q = np.clip(q.astype(np.int64) + rng.integers(-2, 3, q.shape) + (p.astype(np.int64) >> 6), 0, 1023)
        planes.append(np.ascontiguousarray(q.astype(np.uint16)))
    qps = np.array([37, 36, 36], np.uint8)
    out = [p.copy() for p in planes]
    f, keep = dbk_frame_struct(out, d["edges"], 10, list(qps))
    arr = (type(f) * 1)(f)
    po.lib().orc_deblock_frames(ctypes.addressof(arr), 1)
    return dict(in_y=planes[0], in_u=planes[1], in_v=planes[2], edges=d["edges"], qp=qps,
                out_y=out[0], out_u=out[1], out_v=out[2])


def main():
    cases = residual_cases()
    cases["mc_uni_10bit"] = mc_case()
    cases["mc_bi_10bit"] = mc_bi_case()
    cases["deblock_10bit"] = deblock_case()
    for name, arrays in cases.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrays)
        print("%s: %d bytes" % (path, os.path.getsize(path)))


if __name__ == "__main__":
    main()

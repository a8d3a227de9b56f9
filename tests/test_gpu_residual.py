"""GPU residual stage (librmh, through the C ABI) against the oracle, bit-exact."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def _check(acc, po, tus, coeff):
    ref = po.residual_batch(tus, coeff)
    got = acc.residuals(tus, coeff)
    bad = np.nonzero(got != ref)[0]
    if bad.size:
        off = tus["offset"]
        i = int(np.searchsorted(off, bad[0], side="right") - 1)
        raise AssertionError("first mismatch at sample %d in TU %d %s: got %d want %d (%d mismatches)"
                             % (bad[0], i, tus[i], got[bad[0]], ref[bad[0]], bad.size))


def test_full_1080p10_frame_bit_exact(acc, po):
    print("residual 1080p10 frame seed 0", flush=True)
    tus, coeff, _ = W.residual_frame(0)
    _check(acc, po, tus, coeff)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_tool_mix_bit_exact(acc, po, bd):
    print("residual tool mix bd=%d" % bd, flush=True)
    tus, coeff, _ = W.residual_frame(10 + bd, width=480, height=272, bit_depth=bd, with_lfnst=True, with_ts=True,
                                     with_dq=True, qp_range=(4, 51))
    _check(acc, po, tus, coeff)


def test_every_size_type_and_skip(acc, po):
    """Every (w, h) from 1x2 to 64x64 with every supported type pair, extra
    zero-outs, LFNST modes on both layouts, dense and extreme coefficients."""
    rng = np.random.default_rng(5)
    descs, blocks, off = [], [], 0
    for lw in range(0, 7):
        for lh in range(0, 7):
            w, h = 1 << lw, 1 << lh
            if w * h < 2:
                continue
            for _ in range(6):
                t = np.zeros(1, W.TU_DTYPE)[0]
                t["offset"], t["log2_w"], t["log2_h"] = off, lw, lh
                t["bit_depth"] = int(rng.choice([8, 10, 12]))
                t["qp"] = int(rng.integers(0, 64 + 6 * (t["bit_depth"] - 8)))
                ok_h = [0] + ([1, 2] if 4 <= w <= 32 else [])
                ok_v = [0] + ([1, 2] if 4 <= h <= 32 else [])
                t["tr_hor"], t["tr_ver"] = int(rng.choice(ok_h)), int(rng.choice(ok_v))
                if w >= 4 and h >= 4 and rng.random() < 0.4:
                    t["lfnst_idx"], t["lfnst_mode"] = int(rng.integers(1, 3)), int(rng.integers(0, 95))
                if rng.random() < 0.3 and w > 2:
                    t["skip_w"] = int(rng.integers(0, w // 2))
                if rng.random() < 0.3 and h > 2:
                    t["skip_h"] = int(rng.integers(0, h // 2))
                r = rng.random()
                if r < 0.1:
                    t["flags"] = 1
                elif r < 0.25:
                    t["flags"] = 2
                kind = rng.integers(0, 3)
                if kind == 0:
                    c = rng.integers(-32768, 32768, w * h)
                elif kind == 1:
                    c = np.rint(rng.laplace(0, 40, w * h))
                else:
                    c = np.where(rng.random(w * h) < 0.1, rng.integers(-300, 300, w * h), 0)
                descs.append(t)
                blocks.append(np.clip(c, -32768, 32767).astype(np.int16))
                off += w * h
    tus = np.array(descs, dtype=W.TU_DTYPE)
    _check(acc, po, tus, np.concatenate(blocks))


def test_empty_and_all_zero(acc, po):
    tus, coeff, _ = W.residual_frame(3, width=64, height=64)
    z = np.zeros_like(coeff)
    assert (acc.residuals(tus, z) == 0).all()
    assert acc.residuals(tus[:0], coeff).size == coeff.size


def test_untouched_output_outside_tus_and_on_error(acc, rmh):
    tus, coeff, _ = W.residual_frame(4, width=64, height=64)
    out = np.full(coeff.size + 16, 1234, np.int16)
    c2 = np.concatenate([coeff, np.zeros(16, np.int16)])
    acc.residuals(tus, c2, out)
    assert (out[-16:] == 1234).all()
    sentinel = out.copy()
    bad = tus.copy()
    bad[3]["log2_w"] = 7
    with pytest.raises(rmh.InvalidInput):
        acc.residuals(bad, c2, out)
    assert (out == sentinel).all()
    acc.inject_failures(1)
    with pytest.raises(rmh.Execution):
        acc.residuals(tus, c2, out)
    assert (out == sentinel).all()


@pytest.mark.parametrize("seed", [21, 22])
def test_large_tile_classes_extreme_values(acc, po, seed):
    """Tile-eligible TUs of every class with both sides >= 16, every transform
    pair, bit depths 8-12 and QPs 0-63: full-range, saturated and sparse
    coefficients, so that both stages meet their extremes (-32768, 32767,
    intermediates clipped in the first stage, dequantisation saturating)."""
    print("residual large classes seed %d" % seed, flush=True)
    rng = np.random.default_rng(seed)
    descs, blocks, off = [], [], 0
    for lw in range(4, 7):
        for lh in range(4, 7):
            w, h = 1 << lw, 1 << lh
            for i in range(24):
                t = np.zeros(1, W.TU_DTYPE)[0]
                t["offset"], t["log2_w"], t["log2_h"] = off, lw, lh
                t["bit_depth"] = int(rng.choice([8, 10, 12]))
                t["qp"] = int(rng.integers(0, 64 + 6 * (t["bit_depth"] - 8)))
                ok_h = [0] + ([1, 2] if w <= 32 else [])
                ok_v = [0] + ([1, 2] if h <= 32 else [])
                t["tr_hor"], t["tr_ver"] = int(rng.choice(ok_h)), int(rng.choice(ok_v))
                kind = i % 4
                if kind == 0:
                    c = rng.integers(-32768, 32768, w * h)
                elif kind == 1:
                    c = rng.choice([-32768, 32767], w * h)
                elif kind == 2:
                    c = np.rint(rng.laplace(0, 2000, w * h))
                else:
                    c = np.where(rng.random(w * h) < 0.05, rng.integers(-32768, 32768, w * h), 0)
                descs.append(t)
                blocks.append(np.clip(c, -32768, 32767).astype(np.int16))
                off += w * h
    tus = np.array(descs, dtype=W.TU_DTYPE)
    _check(acc, po, tus, np.concatenate(blocks))


def test_dequant_scale_readout_every_class_qp_and_bit_depth(acc, po):
    """One DC level of +-1 and +-37 per TU, for every size class (tiled and
    generic kernels), every qp % 6 over the whole QP range and bit depths
    8..12: the dequantisation scales and shifts of both kernels, odd-area
    (sqrt 2) classes included, against the oracle (whose scales are checked
    against the text of vvc/dct.rs:37-40)."""
    descs, off = [], 0
    for bd in range(8, 13):
        for lw in range(0, 7):
            for lh in range(0, 7):
                w, h = 1 << lw, 1 << lh
                if w * h < 2:
                    continue
                for qp in range(0, 64 + 6 * (bd - 8), 5):
                    t = np.zeros(1, W.TU_DTYPE)[0]
                    t["offset"], t["log2_w"], t["log2_h"], t["bit_depth"], t["qp"] = off, lw, lh, bd, qp
                    descs.append(t)
                    off += w * h
                    off = (off + 3) & ~3
    tus = np.array(descs, W.TU_DTYPE)
    coeff = np.zeros(off, np.int16)
    vals = np.array([1, -1, 37, -37], np.int16)
    coeff[tus["offset"]] = vals[np.arange(len(tus)) % 4]
    print("dequant readout: %d TUs" % len(tus), flush=True)
    _check(acc, po, tus, coeff)


@pytest.mark.parametrize("seed", [31, 32])
def test_dependent_quantisation_in_tiles(acc, po, seed):
    """Dependent quantisation applied by the tiled kernel (TUs whose coded
    region is their whole scan region), mixed in tiles with plain TUs, with
    LFNST on 4x4 and 8x8 blocks, next to TUs the prepass keeps (MTS-32
    zero-out, caller zero-out, LFNST on other sizes): every tile class, bit
    depths 8-12, QPs 0-63, full-range, saturating, sparse and all-odd levels
    (the last make every group's parity XORs non-zero)."""
    print("residual dependent quantisation in tiles seed %d" % seed, flush=True)
    rng = np.random.default_rng(seed)
    descs, blocks, off = [], [], 0
    for lw in range(2, 7):
        for lh in range(2, 7):
            w, h = 1 << lw, 1 << lh
            for i in range(12):
                t = np.zeros(1, W.TU_DTYPE)[0]
                t["offset"], t["log2_w"], t["log2_h"] = off, lw, lh
                t["bit_depth"] = int(rng.choice([8, 10, 12]))
                t["qp"] = int(rng.integers(0, 64 + 6 * (t["bit_depth"] - 8)))
                ok_h = [0] + ([1, 2] if w <= 32 else [])
                ok_v = [0] + ([1, 2] if h <= 32 else [])
                t["tr_hor"], t["tr_ver"] = int(rng.choice(ok_h)), int(rng.choice(ok_v))
                t["flags"] = W.DQ_FLAG if i % 4 != 3 else 0
                if (w, h) in ((4, 4), (8, 8), (8, 16)) and i % 3 == 0:
                    t["lfnst_idx"], t["lfnst_mode"] = int(rng.integers(1, 3)), int(rng.integers(0, 95))
                    t["tr_hor"] = t["tr_ver"] = 0
                if i == 10 and h > 4:
                    t["skip_h"] = 2
                kind = i % 4
                if kind == 0:
                    c = rng.integers(-32768, 32768, w * h)
                elif kind == 1:
                    c = rng.choice([-32768, 32767, -1, 1], w * h)
                elif kind == 2:
                    c = np.where(rng.random(w * h) < 0.15, rng.integers(-60, 61, w * h), 0)
                else:
                    c = 2 * rng.integers(-30, 31, w * h) + 1
                descs.append(t)
                blocks.append(np.clip(c, -32768, 32767).astype(np.int16))
                off += w * h
    tus = np.array(descs, dtype=W.TU_DTYPE)
    _check(acc, po, tus, np.concatenate(blocks))

"""Pins the residual-stage oracle (oracle/orc_residual.c, orc_tables.c).

Known answers are the reference's own tests (crates/rustmpeg-video/src/vvc/
dct.rs:1304-1491, transform.rs:718-842, lfnst.rs:307-389,
decode_general.rs:5741-5750); the rest are properties the mathematics fixes.
"""
import math

import numpy as np
import pytest

from lcg import Lcg

DCT2, DST7, DCT8 = 0, 1, 2


# ---------------------------------------------------------------- matrices

@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64])
def test_dct2_rows_sample_the_scaled_cosine(po, n):
    m = po.transform_matrix(DCT2, n).astype(np.float64)
    assert (m[0] == 64).all()
    k = np.arange(1, n)[:, None]
    j = np.arange(n)[None, :]
    ideal = 64 * math.sqrt(2) * np.cos(np.pi * k * (2 * j + 1) / (2 * n))
    assert np.abs(m[1:] - ideal).max() <= 1.5


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_dst7_and_dct8_sample_the_scaled_sine_and_cosine(po, n):
    amp = 128 * math.sqrt(n / (2 * n + 1))
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    dst7 = amp * np.sin(np.pi * (2 * k + 1) * (j + 1) / (2 * n + 1))
    dct8 = amp * np.cos(np.pi * (2 * k + 1) * (2 * j + 1) / (4 * n + 2))
    assert np.abs(po.transform_matrix(DST7, n) - dst7).max() <= 1.5
    assert np.abs(po.transform_matrix(DCT8, n) - dct8).max() <= 1.5


@pytest.mark.parametrize("t,n", [(DCT2, n) for n in (4, 8, 16, 32, 64)] + [(t, n) for t in (DST7, DCT8) for n in (4, 8, 16, 32)])
def test_matrices_are_nearly_orthogonal(po, t, n):
    m = po.transform_matrix(t, n).astype(np.int64)
    g = m @ m.T
    diag = np.diag(g)
    assert np.abs(diag - 4096 * n).max() < 0.02 * 4096 * n
    off = g - np.diag(diag)
    assert np.abs(off).max() < 0.02 * 4096 * n


def test_unsupported_matrices_are_rejected(po):
    for t, n in [(DST7, 2), (DST7, 64), (DCT8, 64), (DCT2, 3), (DCT2, 128), (DCT8, 1), (5, 4)]:
        assert po.transform_matrix(t, n) is None


# ------------------------------------------------------ reference known answers

def test_constant_block_is_dc_only(po):
    # vvc/dct.rs:1308-1322
    for n in [4, 8, 16, 32]:
        for v in [-40, 0, 37]:
            c = po.forward_transform_square(n, np.full(n * n, v), DCT2, DCT2)
            assert c[0, 0] == v * 128
            assert np.count_nonzero(c.reshape(-1)[1:]) == 0


def test_inverse_of_dc_only_is_constant(po):
    # vvc/dct.rs:1325-1337
    for n in [4, 8, 16, 32]:
        c = np.zeros(n * n, np.int32)
        c[0] = 128 * 64
        r = po.inverse_transform(n, n, c, DCT2, DCT2, bd=8)
        assert (r == (8192 * 64 * 64 + (1 << 18)) >> 19).all()


def test_forward_inverse_round_trip_error_is_bounded(po):
    # vvc/dct.rs:1340-1358
    g = Lcg(0x12345678)
    for n in [4, 8, 16, 32]:
        res = np.array([((g.next_u32() >> 24) % 511) - 255 for _ in range(n * n)], np.int32)
        c = po.forward_transform_square(n, res, DCT2, DCT2)
        rt = po.inverse_transform(n, n, c, DCT2, DCT2, bd=8)
        assert np.abs(rt.reshape(-1) - res).max() <= 2


def test_quantise_dequantise_round_trip_stays_in_range(po):
    # vvc/dct.rs:1360-1376
    for n in [4, 8, 16, 32]:
        res = np.array([((i * 37) % 401) - 200 for i in range(n * n)], np.int32)
        c = po.forward_transform_square(n, res, DCT2, DCT2)
        for qp in [4, 10, 26, 40, 63]:
            lv = po.quant_dct(n, c, qp)
            dq = po.dequant_dct(n, lv, qp)
            assert dq.min() >= -32768 and dq.max() <= 32767


def test_invalid_square_sizes_are_rejected(po):
    # vvc/dct.rs:1378-1387
    for n in [1, 2, 3, 6, 64, 128]:
        z = np.zeros(max(n * n, 1), np.int32)
        assert po.quant_dct(n, z, 26) is None
        assert po.dequant_dct(n, z, 26) is None
        assert po.forward_transform_square(n, z, DCT2, DCT2) is None


def test_mts_matrices_round_trip_within_two_units(po):
    # vvc/dct.rs:1397-1435 (one LCG stream across sizes)
    g = Lcg(0xA5A51234)
    for n in [4, 8, 16, 32]:
        res = np.array([((g.next_u32() >> 24) % 255) - 127 for _ in range(n * n)], np.int32)
        for th, tv in [(DST7, DST7), (DCT8, DST7), (DST7, DCT8), (DCT8, DCT8)]:
            c = po.forward_transform_square(n, res, th, tv)
            rt = po.inverse_transform(n, n, c, th, tv, bd=8) if n < 32 else None
            if n == 32:
                # the square reference inverse takes explicit skips of 0 at 32
                rt = _inverse_square_no_default_skip(po, n, c, th, tv)
            assert np.abs(rt.reshape(-1) - res).max() <= 2, (n, th, tv)
            lv = po.quant_dct(n, c, 26)
            assert lv.min() >= -32768 and lv.max() <= 32767


def _inverse_square_no_default_skip(po, n, c, th, tv):
    """inverse_transform (vvc/dct.rs:427-472) applies no default zero-out; the
    rectangular one does. Compose the square one from the matrices."""
    mh = po.transform_matrix(th, n).astype(np.int64)
    mv = po.transform_matrix(tv, n).astype(np.int64)
    c = np.asarray(c, np.int64).reshape(n, n)
    tmp = np.clip((mv.T @ c + 64) >> 7, -32768, 32767)          # tmp[y][i]
    return np.clip((tmp @ mh + (1 << 11)) >> 12, -32768, 32767)  # out[y][x]


def test_mts_32_point_zeroes_high_frequencies(po):
    # vvc/dct.rs:1437-1453
    res = np.array([max((i % 32) * 3 - 48, -64) for i in range(1024)], np.int32)
    for th, tv in [(DST7, DST7), (DCT8, DST7), (DST7, DCT8), (DCT8, DCT8)]:
        c = po.forward_transform_square(32, res, th, tv, 16, 16)
        assert not c[16:, :].any() and not c[:, 16:].any()


def test_mts_rejects_invalid_zero_out(po):
    # vvc/dct.rs:1455-1461
    z = np.zeros(16, np.int32)
    assert po.forward_transform_square(4, z, DST7, DST7, 4, 0) is None
    assert po.inverse_transform(4, 4, z, DCT8, DCT8, 0, 4) is None
    assert po.forward_transform_square(4, z, DST7, DCT8, 5, 0) is None


def test_dct2_64_constant_block(po):
    # vvc/dct.rs:1463-1475
    c = po.forward_transform(64, 64, np.full(4096, 25), DCT2, DCT2, bd=8)
    assert c[0, 0] == 25 * 128
    assert np.count_nonzero(c.reshape(-1)[1:]) == 0
    r = po.inverse_transform(64, 64, c, DCT2, DCT2, bd=8)
    assert (r == 25).all()


def test_rectangular_transforms_and_dep_quant_round_trip(po):
    # vvc/dct.rs:1477-1490 (identity scan, as the reference test uses)
    for w, h in [(4, 8), (8, 4), (8, 16), (16, 8), (16, 32)]:
        res = np.array([((i * 19) % 101) - 50 for i in range(w * h)], np.int32)
        c = po.forward_transform(w, h, res, DCT2, DCT2, bd=8)
        lv = po.quant_dep_quant(w, h, c, 22, 8)
        dq = po.dequant_dep_quant(w, h, lv, 22, 8)
        r = po.inverse_transform(w, h, dq, DCT2, DCT2, bd=8)
        assert r.shape == (h, w)


def test_transform_skip_dequant_known_answers(po):
    # vvc/transform.rs:796-811
    assert po.dequant_ts(1, 22) == 8
    assert po.dequant_ts(-1, 22) == -8
    assert po.dequant_ts(1000, 63) == 32767
    assert po.dequant_ts(-1000, 63) == -32768
    for v in [-200, -1, 0, 1, 37, 200]:
        lv = po.quant_ts(v, 4)
        assert lv == v and po.dequant_ts(lv, 4) == v


# -------------------------------------------------------------- scans

def test_diagonal_scan_4x4(po):
    # vvc/transform.rs:758-783
    xy = [(r % 4, r // 4) for r in po.coeff_scan(4, 4)]
    assert xy == [(0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0), (0, 3), (1, 2), (2, 1), (3, 0),
                  (1, 3), (2, 2), (3, 1), (2, 3), (3, 2), (3, 3)]


def test_large_scan_covers_top_left_32x32_with_full_stride(po):
    # vvc/transform.rs:721-755
    s = po.coeff_scan(64, 64)
    assert len(s) == 1024 and len(set(s.tolist())) == 1024
    assert ((s % 64) < 32).all() and ((s // 64) < 32).all()
    r = po.coeff_scan(64, 16)
    assert len(r) == 512 and ((r % 64) < 32).all()


def test_grouped_scan_covers_every_coefficient_once(po):
    # vvc/transform.rs:785-795, widened to every rectangular size
    for lw in range(0, 6):
        for lh in range(0, 6):
            w, h = 1 << lw, 1 << lh
            if w * h < 2:
                continue
            s = po.coeff_scan(w, h)
            assert sorted(s.tolist()) == list(range(w * h)), (w, h)


# ------------------------------------------------------------ quantisers

def test_dequant_at_zero_transform_shift_equals_transform_skip(po):
    """At bd=10, 32x32 has transform shift 15-10-5 = 0 and an even size, so the
    rectangular dequantiser must reduce to the transform-skip one."""
    rng = np.random.default_rng(7)
    lv = rng.integers(-3000, 3000, 1024)
    for qp in [4, 17, 22, 33, 40, 51, 63, 75]:
        dq = po.dequant_rect(32, 32, lv, qp, 10).reshape(-1)
        ts = np.array([po.dequant_ts(int(v), qp) for v in lv])
        assert (dq == ts).all(), qp


def test_dequant_doubles_every_six_qp(po):
    rng = np.random.default_rng(8)
    lv = rng.integers(-100, 100, 64)
    for qp in range(4, 40):
        a = po.dequant_rect(8, 8, lv, qp, 8)
        b = po.dequant_rect(8, 8, lv, qp + 6, 8)
        # once the shift is a left shift the scaling is exact
        if 6 - ((15 - 8 - 3) + qp // 6) <= 0:
            assert (b == np.clip(2 * a, -32768, 32767)).all()


def test_dep_quant_states_follow_the_parity_automaton(po):
    """With every level odd, the 4-state machine walks 0 -> 2 -> 1 -> 2 ...;
    the dequantised magnitudes are then (2|l| - (state>>1)) * scale."""
    w = h = 4
    lv = np.ones(16, np.int32)
    out = po.dequant_dep_quant(w, h, lv, 26, 8).reshape(-1)
    scan = po.coeff_scan(4, 4)
    qp = 27
    shift = 7 - qp // 6 - (15 - 8 - 2)
    scale = [40, 45, 51, 57, 64, 72][qp % 6]
    state = 0
    for s in range(15, -1, -1):
        q = 2 - (state >> 1)
        expect = (q * scale * (1 << max(-shift, 0)) + ((1 << (shift - 1)) if shift > 0 else 0)) >> max(shift, 0)
        assert out[scan[s]] == expect
        state = (32040 >> ((state << 2) + 2)) & 3


def test_dep_quant_of_all_zero_block_is_zero(po):
    assert not po.dequant_dep_quant(8, 8, np.zeros(64), 30, 10).any()


# ---------------------------------------------------------------- LFNST

def _lcg_block(n, seed):
    g = Lcg(seed)
    return np.array([((g.next_u32() >> 20) % 81) - 40 for _ in range(n * n)], np.int32)


def test_lfnst_planar_and_dc_never_transpose(po):
    # vvc/lfnst.rs:320-326
    for mode in [0, 1]:
        assert not po.lib().orc_lfnst_transposes(mode)
        assert po.lib().orc_lfnst_set(mode) == 0


def test_lfnst_rejects_invalid_indexes_and_sizes(po):
    # vvc/lfnst.rs:328-335
    z = np.zeros(16, np.int32)
    assert po.forward_lfnst(4, 4, z, 0, 0) is None
    assert po.forward_lfnst(4, 4, z, 3, 0) is None
    assert po.inverse_lfnst(4, 4, z, 0, 0) is None
    assert po.inverse_lfnst(2, 8, np.zeros(16), 1, 0) is None


def test_lfnst_forward_zeroes_outside_low_frequency_region(po):
    # vvc/lfnst.rs:337-357
    scan8 = po.coeff_scan(8, 8)
    for n in [4, 8, 16, 32]:
        for idx in [1, 2]:
            c = po.forward_lfnst(n, n, _lcg_block(n, 0x12345678), idx, 0)
            zero_out = 8 if n in (4, 8) else 16
            scan_len = 48 if n >= 8 else 16
            nz = 0
            for k in range(scan_len):
                x, y = scan8[k] % 8, scan8[k] // 8
                if c[y, x] != 0:
                    assert k < zero_out
                    nz += 1
            assert nz <= zero_out


def test_lfnst_inverse_of_forward_preserves_low_frequencies(po):
    # vvc/lfnst.rs:359-378
    for n in [8, 16, 32]:
        for idx in [1, 2]:
            orig = _lcg_block(n, 0x9ABCDEF0 ^ n ^ idx)
            f = po.forward_lfnst(n, n, orig, idx, 0)
            r = po.inverse_lfnst(n, n, f, idx, 0)
            assert int(((orig.reshape(n, n).astype(np.int64) - r) ** 2).sum()) < 200_000_000


def test_lfnst_transposed_modes_round_trip(po):
    # vvc/lfnst.rs:380-389
    for mode in [35, 50, 66, 81, 94]:
        assert po.lib().orc_lfnst_transposes(mode)
        for n in [4, 8, 16]:
            f = po.forward_lfnst(n, n, _lcg_block(n, 0x55AA0000 ^ mode), 1, mode)
            r = po.inverse_lfnst(n, n, f, 1, mode)
            assert r.min() >= -32768 and r.max() <= 32767


def test_lfnst_kernels_are_nearly_orthonormal(po):
    """Each LFNST kernel is an integer approximation of an orthonormal matrix
    scaled by 128: inverse(forward(x)) reproduces x on the kept subspace."""
    for s in range(4):
        for idx in range(2):
            for big, n_in in [(1, 48), (0, 16)]:
                m = np.array([[po.lib().orc_lfnst_coeff(s, idx, big, i, j) for j in range(n_in)]
                              for i in range(16)], np.int64)
                g = m @ m.T
                assert np.abs(np.diag(g) - 128 * 128).max() < 0.05 * 128 * 128
                assert np.abs(g - np.diag(np.diag(g))).max() < 0.05 * 128 * 128


def test_lfnst_transpose_is_a_transpose(po):
    """A transposed mode with the same matrix set applied to a transposed
    block gives the transposed result (mode 2 -> set 1, mode 66 -> set 1)."""
    rng = np.random.default_rng(3)
    for n in [4, 8, 16]:
        src = np.zeros((n, n), np.int32)
        src.reshape(-1)[:] = 0
        scan8 = po.coeff_scan(8, 8)
        for k in range(16):
            x, y = scan8[k] % 8, scan8[k] // 8
            src[y, x] = rng.integers(-300, 300)
        a = po.inverse_lfnst(n, n, src, 1, 2)
        # the transposed layout reads the same 16 inputs in scan order, so feed them unchanged
        b = po.inverse_lfnst(n, n, src, 1, 66)
        sb = 8 if n >= 8 else 4
        if sb == 4:
            assert (a[:4, :4] == b[:4, :4].T).all()
        else:
            assert (a[:4, :8] == b[:8, :4].T).all()
            assert (a[4:8, :4] == b[:4, 4:8].T).all()


# --------------------------------------------------------- closed forms

def _np_inverse(po, w, h, c, th, tv, sw, sh, bd):
    mh = po.transform_matrix(th, w).astype(np.int64)
    mv = po.transform_matrix(tv, h).astype(np.int64)
    c = np.asarray(c, np.int64).reshape(h, w).copy()
    dw = (16 if (th != DCT2 and w == 32) else max(w - 32, 0))
    dh = (16 if (tv != DCT2 and h == 32) else max(h - 32, 0))
    cw, ch = w - max(dw, sw), h - max(dh, sh)
    c[ch:, :] = 0
    c[:, cw:] = 0
    tmp = np.clip((mv.T @ c + 64) >> 7, -32768, 32767)                # [j][i]
    s2 = 20 - bd
    return np.clip((tmp @ mh + (1 << (s2 - 1))) >> s2, -32768, 32767)  # [y][x]


def test_inverse_transform_equals_matrix_closed_form(po):
    rng = np.random.default_rng(11)
    for _ in range(300):
        lw, lh = rng.integers(1, 7, 2)
        w, h = 1 << int(lw), 1 << int(lh)
        th = int(rng.integers(0, 3)) if 4 <= w <= 32 else DCT2
        tv = int(rng.integers(0, 3)) if 4 <= h <= 32 else DCT2
        bd = int(rng.integers(8, 13))
        c = rng.laplace(0, 400, w * h).astype(np.int64).clip(-32768, 32767)
        sw = int(rng.integers(0, w // 2 + 1)) if rng.random() < 0.3 else 0
        sh = int(rng.integers(0, h // 2 + 1)) if rng.random() < 0.3 else 0
        got = po.inverse_transform(w, h, c, th, tv, sw, sh, bd)
        assert (got == _np_inverse(po, w, h, c, th, tv, sw, sh, bd)).all(), (w, h, th, tv, bd)


def test_one_dimensional_inverse_uses_single_stage(po):
    rng = np.random.default_rng(12)
    for n in [2, 4, 8, 16, 32, 64]:
        for bd in [8, 10]:
            c = rng.integers(-2000, 2000, n)
            m = po.transform_matrix(DCT2, n).astype(np.int64)
            cut = n - max(n - 32, 0)
            cc = c.copy()
            cc[cut:] = 0
            s = 21 - bd
            exp = np.clip((cc @ m + (1 << (s - 1))) >> s, -32768, 32767)
            assert (po.inverse_transform(1, n, c, DCT2, DCT2, bd=bd).reshape(-1) == exp).all()
            assert (po.inverse_transform(n, 1, c, DCT2, DCT2, bd=bd).reshape(-1) == exp).all()


def test_lfnst_skip_table(po):
    # decode_general.rs:5741-5750, exercised through the TU driver: with LFNST,
    # coefficients outside nonZeroW x nonZeroH must not reach the transform.
    from workloads import TU_DTYPE
    cases = {(4, 4): (0, 0), (8, 8): (0, 0), (16, 8): (8, 0), (8, 4): (4, 0), (4, 16): (0, 12),
             (16, 16): (8, 8), (64, 64): (56, 56)}
    for (w, h), (sw, sh) in cases.items():
        tus = np.zeros(1, TU_DTYPE)
        tus[0]["log2_w"], tus[0]["log2_h"] = int(math.log2(w)), int(math.log2(h))
        tus[0]["qp"], tus[0]["bit_depth"], tus[0]["lfnst_idx"] = 30, 8, 1
        base = np.zeros(w * h, np.int16)
        base[0] = 50
        probe = base.copy().reshape(h, w)
        if sw:
            probe[0, w - sw] = 900
        if sh:
            probe[h - sh, 0] = 900
        a = po.residual_batch(tus, base)
        b = po.residual_batch(tus, probe.reshape(-1))
        assert (a == b).all(), (w, h)

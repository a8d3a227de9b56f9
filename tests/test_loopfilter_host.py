"""The host halves of the LMCS and ALF plans, checked on the CPU against the
oracle (whose LMCS tables and ALF filter selection are pinned to the
reference text in test_alf_vs_reference_text.py):

  rmh_lmcs_tables      LmcsReshaper16::from_aps       vvc/filter.rs:2628-2705
  rmh_alf_filter_bank  the per-block coefficients and clips of alf_luma_block
                       (vvc/decode_general.rs:5454-5483) for every fixed set,
                       the APS, every class and every transposition
"""
import importlib

import numpy as np
import pytest

A = importlib.import_module("rustmpeg-next_amd.abi")


def _random_aps(rng, bd):
    lo = int(rng.integers(0, 10))
    hi = int(rng.integers(lo, 16))
    init = 1 << (bd - 4)
    delta = [int(max(d, -init)) for d in rng.integers(-init, init + 1, 16)]
    return A.lmcs_aps(lo, hi, delta, int(rng.integers(-7, 8)))


@pytest.mark.parametrize("bd", [8, 9, 10, 11, 12])
def test_lmcs_tables_equal_oracle(rmh, po, bd):
    rng = np.random.default_rng(100 + bd)
    cases = [A.lmcs_aps(0, 15, [0] * 16, 0), A.lmcs_aps(1, 14, [0, -3, -2, -1, 0, 1, 2, 2, 1, 0, -1, -2, -1, 0, 3, 0], -2),
             A.lmcs_aps(0, 15, [-(1 << (bd - 4))] * 15 + [3], 0),   # codewords fall short of the range
             A.lmcs_aps(5, 5, [0] * 5 + [200] + [0] * 10, 7)]
    cases += [_random_aps(rng, bd) for _ in range(40)]
    for aps in cases:
        got = rmh.lmcs_tables(aps, bd)
        want = po.lmcs_tables(aps, bd)
        for g, w in zip(got, want):
            assert (g == w).all(), (bd, list(aps.bin_cw_delta))
    fwd, inv, _, _ = rmh.lmcs_tables(cases[0], bd)
    assert (fwd == np.arange(1 << bd)).all() and (inv == np.arange(1 << bd)).all()


def test_lmcs_tables_reject_bad_input(rmh):
    with pytest.raises(rmh.InvalidInput):
        rmh.lmcs_tables(A.lmcs_aps(0, 15, [0] * 16, 0), 13)
    bad = A.lmcs_aps(0, 15, [0] * 16, 0)
    bad.reserved = 1
    with pytest.raises(rmh.InvalidInput):
        rmh.lmcs_tables(bad, 10)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_alf_filter_bank_equals_oracle_selection(rmh, po, bd):
    rng = np.random.default_rng(bd)
    n = int(rng.integers(1, 26))
    aps = A.alf_aps([rng.integers(-128, 129, 12).tolist() for _ in range(n)],
                    [rng.integers(0, 4, 12).tolist() for _ in range(n)], True,
                    rng.integers(0, 30, 25).tolist())
    bank = rmh.alf_filter_bank(aps, bd)
    for b in range(17):
        for cls in range(25):
            for t in range(4):
                fc, cl = po.alf_luma_filter(b + 1, cls, t, aps, bd)
                assert bank[b, cls, t, :12].tolist() == fc.tolist(), (b, cls, t)
                assert bank[b, cls, t, 12:].tolist() == cl.tolist(), (b, cls, t)
    empty = rmh.alf_filter_bank(None, bd)
    assert (empty[16, :, :, :12] == 0).all() and (empty[:16] == bank[:16]).all()

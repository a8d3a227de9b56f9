"""The form of dependent dequantisation that the tiled residual kernel uses,
checked against the oracle's dequantize_dep_quant_bd (orc_dequant_dep_quant,
vvc/dct.rs:1191-1248), which runs the reference's 4-state automaton serially
from the last non-zero level.

The kernel never runs the automaton. Its step from state x on a level of
parity p is x' = swap_bits(x) ^ 2p, an affine map over GF(2), so bit 1 of the
state on arrival at scan position s (the only bit the reconstruction offset
uses) is the XOR of the level parities at s + 1, s + 3, s + 5, ... A lane owns
one 4 x 4 coefficient group (16 consecutive scan positions); the XORs over the
later groups of its TU come from two ballots, the rest from a walk over its
own 16 levels. This restates that per-group computation in numpy and
compares it with the automaton on dense, sparse, saturating and all-odd
blocks of every tile size class, every QP class and bit depth.
"""
import numpy as np
import pytest

# The reference's transition table (vvc/dct.rs), 2-bit entries indexed by
# state * 2 + parity, and the affine form the kernel relies on.
_TABLE = 32040


def test_transition_table_is_affine_over_gf2():
    swap = lambda x: ((x & 1) << 1) | (x >> 1)
    for state in range(4):
        for parity in range(2):
            nxt = (_TABLE >> ((state << 2) + (parity << 1))) & 3
            assert nxt == swap(state) ^ (2 * parity)


def _dq_by_groups(w, h, levels, qp, bd, scan):
    """The kernel's computation, group by group, on a raster block."""
    qp = max(qp, 4)
    lsum = int(np.log2(w)) + int(np.log2(h))
    sqrt2 = lsum & 1
    tr_shift = 15 - bd - (lsum >> 1) - sqrt2
    qdq = qp + 1
    shift = 7 - qdq // 6 - tr_shift
    inv = ([57, 64, 72, 80, 90, 102] if sqrt2 else [40, 45, 51, 57, 64, 72])[qdq % 6]
    ls, rs = (-shift, 0) if shift < 0 else (0, shift)
    S = 65536 if ls >= 10 else inv << ls
    rnd = (1 << rs) >> 1
    flat = levels.reshape(-1).astype(np.int64)
    n_cg = len(scan) // 16
    lv = flat[np.asarray(scan)].reshape(n_cg, 16)  # levels in scan order, per group
    par = lv & 1
    xe = np.bitwise_xor.reduce(par[:, 0::2], axis=1)
    xo = np.bitwise_xor.reduce(par[:, 1::2], axis=1)
    out = np.zeros(w * h, np.int64)
    for c in range(n_cg):
        # the ballots: XORs over the later groups
        se = int(np.bitwise_xor.reduce(xe[c + 1:])) if c + 1 < n_cg else 0
        so = int(np.bitwise_xor.reduce(xo[c + 1:])) if c + 1 < n_cg else 0
        for i in range(15, -1, -1):
            level = int(lv[c, i])
            A = se if i & 1 else so
            if level:
                q = 2 * level + (-A if level > 0 else A)
                if ls > 0:
                    q = min(max(q, -32768), 32767)
                prod = q * S
                assert -2 ** 31 <= prod + rnd < 2 ** 31, "the kernel's 32-bit product"
                assert abs(q) < 2 ** 23 and S < 2 ** 17, "24-bit multiplier operands"
                out[scan[16 * c + i]] = min(max((prod + rnd) >> rs, -32768), 32767)
            if i & 1:
                so ^= level & 1
            else:
                se ^= level & 1
    return out.reshape(h, w)


@pytest.mark.parametrize("lw,lh", [(a, b) for a in range(2, 7) for b in range(2, 7)])
def test_group_form_equals_automaton(po, lw, lh):
    w, h = 1 << lw, 1 << lh
    scan = [int(s) for s in po.coeff_scan(w, h)]
    assert len(scan) == min(w, 32) * min(h, 32)
    # the kernel takes a group's origin from the first of its 16 scan entries
    # and walks the 4 x 4 diagonal inside it
    diag4 = [0, 4, 1, 8, 5, 2, 12, 9, 6, 3, 13, 10, 7, 14, 11, 15]
    for c in range(len(scan) // 16):
        x0, y0 = scan[16 * c] % w, scan[16 * c] // w
        assert x0 % 4 == 0 and y0 % 4 == 0
        assert scan[16 * c:16 * c + 16] == [(y0 + p // 4) * w + x0 + p % 4 for p in diag4]
    rng = np.random.default_rng([lw, lh])
    for k in range(24):
        bd = int(rng.choice([8, 10, 12]))
        qp = int(rng.integers(0, 64 + 6 * (bd - 8)))
        kind = k % 4
        if kind == 0:
            lv = rng.integers(-32768, 32768, (h, w))
        elif kind == 1:
            lv = np.where(rng.random((h, w)) < 0.2, rng.integers(-40, 41, (h, w)), 0)
        elif kind == 2:
            lv = rng.choice([-32768, 32767, -1, 1], (h, w))
        else:
            lv = 2 * rng.integers(-20, 21, (h, w)) + 1
        # levels outside the scan region (64-sample sides) take no part
        lv[:, 32:] = 0
        lv[32:, :] = 0
        want = po.dequant_dep_quant(w, h, lv.astype(np.int32), qp, bd)
        got = _dq_by_groups(w, h, lv, qp, bd, scan)
        assert np.array_equal(got, want), (lw, lh, bd, qp, kind)

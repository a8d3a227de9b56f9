"""The even/odd DCT-II butterfly of the tiled residual kernel's uniform passes,
checked against the oracle's DCT-II matrices (orc_transform_matrix, pinned to
vvc/dct.rs in test_tables_vs_reference.py).

For a DCT-II direction the kernel splits a row (or column) of int16 values
into the pairs (v[4q], v[4q + 2]) and (v[4q + 1], v[4q + 3]) and computes,
for each output n < N / 2, an even sum E and an odd sum O against a table of
the matching matrix pairs; outputs n and N - 1 - n are then (E + O + r) >> s
and (E - O + r) >> s. That is exact only because the integer matrices keep
M[k][N - 1 - n] = (-1)^k M[k][n]; this checks the symmetry and the whole
computation, with the kernel's table layout, against the direct product.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64])
def test_dct2_mirror_symmetry(po, n):
    M = po.transform_matrix(0, n).astype(np.int64)
    sign = np.array([(-1) ** k for k in range(n)])[:, None]
    assert np.array_equal(M[:, ::-1], sign * M)


def _eo_table(M):
    """The host's table: per column j, 8 even pairs then 8 odd pairs."""
    n = M.shape[0]
    t = np.zeros((n, 2, 8, 2), np.int64)
    for j in range(n):
        for odd in range(2):
            for q in range(8):
                if 4 * q + 3 < n:
                    t[j, odd, q] = (M[4 * q + odd, j], M[4 * q + 2 + odd, j])
    return t


@pytest.mark.parametrize("n", [8, 16, 32])
def test_butterfly_equals_direct_product(po, n):
    M = po.transform_matrix(0, n).astype(np.int64)
    t = _eo_table(M)
    rng = np.random.default_rng(n)
    for trial in range(200):
        live = 4 * int(rng.integers(1, n // 4 + 1))  # groups of 4 live inputs, as the kernel reads them
        v = np.zeros(n, np.int64)
        v[:live] = rng.integers(-32768, 32768, live)
        s = int(rng.choice([7, 8, 10, 12]))
        r = 1 << (s - 1)
        direct = (v @ M + r) >> s
        nq = live // 4
        ev = np.stack([v[0:live:4], v[2:live:4]], axis=1)
        od = np.stack([v[1:live:4], v[3:live:4]], axis=1)
        got = np.zeros(n, np.int64)
        for j in range(n // 2):
            e = r + (ev * t[j, 0, :nq]).sum()
            o = (od * t[j, 1, :nq]).sum()
            got[j] = (e + o) >> s
            got[n - 1 - j] = (e - o) >> s
        assert np.array_equal(got, direct), (n, trial)

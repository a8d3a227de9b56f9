"""Host-side pieces of the learned CU-split pruning (no device needed)."""
import numpy as np

from conftest import load_package

L = load_package().learned


def test_bd_rate_restatement():
    """bd_rate (scripts/bench/bench.py:777-813): equal curves give 0; a test
    needing 10 % more bits everywhere gives +10 %; non-overlapping curves give None."""
    L = load_package().learned
    a = [(30.0, 1000.0), (33.0, 1800.0), (36.0, 3100.0), (39.0, 5600.0)]
    assert abs(L.bd_rate(a, a)) < 1e-9
    b = [(p, r * 1.1) for p, r in a]
    assert abs(L.bd_rate(a, b) - 10.0) < 1e-6
    c = [(p + 20, r) for p, r in a]
    assert L.bd_rate(a, c) is None
    assert L.bd_rate(a[:2], a[:2]) is None


def test_synthetic_content_is_seeded_and_mixed():
    a = L.synthetic_sequence(4, 2, 320, 192)
    b = L.synthetic_sequence(4, 2, 320, 192)
    assert all((x == y).all() for x, y in zip(a, b))
    assert (a[0] != a[1]).any()  # rectangles move between frames
    assert a[0].dtype == np.uint8 and a[0].shape == (192, 320)
    assert 10 < a[0].std() < 80


def test_node_blocks_and_padding():
    img = np.arange(72 * 100, dtype=np.int64).reshape(72, 100).astype(np.uint8)
    p = L.pad_to_nodes(img)
    assert p.shape == (96, 128) and (p[:72, :100] == img).all() and (p[80, :100] == img[71]).all()
    nb = L.node_blocks(p)
    assert nb.shape == (12, 32, 32)
    assert (nb[5] == p[32:64, 32:64]).all()


def test_soft_targets_follow_the_costs():
    c = np.array([[100.0, 100.5, 103.0, 200.0, 100.0, 150.0]])
    t = L.soft_targets(c, 0.01)
    assert abs(t.sum() - 1) < 1e-6
    assert t[0, 0] == t[0, 4] and t[0, 0] > t[0, 1] > t[0, 2] > t[0, 5] > t[0, 3]

"""Pins the CNN oracle (oracle/cnn_oracle.py): torch autograd in float64,
finite differences, and bf16 rounding against torch's bfloat16."""
import numpy as np
import pytest

import cnn_oracle as CO


def _data(n, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:32, 0:32]
    blocks = []
    for i in range(n):
        img = 128 + rng.normal(0, 20) * (xx - 16) / 16 + rng.normal(0, 20) * (yy - 16) / 16 + rng.normal(0, 10, (32, 32))
        blocks.append(np.clip(img, 0, 255))
    return (np.array(blocks).astype(np.uint8), rng.uniform(0.3, 0.8, n).astype(np.float32),
            rng.integers(0, 6, n))


def _torch_grads(flat, blocks, qp, labels):
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    P = {k: torch.tensor(np.asarray(v, np.float64), requires_grad=True) for k, v in CO.unpack(flat).items()}
    x = torch.tensor((blocks.astype(np.float64) - 128.0) / 64.0)[:, None]
    def conv(a, w, b):
        return F.relu(F.conv2d(a, w.permute(0, 3, 1, 2), b, padding=1))
    a = conv(conv(conv(x, P["W1"], P["b1"]), P["W2"], P["b2"]), P["W3"], P["b3"])
    pooled = a.mean(dim=(2, 3))
    inp = torch.cat([pooled, torch.tensor(qp.astype(np.float64))[:, None]], dim=1)
    z = F.linear(F.relu(F.linear(inp, P["W4"], P["b4"])), P["W5"], P["b5"])
    loss = F.cross_entropy(z, torch.tensor(labels))
    loss.backward()
    return CO.pack({k: v.grad.numpy() for k, v in P.items()}), float(loss.detach())


def test_oracle_gradients_equal_torch_autograd_float64():
    blocks, qp, labels = _data(6, 1)
    flat = CO.he_init(3).astype(np.float64)
    g, losses = CO.gradients(flat, blocks, qp, labels, emulate_bf16=False, dtype=np.float64)
    gt, lt = _torch_grads(flat, blocks, qp, labels)
    assert abs(losses.mean() - lt) < 1e-12
    for name, shape in CO.SHAPES:
        a, b = CO.unpack(g)[name], CO.unpack(gt)[name]
        assert np.allclose(a, b, rtol=1e-9, atol=1e-12), name


def test_oracle_gradients_match_finite_differences():
    blocks, qp, labels = _data(3, 2)
    flat = CO.he_init(4).astype(np.float64)
    g, _ = CO.gradients(flat, blocks, qp, labels, emulate_bf16=False, dtype=np.float64)
    rng = np.random.default_rng(0)
    offs, o = {}, 0
    for name, shape in CO.SHAPES:
        offs[name] = (o, int(np.prod(shape)))
        o += int(np.prod(shape))
    picks = [offs[n][0] + int(rng.integers(0, offs[n][1])) for n, _ in CO.SHAPES for _ in range(2)]
    eps = 1e-6
    for i in picks:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += eps
        fm[i] -= eps
        lp = CO.gradients(fp, blocks, qp, labels, False, np.float64)[1].mean()
        lm = CO.gradients(fm, blocks, qp, labels, False, np.float64)[1].mean()
        num = (lp - lm) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, num, g[i])


def test_bf16_rounding_equals_torch_bfloat16():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.normal(0, 1, 10000), rng.normal(0, 1e-3, 1000), [0.0, -0.0, 1.0, 3.0, 1 + 2 ** -8,
                        1 + 3 * 2 ** -8, 65504.0]]).astype(np.float32)
    want = torch.tensor(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert (CO.bf16(v) == want).all()


def test_bf16_emulation_stays_close_to_float64():
    blocks, qp, labels = _data(4, 3)
    flat = CO.he_init(5)
    g32, l32 = CO.gradients(flat, blocks, qp, labels, emulate_bf16=True)
    g64, l64 = CO.gradients(flat.astype(np.float64), blocks, qp, labels, emulate_bf16=False, dtype=np.float64)
    assert np.abs(l32 - l64).max() < 2e-2
    cos = float(g32 @ g64 / np.linalg.norm(g32) / np.linalg.norm(g64))
    assert cos > 0.99

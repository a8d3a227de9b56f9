"""The CU-split CNN on the GPU (BF16 MFMA forward and backward kernels)
against the CPU oracle with the same bf16 storage points."""
import numpy as np
import pytest

import cnn_oracle as CO
from conftest import load_package

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:32, 0:32]
    out = []
    for i in range(n):
        img = 128 + rng.normal(0, 30) * (xx - 16) / 16 + rng.normal(0, 30) * (yy - 16) / 16
        if i % 3 == 0:
            img[rng.integers(0, 16):rng.integers(16, 32), rng.integers(0, 16):rng.integers(16, 32)] += rng.normal(0, 60)
        img = img + rng.normal(0, rng.uniform(1, 20), (32, 32))
        out.append(np.clip(img, 0, 255))
    return np.array(out).astype(np.uint8), rng.uniform(22, 37, n).astype(np.float32) / 51.0, rng.integers(0, 6, n).astype(np.uint8)


def _upload(hip, *arrays):
    return [hip.DeviceBuffer.from_array(np.ascontiguousarray(a)) for a in arrays]


def test_forward_matches_oracle(acc):
    rmh = load_package()
    net = rmh.Cnn(acc, max_batch=64, seed=7)
    blocks, qp, _ = _data(40, 1)
    flat = net.params()
    assert flat.size == CO.N_PARAMS
    want = CO.forward(flat, blocks, qp)["probs"]
    got = net.predict(blocks, qp)
    assert np.abs(got - want).max() < 2e-3, np.abs(got - want).max()
    # the device-buffer entry point gives the same probabilities
    hip = rmh.hip
    db, dq = _upload(hip, blocks, qp)
    dp = hip.DeviceBuffer(40 * 6 * 4)
    net.forward(db.ptr, dq.ptr, 40, dp.ptr)
    hip.device_synchronize()
    assert (dp.download(np.float32).reshape(40, 6) == got).all()
    net.close()


def test_gradients_match_oracle(acc):
    rmh = load_package()
    hip = rmh.hip
    net = rmh.Cnn(acc, max_batch=32, seed=11)
    blocks, qp, labels = _data(24, 2)
    # a trained-looking scale: perturb the init so every layer's gradient is non-trivial
    flat = net.params()
    want, wl = CO.gradients(flat, blocks, qp, labels)
    db, dq, dl = _upload(hip, blocks, qp, labels)
    dg = hip.DeviceBuffer(flat.size * 4)
    dloss = hip.DeviceBuffer(24 * 4)
    net.gradients(db.ptr, dq.ptr, dl.ptr, 24, dg.ptr, dloss.ptr)
    hip.device_synchronize()
    got = dg.download(np.float32)
    assert np.abs(dloss.download(np.float32) - wl).max() < 5e-3
    G, Wt = CO.unpack(got), CO.unpack(want)
    for name, _ in CO.SHAPES:
        a, b = G[name].reshape(-1), Wt[name].reshape(-1)
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        assert rel < 3e-2 and cos > 0.999, (name, rel, cos)
    net.close()


def test_training_lowers_the_loss_and_set_params_round_trips(acc):
    rmh = load_package()
    hip = rmh.hip
    net = rmh.Cnn(acc, max_batch=64, seed=3)
    blocks, qp, _ = _data(64, 3)
    # a learnable rule: label from the block's dominant gradient direction
    gx = np.abs(np.diff(blocks.astype(np.int32), axis=2)).mean(axis=(1, 2))
    gy = np.abs(np.diff(blocks.astype(np.int32), axis=1)).mean(axis=(1, 2))
    labels = np.where(gx > gy, 3, 2).astype(np.uint8)
    db, dq, dl = _upload(hip, blocks, qp, labels)
    dloss = hip.DeviceBuffer(64 * 4)
    losses = []
    for step in range(60):
        net.train_step(db.ptr, dq.ptr, dl.ptr, 64, 2e-3, dloss.ptr)
        if step % 10 == 0 or step == 59:
            hip.device_synchronize()
            losses.append(float(dloss.download(np.float32).mean()))
    assert losses[-1] < 0.5 * losses[0], losses
    p = net.params()
    net2 = rmh.Cnn(acc, max_batch=64, seed=99)
    net2.set_params(p)
    assert (net2.params() == p).all()
    assert (net2.predict(blocks, qp) == net.predict(blocks, qp)).all()
    net.close()
    net2.close()

"""CPU oracle of the CU-split CNN: forward pass and gradients in numpy.

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's checks use it; librmh never
does). The network has no counterpart in the reference, which has no learned
component; the six outputs follow CodingTreeSplit
(crates/rustmpeg-video/src/vvc/transform.rs:346-359). The oracle is pinned by
torch autograd in float64 and by finite differences (tests/test_cnn_oracle.py).

With emulate_bf16=True the oracle rounds to bf16 (nearest even) exactly where
the GPU stores bf16: the network input, the three convolution weights, the
three activations and the three convolution output gradients. Sums run in
the array dtype (float32 by default; float64 for the pins).

Parameter layout (flat, float32): W1[32][3][3][1], b1[32], W2[32][3][3][32],
b2[32], W3[32][3][3][32], b3[32], W4[64][33], b4[64], W5[6][64], b5[6].
"""
import numpy as np

SHAPES = [("W1", (32, 3, 3, 1)), ("b1", (32,)), ("W2", (32, 3, 3, 32)), ("b2", (32,)),
          ("W3", (32, 3, 3, 32)), ("b3", (32,)), ("W4", (64, 33)), ("b4", (64,)), ("W5", (6, 64)), ("b5", (6,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + k]).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name]).reshape(-1) for name, _ in SHAPES])


def bf16(a):
    """Round float32 values to the nearest bf16 (ties to even), returned as float32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _patches(x):
    """NHWC x -> (n, 32, 32, 9 * c) zero-padded 3x3 patches, k = (ky, kx, c)."""
    n, h, w, c = x.shape
    xp = np.zeros((n, h + 2, w + 2, c), x.dtype)
    xp[:, 1:-1, 1:-1, :] = x
    cols = [xp[:, ky:ky + h, kx:kx + w, :] for ky in range(3) for kx in range(3)]
    return np.concatenate(cols, axis=3)


def _conv(x, w):
    """3x3 zero-padded convolution (cross-correlation), NHWC, w[co][ky][kx][ci]."""
    n = x.shape[0]
    p = _patches(x).reshape(n * 1024, -1)
    return (p @ w.reshape(w.shape[0], -1).T.astype(x.dtype)).reshape(n, 32, 32, w.shape[0])


def _flip(w):
    """[co][ky][kx][ci] -> [ci][2-ky][2-kx][co]: the input-gradient weights."""
    return np.ascontiguousarray(w[:, ::-1, ::-1, :].transpose(3, 1, 2, 0))


def forward(flat, blocks, qp, emulate_bf16=True, dtype=np.float32):
    """blocks: uint8 (n, 32, 32); qp: (n,) normalised QP. Returns a dict of the
    intermediate values and 'probs' (n, 6)."""
    r = (lambda a: bf16(a).astype(dtype)) if emulate_bf16 else (lambda a: np.asarray(a, dtype))
    P = {k: np.asarray(v, dtype) for k, v in unpack(np.asarray(flat, dtype)).items()}
    Wc = {k: r(P[k]) for k in ("W1", "W2", "W3")}
    x = r((blocks.astype(dtype) - 128.0) / 64.0)[..., None]
    a1 = r(np.maximum(_conv(x, Wc["W1"]) + P["b1"], 0))
    a2 = r(np.maximum(_conv(a1, Wc["W2"]) + P["b2"], 0))
    a3 = r(np.maximum(_conv(a2, Wc["W3"]) + P["b3"], 0))
    pooled = a3.reshape(len(blocks), 1024, 32).mean(axis=1)
    inp = np.concatenate([pooled, np.asarray(qp, dtype)[:, None]], axis=1)
    hpre = inp @ P["W4"].T + P["b4"]
    h = np.maximum(hpre, 0)
    z = h @ P["W5"].T + P["b5"]
    e = np.exp(z - z.max(axis=1, keepdims=True))
    probs = e / e.sum(axis=1, keepdims=True)
    return dict(x=x, a1=a1, a2=a2, a3=a3, pooled=pooled, inp=inp, h=h, z=z, probs=probs, W=Wc, P=P)


def gradients(flat, blocks, qp, labels, emulate_bf16=True, dtype=np.float32):
    """Gradient of the mean softmax cross-entropy; returns (flat grads, per-sample losses)."""
    r = (lambda a: bf16(a).astype(dtype)) if emulate_bf16 else (lambda a: np.asarray(a, dtype))
    f = forward(flat, blocks, qp, emulate_bf16, dtype)
    n = len(blocks)
    labels = np.asarray(labels, np.int64)
    z = f["z"]
    lse = np.log(np.exp(z - z.max(axis=1, keepdims=True)).sum(axis=1)) + z.max(axis=1)
    losses = lse - z[np.arange(n), labels]
    dz = f["probs"].copy()
    dz[np.arange(n), labels] -= 1.0
    dz /= n
    P, W = f["P"], f["W"]
    g = {}
    g["W5"] = dz.T @ f["h"]
    g["b5"] = dz.sum(axis=0)
    dh = (dz @ P["W5"]) * (f["h"] > 0)
    g["W4"] = dh.T @ f["inp"]
    g["b4"] = dh.sum(axis=0)
    dpool = dh @ P["W4"][:, :32]
    dz3 = r((f["a3"] > 0) * (dpool[:, None, None, :] / 1024.0))
    g["W3"] = (dz3.reshape(-1, 32).T @ _patches(f["a2"]).reshape(n * 1024, -1)).reshape(32, 3, 3, 32)
    g["b3"] = dz3.reshape(-1, 32).sum(axis=0)
    dz2 = r((f["a2"] > 0) * _conv(dz3, _flip(W["W3"])))
    g["W2"] = (dz2.reshape(-1, 32).T @ _patches(f["a1"]).reshape(n * 1024, -1)).reshape(32, 3, 3, 32)
    g["b2"] = dz2.reshape(-1, 32).sum(axis=0)
    dz1 = r((f["a1"] > 0) * _conv(dz2, _flip(W["W2"])))
    g["W1"] = (dz1.reshape(-1, 32).T @ _patches(f["x"]).reshape(n * 1024, -1)).reshape(32, 3, 3, 1)
    g["b1"] = dz1.reshape(-1, 32).sum(axis=0)
    return pack(g), losses


def he_init(seed):
    """A seeded initialisation for oracle-only tests (not the library's)."""
    rng = np.random.default_rng(seed)
    d = {}
    fan = {"W1": 9, "W2": 288, "W3": 288, "W4": 33, "W5": 128}
    for name, shape in SHAPES:
        d[name] = (rng.normal(0, np.sqrt(2.0 / fan[name]), shape) if name in fan
                   else rng.normal(0, 0.05, shape)).astype(np.float32)
    return pack(d)

"""CPU restatement (numpy, float64) of the picker network and its loss, the
checker of the GPU forward/backward kernels.  TEST INFRASTRUCTURE ONLY.

Network: 16 -> 64 -> 64 (ReLU) -> 6; outputs 0..4 are logits over
ncrit in {16, 32, 64, 128, 256}, output 5 maps to theta = 0.2 + 0.4 sigmoid.
Loss per sample: cross entropy + (theta - label)^2 / 0.04.
Parameter layout: W1[64][16] b1[64] W2[64][64] b2[64] W3[6][64] b3[6].
The picker has no counterpart in the reference; this restatement is pinned
by finite differences (tests/test_picker_oracle.py).
"""
import numpy as np

SHAPES = [("W1", (64, 16)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (6, 64)),
          ("b3", (6,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(w):
    w = np.asarray(w, np.float64)
    out, off = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = w[off:off + n].reshape(shp)
        off += n
    return out


def forward(w, X):
    p = unpack(w)
    X = np.asarray(X, np.float64)
    a1 = X @ p["W1"].T + p["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(a2, 0)
    o = h2 @ p["W3"].T + p["b3"]
    return o, (X, a1, h1, a2, h2)


def loss_and_grad(w, X, cls, theta):
    """Per-sample losses and the mean gradient over the batch."""
    p = unpack(w)
    o, (X, a1, h1, a2, h2) = forward(w, X)
    n = len(X)
    logits = o[:, :5]
    mx = logits.max(1, keepdims=True)
    e = np.exp(logits - mx)
    se = e.sum(1, keepdims=True)
    cls = np.asarray(cls)
    logp = logits - mx - np.log(se)
    sg = 1.0 / (1.0 + np.exp(-o[:, 5]))
    th = 0.2 + 0.4 * sg
    dth = th - np.asarray(theta, np.float64)
    loss = -logp[np.arange(n), cls] + dth * dth / 0.04
    do = np.zeros_like(o)
    do[:, :5] = e / se
    do[np.arange(n), cls] -= 1.0
    do[:, 5] = 2.0 * dth / 0.04 * 0.4 * sg * (1.0 - sg)
    g = {}
    g["W3"] = do.T @ h2 / n
    g["b3"] = do.mean(0)
    dh2 = (do @ p["W3"]) * (a2 > 0)
    g["W2"] = dh2.T @ h1 / n
    g["b2"] = dh2.mean(0)
    dh1 = (dh2 @ p["W2"]) * (a1 > 0)
    g["W1"] = dh1.T @ X / n
    g["b1"] = dh1.mean(0)
    grad = np.concatenate([g[name].ravel() for name, _ in SHAPES])
    return loss, grad


def decode(o):
    o = np.atleast_2d(o)
    return np.array([16, 32, 64, 128, 256])[np.argmax(o[:, :5], 1)], 0.2 + 0.4 / (
        1.0 + np.exp(-o[:, 5]))

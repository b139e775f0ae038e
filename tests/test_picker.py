"""The learned parameter picker: the CPU restatement of the network pinned by
finite differences, the GPU forward/backward kernels against it, and GPU
training."""
import numpy as np
import pytest

import mlp_oracle


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 16))
    cls = rng.integers(0, 5, n)
    theta = rng.uniform(0.2, 0.6, n)
    return X, cls, theta


def _weights(seed):
    return np.random.default_rng(seed).normal(scale=0.3, size=mlp_oracle.NPARAM)


def test_oracle_gradient_matches_finite_differences():
    X, cls, theta = _batch(7, 1)
    w = _weights(2)
    _, g = mlp_oracle.loss_and_grad(w, X, cls, theta)
    rng = np.random.default_rng(3)
    idx = rng.choice(mlp_oracle.NPARAM, 60, replace=False)
    h = 1e-6
    for i in idx:
        wp = w.copy()
        wp[i] += h
        wm = w.copy()
        wm[i] -= h
        fd = (mlp_oracle.loss_and_grad(wp, X, cls, theta)[0].mean() -
              mlp_oracle.loss_and_grad(wm, X, cls, theta)[0].mean()) / (2 * h)
        assert abs(g[i] - fd) < 1e-6 * max(1.0, abs(fd)), (i, g[i], fd)


def test_oracle_loss_closed_form():
    """Zero weights: uniform logits (CE = log 5) and theta = 0.4."""
    X, cls, theta = _batch(4, 4)
    loss, _ = mlp_oracle.loss_and_grad(np.zeros(mlp_oracle.NPARAM), X, cls, theta)
    assert np.allclose(loss, np.log(5) + (0.4 - theta) ** 2 / 0.04)


@pytest.mark.gpu
def test_gpu_gradient_and_forward_match_oracle():
    import tfmm
    pk = tfmm.Picker(0, seed=5)
    X, cls, theta = _batch(33, 6)
    w = pk.weights().astype(np.float64)
    g_gpu, loss_gpu = pk.gradient(X, cls, theta)
    loss_ref, g_ref = mlp_oracle.loss_and_grad(w, X.astype(np.float32), cls, theta.astype(np.float32))
    assert np.allclose(loss_gpu, loss_ref, rtol=1e-4, atol=1e-5)
    assert np.abs(g_gpu - g_ref).max() < 1e-4 * max(1.0, np.abs(g_ref).max())
    out = pk.predict(X)
    o_ref, _ = mlp_oracle.forward(w, X.astype(np.float32))
    assert np.allclose(out, o_ref, rtol=1e-4, atol=1e-5)


@pytest.mark.gpu
def test_gpu_training_fits_a_separable_task():
    import tfmm
    pk = tfmm.Picker(0, seed=7)
    rng = np.random.default_rng(8)
    X = rng.normal(size=(200, 16))
    cls = (X[:, 0] > 0).astype(np.int32) * 2 + (X[:, 1] > 0).astype(np.int32)
    theta = 0.4 + 0.15 * np.tanh(X[:, 2])
    loss = pk.train(X, cls, theta, steps=300, lr=1e-2)
    assert loss[-1] < 0.25 * loss[0]
    nc, th = tfmm.decode(pk.predict(X))
    acc = np.mean(nc == np.array(tfmm.NCRIT_CLASSES)[cls])
    assert acc > 0.9 and np.abs(th - theta).mean() < 0.03

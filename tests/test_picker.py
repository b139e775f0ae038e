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


def test_trained_weights_file_is_consistent():
    """The committed weights: right size, finite, theta cap inside the head's
    range, and the held-out record written with them meets the target."""
    import json
    import os
    here = os.path.join(os.path.dirname(__file__), "..", "tapas-reduce_amd", "picker")
    d = np.load(os.path.join(here, "weights.npz"), allow_pickle=False)
    assert d["weights"].shape == (mlp_oracle.NPARAM,) and np.isfinite(d["weights"]).all()
    assert 0.2 <= float(d["theta_max"]) <= 0.6
    s = json.load(open(os.path.join(here, "summary.json")))["summary"]
    assert s["runs"] >= 2000 and s["heldout"] >= 10
    assert s["heldout_meets_target"] == 1.0
    assert s["heldout_geomean_speedup_vs_fallback"] > 1.0


@pytest.mark.gpu
def test_gpu_learned_parameters_beat_the_fallback():
    """On problems outside the training set, the trained picker's (ncrit,
    theta) meets the 1e-4 error target and is faster than the fallback
    ncrit = 64, theta = 0.4 (best of 3 timed evaluations each)."""
    import os
    import time
    import datasets
    import tfmm
    path = os.path.join(os.path.dirname(__file__), "..", "tapas-reduce_amd", "picker",
                        "weights.npz")
    pk = tfmm.Picker.load(path)
    ratios = []
    with tfmm.Context(0) as ctx:
        for dist, n, seed in (("plummer", 1 << 20, 5001), ("two_plummer", 1 << 21, 5002)):
            x, y, z, q = datasets.make(dist, n, seed=seed)
            ctx.set_bodies(x, y, z, q)
            rng = np.random.default_rng(seed)
            targets = np.sort(rng.choice(n, size=1024, replace=False))
            direct = ctx.direct(targets)
            picked = ctx.pick_params(pk, tfmm.MIXED)
            times = {}
            for name, prm in (("picked", tfmm.make_params(int(picked.ncrit), float(picked.theta),
                                                          tfmm.MIXED)),
                              ("fallback", tfmm.make_params(64, 0.4, tfmm.MIXED))):
                ctx.evaluate(prm)
                best = 1e30
                for _ in range(3):
                    t0 = time.perf_counter()
                    ctx.evaluate(prm)
                    ctx.sync()
                    best = min(best, time.perf_counter() - t0)
                times[name] = best
                f = ctx.fields()[targets]
                ep = np.sqrt(((f[:, 0] - direct[:, 0]) ** 2).sum() / (direct[:, 0] ** 2).sum())
                ea = np.sqrt(((f[:, 1:] - direct[:, 1:]) ** 2).sum() / (direct[:, 1:] ** 2).sum())
                assert max(ep, ea) <= 1e-4, (dist, name, ep, ea)
            print("%s n=%d picked (%d, %.3f): %.2f ms vs fallback %.2f ms" % (
                dist, n, picked.ncrit, picked.theta, 1e3 * times["picked"],
                1e3 * times["fallback"]))
            ratios.append(times["fallback"] / times["picked"])
    pk.close()
    assert np.exp(np.mean(np.log(ratios))) > 1.05, ratios

"""The workloads of the benchmark at their full sizes, through the C ABI.

c2: 4,194,304-body Plummer sphere, mixed precision, learned parameters:
    fields within relative L2 1e-4 of the FP64 direct sum on 4096 targets and
    within 1e-5 of the FP64 FMM of the same bodies and parameters, with the
    same interaction counts in both precisions.
c3: 256 independent problems (N log-uniform in [2^10, 2^16], cube / Plummer /
    1:0.3:0.1 ellipsoid) as one forest: every problem's keys, body order and
    counts equal its own evaluation's, fields within 1e-6 (mixed); the FP64
    forest within 1e-10 of the CPU oracle on its smallest problems.
c4: 8,388,608-body two-cluster Plummer: the learned parameters meet 1e-4
    against the FP64 direct sum and beat the fallback ncrit 64 / theta 0.4.
"""
import os
import time

import numpy as np
import pytest

import datasets
import oracle
import tfmm

pytestmark = pytest.mark.gpu

PICKER = os.path.join(os.path.dirname(__file__), "..", "tapas-reduce_amd", "picker",
                      "weights.npz")


@pytest.fixture(scope="module")
def ctx():
    c = tfmm.Context(0)
    yield c
    c.close()


def rel(a, b):
    return oracle.rel_l2(a, b)


def best_of(ctx, prm, reps=3):
    ctx.evaluate(prm)
    best = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.evaluate(prm)
        ctx.sync()
        best = min(best, time.perf_counter() - t0)
    return best


def test_gpu_c2_plummer_4m_mixed_learned(ctx):
    n = 4194304
    x, y, z, q = datasets.plummer(n, seed=0)
    print("c2: plummer n=%d" % n, flush=True)
    ctx.set_bodies(x, y, z, q)
    pk = tfmm.Picker.load(PICKER)
    p = ctx.pick_params(pk, tfmm.MIXED)
    pk.close()
    ncrit, theta = int(p.ncrit), float(p.theta)
    print("c2: picked ncrit %d theta %.4f" % (ncrit, theta), flush=True)
    ctx.evaluate(tfmm.make_params(ncrit, theta, tfmm.MIXED))
    f32, st32 = ctx.fields(), ctx.stats()
    ctx.evaluate(tfmm.make_params(ncrit, theta, tfmm.FP64))
    f64, st64 = ctx.fields(), ctx.stats()
    for k in ("n_m2l", "n_p2p_calls", "n_p2p_leafpairs", "ncells"):
        assert st32[k] == st64[k], k
    t = np.sort(np.random.default_rng(17).choice(n, 4096, replace=False))
    d = ctx.direct(t)
    e32, e64, emix = rel(f32[t], d), rel(f64[t], d), rel(f32, f64)
    print("c2: mixed vs direct %.2e %.2e, fp64 vs direct %.2e %.2e, mixed vs fp64 %.2e %.2e"
          % (*e32, *e64, *emix))
    assert max(e32) <= 1e-4
    assert max(emix) <= 1e-5


def test_gpu_c3_batch256_forest(ctx):
    probs, kinds = datasets.batch(256, seed=0)
    off = np.concatenate([[0], np.cumsum([len(p[0]) for p in probs])])
    print("c3: %d problems, %d bodies" % (len(probs), off[-1]), flush=True)
    prm = tfmm.make_params(128, 0.4, tfmm.MIXED)
    ctx.set_batch(probs)
    assert np.array_equal(ctx.problem_offsets(), off)
    ctx.evaluate(prm)
    fb, sb = ctx.fields(), ctx.stats()
    keys_b, perm_b = ctx.sorted()
    assert sb["n_problems"] == 256
    tot = {"ncells": 0, "n_m2l": 0, "n_p2p_calls": 0}
    worst = 0.0
    for i, p in enumerate(probs):
        ctx.set_bodies(*p)
        ctx.evaluate(prm)
        f, s = ctx.fields(), ctx.stats()
        k, perm = ctx.sorted()
        assert np.array_equal(keys_b[off[i]:off[i + 1]], k), i
        assert np.array_equal(np.sort(perm_b[off[i]:off[i + 1]] - off[i]), np.sort(perm)), i
        e = max(rel(fb[off[i]:off[i + 1]], f))
        worst = max(worst, e)
        assert e < 1e-6, (i, kinds[i], len(p[0]), e)
        for key in tot:
            tot[key] += s[key]
    print("c3: worst per-problem mixed batch vs single %.2e" % worst)
    for key in tot:
        assert sb[key] == tot[key], key
    # FP64 forest against the oracle on the smallest problems
    ctx.set_batch(probs)
    ctx.evaluate(tfmm.make_params(64, 0.4, tfmm.FP64))
    f64 = ctx.fields()
    for i in np.argsort([len(p[0]) for p in probs])[:4]:
        o = oracle.Run(*probs[i], 64, 0.4)
        ep, ea = rel(f64[off[i]:off[i + 1]], o.fields())
        print("c3: problem %d (%s, n=%d) fp64 vs oracle %.2e %.2e" % (i, kinds[i],
                                                                     len(probs[i][0]), ep, ea))
        assert ep < 1e-10 and ea < 1e-10


def test_gpu_c4_two_plummer_8m_learned(ctx):
    n = 8388608
    x, y, z, q = datasets.two_plummer(n, seed=0)
    print("c4: two-cluster plummer n=%d" % n, flush=True)
    ctx.set_bodies(x, y, z, q)
    pk = tfmm.Picker.load(PICKER)
    p = ctx.pick_params(pk, tfmm.MIXED)
    pk.close()
    t = np.sort(np.random.default_rng(23).choice(n, 4096, replace=False))
    d = ctx.direct(t)
    times = {}
    for name, prm in (("learned", tfmm.make_params(int(p.ncrit), float(p.theta), tfmm.MIXED)),
                      ("fallback", tfmm.make_params(64, 0.4, tfmm.MIXED))):
        times[name] = best_of(ctx, prm)
        e = rel(ctx.fields()[t], d)
        print("c4: %s (%d, %.4f): %.2f ms, error %.2e %.2e" % (
            name, prm.ncrit, prm.theta, 1e3 * times[name], *e), flush=True)
        assert max(e) <= 1e-4, (name, e)
    assert times["learned"] < times["fallback"]

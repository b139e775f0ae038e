"""The Cartesian basis on the GPU (tfmm_set_basis, TFMM_BASIS_CARTESIAN),
through the C ABI, against the CPU oracle's restatement of ExaFMM's
LaplaceCartesianCPU.cxx (itself pinned to the reference's own Cartesian
kernels and by the mathematics, tests/test_oracle_cartesian.py), on both
routes of the Cartesian M2L (TFMM_TUNE_CART_M2L: 0 the FP64 tile kernel
between exact basis changes, 1 the direct Cartesian kernel):

* FP64, mutual 0 and 1: fields within relative L2 1e-10 of the oracle's
  Cartesian FMM; M of every cell and L of every non-empty cell within 1e-10
  (max norm); numM2L / numP2P equal the oracle's; the fields equal the
  spherical basis's to 1e-10 (both truncate the same series);
* mixed precision (FP32 P2P, FP64 Cartesian expansions) within 1e-5 of FP64;
* the phase-wise calls, tfmm_evaluate and its captured numeric phase give
  the same fields bit for bit; TFMM_TUNE_GRAPH 2 runs the ordinary way;
* a forest of problems, and periodic images 1 and 2, against the oracle;
* switching the basis keeps the tree and gives a fresh context's fields;
* at 1M bodies (cube, ncrit 64) the two bases agree to 1e-10.
"""
import numpy as np
import pytest

import datasets
import oracle
import tfmm

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    c = tfmm.Context(0)
    yield c
    c.close()


def rel_max(a, b):
    scale = np.abs(b).max()
    return np.abs(a - b).max() / scale if scale > 0 else np.abs(a - b).max()


def run(ctx, b, ncrit, theta, precision=tfmm.FP64, mutual=0, basis=tfmm.BASIS_CARTESIAN,
        route=0):
    print("basis %d route %d n=%d ncrit=%d theta=%g precision=%d mutual=%d" % (
        basis, route, len(b[0]), ncrit, theta, precision, mutual), flush=True)
    ctx.set_basis(basis)
    ctx.set_tuning(tfmm.TUNE_CART_M2L, route)
    ctx.set_bodies(*b)
    ctx.evaluate(tfmm.make_params(ncrit, theta, precision, mutual))
    return ctx.fields(), ctx.stats()


@pytest.mark.parametrize("route", [0, 1])
@pytest.mark.parametrize("name,ncrit,theta,mutual", [
    ("cube", 32, 0.4, 0), ("plummer", 64, 0.5, 0), ("cube", 32, 0.4, 1),
    ("plummer", 32, 0.5, 1)])
def test_gpu_cartesian_fp64_matches_oracle(ctx, name, ncrit, theta, mutual, route):
    b = datasets.make(name, 6000, seed=11)
    f, st = run(ctx, b, ncrit, theta, mutual=mutual, route=route)
    o = oracle.Run(*b, ncrit, theta, mutual=mutual, basis=oracle.CARTESIAN)
    ep, ea = oracle.rel_l2(f, o.fields())
    assert ep < TOL and ea < TOL, (ep, ea)
    cnt = o.counts()
    assert st["n_m2l"] == cnt["m2l"]
    assert st["n_p2p_calls"] == cnt["p2p_bodypairs"] + cnt["p2p_selfpairs"]
    if route == 1:
        assert st["n_m2l_groups"] == 0 and st["n_m2l_items"] == 0 and st["n_m2l_fallback"] == 0
    else:
        assert st["n_m2l_items"] > 0
    # expansions of every cell (cells compared in key order)
    M, L = ctx.expansions()
    assert M.shape == (st["ncells"], tfmm.CART_TERMS)
    Mo, Lo = o.expansions()
    g = np.argsort(ctx.cells()["key"])
    oc = o.cells()
    og = np.argsort(oc["key"])
    assert rel_max(M[g], Mo[og]) < TOL
    ne = oc["nb"][og] > 0  # M2L into an empty cell is not executed (its L has no use)
    assert rel_max(L[g][ne], Lo[og][ne]) < TOL
    # the spherical basis: the same truncated series
    fs, _ = run(ctx, b, ncrit, theta, mutual=mutual, basis=tfmm.BASIS_SPHERICAL)
    ep, ea = oracle.rel_l2(f, fs)
    assert ep < TOL and ea < TOL, (ep, ea)


def test_gpu_cartesian_mixed(ctx):
    b = datasets.plummer(20000, seed=3, charges="uniform")
    f64, _ = run(ctx, b, 64, 0.4)
    f32, _ = run(ctx, b, 64, 0.4, precision=tfmm.MIXED)
    ep, ea = oracle.rel_l2(f32, f64)
    assert ep < 1e-5 and ea < 1e-5, (ep, ea)
    t = np.arange(0, 20000, 97)
    d = ctx.direct(t)
    ep, ea = oracle.rel_l2(f32[t], d)
    assert ep < 1e-4 and ea < 1e-3, (ep, ea)


@pytest.mark.parametrize("route", [0, 1])
def test_gpu_cartesian_paths_agree(ctx, route):
    """phase-wise calls, tfmm_evaluate and its captured numeric phase: the
    same kernels on the same lists, bit for bit"""
    b = datasets.cube(30000, seed=5)
    prm = tfmm.make_params(64, 0.4, tfmm.FP64)
    f_eval, s_eval = run(ctx, b, 64, 0.4, route=route)
    ctx.build_tree(prm)
    ctx.upward()
    ctx.traverse()
    ctx.downward()
    assert np.array_equal(ctx.fields(), f_eval)
    ctx.set_tuning(tfmm.TUNE_GRAPH, 1)
    try:
        for rep in range(3):
            print("captured numeric phase, launch %d" % rep, flush=True)
            ctx.evaluate(prm)
            assert np.array_equal(ctx.fields(), f_eval), rep
            t = np.array([0, 15000, 29999])
            assert np.all(np.isfinite(ctx.direct(t)))  # eager work between the launches
        ctx.set_tuning(tfmm.TUNE_GRAPH, 2)
        for rep in range(2):
            ctx.evaluate(prm)
            assert np.array_equal(ctx.fields(), f_eval), rep
            assert ctx.stats()["one_graph"] == 0
    finally:
        ctx.set_tuning(tfmm.TUNE_GRAPH, 0)


@pytest.mark.parametrize("route", [0, 1])
def test_gpu_cartesian_batch(ctx, route):
    probs, _ = datasets.batch(5, seed=8, nmin=500, nmax=4000)
    probs.append((np.array([0.3]), np.array([0.2]), np.array([0.9]), np.array([1.0])))
    ctx.set_basis(tfmm.BASIS_CARTESIAN)
    ctx.set_tuning(tfmm.TUNE_CART_M2L, route)
    ctx.set_batch(probs)
    print("forest of %d problems" % len(probs), flush=True)
    ctx.evaluate(tfmm.make_params(64, 0.4, tfmm.FP64))
    fb = ctx.fields()
    off = ctx.problem_offsets()
    for i, (x, y, z, q) in enumerate(probs):
        seg = fb[off[i]:off[i + 1]]
        if len(x) == 1:
            assert not np.any(seg)
            continue
        o = oracle.Run(x, y, z, q, 64, 0.4, basis=oracle.CARTESIAN)
        ep, ea = oracle.rel_l2(seg, o.fields())
        assert ep < TOL and ea < TOL, (i, ep, ea)


@pytest.mark.parametrize("route", [0, 1])
@pytest.mark.parametrize("images", [1, 2])
def test_gpu_cartesian_periodic(ctx, images, route):
    r = np.random.default_rng(images)
    b = tuple(r.uniform(-np.pi, np.pi, 3000) for _ in range(3)) + (r.random(3000) - 0.5,)
    cyc = 2 * np.pi
    ctx.set_periodic(images, cyc)
    try:
        f, _ = run(ctx, b, 32, 0.4, route=route)
        o = oracle.Run(*b, 32, 0.4, images=images, cycle=cyc, basis=oracle.CARTESIAN)
        ep, ea = oracle.rel_l2(f, o.fields())
        assert ep < TOL and ea < TOL, (images, ep, ea)
        fs, _ = run(ctx, b, 32, 0.4, basis=tfmm.BASIS_SPHERICAL)
        ep, ea = oracle.rel_l2(f, fs)
        assert ep < TOL and ea < TOL, (images, ep, ea)
    finally:
        ctx.set_periodic(0, 0.0)


def test_gpu_cartesian_switch_keeps_tree(ctx):
    b = datasets.plummer(15000, seed=9, charges="uniform")
    prm = tfmm.make_params(32, 0.4, tfmm.FP64)
    ctx.set_basis(tfmm.BASIS_SPHERICAL)
    ctx.set_bodies(*b)
    ctx.build_tree(prm)
    ctx.upward()
    ctx.traverse()
    ctx.downward()
    fs = ctx.fields()
    ctx.set_basis(tfmm.BASIS_CARTESIAN)
    assert ctx.basis() == tfmm.BASIS_CARTESIAN
    with pytest.raises(tfmm.TfmmError):
        ctx.downward()  # the expansions must be formed again
    ctx.upward()
    ctx.traverse()
    ctx.downward()
    fc = ctx.fields()
    fresh = tfmm.Context(0)
    try:
        fresh.set_basis(tfmm.BASIS_CARTESIAN)
        fresh.set_bodies(*b)
        fresh.evaluate(prm)
        assert np.array_equal(fresh.fields(), fc)
    finally:
        fresh.close()
    ep, ea = oracle.rel_l2(fc, fs)
    assert ep < TOL and ea < TOL
    with pytest.raises(tfmm.TfmmError):
        ctx.set_basis(2)
    ctx.set_basis(tfmm.BASIS_SPHERICAL)


def test_gpu_cartesian_1m_equals_spherical(ctx):
    b = datasets.cube(1 << 20, seed=0)
    fs, ss = run(ctx, b, 64, 0.4, basis=tfmm.BASIS_SPHERICAL)
    for route in (0, 1):
        fc, st = run(ctx, b, 64, 0.4, route=route)
        assert st["n_m2l"] == ss["n_m2l"]
        ep, ea = oracle.rel_l2(fc, fs)
        assert ep < TOL and ea < TOL, (route, ep, ea)
        print("1M cube: cartesian route %d %.2f ms, spherical %.2f ms" % (
            route, st["ms_total"], ss["ms_total"]), flush=True)
    ctx.set_tuning(tfmm.TUNE_CART_M2L, 0)


def test_gpu_cartesian_route_knob(ctx):
    assert ctx.get_tuning(tfmm.TUNE_CART_M2L) == 0
    with pytest.raises(tfmm.TfmmError):
        ctx.set_tuning(tfmm.TUNE_CART_M2L, 2)

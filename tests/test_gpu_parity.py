"""Parity of the HIP library (through the C ABI) with the reference's outputs
(tests/golden/) and with the CPU oracle, FP64 path.

Bars: keys, body order (up to ties), cells, interaction counters and the
traversal lists bit-exact; multipole/local expansions and fields within
relative 1e-10 (max-norm for expansions, relative L2 for fields).
"""
import glob
import hashlib
import os

import numpy as np
import pytest

import datasets
import oracle
import tfmm
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

# fixtures of the whole pipeline (the 1M-body digest and the Cartesian
# operator vectors have tests of their own)
FULL = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
              if not os.path.basename(p).startswith(("c1_", "cartesian_ops")))
TOL = 1e-10


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def ctx():
    c = tfmm.Context(0)
    yield c
    c.close()


def evaluate(ctx, x, y, z, q, ncrit, theta, precision=tfmm.FP64, mutual=0):
    print("case n=%d ncrit=%d theta=%g precision=%d mutual=%d" % (len(x), ncrit, theta,
                                                                  precision, mutual), flush=True)
    ctx.set_bodies(x, y, z, q)
    ctx.evaluate(tfmm.make_params(ncrit, theta, precision, mutual))
    return ctx.fields(), ctx.stats()


def mutual_of(g):
    return int(g["mutual"]) if "mutual" in g.files else 0


def rel_max(a, b):
    scale = np.abs(b).max()
    return np.abs(a - b).max() / scale if scale > 0 else np.abs(a - b).max()


def same_key_groups(keys, perm_a, perm_b):
    if np.array_equal(perm_a, perm_b):
        return True
    bounds = np.flatnonzero(np.diff(keys)) + 1
    return all(np.array_equal(np.sort(a), np.sort(b))
               for a, b in zip(np.split(perm_a, bounds), np.split(perm_b, bounds)))


@pytest.mark.parametrize("name", FULL)
def test_gpu_matches_reference(ctx, name):
    g = load(name)
    f, st = evaluate(ctx, g["x"], g["y"], g["z"], g["q"], int(g["ncrit"]), float(g["theta"]),
                     mutual=mutual_of(g))
    assert np.array_equal(st["region_min"], g["rmin"]) and np.array_equal(st["region_max"], g["rmax"])
    keys, perm = ctx.sorted()
    assert np.array_equal(keys, g["keys"])
    assert same_key_groups(keys, perm.astype(np.int64), g["ibody"])
    c = ctx.cells()
    order = np.argsort(c["key"])
    assert st["ncells"] == int(g["ncells"]) and st["max_depth"] == int(g["max_depth"])
    assert np.array_equal(c["key"][order], g["cell_key"])
    assert np.array_equal(c["nb"][order], g["cell_nb"])
    assert np.array_equal(c["leaf"][order], g["cell_leaf"])
    nel = (g["cell_leaf"] == 1) & (g["cell_nb"] > 0)
    assert np.array_equal(c["beg"][order][nel], g["cell_bid"][nel])
    assert np.array_equal(c["center"][order], g["cell_center"])
    assert np.array_equal(c["width"][order], g["cell_width"].max(axis=1))
    assert st["n_m2l"] == int(g["n_m2l"])
    assert st["n_p2p_calls"] == int(g["n_p2p_calls"])
    if "M" in g.files:
        M, L = ctx.expansions()
        assert rel_max(M[order], g["M"]) < TOL
        # the local expansion of an empty leaf has no consumer (L2P is skipped
        # for it, tapas_exafmm.cxx:133), so M2L into empty cells is not
        # executed and their L is left at zero; compare L on the non-empty cells
        ne = g["cell_nb"] > 0
        if np.abs(g["L"][ne]).max() > 0:
            assert rel_max(L[order][ne], g["L"][ne]) < TOL
    ep, ea = oracle.rel_l2(f, g["trg"])
    assert ep < TOL and ea < TOL, (ep, ea)


@pytest.mark.parametrize("mutual", [0, 1])
def test_gpu_lists_match_oracle(ctx, mutual):
    """Executed M2L and P2P lists equal the oracle's (as sets of cell-key
    pairs; with mutual, every unordered pair in both directions)."""
    g = load("plummer3000_c32_t05")
    evaluate(ctx, g["x"], g["y"], g["z"], g["q"], int(g["ncrit"]), float(g["theta"]),
             mutual=mutual)
    c = ctx.cells()
    (mo, ms), (po, ps) = ctx.lists()
    o = oracle.Run(g["x"], g["y"], g["z"], g["q"], int(g["ncrit"]), float(g["theta"]), phases=2,
                   record_lists=True, mutual=mutual)
    oc = o.cells()
    om, op = o.lists()
    onb = oc["nb"]

    def gpu_pairs(off, src):
        tgt = np.repeat(np.arange(len(off) - 1), np.diff(off))
        return set(zip(c["key"][tgt].tolist(), c["key"][src].tolist()))

    def orc_pairs(lst, nonempty):
        if nonempty:
            lst = lst[(onb[lst[:, 0]] > 0) & (onb[lst[:, 1]] > 0)]
        return set(zip(oc["key"][lst[:, 0]].tolist(), oc["key"][lst[:, 1]].tolist()))

    assert gpu_pairs(mo, ms) == orc_pairs(om, True)
    assert gpu_pairs(po, ps) == orc_pairs(op, False)
    # sorted by target, then source: deterministic order
    for off, src in ((mo, ms), (po, ps)):
        tgt = np.repeat(np.arange(len(off) - 1), np.diff(off))
        assert np.all(np.diff(tgt.astype(np.int64) * (1 << 32) + src) > 0)


def test_gpu_long_p2p_list_matches_oracle(ctx):
    """A large leaf beside a dense cluster of small leaves gets more P2P
    sources (2,306) than one wave's sort stages in LDS (2,048), so its list is
    ranked from global memory: the lists still equal the oracle's, in
    (target, source) order, and the fields match."""
    rng = np.random.default_rng(7)
    m = 8000
    cl = 0.47 + 0.03 * rng.random((m, 3))
    x = np.concatenate([cl[:, 0], [0.9, 0.0]])
    y = np.concatenate([cl[:, 1], [0.9, 0.0]])
    z = np.concatenate([cl[:, 2], [0.9, 0.0]])
    q = rng.uniform(-1, 1, len(x))
    f, _ = evaluate(ctx, x, y, z, q, 4, 0.5)
    c = ctx.cells()
    (mo, ms), (po, ps) = ctx.lists()
    assert np.diff(po).max() > 2048
    o = oracle.Run(x, y, z, q, 4, 0.5, phases=3, record_lists=True)
    oc = o.cells()
    om, op = o.lists()
    tgt = np.repeat(np.arange(len(po) - 1), np.diff(po))
    assert set(zip(c["key"][tgt].tolist(), c["key"][ps].tolist())) == \
        set(zip(oc["key"][op[:, 0]].tolist(), oc["key"][op[:, 1]].tolist()))
    assert np.all(np.diff(tgt.astype(np.int64) * (1 << 32) + ps) > 0)
    ep, ea = oracle.rel_l2(f, o.fields())
    assert ep < TOL and ea < TOL, (ep, ea)


def test_gpu_very_long_p2p_lists(ctx):
    """Caller-recorded P2P lists with targets of ~3,000, ~15,000 and every
    nonempty leaf (~30,000) as sources: the per-target sort queues lists
    longer than one wave's 2,048 for the workgroup sort (bitonic in LDS up
    to 16,384, ranked by the workgroup beyond).  The lists come back sorted
    and equal to the given sets, and a target whose sources are every leaf
    gets the direct sum over all bodies."""
    rng = np.random.default_rng(21)
    n = 30000
    x, y, z = [rng.random(n) for _ in range(3)]
    q = rng.uniform(-1, 1, n)
    ctx.set_bodies(x, y, z, q)
    ctx.build_tree(tfmm.make_params(1, 0.4, tfmm.FP64))
    c = ctx.cells()
    leaves = np.flatnonzero((c["leaf"] == 1) & (c["nb"] > 0)).astype(np.int32)
    assert len(leaves) > 16384
    t0, t1, t2 = leaves[0], leaves[1], leaves[2]
    src = {t0: leaves, t1: leaves[::2], t2: leaves[::10]}
    pairs = [np.stack([np.full(len(s), t, np.int32), s], axis=1) for t, s in src.items()]
    others = np.setdiff1d(leaves, [t0, t1, t2])
    pairs.append(np.stack([others, others], axis=1))  # every other leaf with itself
    p2p = np.concatenate(pairs)
    rng.shuffle(p2p)
    print("long lists: %d, %d, %d sources" % tuple(len(s) for s in src.values()), flush=True)
    ctx.upward()
    ctx.traverse_lists(np.zeros((0, 2), np.int32), p2p)
    ctx.downward()
    (_, _), (po, ps) = ctx.lists()
    for t, s in src.items():
        got = ps[po[t]:po[t + 1]]
        assert np.array_equal(got, np.sort(s)), t
    assert po[others[0] + 1] - po[others[0]] == 1
    # target t0 sees every body through P2P (no M2L given): the direct sum
    _, perm = ctx.sorted()
    mine = perm[c["beg"][t0]:c["beg"][t0] + c["nb"][t0]].astype(np.int64)
    f = ctx.fields()
    ep, ea = oracle.rel_l2(f[mine], ctx.direct(mine))
    assert ep < 1e-12 and ea < 1e-12, (ep, ea)


def test_gpu_c1_matches_reference(ctx):
    """1M uniform bodies: key/cell digests, counters and 4096 sampled fields
    against the reference, and the FMM against the GPU FP64 direct sum."""
    g = load("c1_cube1m")
    x, y, z, q = datasets.cube(1 << 20, seed=0)
    h = hashlib.sha256()
    for a in (x, y, z, q):
        h.update(a.tobytes())
    assert h.hexdigest() == g["input_sha"].item().decode()
    f, st = evaluate(ctx, x, y, z, q, 64, 0.4)
    keys, _ = ctx.sorted()
    assert hashlib.sha256(keys.tobytes()).hexdigest() == g["keys_sha"].item().decode()
    c = ctx.cells()
    order = np.argsort(c["key"])
    h = hashlib.sha256()
    for a in (c["key"][order], c["nb"][order].astype(np.int64), c["leaf"][order]):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == g["cells_sha"].item().decode()
    assert st["n_m2l"] == int(g["n_m2l"]) == 5656496
    assert st["n_p2p_leafpairs"] == 2421064
    assert st["n_p2p_calls"] == int(g["n_p2p_calls"])
    s = g["sample"]
    ep, ea = oracle.rel_l2(f[s], g["trg_sample"])
    assert ep < TOL and ea < TOL, (ep, ea)
    d = ctx.direct(s)
    ep, ea = oracle.rel_l2(f[s], d)
    assert ep < 1e-4 and ea < 1e-4, (ep, ea)


def test_gpu_direct_matches_oracle(ctx):
    x, y, z, q = datasets.cube(3000, seed=3)
    ctx.set_bodies(x, y, z, q)
    t = np.arange(0, 3000, 11)
    got = ctx.direct(t)
    ref = oracle.direct(x, y, z, q, t)
    ep, ea = oracle.rel_l2(got, ref)
    assert ep < 1e-13 and ea < 1e-13


@pytest.mark.parametrize("n,ncrit", [(1, 64), (7, 64), (64, 64), (100, 1)])
def test_gpu_small_and_root_leaf(ctx, n, ncrit):
    """Root-leaf trees (n <= ncrit) and ncrit = 1; checked against the oracle."""
    x, y, z, q = datasets.cube(n, seed=n)
    if n == 1:
        q = np.ones(1)  # (zero-net charges would be q = 0, and L2P divides by q)
    f, st = evaluate(ctx, x, y, z, q, ncrit, 0.4)
    o = oracle.Run(x, y, z, q, ncrit, 0.4)
    assert st["ncells"] == o.ncells
    assert st["n_m2l"] == o.counts()["m2l"]
    ep, ea = oracle.rel_l2(f, o.fields())
    if n == 1:
        assert np.all(f == 0)
    else:
        assert ep < TOL and ea < TOL


def test_gpu_repeat_is_deterministic(ctx):
    x, y, z, q = datasets.plummer(20000, seed=4, charges="uniform")
    f1, _ = evaluate(ctx, x, y, z, q, 32, 0.4)
    f2, _ = evaluate(ctx, x, y, z, q, 32, 0.4)
    assert np.array_equal(f1, f2)


def test_gpu_phases_in_order_and_errors(ctx):
    x, y, z, q = datasets.cube(500, seed=6)
    ctx.set_bodies(x, y, z, q)
    with pytest.raises(tfmm.TfmmError):
        ctx.upward()  # before build_tree
    with pytest.raises(tfmm.TfmmError):
        ctx.build_tree(tfmm.make_params(64, 0.4, mutual=2))
    with pytest.raises(tfmm.TfmmError):
        ctx.build_tree(tfmm.make_params(0, 0.4))
    with pytest.raises(tfmm.TfmmError):
        ctx.build_tree(tfmm.make_params(64, -1.0))
    p = tfmm.make_params(16, 0.4)
    ctx.build_tree(p)
    ctx.upward()
    ctx.traverse()
    ctx.downward()
    o = oracle.Run(x, y, z, q, 16, 0.4)
    ep, ea = oracle.rel_l2(ctx.fields(), o.fields())
    assert ep < TOL and ea < TOL


# ---------------------------------------------------------------------------
# mixed precision (FP32 P2P with compensated sums, FP32 MFMA M2L)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("dist,n,ncrit", [("plummer", 200000, 64), ("cube", 100000, 32),
                                          ("two_plummer", 150000, 64)])
def test_gpu_mixed_error_against_direct(ctx, dist, n, ncrit):
    """Mixed precision within rel-L2 1e-4 of the FP64 direct sum, and within
    1e-5 of the FP64 FMM (the part of the error the reduced precision adds)."""
    x, y, z, q = datasets.make(dist, n, seed=21)
    f64, st64 = evaluate(ctx, x, y, z, q, ncrit, 0.4, tfmm.FP64)
    f32, st32 = evaluate(ctx, x, y, z, q, ncrit, 0.4, tfmm.MIXED)
    assert st32["n_m2l"] == st64["n_m2l"] and st32["n_p2p_calls"] == st64["n_p2p_calls"]
    t = np.sort(np.random.default_rng(1).choice(n, 2048, replace=False))
    d = ctx.direct(t)
    e64 = oracle.rel_l2(f64[t], d)
    e32 = oracle.rel_l2(f32[t], d)
    emix = oracle.rel_l2(f32, f64)
    print("fp64 vs direct", e64, "mixed vs direct", e32, "mixed vs fp64", emix)
    assert max(e32) < 1e-4
    assert max(emix) < 1e-5


def test_gpu_m2l_groups_cover_pairs(ctx):
    """Every executed M2L pair is either in an offset group or in the per-pair
    fallback list; uniform trees need no fallback."""
    x, y, z, q = datasets.cube(1 << 17, seed=3)
    _, st = evaluate(ctx, x, y, z, q, 64, 0.4)
    assert st["n_m2l_fallback"] == 0
    assert st["n_m2l_items"] * 16 >= st["n_m2l_exec"]
    x, y, z, q = datasets.plummer(1 << 17, seed=3)
    _, st = evaluate(ctx, x, y, z, q, 64, 0.4)
    assert st["n_m2l_items"] * 16 + st["n_m2l_fallback"] >= st["n_m2l_exec"]
    print("plummer: %d of %d M2L pairs outside the offset groups" % (st["n_m2l_fallback"],
                                                                     st["n_m2l_exec"]))
    assert st["n_m2l_fallback"] < 0.05 * st["n_m2l_exec"]


@pytest.mark.parametrize("ncrit", [24, 48, 100, 256])
def test_gpu_fp64_leaf_sizes_match_oracle(ctx, ncrit):
    """FP64 fields equal the oracle's within 1e-10 on a Plummer sphere whose
    leaves span every lane layout of the P2P kernel (several lanes per target,
    one lane per target, several passes per leaf)."""
    x, y, z, q = datasets.plummer(30000, seed=5, charges="uniform")
    f, st = evaluate(ctx, x, y, z, q, ncrit, 0.4)
    o = oracle.Run(x, y, z, q, ncrit, 0.4)
    c = o.counts()
    assert st["n_m2l"] == c["m2l"] and st["n_p2p_leafpairs"] == c["p2p_leafpairs"]
    ep, ea = oracle.rel_l2(f, o.fields())
    print("ncrit %d: rel-L2 pot %.2e acc %.2e" % (ncrit, ep, ea))
    assert ep < TOL and ea < TOL


@pytest.mark.parametrize("ncrit", [24, 48, 100, 256])
def test_gpu_mixed_leaf_sizes_against_fp64(ctx, ncrit):
    """Mixed precision within 1e-5 of the FP64 FMM for the same lane layouts."""
    x, y, z, q = datasets.plummer(60000, seed=6)
    f64, st64 = evaluate(ctx, x, y, z, q, ncrit, 0.4, tfmm.FP64)
    f32, st32 = evaluate(ctx, x, y, z, q, ncrit, 0.4, tfmm.MIXED)
    assert st32["n_p2p_calls"] == st64["n_p2p_calls"]
    emix = oracle.rel_l2(f32, f64)
    print("ncrit %d: mixed vs fp64 %.2e %.2e" % (ncrit, *emix))
    assert max(emix) < 1e-5


def test_gpu_mixed_coincident_across_leaf_boundary(ctx):
    """Two bodies 2e-12 apart on either side of the plane x = 0.5 (a cell
    boundary at every depth of the [0, 1] cube) have equal leaf-relative FP32
    coordinates, so their FP32 distance is exactly zero.  The mixed P2P must
    drop that pair (as it drops a body's pair with itself) and stay finite; all
    other bodies keep the mixed-precision accuracy against the FP64 FMM."""
    rng = np.random.default_rng(8)
    n = 20000
    x, y, z = rng.random((3, n))
    q = rng.uniform(-1, 1, n)
    x[0], y[0], z[0] = 0.5, 0.5, 0.5  # body 0 (its x/y enter the bbox specially) inside
    x[1], y[1], z[1] = 0.0, 0.0, 0.0  # region [0, 1]^3 exactly
    x[2], y[2], z[2] = 1.0, 1.0, 1.0
    a, b = 3, 4
    x[a], x[b] = 0.5 - 1e-12, 0.5 + 1e-12
    y[a] = y[b] = 0.3
    z[a] = z[b] = 0.7
    f64, _ = evaluate(ctx, x, y, z, q, 16, 0.4, tfmm.FP64)
    f32, _ = evaluate(ctx, x, y, z, q, 16, 0.4, tfmm.MIXED)
    assert np.isfinite(f64).all()
    assert np.isfinite(f32).all(), np.argwhere(~np.isfinite(f32))
    keep = np.ones(n, bool)
    keep[[a, b]] = False
    emix = oracle.rel_l2(f32[keep], f64[keep])
    print("mixed vs fp64 without the close pair: %.2e %.2e" % emix)
    assert max(emix) < 1e-5


@pytest.mark.parametrize("dist,n,ncrit", [("plummer", 100000, 64), ("cube", 60000, 32)])
def test_gpu_mutual_against_oracle_and_mixed(ctx, dist, n, ncrit):
    """Mutual interaction at a size beyond the fixtures: FP64 fields within
    1e-10 of the oracle's mutual FMM, the counters equal, and the mixed
    precision form within 1e-5 of the FP64 form."""
    x, y, z, q = datasets.make(dist, n, seed=29)
    f64, st = evaluate(ctx, x, y, z, q, ncrit, 0.4, tfmm.FP64, mutual=1)
    o = oracle.Run(x, y, z, q, ncrit, 0.4, mutual=1)
    c = o.counts()
    assert st["n_m2l"] == c["m2l"] and st["n_p2p_leafpairs"] == c["p2p_leafpairs"]
    assert st["n_p2p_calls"] == c["p2p_bodypairs"] + c["p2p_selfpairs"]
    ep, ea = oracle.rel_l2(f64, o.fields())
    print("mutual fp64 vs oracle %.2e %.2e" % (ep, ea))
    assert ep < TOL and ea < TOL
    f32, _ = evaluate(ctx, x, y, z, q, ncrit, 0.4, tfmm.MIXED, mutual=1)
    emix = oracle.rel_l2(f32, f64)
    print("mutual mixed vs fp64 %.2e %.2e" % emix)
    assert max(emix) < 1e-5

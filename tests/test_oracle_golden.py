"""The CPU oracle against the reference's own outputs (tests/golden/, written
by tests/golden/make_golden.py from the compiled reference pipeline).

Pins, for every fixture: the scalar-cube bounding box, the sorted keys and
body order, the cell set (key, body count, leaf flag, first body, centre,
width), the interaction counters numM2L/numP2P, the multipole and local
expansions of every cell and the fields of every body.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

# fixtures of the whole pipeline (the 1M-body digest and the Cartesian
# operator vectors have tests of their own)
FULL = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
              if not os.path.basename(p).startswith(("c1_", "cartesian_ops")))


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def mutual_of(g):
    return int(g["mutual"]) if "mutual" in g.files else 0


def run_oracle(g, **kw):
    return oracle.Run(g["x"], g["y"], g["z"], g["q"], int(g["ncrit"]), float(g["theta"]),
                      mutual=mutual_of(g), **kw)


def same_key_groups(keys, perm_a, perm_b):
    """Equal sorted keys; body order equal up to permutations among equal keys."""
    if np.array_equal(perm_a, perm_b):
        return True
    bounds = np.flatnonzero(np.diff(keys)) + 1
    for a, b in zip(np.split(perm_a, bounds), np.split(perm_b, bounds)):
        if not np.array_equal(np.sort(a), np.sort(b)):
            return False
    return True


def rel_max(a, b):
    scale = np.abs(b).max()
    return np.abs(a - b).max() / scale if scale > 0 else np.abs(a - b).max()


@pytest.mark.parametrize("name", FULL)
def test_oracle_tree_matches_reference(name):
    g = load(name)
    o = run_oracle(g)
    mn, mx = o.region()
    assert np.array_equal(mn, g["rmin"]) and np.array_equal(mx, g["rmax"])
    keys, perm = o.sorted()
    assert np.array_equal(keys, g["keys"])
    assert same_key_groups(keys, perm, g["ibody"])
    c = o.cells()
    order = np.argsort(c["key"])
    assert o.ncells == int(g["ncells"])
    assert o.max_depth == int(g["max_depth"])
    assert np.array_equal(c["key"][order], g["cell_key"])
    assert np.array_equal(c["nb"][order], g["cell_nb"])
    assert np.array_equal(c["leaf"][order], g["cell_leaf"])
    nonempty_leaf = (g["cell_leaf"] == 1) & (g["cell_nb"] > 0)
    assert np.array_equal(c["bid"][order][nonempty_leaf], g["cell_bid"][nonempty_leaf])
    assert np.array_equal(c["center"][order], g["cell_center"])
    assert np.array_equal(c["width"][order], g["cell_width"])


@pytest.mark.parametrize("name", FULL)
def test_oracle_counts_and_fields_match_reference(name):
    g = load(name)
    o = run_oracle(g)
    cnt = o.counts()
    assert cnt["m2l"] == int(g["n_m2l"])
    assert cnt["p2p_bodypairs"] + cnt["p2p_selfpairs"] == int(g["n_p2p_calls"])
    if "M" in g.files:
        M, L = o.expansions()
        order = np.argsort(o.cells()["key"])
        assert rel_max(M[order], g["M"]) < 1e-13
        if np.abs(g["L"]).max() > 0:
            assert rel_max(L[order], g["L"]) < 1e-13
    ep, ea = oracle.rel_l2(o.fields(), g["trg"])
    assert ep < 1e-13 and ea < 1e-13, (ep, ea)


def c1_bodies(g):
    """The 1M-body input, regenerated from the seeded generator and checked
    against the digest the fixture keeps."""
    import hashlib
    import datasets
    x, y, z, q = datasets.cube(1 << 20, seed=0)
    h = hashlib.sha256()
    for a in (x, y, z, q):
        h.update(a.tobytes())
    assert h.hexdigest() == g["input_sha"].item().decode()
    return x, y, z, q


def test_oracle_c1_matches_reference():
    """1M uniform bodies (P=10, ncrit=64, theta=0.4): key and cell digests,
    counters and the fields of 4096 sampled bodies."""
    import hashlib
    g = load("c1_cube1m")
    x, y, z, q = c1_bodies(g)
    o = oracle.Run(x, y, z, q, 64, 0.4)
    keys, _ = o.sorted()
    assert hashlib.sha256(keys.tobytes()).hexdigest() == g["keys_sha"].item().decode()
    c = o.cells()
    order = np.argsort(c["key"])
    h = hashlib.sha256()
    for a in (c["key"][order], c["nb"][order], c["leaf"][order]):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == g["cells_sha"].item().decode()
    cnt = o.counts()
    assert cnt["m2l"] == int(g["n_m2l"]) == 5656496
    assert cnt["p2p_leafpairs"] == 2421064
    assert cnt["p2p_bodypairs"] + cnt["p2p_selfpairs"] == int(g["n_p2p_calls"])
    ep, ea = oracle.rel_l2(o.fields()[g["sample"]], g["trg_sample"])
    assert ep < 1e-13 and ea < 1e-13, (ep, ea)


def test_bbox_quirk_excludes_body0_xy():
    """The root box is a cube over every coordinate except body 0's x and y."""
    g = load("bbox_quirk500")
    x, y, z = g["x"], g["y"], g["z"]
    allc = np.concatenate([[z[0]], x[1:], y[1:], z[1:]])
    assert np.all(g["rmin"] == allc.min()) and np.all(g["rmax"] == allc.max())
    assert x[0] < g["rmin"][0] and y[0] > g["rmax"][1]


def test_mutual_fixtures_are_the_symmetric_interaction_of_the_same_tree():
    """Closed-form checks of the mutual fixtures that do not need the oracle:
    the tree does not depend on mutual; on the complete c0 tree every M2L
    call is one unordered pair (half the ordered calls of mutual = 0), and
    numP2P counts the pairs j <= i of a leaf with itself:
    (numP2P(mutual 0) + N) / 2."""
    g0, g1 = load("c0_cube16k"), load("c0_cube16k_mutual")
    for k in ("keys", "cell_key", "cell_nb", "cell_center"):
        assert np.array_equal(g0[k], g1[k])
    assert int(g1["n_m2l"]) * 2 == int(g0["n_m2l"])
    assert int(g1["n_p2p_calls"]) * 2 == int(g0["n_p2p_calls"]) + len(g0["x"])
    # both are FMM approximations of the same sums
    e = oracle.rel_l2(g1["trg"], g0["trg"])
    assert max(e) < 1e-4


@pytest.mark.parametrize("name", ["cube_n200_c16", "plummer3000_c32_t05"])
def test_mutual_oracle_lists_are_symmetric(name):
    """The oracle's mutual traversal records every M2L call in both
    directions, and a leaf pair of two different leaves too; the M2L of
    (i <- j) and (j <- i) use the exactly negated offset."""
    g = load(name + "_mutual")
    o = run_oracle(g, phases=2, record_lists=True)
    m, p = o.lists()
    cnt = o.counts()
    assert len(m) == 2 * cnt["m2l"]
    ms = set(map(tuple, m.tolist()))
    assert all((j, i) in ms for i, j in ms)
    ps = set(map(tuple, p.tolist()))
    assert all((j, i) in ps for i, j in ps)
    nself = sum(1 for i, j in ps if i == j)
    assert len(p) == 2 * cnt["p2p_leafpairs"] - nself

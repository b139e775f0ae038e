"""The CPU oracle pinned by what the mathematics fixes, independent of the
reference's outputs: key-algebra known answers (the reference's own unit
tests, cpp/tests/test_morton_key.cpp), closed forms of the solid harmonics,
exactness of the M2M/L2L translations, convergence of P2M-M2L-L2P to the
direct sum, the gradient returned by L2P, the traversal's coverage invariant,
and the P2P / direct-sum / error-metric closed forms.
"""
import math

import numpy as np
import pytest
import scipy.special as sp

import datasets
import oracle

P = 10


def tri(n, m):
    return n * (n + 1) // 2 + m


def gen_key(path):
    k = 0
    for idx in path:
        k = int(oracle.lib().orc_key_child(np.uint64(k), int(idx)))
    return k


def setup_module(_):
    L = oracle.lib()
    import ctypes
    L.orc_key_child.restype = ctypes.c_uint64
    L.orc_key_child.argtypes = [ctypes.c_uint64, ctypes.c_int]
    L.orc_key_next.restype = ctypes.c_uint64
    L.orc_key_next.argtypes = [ctypes.c_uint64]


# --- key algebra (cpp/tests/test_morton_key.cpp:24-362) ---------------------

def test_key_next_known_answers():
    nxt = oracle.lib().orc_key_next
    assert nxt(gen_key([0, 1, 1])) == gen_key([0, 1, 2])
    assert nxt(gen_key([0, 1, 7])) == gen_key([0, 2, 0])


def test_key_dirs_known_answers():
    """Digits are z,y,x with x least significant: key {7,3,2} = 111-011-010."""
    k = gen_key([7, 3, 2])
    body = k >> 5
    dirs = [[(body >> ((19 - dep) * 3 + dim)) & 1 for dep in (1, 2, 3)] for dim in range(3)]
    assert dirs == [[1, 1, 0], [1, 1, 1], [1, 0, 0]]
    assert k & 31 == 3


def test_cell_region_known_answers():
    """3-D form of Test_Morton_CalcRegion: root [-1,1]^3."""
    rmin, rmax = np.array([-1.0] * 3), np.array([1.0] * 3)
    cases = {
        (): ([-1, -1, -1], [1, 1, 1]),
        (7,): ([0, 0, 0], [1, 1, 1]),
        (0,): ([-1, -1, -1], [0, 0, 0]),
        (7, 7): ([0.5, 0.5, 0.5], [1, 1, 1]),
        (0, 6): ([-1, -0.5, -0.5], [-0.5, 0, 0]),
    }
    for path, (mn_ans, mx_ans) in cases.items():
        mn, mx = oracle.cell_region(gen_key(list(path)), rmin, rmax)
        assert np.array_equal(mn, mn_ans) and np.array_equal(mx, mx_ans), path


def test_finest_key_is_child_path():
    """CalcFinestKey == the depth-19 key reached by Child() along the anchor's
    digits, most significant level first."""
    rng = np.random.default_rng(0)
    for _ in range(50):
        a = rng.integers(0, 1 << 19, 3)
        path = [(((a[2] >> b) & 1) << 2) | (((a[1] >> b) & 1) << 1) | ((a[0] >> b) & 1)
                for b in range(18, -1, -1)]
        assert oracle.finest_key(a) == gen_key(path)


def test_keys_sort_like_cells():
    """A body lies in the cell whose key prefix its key shares: its region
    contains the body (for bodies inside the root box)."""
    x, y, z, q = datasets.cube(64, seed=3)
    mn, mx = oracle.bbox(x, y, z)
    k = oracle.keys(x, y, z, mn, mx)
    for i in range(64):
        cmin, cmax = oracle.cell_region(int(k[i]), mn, mx)
        p = np.array([x[i], y[i], z[i]])
        assert np.all(p >= cmin - 1e-15) and np.all(p <= cmax + 1e-15)


def test_bbox_scalar_cube():
    x, y, z, q = datasets.cube(100, seed=1)
    mn, mx = oracle.bbox(x, y, z)
    allc = np.concatenate([[z[0]], x[1:], y[1:], z[1:]])
    assert np.all(mn == allc.min()) and np.all(mx == allc.max())


# --- solid harmonics ----------------------------------------------------------

ANGLES = [(0.7, 1.1, -2.0), (1.3, 0.3, 0.4), (2.5, 2.9, 3.0), (0.05, 1.7, 1.0)]


@pytest.mark.parametrize("rho,alpha,beta", ANGLES)
def test_multipole_closed_form(rho, alpha, beta):
    """r^n Y_n^m = (-1)^n rho^n P_n^m(cos a) e^{i m b} / (n+m)!  (Condon-Shortley P)."""
    Y, _ = oracle.eval_multipole(rho, alpha, beta)
    for n in range(P):
        for m in range(-n, n + 1):
            am = abs(m)
            ref = ((-1) ** n * rho ** n * sp.lpmv(am, n, math.cos(alpha)) *
                   np.exp(1j * am * beta) / math.factorial(n + am))
            if m < 0:
                ref = np.conj(ref)
            got = Y[n * n + n + m]
            assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (n, m, got, ref)


@pytest.mark.parametrize("rho,alpha,beta", ANGLES)
def test_multipole_theta_derivative(rho, alpha, beta):
    _, T = oracle.eval_multipole(rho, alpha, beta)
    h = 1e-6
    Yp, _ = oracle.eval_multipole(rho, alpha + h, beta)
    Ym, _ = oracle.eval_multipole(rho, alpha - h, beta)
    for n in range(P):
        for m in range(0, n + 1):
            i = n * n + n + m
            fd = (Yp[i] - Ym[i]) / (2 * h)
            assert abs(T[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (n, m)


@pytest.mark.parametrize("rho,alpha,beta", ANGLES)
def test_local_closed_form(rho, alpha, beta):
    """r^{-n-1} Y_n^m = (-1)^n (n-m)! P_n^m(cos a) e^{i m b} / rho^{n+1}."""
    Y = oracle.eval_local(rho, alpha, beta)
    for n in range(P):
        for m in range(-n, n + 1):
            am = abs(m)
            ref = ((-1) ** n * math.factorial(n - am) * sp.lpmv(am, n, math.cos(alpha)) *
                   np.exp(1j * am * beta) / rho ** (n + 1))
            if m < 0:
                ref = np.conj(ref)
            got = Y[n * n + n + m]
            assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (n, m, got, ref)


# --- translations ------------------------------------------------------------

def _cluster(rng, n, center, size):
    pts = center + size * (rng.random((n, 3)) - 0.5)
    return pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), rng.uniform(-1, 1, n)


def test_m2m_of_p2m_is_p2m_at_parent():
    rng = np.random.default_rng(1)
    child = np.array([0.25, -0.25, 0.25])
    parent = np.zeros(3)
    x, y, z, q = _cluster(rng, 30, child, 0.5)
    Mc = oracle.p2m(child, x, y, z, q)
    Mp = oracle.m2m(Mc, parent - child)
    Md = oracle.p2m(parent, x, y, z, q)
    # exact for the low degrees; the shift of a truncated expansion only drops
    # terms of degree >= P, so every coefficient must agree
    assert np.abs(Mp - Md).max() <= 1e-13 * np.abs(Md).max()


def test_l2l_is_exact_reexpansion():
    rng = np.random.default_rng(2)
    xs, ys, zs, qs = _cluster(rng, 20, np.array([6.0, 1.0, -2.0]), 1.0)
    Mj = oracle.p2m(np.array([6.0, 1.0, -2.0]), xs, ys, zs, qs)
    a = np.zeros(3)
    La = oracle.m2l(Mj, a - np.array([6.0, 1.0, -2.0]))
    b = np.array([0.2, -0.1, 0.15])
    Lb = oracle.l2l(La, b - a)
    t = np.array([0.3, -0.05, 0.1])
    pa = oracle.l2p(np.zeros(4), t, 1.0, La, a)
    pb = oracle.l2p(np.zeros(4), t, 1.0, Lb, b)
    assert np.abs(pa - pb).max() <= 1e-12 * np.abs(pa).max()


def _direct_field(t, xs, ys, zs, qs):
    d = t[None, :] - np.stack([xs, ys, zs], 1)
    r = np.sqrt((d * d).sum(1))
    pot = (qs / r).sum()
    acc = -(qs[:, None] * d / r[:, None] ** 3).sum(0)
    return np.concatenate([[pot], acc])


def _m2l_error(dist, seed=3):
    rng = np.random.default_rng(seed)
    src = np.array([dist, 0.3, -0.2])
    xs, ys, zs, qs = _cluster(rng, 40, src, 1.0)
    M = oracle.p2m(src, xs, ys, zs, qs)
    tc = np.zeros(3)
    L = oracle.m2l(M, tc - src)
    errs = []
    for _ in range(10):
        t = (rng.random(3) - 0.5)
        got = oracle.l2p(np.zeros(4), t, 1.0, L, tc)
        ref = _direct_field(t, xs, ys, zs, qs)
        errs.append(np.abs(got - ref).max() / np.abs(ref).max())
    return max(errs)


def test_p2m_m2l_l2p_converges_to_direct():
    """Error of the truncated expansions: below ((rs+rt)/d)^P / (1 - (rs+rt)/d)
    (rs, rt the source and target radii, d the centre distance), and falling
    by about 2^P when d doubles."""
    rs = rt = np.sqrt(3) / 2
    errs = {}
    for d in (4.0, 8.0, 16.0):
        ratio = (rs + rt) / np.hypot(d, np.hypot(0.3, 0.2))
        errs[d] = _m2l_error(d)
        assert errs[d] < 3 * ratio ** P / (1 - ratio), (d, errs[d])
    assert errs[8.0] < errs[4.0] / 100 and errs[16.0] < errs[8.0] / 100, errs


def test_l2p_acceleration_is_potential_gradient():
    rng = np.random.default_rng(4)
    src = np.array([5.0, -1.0, 2.0])
    xs, ys, zs, qs = _cluster(rng, 10, src, 0.5)
    L = oracle.m2l(oracle.p2m(src, xs, ys, zs, qs), -src)
    t = np.array([0.21, -0.13, 0.07])
    f = oracle.l2p(np.zeros(4), t, 1.0, L, np.zeros(3))
    h = 1e-6
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        fd = (oracle.l2p(np.zeros(4), t + e, 1.0, L, np.zeros(3))[0] -
              oracle.l2p(np.zeros(4), t - e, 1.0, L, np.zeros(3))[0]) / (2 * h)
        assert abs(f[1 + d] - fd) < 1e-6 * np.abs(f[1:]).max()


def test_l2p_divides_p2p_sum_by_charge():
    trg = np.array([2.0, 4.0, -6.0, 8.0])
    out = oracle.l2p(trg, np.array([0.1, 0.2, 0.3]), 2.0, np.zeros(55, np.complex128),
                     np.zeros(3))
    assert np.array_equal(out, trg / 2.0)


# --- traversal and near field --------------------------------------------------

@pytest.mark.parametrize("dist,n,ncrit,theta", [("cube", 300, 8, 0.4), ("plummer", 400, 4, 0.6),
                                                ("ellipsoid", 300, 16, 0.3)])
def test_traversal_covers_every_pair_once(dist, n, ncrit, theta):
    x, y, z, q = datasets.make(dist, n, seed=5)
    o = oracle.Run(x, y, z, q, ncrit, theta, phases=2, record_lists=True)
    c = o.cells()
    keys, perm = o.sorted()
    m2l, p2p = o.lists()
    # body ranges per cell (sorted positions): [bid, bid+nb) for leaves;
    # for internal cells the range of its keys
    nc = o.ncells
    lo = np.zeros(nc, np.int64)
    hi = np.zeros(nc, np.int64)
    for i in range(nc):
        k = int(c["key"][i])
        d = k & 31
        nxt = k + (1 << (3 * (19 - d) + 5))
        lo[i] = np.searchsorted(keys, np.uint64(k), "left")
        hi[i] = np.searchsorted(keys, np.uint64(nxt), "left") if nxt < (1 << 64) else n
    cover = np.zeros((n, n), np.int32)
    for a, b in m2l:
        cover[lo[a]:hi[a], lo[b]:hi[b]] += 1
    for a, b in p2p:
        cover[lo[a]:hi[a], lo[b]:hi[b]] += 1
    assert np.all(cover == 1)


def test_p2p_two_bodies_closed_form():
    x = np.array([0.0, 0.3])
    y = np.array([0.0, 0.4])
    z = np.array([0.0, 0.0])
    q = np.array([2.0, -3.0])
    out = oracle.direct(x, y, z, q, [0, 1])
    r = 0.5
    # pot_i = q_j / r, acc_i = q_j (x_j - x_i) / r^3 (gradient of pot_i)
    assert np.allclose(out[0], [-3.0 / r, -3.0 * 0.3 / r ** 3, -3.0 * 0.4 / r ** 3, 0],
                       rtol=1e-15, atol=0)
    assert np.allclose(out[1], [2.0 / r, -2.0 * 0.3 / r ** 3, -2.0 * 0.4 / r ** 3, 0],
                       rtol=1e-15, atol=0)


def test_direct_matches_numpy():
    x, y, z, q = datasets.cube(300, seed=7)
    t = np.arange(0, 300, 7)
    got = oracle.direct(x, y, z, q, t)
    for k, i in enumerate(t):
        mask = np.arange(300) != i
        ref = _direct_field(np.array([x[i], y[i], z[i]]), x[mask], y[mask], z[mask], q[mask])
        assert np.allclose(got[k], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_fmm_error_against_direct_small():
    """Full oracle FMM within the FMM truncation error of the direct sum."""
    x, y, z, q = datasets.cube(4000, seed=8)
    o = oracle.Run(x, y, z, q, 32, 0.4)
    t = np.arange(0, 4000, 13)
    ep, ea = oracle.rel_l2(o.fields()[t], oracle.direct(x, y, z, q, t))
    assert ep < 1e-4 and ea < 1e-4


def test_rel_l2_closed_form():
    ref = np.random.default_rng(0).normal(size=(50, 4))
    ep, ea = oracle.rel_l2(ref * (1 + 1e-3), ref)
    assert abs(ep - 1e-3) < 1e-12 and abs(ea - 1e-3) < 1e-12


def test_root_leaf_is_direct_sum():
    """n <= ncrit: the root is the only leaf and the FMM is the direct sum
    (the reference aborts in this case; see the golden script)."""
    x, y, z, q = datasets.cube(50, seed=9)
    o = oracle.Run(x, y, z, q, 64, 0.4)
    assert o.ncells == 1
    ref = oracle.direct(x, y, z, q, np.arange(50))
    assert np.allclose(o.fields(), ref, rtol=1e-13, atol=1e-13)


def test_batch_generator_follows_its_definition():
    """The batched workload: sizes log-uniform within [2^10, 2^16], the three
    kinds, charges in [-1, 1], ellipsoid semi-axes 1 : 0.3 : 0.1."""
    import datasets
    probs, kinds = datasets.batch(64, seed=3)
    sizes = np.array([len(p[0]) for p in probs])
    assert sizes.min() >= 1 << 10 and sizes.max() <= 1 << 16
    assert set(kinds) <= {"cube", "plummer", "ellipsoid"} and len(set(kinds)) == 3
    for (x, y, z, q), k in zip(probs, kinds):
        assert np.all(np.abs(q) <= 1)
        if k == "ellipsoid":
            assert np.all((x / 1.0) ** 2 + (y / 0.3) ** 2 + (z / 0.1) ** 2 <= 1 + 1e-12)
        if k == "cube":
            assert x.min() >= 0 and x.max() < 1
    # log-uniform: the median size is near the geometric mean 2^13
    assert 2 ** 11.5 < np.median(sizes) < 2 ** 14.5

"""The oracle's Cartesian Taylor basis (the reference's BASIS = Cartesian
build, sample/exafmm-dev-13274dd4ac68/kernels/LaplaceCartesianCPU.cxx)
pinned three ways:

* golden vectors of the reference's own kernels (tests/golden/
  cartesian_ops.npz, written by tests/golden/make_golden.py through
  oracle/_ref/ref_cartesian): P2M, M2M, M2L with mutual = 1 and in a periodic
  image, L2L and L2P on fixed cells and bodies, to 1e-13;
* what the mathematics fixes: closed forms of the low derivatives of 1/r,
  a central difference of a degree-8 term, the harmonic identity
  sum_i D^{k+2e_i}(1/r) = 0 (a wrong coefficient or index in the recurrence
  breaks it), exactness of M2M and L2L (a multipole shifted in two steps is
  the one made at the far centre; a local polynomial re-expanded about a
  child gives the same values), and P2M-M2L-L2P convergence to the direct sum;
* the two bases' truncations are the same (total degree P - 1 of the double
  Taylor series), so the whole Cartesian FMM gives the spherical FMM's
  fields to rounding, with mutual = 0 and 1 and with periodic images; the
  spherical oracle is itself pinned to the reference's TapasFMM outputs.
"""
import os

import numpy as np
import pytest

import datasets
import oracle
from conftest import GOLDEN

P = oracle.P
CN = oracle.CN
TERMS = oracle.cart_terms()


def relmax(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def ops():
    return np.load(os.path.join(GOLDEN, "cartesian_ops.npz"))


def test_index_order():
    """the reference's term order: degree by degree, x falling, then y"""
    assert CN == 220 and oracle.lib().orc_cart_terms() == CN
    assert TERMS[:10] == [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0),
                          (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]
    for i, m in enumerate(TERMS):
        assert oracle.cart_index(*m) == i


def test_reference_p2m_m2m(ops):
    c, M = ops["center"], ops["M"]
    ib, nb = ops["cell_ibody"], ops["cell_nbody"]
    x, q = ops["x"], ops["q"]
    for ci in range(len(M)):
        if nb[ci] == 0:
            continue
        s = slice(ib[ci], ib[ci] + nb[ci])
        assert relmax(oracle.cart_p2m(c[ci], x[s, 0], x[s, 1], x[s, 2], q[s]), M[ci]) < 1e-13
    Mp = np.zeros(CN)
    for ch in range(ops["cell_ichild"][0], ops["cell_ichild"][0] + ops["cell_nchild"][0]):
        Mp = oracle.cart_m2m(M[ch], c[0] - c[ch], Mp)
    assert relmax(Mp, M[0]) < 1e-13


def test_reference_m2l(ops):
    c, M, L, xp = ops["center"], ops["M"], ops["L"], ops["xperiodic"]
    Li, Lj = oracle.cart_m2l_mutual(M[9], M[10], c[9] - c[10] - xp[0])
    assert relmax(Li, L[9]) < 1e-13
    assert relmax(Lj, L[10]) < 1e-13
    assert relmax(oracle.cart_m2l(M[10], c[11] - c[10] - xp[1]), L[11]) < 1e-13


def test_reference_l2l_l2p(ops):
    c, L = ops["center"], ops["L"]
    x, q, tin, tout = ops["x"], ops["q"], ops["trg_in"], ops["trg_out"]
    par = 12
    for ch in range(ops["cell_ichild"][par], ops["cell_ichild"][par] + ops["cell_nchild"][par]):
        assert ops["cell_iparent"][ch] == par
        assert relmax(oracle.cart_l2l(L[par], c[ch] - c[par]), L[ch]) < 1e-13
        for i in range(ops["cell_ibody"][ch], ops["cell_ibody"][ch] + ops["cell_nbody"][ch]):
            t = oracle.cart_l2p(tin[i], x[i], q[i], L[ch], c[ch])
            assert relmax(t, tout[i]) < 1e-13, i


def test_derivative_closed_forms():
    dX = np.array([0.3, -0.7, 0.5])
    x, y, z = dX
    r = np.sqrt(dX @ dX)
    C = oracle.cart_derivs(dX)
    I = oracle.cart_index
    known = {
        (0, 0, 0): 1 / r,
        (1, 0, 0): -x / r**3,
        (0, 0, 1): -z / r**3,
        (2, 0, 0): 3 * x * x / r**5 - 1 / r**3,
        (1, 1, 0): 3 * x * y / r**5,
        (1, 1, 1): -15 * x * y * z / r**7,
        (3, 0, 0): -15 * x**3 / r**7 + 9 * x / r**5,
        (2, 0, 1): -15 * x * x * z / r**7 + 3 * z / r**5,
    }
    for k, v in known.items():
        assert abs(C[I(*k)] - v) <= 1e-13 * abs(v), k


def test_derivative_central_difference():
    """D^{k+e_x} = d/dx D^k for a degree-8 term (fourth-order difference)"""
    dX = np.array([0.9, -0.4, 0.6])
    h = 1e-3
    k = (3, 2, 3)
    I = oracle.cart_index
    f = [oracle.cart_derivs(dX + s * np.array([h, 0, 0]))[I(*k)] for s in (-2, -1, 1, 2)]
    fd = (f[0] - 8 * f[1] + 8 * f[2] - f[3]) / (12 * h)
    exact = oracle.cart_derivs(dX)[I(k[0] + 1, k[1], k[2])]
    assert abs(fd - exact) < 1e-7 * abs(exact)


@pytest.mark.parametrize("dX", [(0.3, -0.7, 0.5), (2.0, 0.1, -1.3), (-0.05, 0.2, 0.01)])
def test_derivatives_harmonic(dX):
    C = oracle.cart_derivs(np.array(dX))
    I = oracle.cart_index
    for (a, b, c) in TERMS:
        if a + b + c > P - 3:
            continue
        parts = [C[I(a + 2, b, c)], C[I(a, b + 2, c)], C[I(a, b, c + 2)]]
        assert abs(sum(parts)) <= 1e-12 * max(abs(p) for p in parts), (a, b, c)


def test_m2m_exact():
    """P2M about a child centre then M2M to the parent = P2M about the parent"""
    r = np.random.default_rng(3)
    x, y, z = [r.uniform(-0.2, 0.2, 30) for _ in range(3)]
    q = r.uniform(-1, 1, 30)
    child = np.array([0.05, -0.1, 0.02])
    parent = np.array([0.3, 0.2, -0.25])
    Mc = oracle.cart_p2m(child, x, y, z, q)
    Mp = oracle.cart_m2m(Mc, parent - child)
    assert relmax(Mp, oracle.cart_p2m(parent, x, y, z, q)) < 1e-13


def test_l2l_exact():
    """a local polynomial re-expanded about a child centre has the same values"""
    r = np.random.default_rng(4)
    Lp = r.standard_normal(CN)
    parent = np.array([0.0, 0.0, 0.0])
    child = np.array([0.12, -0.07, 0.2])
    Lc = oracle.cart_l2l(Lp, child - parent)
    for _ in range(5):
        b = child + r.uniform(-0.1, 0.1, 3)
        t0 = oracle.cart_l2p(np.zeros(4), b, 1.0, Lp, parent)
        t1 = oracle.cart_l2p(np.zeros(4), b, 1.0, Lc, child)
        assert relmax(t1, t0) < 1e-12


def test_m2l_converges_to_direct():
    r = np.random.default_rng(5)
    n = 40
    src = np.array([0.0, 0.0, 0.0])
    x, y, z = [r.uniform(-0.25, 0.25, n) for _ in range(3)]
    q = r.uniform(0.1, 1.0, n)
    M = oracle.cart_p2m(src, x, y, z, q)
    errs = []
    for dist in (1.5, 3.0, 6.0):
        tgt = np.array([dist, 0.3 * dist, -0.2 * dist])
        L = oracle.cart_m2l(M, tgt - src)
        b = tgt + np.array([0.1, -0.05, 0.08])
        f = oracle.cart_l2p(np.zeros(4), b, 1.0, L, tgt)
        d = np.zeros(4)
        for j in range(n):
            dx = b - np.array([x[j], y[j], z[j]])
            rr = np.sqrt(dx @ dx)
            d[0] += q[j] / rr
            d[1:] -= q[j] * dx / rr**3
        errs.append(relmax(f, d))
    assert errs[0] < 1e-5 and errs[2] < 1e-9
    assert errs[1] < errs[0] / 50 and errs[2] < errs[1] / 50, errs


@pytest.mark.parametrize("mutual", [0, 1])
def test_cartesian_fmm_equals_spherical(mutual):
    x, y, z, q = datasets.plummer(3000, seed=5, charges="uniform")
    fs = oracle.Run(x, y, z, q, 32, 0.5, mutual=mutual).fields()
    fc = oracle.Run(x, y, z, q, 32, 0.5, mutual=mutual, basis=oracle.CARTESIAN).fields()
    ep, ea = oracle.rel_l2(fc, fs)
    assert ep < 1e-12 and ea < 1e-12, (ep, ea)
    t = np.arange(0, 3000, 13)
    dp, da = oracle.rel_l2(fc[t], oracle.direct(x, y, z, q, t))
    assert dp < 1e-3 and da < 1e-3  # the reference driver's gate


def test_cartesian_periodic_equals_spherical():
    r = np.random.default_rng(1)
    x, y, z = [r.uniform(-np.pi, np.pi, 400) for _ in range(3)]
    q = r.random(400) - 0.5
    fs = oracle.Run(x, y, z, q, 32, 0.4, images=2, cycle=2 * np.pi).fields()
    fc = oracle.Run(x, y, z, q, 32, 0.4, images=2, cycle=2 * np.pi,
                    basis=oracle.CARTESIAN).fields()
    ep, ea = oracle.rel_l2(fc, fs)
    assert ep < 1e-12 and ea < 1e-12, (ep, ea)


def test_cartesian_expansions_shape():
    x, y, z, q = datasets.cube(300, seed=2)
    r = oracle.Run(x, y, z, q, 16, 0.4, basis=oracle.CARTESIAN)
    M, L = r.expansions()
    assert M.shape == (r.ncells, CN) and L.shape == (r.ncells, CN)
    # the root's monopole is the total charge
    assert abs(M[0, 0] - q.sum()) < 1e-12 * np.abs(q).sum()

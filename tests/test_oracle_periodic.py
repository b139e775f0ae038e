"""The oracle's periodic images (ExaFMM's images, EX/include/traversal.h:
126-185 and 247-268), pinned by mathematics: the periodic FMM against the
direct sum over the same (3^images)^3 copies of the bodies (the reference's
CheckResult with prange, tapas_exafmm.cxx:314-345).

* images = 1 with a MAC that accepts nothing (theta 0.01): every pair is a
  P2P of the 27 first images, so FMM and direct sum agree to rounding;
* images = 1, 2, 3 at theta 0.4: within the expansion error (the far field
  of images >= 2 through multipoles of blocks of copies);
* a planted mistake -- the block multipole without its centre copy, which is
  what resetting the parent before the M2M (as the vendored reference loop
  does) gives -- fails the images = 3 check by four orders of magnitude
  (3e-2 against 8e-8 in potential, measured while writing this test);
* translation invariance: moving every body by the same vector (inside one
  period) leaves the periodic fields unchanged up to the expansion error.
Bodies: the reference's cube (U[-pi, pi)^3, period 2 pi) with charges
U[-1/2, 1/2) -- not neutral, so the far copies' monopoles weigh in."""
import numpy as np
import pytest

import oracle

CYCLE = 2 * np.pi


def bodies(n=400, seed=1, half=np.pi):
    r = np.random.default_rng(seed)
    x, y, z = [r.uniform(-half, half, n) for _ in range(3)]
    return x, y, z, r.random(n) - 0.5


def fields(b, images, theta=0.4, ncrit=32):
    return oracle.Run(*b, ncrit, theta, images=images, cycle=CYCLE).fields()


def test_first_images_all_p2p_equal_direct():
    b = bodies()
    t = np.arange(0, 400, 3)
    f = fields(b, 1, theta=0.01)
    d = oracle.direct_periodic(*b, t, 1, CYCLE)
    ep, ea = oracle.rel_l2(f[t], d)
    assert ep < 1e-13 and ea < 1e-13, (ep, ea)


@pytest.mark.parametrize("images", [1, 2, 3])
def test_periodic_fmm_matches_periodic_direct(images):
    b = bodies()
    t = np.arange(0, 400, 7)
    f = fields(b, images)
    d = oracle.direct_periodic(*b, t, images, CYCLE)
    ep, ea = oracle.rel_l2(f[t], d)
    assert ep < 1e-5 and ea < 1e-5, (images, ep, ea)
    if images > 1:  # the far copies matter: fewer images is far off
        ep1, _ = oracle.rel_l2(fields(b, images - 1)[t], d)
        assert ep1 > 100 * ep, (images, ep1, ep)


def test_periodic_translation_invariance():
    x, y, z, q = bodies(300, seed=4, half=2.5)  # room to move inside one period
    sh = np.array([0.3, -0.2, 0.1])
    f0 = fields((x, y, z, q), 2)
    f1 = fields((x + sh[0], y + sh[1], z + sh[2], q), 2)
    ep, ea = oracle.rel_l2(f1, f0)
    # two trees, two expansion errors (each about 1e-5 in acceleration here)
    assert ep < 5e-5 and ea < 5e-5, (ep, ea)


def test_periodic_region_is_the_period():
    b = bodies(200, seed=2)
    r = oracle.Run(*b, 32, 0.4, images=1, cycle=CYCLE)
    mn, mx = r.region()
    assert np.allclose(mx - mn, CYCLE)
    x, y, z, q = b
    lo = min(x.min(), y.min(), z.min())
    hi = max(x.max(), y.max(), z.max())
    assert np.allclose(0.5 * (mn + mx), 0.5 * (lo + hi))
    with pytest.raises(ValueError):  # bodies spanning more than one period
        oracle.Run(x * 2, y, z, q, 32, 0.4, images=1, cycle=CYCLE)

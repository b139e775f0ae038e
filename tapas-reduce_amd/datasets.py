"""Seeded body generators for the FMM workloads.

Each generator returns float64 arrays ``(x, y, z, q)`` of length ``n``.  The
reference ships its own generators (sample/exafmm-dev-13274dd4ac68/include/
dataset.h:131-310, srand48-based); the workloads measured here use different
distributions, so both the HIP library and the CPU checkers are fed the same
bodies from these functions.
"""
import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def zero_net_charges(rng, n):
    """Charges uniform in [-1, 1], shifted so that they sum to zero."""
    q = rng.uniform(-1.0, 1.0, n)
    return q - q.mean()


def cube(n, seed=0):
    """Uniform i.i.d. positions in [0, 1)^3, zero-net charges in [-1, 1]."""
    rng = _rng(seed)
    pos = rng.random((3, n))
    q = zero_net_charges(rng, n)
    return pos[0].copy(), pos[1].copy(), pos[2].copy(), q


def _plummer_radii(rng, n, a, rmax):
    out = np.empty(n)
    filled = 0
    while filled < n:
        m = int((n - filled) * 1.05) + 16
        u = rng.random(m)
        u = u[u > 0]
        r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = r[r <= rmax]
        take = min(n - filled, r.size)
        out[filled:filled + take] = r[:take]
        filled += take
    return out


def _isotropic(rng, n):
    cz = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    s = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
    return s * np.cos(ph), s * np.sin(ph), cz


def plummer(n, seed=0, a=1.0, rtrunc=20.0, charges="equal"):
    """Plummer sphere with scale radius ``a``, radii truncated at ``rtrunc*a``,
    isotropic directions.  ``charges='equal'`` gives q = 1/n."""
    rng = _rng(seed)
    r = _plummer_radii(rng, n, a, rtrunc * a)
    ux, uy, uz = _isotropic(rng, n)
    if charges == "equal":
        q = np.full(n, 1.0 / n)
    else:
        q = rng.uniform(-1.0, 1.0, n)
    return r * ux, r * uy, r * uz, q


def ellipsoid(n, seed=0, axes=(1.0, 0.3, 0.1), charges="uniform"):
    """Uniform density inside an ellipsoid with the given semi-axes."""
    rng = _rng(seed)
    ux, uy, uz = _isotropic(rng, n)
    r = rng.random(n) ** (1.0 / 3.0)
    q = rng.uniform(-1.0, 1.0, n) if charges == "uniform" else np.full(n, 1.0 / n)
    return axes[0] * r * ux, axes[1] * r * uy, axes[2] * r * uz, q


def two_plummer(n, seed=0, a=1.0, separation=10.0, mass_ratio=(1, 3), rtrunc=20.0):
    """Two Plummer clusters ``separation*a`` apart along x, body counts split
    by ``mass_ratio`` (equal charges 1/n, so the cluster masses follow it)."""
    rng = _rng(seed)
    w0, w1 = mass_ratio
    n0 = int(round(n * w0 / float(w0 + w1)))
    n1 = n - n0
    xs, ys, zs = [], [], []
    for cnt, cx in ((n0, -0.5 * separation * a), (n1, 0.5 * separation * a)):
        r = _plummer_radii(rng, cnt, a, rtrunc * a)
        ux, uy, uz = _isotropic(rng, cnt)
        xs.append(r * ux + cx)
        ys.append(r * uy)
        zs.append(r * uz)
    perm = rng.permutation(n)
    x = np.concatenate(xs)[perm]
    y = np.concatenate(ys)[perm]
    z = np.concatenate(zs)[perm]
    return x, y, z, np.full(n, 1.0 / n)


GENERATORS = {
    "cube": cube,
    "plummer": plummer,
    "ellipsoid": ellipsoid,
    "two_plummer": two_plummer,
}


def make(dist, n, seed=0, **kw):
    return GENERATORS[dist](n, seed=seed, **kw)


BATCH_KINDS = ("cube", "plummer", "ellipsoid")


def batch(nprob, seed=0, nmin=1 << 10, nmax=1 << 16):
    """Independent small problems: N log-uniform in [nmin, nmax], each drawn
    from the uniform cube, a Plummer sphere (a = 1) or the 1:0.3:0.1
    ellipsoid (uniform random choice), charges uniform in [-1, 1].
    Returns a list of (x, y, z, q) and the list of kinds."""
    rng = _rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nprob)).astype(np.int64)
    sizes = np.clip(sizes, nmin, nmax)
    kinds = [BATCH_KINDS[k] for k in rng.integers(0, len(BATCH_KINDS), nprob)]
    seeds = rng.integers(0, 2 ** 62, nprob)
    out = []
    for n, kind, s in zip(sizes, kinds, seeds):
        if kind == "cube":
            r = _rng(int(s))
            pos = r.random((3, int(n)))
            p = (pos[0].copy(), pos[1].copy(), pos[2].copy(), r.uniform(-1.0, 1.0, int(n)))
        elif kind == "plummer":
            p = plummer(int(n), seed=int(s), charges="uniform")
        else:
            p = ellipsoid(int(n), seed=int(s))
        out.append(p)
    return out, kinds

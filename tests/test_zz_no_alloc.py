"""tfmm_direct_noalloc (the direct sum staged in the library's static device
array, no device allocation) against tfmm_direct, through the C ABI: bit for
bit on two calls in a row (the second smaller), in a batch whose targets
alternate between problems, and its size limit."""
import numpy as np
import pytest

import datasets
import tfmm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = tfmm.Context(0)
    yield c
    c.close()


def test_noalloc_direct_equals_direct(ctx):
    x, y, z, q = datasets.plummer(50000, seed=31, charges="uniform")
    ctx.set_bodies(x, y, z, q)
    rng = np.random.default_rng(4)
    for nt in (5000, 700):  # second call smaller: stale entries of the first must not leak
        t = rng.choice(len(x), nt, replace=False)
        print("noalloc direct: n=%d targets=%d" % (len(x), nt), flush=True)
        a = ctx.direct_noalloc(t)
        b = ctx.direct(t)
        assert np.array_equal(a, b)
        assert np.isfinite(a).all()


def test_noalloc_direct_batch_and_limit(ctx):
    probs = [datasets.cube(3000, seed=1), datasets.plummer(2000, seed=2, charges="uniform"),
             datasets.ellipsoid(1500, seed=3)]
    ctx.set_batch(probs)
    off = ctx.problem_offsets()
    rng = np.random.default_rng(5)
    t = rng.integers(0, int(off[-1]), 4000)  # interleaved problems: one launch per run
    print("noalloc direct batch: %d problems, targets=%d" % (len(probs), len(t)), flush=True)
    assert np.array_equal(ctx.direct_noalloc(t), ctx.direct(t))
    t2 = np.sort(t[:1000])
    assert np.array_equal(ctx.direct_noalloc(t2), ctx.direct(t2))
    with pytest.raises(tfmm.TfmmError):
        ctx.direct_noalloc(np.zeros(tfmm.DIRECT_NOALLOC_MAX + 1, np.int64))
    with pytest.raises(tfmm.TfmmError):
        ctx.direct_noalloc(np.array([int(off[-1])], np.int64))

"""Periodic images on the GPU (tfmm_set_periodic), through the C ABI, against
the CPU oracle's restatement of ExaFMM's periodic traversal
(EX/include/traversal.h:126-185, 247-268) and the periodic direct sum:

* FP64: fields within relative L2 1e-10 of the oracle for images 1, 2 and 3;
  numM2L and numP2P equal; the interaction lists equal as sets of (target,
  source, image); the FMM within 2e-4 (its expansion error) of the direct
  sum over the same (3^images)^3 copies, and tfmm_direct equal to the
  oracle's periodic direct sum to 1e-12;
* mixed precision within 1e-4 of the FP64 periodic direct sum;
* misuse is refused (a batch, mutual interaction, bodies wider than the
  period, recorded lists, bad arguments), and images 0 gives back the
  ordinary evaluation bit for bit."""
import numpy as np
import pytest

import oracle
import tfmm

pytestmark = pytest.mark.gpu

CYCLE = 2 * np.pi
MASK = (1 << 27) - 1


@pytest.fixture(scope="module")
def ctx():
    c = tfmm.Context(0)
    yield c
    c.close()


def bodies(n, seed, half=np.pi):
    r = np.random.default_rng(seed)
    x, y, z = [r.uniform(-half, half, n) for _ in range(3)]
    return x, y, z, r.random(n) - 0.5


@pytest.mark.parametrize("images", [1, 2, 3])
def test_gpu_periodic_fp64_matches_oracle(ctx, images):
    b = bodies(3000, seed=images)
    print("periodic images %d" % images, flush=True)
    ctx.set_periodic(images, CYCLE)
    try:
        ctx.set_bodies(*b)
        ctx.evaluate(tfmm.make_params(32, 0.4, tfmm.FP64))
        f = ctx.fields()
        st = ctx.stats()
        o = oracle.Run(*b, 32, 0.4, images=images, cycle=CYCLE, record_lists=True)
        ep, ea = oracle.rel_l2(f, o.fields())
        assert ep < 1e-10 and ea < 1e-10, (images, ep, ea)
        cnt = o.counts()
        assert st["n_m2l"] == cnt["m2l"]
        assert st["n_p2p_calls"] == cnt["p2p_bodypairs"] + cnt["p2p_selfpairs"]
        # lists as sets of (target key, source key, image)
        c = ctx.cells()
        (mo, ms), (po, ps) = ctx.lists()
        oc = o.cells()
        om, op = o.lists()
        onb = oc["nb"]

        def gpu_set(off, src):
            tgt = np.repeat(np.arange(len(off) - 1), np.diff(off))
            s = src.astype(np.int64) & 0xFFFFFFFF
            return set(zip(c["key"][tgt].tolist(), c["key"][s & MASK].tolist(),
                           (s >> 27).tolist()))

        def orc_set(lst, nonempty):
            cell = lst[:, 1] & MASK
            if nonempty:
                keep = (onb[lst[:, 0]] > 0) & (onb[cell] > 0)
                lst, cell = lst[keep], cell[keep]
            return set(zip(oc["key"][lst[:, 0]].tolist(), oc["key"][cell].tolist(),
                           (lst[:, 1] >> 27).tolist()))

        assert gpu_set(mo, ms) == orc_set(om, True)
        assert gpu_set(po, ps) == orc_set(op, False)
        # the periodic direct sum
        t = np.arange(0, 3000, 97 if images < 3 else 300)  # 3^(3 images) copies per body
        d = ctx.direct(t)
        od = oracle.direct_periodic(*b, t, images, CYCLE)
        assert np.allclose(d, od, rtol=1e-12, atol=1e-12 * np.abs(od).max())
        ep, ea = oracle.rel_l2(f[t], d)  # the expansion error (P = 10, theta 0.4)
        assert ep < 2e-4 and ea < 2e-4, (images, ep, ea)
    finally:
        ctx.set_periodic(0)


def test_gpu_periodic_mixed(ctx):
    b = bodies(20000, seed=7)
    ctx.set_periodic(2, CYCLE)
    try:
        ctx.set_bodies(*b)
        ctx.evaluate(tfmm.make_params(64, 0.4, tfmm.MIXED))
        f = ctx.fields()
        t = np.arange(0, 20000, 211)
        d = ctx.direct(t)
        ep, ea = oracle.rel_l2(f[t], d)
        assert ep < 1e-4 and ea < 1e-4, (ep, ea)
    finally:
        ctx.set_periodic(0)


def test_gpu_periodic_refusals_and_off(ctx):
    b = bodies(2000, seed=3)
    prm = tfmm.make_params(32, 0.4, tfmm.FP64)
    ctx.set_bodies(*b)
    ctx.evaluate(prm)
    plain = ctx.fields()
    for images, cyc in ((-1, CYCLE), (tfmm.PERIODIC_MAX_IMAGES + 1, CYCLE), (1, 0.0),
                        (1, float("nan")), (1, -1.0)):
        with pytest.raises(tfmm.TfmmError):
            ctx.set_periodic(images, cyc)
    ctx.set_periodic(1, CYCLE)
    try:
        with pytest.raises(tfmm.TfmmError):  # mutual interaction
            ctx.evaluate(tfmm.make_params(32, 0.4, tfmm.FP64, mutual=1))
        x, y, z, q = b
        ctx.set_bodies(2.5 * x, y, z, q)  # wider than the period
        with pytest.raises(tfmm.TfmmError):
            ctx.evaluate(prm)
        ctx.set_batch([b, b])
        with pytest.raises(tfmm.TfmmError):  # a batch
            ctx.evaluate(prm)
        ctx.set_bodies(*b)
        ctx.evaluate(prm)
        (mo, ms), (po, ps) = ctx.lists()
        nm = int(mo[-1])
        tgt = np.repeat(np.arange(len(mo) - 1), np.diff(mo)).astype(np.int32)
        with pytest.raises(tfmm.TfmmError):  # recorded lists
            ctx.traverse_lists(np.stack([tgt, ms[:nm]], 1), np.zeros((0, 2), np.int32))
    finally:
        ctx.set_periodic(0)
    ctx.set_bodies(*b)
    ctx.evaluate(prm)
    assert np.array_equal(ctx.fields(), plain)

"""Batched independent problems evaluated as one forest (through the C ABI):
the batch gives, problem by problem, what evaluating each problem alone
gives -- keys, cells and interaction counts bit for bit, fields up to the
summation order of the M2L work units (which depends on how many items the
whole forest has per tile): relative L2 1e-12 in FP64, 1e-6 in mixed
precision -- and the FP64 batch matches the CPU oracle."""
import numpy as np
import pytest

import datasets
import oracle
import tfmm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = tfmm.Context(0)
    yield c
    c.close()


def problems():
    probs, kinds = datasets.batch(9, seed=5, nmin=200, nmax=20000)
    # edge cases: one body, a root leaf, a tight pair of clumps
    probs.append((np.array([0.3]), np.array([0.2]), np.array([0.9]), np.array([1.0])))
    x, y, z, q = datasets.cube(40, seed=2)
    probs.append((x, y, z, q))
    r = np.random.default_rng(3)
    x = np.repeat([0.1, 0.7], 35) + r.normal(0, 1e-3, 70)
    probs.append((x, r.random(70), r.random(70), r.uniform(-1, 1, 70)))
    return probs


@pytest.mark.parametrize("precision", [tfmm.FP64, tfmm.MIXED])
def test_gpu_batch_matches_single_problems(ctx, precision):
    probs = problems()
    ncrit, theta = 32, 0.4
    prm = tfmm.make_params(ncrit, theta, precision)
    print("batch of %d problems, precision %d" % (len(probs), precision), flush=True)
    ctx.set_batch(probs)
    ctx.evaluate(prm)
    fb = ctx.fields()
    sb = ctx.stats()
    keys_b, perm_b = ctx.sorted()
    off = ctx.problem_offsets()
    assert sb["n_problems"] == len(probs)
    assert np.array_equal(off, np.concatenate([[0], np.cumsum([len(p[0]) for p in probs])]))
    ncells = nm2l = np2p = 0
    for i, p in enumerate(probs):
        print("problem %d n=%d" % (i, len(p[0])), flush=True)
        ctx.set_bodies(*p)
        ctx.evaluate(prm)
        f = ctx.fields()
        s = ctx.stats()
        k, perm = ctx.sorted()
        seg = fb[off[i]:off[i + 1]]
        if not np.any(f):  # a lone body feels nothing
            assert not np.any(seg)
        else:
            ep, ea = oracle.rel_l2(seg, f)
            tol = 1e-12 if precision == tfmm.FP64 else 1e-6
            assert ep < tol and ea < tol, ("fields of problem %d" % i, ep, ea)
        assert np.array_equal(keys_b[off[i]:off[i + 1]], k)
        assert np.array_equal(np.sort(perm_b[off[i]:off[i + 1]] - off[i]), np.arange(len(k)))
        ncells += s["ncells"]
        nm2l += s["n_m2l"]
        np2p += s["n_p2p_calls"]
    assert sb["ncells"] == ncells and sb["n_m2l"] == nm2l and sb["n_p2p_calls"] == np2p


def test_gpu_batch_against_oracle(ctx):
    """FP64 batch within relative L2 1e-10 of the CPU oracle, per problem, and
    the per-problem direct sum agrees with the oracle's."""
    probs, _ = datasets.batch(4, seed=8, nmin=500, nmax=4000)
    ctx.set_batch(probs)
    ctx.evaluate(tfmm.make_params(64, 0.4, tfmm.FP64))
    fb = ctx.fields()
    off = ctx.problem_offsets()
    t = np.array([0, 17, off[1] + 3, off[2], off[3] + 5, off[4] - 1], np.int64)
    d = ctx.direct(t)
    for i, (x, y, z, q) in enumerate(probs):
        o = oracle.Run(x, y, z, q, 64, 0.4)
        ep, ea = oracle.rel_l2(fb[off[i]:off[i + 1]], o.fields())
        assert ep < 1e-10 and ea < 1e-10, (i, ep, ea)
        sel = (t >= off[i]) & (t < off[i + 1])
        dd = oracle.direct(x, y, z, q, t[sel] - off[i])
        assert np.allclose(d[sel], dd, rtol=1e-11, atol=1e-11 * np.abs(dd).max())


def test_gpu_batch_errors(ctx):
    with pytest.raises(tfmm.TfmmError):
        ctx.set_batch([])
    x, y, z, q = datasets.cube(100, seed=1)
    ctx.set_batch([(x, y, z, q), (x, y, z, q)])
    pk = tfmm.Picker(0)
    try:
        with pytest.raises(tfmm.TfmmError):
            ctx.pick_params(pk, tfmm.MIXED)  # the picker describes one problem
    finally:
        pk.close()


def test_gpu_batch_degenerate_problems(ctx):
    """A forest holding a problem whose bodies all coincide (a zero-width
    root split down to depth 19) and a single-body problem beside ordinary
    ones: those fields are zero, the others match their single-problem runs."""
    r = np.random.default_rng(9)
    same = (np.full(100, 0.25), np.full(100, -0.5), np.full(100, 2.0), r.uniform(-1, 1, 100))
    a = datasets.cube(3000, seed=21)
    b = datasets.plummer(5000, seed=22, charges="uniform")
    lone = (np.array([1.0]), np.array([2.0]), np.array([3.0]), np.array([0.5]))
    probs = [a, same, b, lone]
    prm = tfmm.make_params(64, 0.4, tfmm.MIXED)
    ctx.set_batch(probs)
    ctx.evaluate(prm)
    fb = ctx.fields()
    off = ctx.problem_offsets()
    assert np.all(np.isfinite(fb))
    assert not np.any(fb[off[1]:off[2]]) and not np.any(fb[off[3]:off[4]])
    for i in (0, 2):
        ctx.set_bodies(*probs[i])
        ctx.evaluate(prm)
        ep, ea = oracle.rel_l2(fb[off[i]:off[i + 1]], ctx.fields())
        assert ep < 1e-6 and ea < 1e-6, (i, ep, ea)


def test_gpu_batch_device_matches_host_copy(ctx):
    """tfmm_set_batch_device (device SoA arrays + offsets) gives the fields of
    tfmm_set_batch (host AoS copies) bit for bit; bad offsets are rejected."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")  # the runtime the library already uses
    probs, _ = datasets.batch(6, seed=31, nmin=300, nmax=9000)
    prm = tfmm.make_params(32, 0.4, tfmm.FP64)
    ctx.set_batch(probs)
    ctx.evaluate(prm)
    ref = ctx.fields()
    off = np.concatenate([[0], np.cumsum([len(p[0]) for p in probs])]).astype(np.int64)
    ptrs = []
    try:
        for k in range(4):
            host = np.ascontiguousarray(np.concatenate([p[k] for p in probs]))
            d = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(host.nbytes)) == 0
            ptrs.append(d)
            assert hip.hipMemcpy(d, host.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_size_t(host.nbytes), 1) == 0  # host to device
        ctx.set_batch_device(off, *[d.value for d in ptrs])
        ctx.evaluate(prm)
        assert np.array_equal(ctx.fields(), ref)
        bad = off.copy()
        bad[2] = bad[1]  # an empty problem
        with pytest.raises(tfmm.TfmmError):
            ctx.set_batch_device(bad, *[d.value for d in ptrs])
        bad = off + 1  # offsets[0] != 0
        with pytest.raises(tfmm.TfmmError):
            ctx.set_batch_device(bad, *[d.value for d in ptrs])
        ctx.set_batch(probs)  # the context no longer borrows the arrays
    finally:
        ctx.sync()
        for d in ptrs:
            hip.hipFree(d)


def test_gpu_body_descriptor_offsets_checked_without_overflow(ctx):
    """Offsets near SIZE_MAX do not wrap around the stride check: the call is
    refused before any body record is read."""
    import ctypes
    body = np.zeros((4, 12), np.float64)
    for off_x, off_q in ((2 ** 64 - 8, 24), (0, 2 ** 64 - 4), (80, 24)):
        d = tfmm.BodyDesc(4, 96, off_x, off_q)
        rc = tfmm.lib().tfmm_set_bodies(ctx._h, body.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.byref(d))
        assert rc == -22, (off_x, off_q, rc)


def test_gpu_device_bodies_must_be_finite(ctx):
    """Device-resident bodies with a NaN coordinate or an infinite charge are
    refused as the host path refuses them (TFMM_EINVAL)."""
    import torch
    x, y, z, q = datasets.cube(5000, seed=3)
    for arr, val in ((1, float("nan")), (3, float("inf"))):
        a = [np.array(v) for v in (x, y, z, q)]
        a[arr][1234] = val
        t = [torch.from_numpy(v).to("cuda:0") for v in a]
        torch.cuda.synchronize()
        with pytest.raises(tfmm.TfmmError, match="not finite"):
            ctx.set_bodies_device(len(x), *[v.data_ptr() for v in t])
        with pytest.raises(tfmm.TfmmError, match="not finite"):
            ctx.set_batch_device(np.array([0, 2000, len(x)], np.int64),
                                 *[v.data_ptr() for v in t])
    t = [torch.from_numpy(np.array(v)).to("cuda:0") for v in (x, y, z, q)]
    torch.cuda.synchronize()
    ctx.set_bodies_device(len(x), *[v.data_ptr() for v in t])
    ctx.evaluate(tfmm.make_params(64, 0.4))
    assert np.isfinite(ctx.fields()).all()


def test_gpu_allocation_failure_is_reported_and_recoverable():
    """A device-memory limit below what an evaluation needs makes it fail
    with TFMM_ENOMEM (no abort, no partial state used later); with the limit
    lifted the same context evaluates and matches the oracle."""
    x, y, z, q = datasets.plummer(40000, seed=12, charges="uniform")
    c = tfmm.Context(0)
    try:
        c.set_bodies(x, y, z, q)
        # the limit counts every context's buffers: 1 MiB is below what the
        # process already holds, so the evaluation's first new buffer fails
        c.set_tuning(tfmm.TUNE_DEVICE_LIMIT, 1 << 20)
        assert c.get_tuning(tfmm.TUNE_DEVICE_LIMIT) == 1 << 20
        with pytest.raises(tfmm.TfmmError) as e:
            c.evaluate(tfmm.make_params(32, 0.4))
        print("with a 1 MiB limit:", e.value)
        assert "(-12)" in str(e.value) or "out of memory" in str(e.value).lower()
        c.set_tuning(tfmm.TUNE_DEVICE_LIMIT, 0)
        c.evaluate(tfmm.make_params(32, 0.4))
        o = oracle.Run(x, y, z, q, 32, 0.4)
        ep, ea = oracle.rel_l2(c.fields(), o.fields())
        assert ep < 1e-10 and ea < 1e-10
    finally:
        c.set_tuning(tfmm.TUNE_DEVICE_LIMIT, 0)
        c.close()

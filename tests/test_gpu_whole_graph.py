"""tfmm_evaluate as one graph launch (TFMM_TUNE_GRAPH = 2), through the C ABI.

The first evaluation of a context runs the ordinary way and leaves a plan
(capacities from its sizes); later evaluations with the same body count,
parameters and knobs are one hipGraph launch whose kernels read the sizes of
the evaluation on the device.  Every such evaluation gives the fields of the
ordinary evaluation of the same bodies bit for bit (same trees, lists, work
units and summation order), with the same counters; bodies that outgrow the
plan are evaluated again the ordinary way.  The FP64 forest matches the CPU
oracle within 1e-10 per problem.  Replays are separated by eager work and
synchronisations."""
import numpy as np
import pytest

import datasets
import oracle
import tfmm

pytestmark = pytest.mark.gpu

COUNTERS = ("ncells", "nleaves", "nlevels", "n_m2l", "n_m2l_exec", "n_p2p_leafpairs",
            "n_p2p_calls", "n_p2p_self", "n_m2l_groups", "n_m2l_items", "n_m2l_fallback")


@pytest.fixture(scope="module")
def ctx():
    c = tfmm.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ref():
    c = tfmm.Context(0)
    yield c
    c.close()


def ordinary(ref, bodies, prm, batch=False):
    """fields and counters of the ordinary evaluation (a context of its own)"""
    if batch:
        ref.set_batch(bodies)
    else:
        ref.set_bodies(*bodies)
    ref.evaluate(prm)
    return ref.fields(), ref.stats()


def check(ctx, f_ref, s_ref, label):
    f = ctx.fields()
    s = ctx.stats()
    assert np.array_equal(f, f_ref), label
    for k in COUNTERS:
        assert s[k] == s_ref[k], (label, k, s[k], s_ref[k])
    assert s["ms_total"] > 0
    return s


@pytest.mark.parametrize("precision,mutual", [(tfmm.FP64, 0), (tfmm.MIXED, 0), (tfmm.MIXED, 1)])
def test_gpu_whole_graph_matches_ordinary(ctx, ref, precision, mutual):
    prm = tfmm.make_params(32, 0.4, precision, mutual=mutual)
    cases = [("plummer", datasets.plummer(60000, seed=3, charges="uniform")),
             ("cube", datasets.cube(40000, seed=4))]
    ctx.set_tuning(tfmm.TUNE_GRAPH, 2)
    try:
        for name, b in cases:
            print("case %s n=%d precision %d mutual %d" % (name, len(b[0]), precision, mutual),
                  flush=True)
            f_ref, s_ref = ordinary(ref, b, prm)
            ctx.set_bodies(*b)
            ctx.evaluate(prm)  # ordinary, leaves the plan
            s = check(ctx, f_ref, s_ref, (name, "first"))
            assert s["one_graph"] == 0
            for rep in range(3):
                ctx.evaluate(prm)
                s = check(ctx, f_ref, s_ref, (name, rep))
                assert s["one_graph"] == 1, (name, rep)
                t = np.array([0, len(b[0]) // 2, len(b[0]) - 1])
                assert np.all(np.isfinite(ctx.direct(t)))  # eager work between the launches
    finally:
        ctx.set_tuning(tfmm.TUNE_GRAPH, 0)


def test_gpu_whole_graph_new_bodies(ctx, ref):
    """The plan is kept for new bodies of the same count: bodies of the same
    kind fit it (one graph launch), bodies of another shape may not (then the
    evaluation runs again the ordinary way); the fields are the ordinary
    evaluation's either way."""
    prm = tfmm.make_params(64, 0.4, tfmm.MIXED)
    n = 50000
    first = datasets.plummer(n, seed=1, charges="uniform")
    same_kind = datasets.plummer(n, seed=2, charges="uniform")
    other = datasets.make("two_plummer", n, seed=3)
    ctx.set_tuning(tfmm.TUNE_GRAPH, 2)
    try:
        ctx.set_bodies(*first)
        ctx.evaluate(prm)
        for label, b, must_fit in (("same kind", same_kind, True), ("other", other, False),
                                   ("same kind again", same_kind, False)):
            print("bodies: %s" % label, flush=True)
            f_ref, s_ref = ordinary(ref, b, prm)
            ctx.set_bodies(*b)
            ctx.evaluate(prm)
            s = check(ctx, f_ref, s_ref, label)
            if must_fit:
                assert s["one_graph"] == 1, label
            ctx.evaluate(prm)  # a plan from these bodies: one launch now
            s = check(ctx, f_ref, s_ref, label)
            assert s["one_graph"] == 1, label
    finally:
        ctx.set_tuning(tfmm.TUNE_GRAPH, 0)


def test_gpu_whole_graph_forest_against_oracle(ctx, ref):
    """A forest of small problems (one body, a root leaf and ordinary ones) as
    one graph launch: bit for bit the ordinary evaluation, and in FP64 within
    relative L2 1e-10 of the CPU oracle per problem."""
    probs, _ = datasets.batch(5, seed=8, nmin=500, nmax=4000)
    probs.append((np.array([0.3]), np.array([0.2]), np.array([0.9]), np.array([1.0])))
    x, y, z, q = datasets.cube(40, seed=2)
    probs.append((x, y, z, q))
    prm = tfmm.make_params(64, 0.4, tfmm.FP64)
    f_ref, s_ref = ordinary(ref, probs, prm, batch=True)
    ctx.set_tuning(tfmm.TUNE_GRAPH, 2)
    try:
        ctx.set_batch(probs)
        ctx.evaluate(prm)
        for rep in range(2):
            print("forest of %d problems, launch %d" % (len(probs), rep), flush=True)
            ctx.evaluate(prm)
            s = check(ctx, f_ref, s_ref, ("forest", rep))
            assert s["one_graph"] == 1
        fb = ctx.fields()
        off = ctx.problem_offsets()
        for i, (x, y, z, q) in enumerate(probs):
            seg = fb[off[i]:off[i + 1]]
            if len(x) == 1:
                assert not np.any(seg)
                continue
            o = oracle.Run(x, y, z, q, 64, 0.4)
            ep, ea = oracle.rel_l2(seg, o.fields())
            assert ep < 1e-10 and ea < 1e-10, (i, ep, ea)
    finally:
        ctx.set_tuning(tfmm.TUNE_GRAPH, 0)


def test_gpu_whole_graph_plan_follows_parameters(ctx, ref):
    """A change of theta, leaf size or knob makes a new plan (one ordinary
    evaluation), after which the launches are one graph again."""
    b = datasets.plummer(30000, seed=5, charges="uniform")
    ctx.set_tuning(tfmm.TUNE_GRAPH, 2)
    try:
        ctx.set_bodies(*b)
        for ncrit, theta, unit in ((32, 0.4, 0.0), (64, 0.4, 0.0), (64, 0.45, 0.0),
                                   (64, 0.45, 1.5)):
            print("ncrit %d theta %.2f unit %.1f" % (ncrit, theta, unit), flush=True)
            prm = tfmm.make_params(ncrit, theta, tfmm.MIXED)
            ctx.set_tuning(tfmm.TUNE_M2L_UNIT, unit)
            ref.set_tuning(tfmm.TUNE_M2L_UNIT, unit)
            f_ref, s_ref = ordinary(ref, b, prm)
            ctx.evaluate(prm)
            assert ctx.stats()["one_graph"] == 0
            ctx.evaluate(prm)
            s = check(ctx, f_ref, s_ref, (ncrit, theta, unit))
            assert s["one_graph"] == 1
    finally:
        ctx.set_tuning(tfmm.TUNE_GRAPH, 0)
        ctx.set_tuning(tfmm.TUNE_M2L_UNIT, 0)
        ref.set_tuning(tfmm.TUNE_M2L_UNIT, 0)


def test_gpu_pairs_outside_the_groups(ctx, ref):
    """An adaptive problem whose M2L pairs partly fall outside the offset
    groups (level difference above 3, or an offset beyond the table) runs the
    split into grouped keys and the per-pair list: FP64 within relative L2
    1e-10 of the CPU oracle, and as one graph launch bit for bit the ordinary
    evaluation, with the same counters."""
    prm = tfmm.make_params(16, 0.4, tfmm.FP64)
    found = None
    for name, n, seed in (("two_plummer", 20000, 3), ("plummer", 30000, 7),
                          ("two_plummer", 60000, 5)):
        b = datasets.make(name, n, seed=seed)
        f_ref, s_ref = ordinary(ref, b, prm)
        print("%s n=%d: %d of %d M2L pairs outside the groups" % (
            name, n, s_ref["n_m2l_fallback"], s_ref["n_m2l_exec"]), flush=True)
        if s_ref["n_m2l_fallback"] > 0:
            found = (b, f_ref, s_ref)
            break
    assert found is not None, "no case with pairs outside the groups"
    b, f_ref, s_ref = found
    o = oracle.Run(*b, 16, 0.4)
    ep, ea = oracle.rel_l2(f_ref, o.fields())
    assert ep < 1e-10 and ea < 1e-10, (ep, ea)
    ctx.set_tuning(tfmm.TUNE_GRAPH, 2)
    try:
        ctx.set_bodies(*b)
        ctx.evaluate(prm)
        for rep in range(2):
            print("pairs outside the groups, graph launch %d" % rep, flush=True)
            ctx.evaluate(prm)
            s = check(ctx, f_ref, s_ref, ("outside", rep))
            assert s["one_graph"] == 1
    finally:
        ctx.set_tuning(tfmm.TUNE_GRAPH, 0)

"""Tuning knobs of the batched M2L kernel and the autotuner (through the C
ABI): every setting gives the same fields up to summation order (relative
L2 1e-12 in FP64, 1e-6 in mixed precision), the autotuner returns one of its
candidates and leaves it set, and invalid settings are rejected."""
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


@pytest.mark.parametrize("precision,tol", [(tfmm.FP64, 1e-12), (tfmm.MIXED, 1e-6)])
def test_gpu_knobs_do_not_change_results(ctx, precision, tol):
    x, y, z, q = datasets.plummer(120000, seed=9, charges="uniform")
    print("plummer n=%d precision %d" % (len(x), precision), flush=True)
    ctx.set_bodies(x, y, z, q)
    prm = tfmm.make_params(32, 0.4, precision)
    ctx.set_tuning(tfmm.TUNE_M2L_UNIT, 0)
    ctx.set_tuning(tfmm.TUNE_M2L_XCD_RUN, 32)
    ctx.evaluate(prm)
    ref = ctx.fields()
    for unit, run, sched in ((0.5, 8, 1), (1.0, 1, 0), (1.5, 64, 1), (3.0, 16, 0)):
        ctx.set_tuning(tfmm.TUNE_M2L_UNIT, unit)
        ctx.set_tuning(tfmm.TUNE_M2L_XCD_RUN, run)
        ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, sched)
        assert ctx.get_tuning(tfmm.TUNE_M2L_UNIT) == unit
        assert ctx.get_tuning(tfmm.TUNE_M2L_XCD_RUN) == run
        ctx.evaluate(prm)
        f = ctx.fields()
        ep, ea = oracle.rel_l2(f, ref)
        assert ep < tol and ea < tol, (unit, run, ep, ea)
        ctx.evaluate(prm)  # the same setting twice: bit for bit
        assert np.array_equal(ctx.fields(), f)
    ctx.set_tuning(tfmm.TUNE_M2L_UNIT, 0)
    ctx.set_tuning(tfmm.TUNE_M2L_XCD_RUN, 32)
    ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, 0)


def test_gpu_autotune(ctx):
    x, y, z, q = datasets.cube(60000, seed=2)
    ctx.set_bodies(x, y, z, q)
    prm = tfmm.make_params(64, 0.4, tfmm.MIXED)
    f0 = None
    try:
        ctx.evaluate(prm)
        f0 = ctx.fields()
        unit, run, sched, ms = ctx.autotune(prm, reps=1)
        assert unit in (0.0, 1.5, 2.0, 3.0) and run in (8, 16, 32, 64)
        assert ms > 0
        assert ctx.get_tuning(tfmm.TUNE_M2L_UNIT) == unit
        assert ctx.get_tuning(tfmm.TUNE_M2L_XCD_RUN) == run
        assert ctx.get_tuning(tfmm.TUNE_P2P_SCHEDULE) == sched
        assert sched in (0, 1, 2) and ctx.get_tuning(tfmm.TUNE_P2P_WAVES) in (0, 3, 4)
        ctx.evaluate(prm)  # the chosen setting: the same fields up to summation order
        ep, ea = oracle.rel_l2(ctx.fields(), f0)
        assert ep < 1e-6 and ea < 1e-6, (ep, ea)
    finally:
        ctx.set_tuning(tfmm.TUNE_M2L_UNIT, 0)
        ctx.set_tuning(tfmm.TUNE_M2L_XCD_RUN, 32)
        ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, 0)
        ctx.set_tuning(tfmm.TUNE_P2P_WAVES, 0)


def test_gpu_invalid_knobs(ctx):
    for knob, value in ((tfmm.TUNE_M2L_UNIT, -1.0), (tfmm.TUNE_M2L_XCD_RUN, 0), (tfmm.TUNE_GRAPH, 3), (tfmm.TUNE_P2P_WAVES, 9), (tfmm.TUNE_P2P_SCHEDULE, 5),
                        (tfmm.TUNE_M2L_XCD_RUN, 2.5),
                        (99, 1.0)):
        with pytest.raises(tfmm.TfmmError):
            ctx.set_tuning(knob, value)


@pytest.mark.parametrize("precision", [tfmm.FP64, tfmm.MIXED])
def test_gpu_graph_replays_match_eager(ctx, precision):
    """tfmm_evaluate's numeric phase as a hipGraph: every replay (the graph
    updated in place, or instantiated again when the sizes change) gives bit
    for bit the fields of the same kernels launched one by one, with eager
    work and synchronisations between the replays."""
    prm = tfmm.make_params(32, 0.4, precision)
    cases = [datasets.plummer(50000, seed=3, charges="uniform"), datasets.cube(30000, seed=4),
             datasets.plummer(50000, seed=3, charges="uniform")]
    for i, (x, y, z, q) in enumerate(cases):
        print("case %d n=%d precision %d" % (i, len(x), precision), flush=True)
        ctx.set_bodies(x, y, z, q)
        ctx.set_tuning(tfmm.TUNE_GRAPH, 0)
        ctx.evaluate(prm)
        eager = ctx.fields()
        st_e = ctx.stats()
        ctx.set_tuning(tfmm.TUNE_GRAPH, 1)
        for rep in range(3):
            ctx.evaluate(prm)
            f = ctx.fields()
            assert np.array_equal(f, eager), (i, rep)
            st = ctx.stats()
            assert st["n_m2l"] == st_e["n_m2l"] and st["n_p2p_calls"] == st_e["n_p2p_calls"]
            assert st["ms_total"] > 0
            t = np.array([0, len(x) // 2, len(x) - 1])
            d = ctx.direct(t)  # eager work between the replays
            assert np.all(np.isfinite(d))
    ctx.set_tuning(tfmm.TUNE_GRAPH, 0)


def test_gpu_evaluate_matches_phase_calls(ctx):
    """tfmm_evaluate's order (traversal first, numeric phase after) gives the
    fields of the four separate phase calls bit for bit."""
    x, y, z, q = datasets.plummer(40000, seed=6, charges="uniform")
    ctx.set_bodies(x, y, z, q)
    prm = tfmm.make_params(64, 0.4, tfmm.MIXED)
    ctx.evaluate(prm)
    f1 = ctx.fields()
    ctx.build_tree(prm)
    ctx.upward()
    ctx.traverse()
    ctx.downward()
    assert np.array_equal(ctx.fields(), f1)


@pytest.mark.parametrize("schedule", [0, 1, 2, 3, 4])
def test_gpu_resident_p2p_and_schedules(ctx, schedule):
    """The FP32 P2P as resident waves taking target leaves from a counter, and
    every P2P schedule (3: the resident waves drained by a second launch on
    the same counter after the tiles), give bit for bit the fields of the
    default launch, also inside both graph modes."""
    x, y, z, q = datasets.plummer(60000, seed=12, charges="uniform")
    ctx.set_bodies(x, y, z, q)
    prm = tfmm.make_params(32, 0.4, tfmm.MIXED)
    ctx.set_tuning(tfmm.TUNE_P2P_WAVES, 0)
    ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, 0)
    ctx.evaluate(prm)
    ref = ctx.fields()
    ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, schedule)
    try:
        for waves in (0, 1, 3):
            print("schedule %d waves %d" % (schedule, waves), flush=True)
            ctx.set_tuning(tfmm.TUNE_P2P_WAVES, waves)
            for _ in range(2):
                ctx.evaluate(prm)
                assert np.array_equal(ctx.fields(), ref), (schedule, waves)
        if schedule >= 2:
            ctx.set_tuning(tfmm.TUNE_P2P_WAVES, 2)
            for graph in (1, 2):
                ctx.set_tuning(tfmm.TUNE_GRAPH, graph)
                for rep in range(3):
                    print("schedule %d waves 2 graph %d launch %d" % (schedule, graph, rep),
                          flush=True)
                    ctx.evaluate(prm)
                    assert np.array_equal(ctx.fields(), ref), (graph, rep)
                    assert np.all(np.isfinite(ctx.direct(np.array([0, 30000]))))
                ctx.set_tuning(tfmm.TUNE_GRAPH, 0)
    finally:
        ctx.set_tuning(tfmm.TUNE_GRAPH, 0)
        ctx.set_tuning(tfmm.TUNE_P2P_WAVES, 0)
        ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, 0)


def test_gpu_resident_p2p_fp64(ctx):
    """The FP64 P2P as resident waves (TFMM_TUNE_P2P_WAVES > 0) gives bit for
    bit the fields of its one-workgroup-per-cell launch, started early or
    late, and inside the one-launch graph."""
    x, y, z, q = datasets.plummer(40000, seed=13, charges="uniform")
    ctx.set_bodies(x, y, z, q)
    prm = tfmm.make_params(32, 0.4, tfmm.FP64)
    ctx.set_tuning(tfmm.TUNE_P2P_WAVES, 0)
    ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, 0)
    ctx.evaluate(prm)
    ref = ctx.fields()
    try:
        for schedule in (0, 2, 3):
            for waves in (1, 3):
                print("fp64 schedule %d waves %d" % (schedule, waves), flush=True)
                ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, schedule)
                ctx.set_tuning(tfmm.TUNE_P2P_WAVES, waves)
                ctx.evaluate(prm)
                assert np.array_equal(ctx.fields(), ref), (schedule, waves)
        ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, 2)
        ctx.set_tuning(tfmm.TUNE_P2P_WAVES, 2)
        ctx.set_tuning(tfmm.TUNE_GRAPH, 2)
        for rep in range(3):
            print("fp64 waves 2 one graph, launch %d" % rep, flush=True)
            ctx.evaluate(prm)
            assert np.array_equal(ctx.fields(), ref), rep
    finally:
        ctx.set_tuning(tfmm.TUNE_GRAPH, 0)
        ctx.set_tuning(tfmm.TUNE_P2P_WAVES, 0)
        ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, 0)

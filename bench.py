"""Benchmark of the MI355X TapasFMM path.

One step = one full FMM evaluation (tree build + upward + traversal with M2L
and P2P + downward) of the workload's bodies, which are resident in HBM
before the timed region starts.  Prints one JSON line (rank 0).

Workload of the line (default): 4,194,304 bodies from a Plummer sphere
(a = 1, radii truncated at 20a, isotropic), equal charges 1/N, expansion
order P = 10, mixed precision, ncrit/theta from the learned picker.  The
error of the evaluated fields is measured against the FP64 direct sum on
4096 random targets (outside the timed region).  The plain FP64 form of the
same method (ncrit 64, theta 0.4) is timed beside it.  The CPU baseline is
the reference's own pipeline (oracle/_ref, 1 MPI rank) timed on a bounded
sample of the same distribution.

The default run also measures, each in a block of its own under "configs":
  c1  1,048,576-body uniform cube, FP64, ncrit 64, theta 0.4;
  c3  256 independent problems (N log-uniform in [2^10, 2^16], cube /
      Plummer / 1:0.3:0.1 ellipsoid, charges uniform in [-1, 1]) evaluated
      as one forest, mixed precision, theta 0.4;
  c4  8,388,608-body two-cluster Plummer (separation 10a, mass 1:3), mixed
      precision, learned parameters, beside the fallback ncrit 64 / theta 0.4.
  c1_cartesian  c1 in the Cartesian Taylor basis (ExaFMM's
      LaplaceCartesianCPU.cxx; FP64 expansions, the M2L through the FP64
      tile kernel between exact basis changes).
      --cart-m2l 1: the direct Cartesian M2L kernel instead.
--only-headline skips them; --batch K / --dist / --n measure one workload
(--basis cartesian: in the Cartesian basis).

roofline: the P2P kernel (20 flops per evaluated body pair against the FP32
vector peak; the FP64 kernel against the FP64 vector peak), timed by HIP
events on the stream it runs on; the M2L tile kernel is reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tapas-reduce_amd"))

PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
P2P_FLOPS_PER_PAIR = 20
# complex multiply-adds of one M2L pair at P = 10 (sum_j (j+1)(P-j)^2 = 1210) x 8
M2L_FLOPS_PER_PAIR = 9680
# one Cartesian M2L pair at P = 10: the contraction's 5005 multiply-adds (x 2),
# the derivative recurrence's 219 terms (10 flops each) and the k! scaling (220)
CART_M2L_FLOPS_PER_PAIR = 2 * 5005 + 10 * 219 + 220
HEADLINE_METRIC = "FMM bodies/s at rel-L2 err <=1e-4, 4M Plummer; P2P % FP32 peak"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=None, help="bodies (default 4,194,304)")
    ap.add_argument("--dist", default=None, help="distribution (default plummer)")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--picker", default=os.path.join(ROOT, "tapas-reduce_amd", "picker",
                                                     "weights.npz"),
                    help="trained picker weights; the fallback ncrit/theta is used if absent")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=3,
                    help="steps of the plain FP64 form timed beside the headline (0: skip)")
    ap.add_argument("--ncrit", type=int, default=None,
                    help="leaf size (default 64; for --batch, the fastest of 64/128/256 at "
                         "theta 0.4, timed before the steps)")
    ap.add_argument("--theta", type=float, default=0.4)
    ap.add_argument("--targets", type=int, default=4096)
    ap.add_argument("--cpu-sample", type=int, default=524288)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-headline", action="store_true",
                    help="skip the c1 / c3 / c4 blocks of the default run")
    ap.add_argument("--config-steps", type=int, default=5,
                    help="timed steps of each c1 / c3 / c4 block")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile-kernel knobs instead of autotuning them")
    ap.add_argument("--unit-factor", type=float, default=None,
                    help="M2L work-unit size in mean items per tile (0: one unit per tile)")
    ap.add_argument("--xcd-run", type=int, default=None,
                    help="M2L work units dealt to one XCD in a row")
    ap.add_argument("--p2p-schedule", type=int, default=None, choices=[0, 1, 2, 3, 4],
                    help="0: P2P from the numeric phase's start, 1: beside the M2L tiles, "
                         "2: from phase A, beside the M2L batch building, 3: as 2, drained "
                         "by a full-occupancy launch after the tiles, 4: serially, alone")
    ap.add_argument("--p2p-waves", type=int, default=None, choices=[0, 1, 2, 3, 4, 5, 6, 7, 8],
                    help="fix the FP32 P2P's resident waves per SIMD (0: one workgroup per cell)")
    ap.add_argument("--graph", action="store_true",
                    help="launch every evaluation as one hipGraph (after a first evaluation "
                         "that sets its capacities)")
    ap.add_argument("--basis", default="spherical", choices=["spherical", "cartesian"],
                    help="expansion basis of a single-workload run")
    ap.add_argument("--cart-m2l", type=int, default=0, choices=[0, 1],
                    help="route of the Cartesian M2L: 0 the FP64 tile kernel between basis "
                         "changes, 1 the direct Cartesian kernel")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 256 MiB device buffer three times (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    ap.add_argument("--batch", type=int, default=0,
                    help="evaluate this many independent small problems as one forest")
    args = ap.parse_args()
    args.single = args.batch > 0 or args.n is not None or args.dist is not None
    args.n = 4194304 if args.n is None else args.n
    args.dist = "plummer" if args.dist is None else args.dist
    return args


def cpu_baseline(dist, n_sample, ncrit, theta):
    """Reference pipeline (oracle/_ref/tapas_ref) on a bounded sample, at the
    workload's ncrit and theta."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import datasets
    exe = oracle.ref_binary(count=False)
    if exe is None:
        return None
    x, y, z, q = datasets.make(dist, n_sample, seed=1)
    t0 = time.time()
    r = oracle.run_ref(x, y, z, q, ncrit, theta, count=False, timeout=600)
    wall = time.time() - t0
    ph = r["times"]
    t = ph["tree"] + ph["upward"] + ph["traverse"] + ph["downward"]
    return {"value": n_sample / t, "unit": "bodies/s", "cores": 1, "kind": "reference",
            "sample": "%d-body %s (same generator, ncrit %d, theta %.4g), reference TapasFMM "
                      "pipeline, 1 MPI rank; phases tree %.2fs upward %.2fs traverse %.2fs "
                      "downward %.2fs (upward includes the reference's double upward pass "
                      "on one rank); process wall %.1fs" % (n_sample, dist, ncrit, theta,
                                                             ph["tree"], ph["upward"],
                                                             ph["traverse"], ph["downward"],
                                                             wall)}


def cpu_baseline_batch(probs, ncrit, theta, budget_bodies):
    """Reference pipeline on the batch's first problems (one process each,
    1 MPI rank) until budget_bodies bodies are covered."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if oracle.ref_binary(count=False) is None:
        return None
    done = nb = 0
    t = 0.0
    for x, y, z, q in probs:
        r = oracle.run_ref(x, y, z, q, ncrit, theta, count=False, timeout=600)
        ph = r["times"]
        t += ph["tree"] + ph["upward"] + ph["traverse"] + ph["downward"]
        nb += len(x)
        done += 1
        if nb >= budget_bodies:
            break
    return {"value": nb / t, "unit": "bodies/s", "cores": 1, "kind": "reference",
            "sample": "the first %d problems of the batch (%d bodies), each a separate reference "
                      "TapasFMM run (ncrit %d, theta %.4g, 1 MPI rank); summed phase times %.2fs"
                      % (done, nb, ncrit, theta, t)}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from a committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or
    None when no summary covers this workload."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.txt"))):
        fields = {}
        with open(path) as f:
            for line in f:
                if ":" in line and not line.startswith(" "):
                    k, v = line.split(":", 1)
                    fields[k.strip()] = v.strip()
        if fields.get("workload") == workload and fields.get("kernel") == kernel:
            return {"bytes": float(fields["traffic_bytes_per_launch"]),
                    "source": os.path.relpath(path, ROOT), "knobs": fields.get("knobs")}
    return None


def time_config(ctx, prm, steps, warmup):
    import numpy as np
    for _ in range(warmup):
        ctx.evaluate(prm)
    ctx.sync()
    phase = {k: [] for k in ("ms_tree", "ms_upward", "ms_traverse", "ms_downward", "ms_total",
                             "ms_p2p", "ms_m2l", "ms_dtt", "ms_m2l_tiles")}
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.evaluate(prm)
        st = ctx.stats()  # waits for the step; phase times from HIP events
        for k in phase:
            phase[k].append(st[k])
    ctx.sync()
    elapsed = time.perf_counter() - t0
    return elapsed * 1e3 / steps, {k: float(np.mean(v)) for k, v in phase.items()}, st


def one_graph(ctx, prm, steps, n):
    """The same evaluation as one hipGraph launch per step (TFMM_TUNE_GRAPH =
    2): one ordinary evaluation makes the plan, then `steps` launches are
    timed; their fields must equal the ordinary evaluation's bit for bit.
    With resident P2P waves the P2P's start is picked again for the graph,
    among the autotuned setting, schedule 3 with 2 waves and schedule 2 with
    3 (5 launches each, outside the timed steps): the graph's branches
    overlap differently from two streams (the P2P setting changes no
    field)."""
    import numpy as np
    import tfmm
    ref = ctx.fields()
    was = ctx.get_tuning(tfmm.TUNE_GRAPH)
    knobs = (int(ctx.get_tuning(tfmm.TUNE_P2P_SCHEDULE)), int(ctx.get_tuning(tfmm.TUNE_P2P_WAVES)))
    cands = [knobs]
    if knobs[1] > 0:
        cands += [k for k in ((3, 2), (2, 3)) if k != knobs]
    picked = {}
    ctx.set_tuning(tfmm.TUNE_GRAPH, 2)
    try:
        if len(cands) > 1:
            for sched, waves in cands:
                ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, sched)
                ctx.set_tuning(tfmm.TUNE_P2P_WAVES, waves)
                ctx.evaluate(prm)  # the plan for this setting
                picked["%d/%d" % (sched, waves)] = time_config(ctx, prm, 5, 2)[0]
            best = min(cands, key=lambda k: picked["%d/%d" % k])
            ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, best[0])
            ctx.set_tuning(tfmm.TUNE_P2P_WAVES, best[1])
        ctx.evaluate(prm)  # the plan
        ms, phases, st = time_config(ctx, prm, steps, 1)
        same = bool(np.array_equal(ctx.fields(), ref))
        used = (int(ctx.get_tuning(tfmm.TUNE_P2P_SCHEDULE)),
                int(ctx.get_tuning(tfmm.TUNE_P2P_WAVES)))
    finally:
        ctx.set_tuning(tfmm.TUNE_GRAPH, was)
        ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, knobs[0])
        ctx.set_tuning(tfmm.TUNE_P2P_WAVES, knobs[1])
    return {"value": n / (ms * 1e-3), "ms_per_step": ms, "steps": steps,
            "launches_one_graph": bool(st["one_graph"]), "fields_equal_ordinary": same,
            "p2p_schedule": used[0], "p2p_waves": used[1],
            "p2p_sweep_ms": picked or None,
            "note": "every step one hipGraph launch of the whole evaluation (tree, traversal, "
                    "lists, M2L batches, numeric phase), sizes read on the device; no host "
                    "read-back inside a step"}


def p2p_isolated(ctx, prm, st, roof):
    """The P2P kernel's own rate: three more evaluations with the P2P at full
    occupancy (no resident-wave cap) on the context stream with nothing
    beside it (TFMM_TUNE_P2P_SCHEDULE 4), outside the timed steps.  The
    headline's figure is over the P2P's span while the upward pass, the M2L
    batching and the tiles share the GPU with it."""
    import tfmm
    was = ctx.get_tuning(tfmm.TUNE_P2P_SCHEDULE), ctx.get_tuning(tfmm.TUNE_P2P_WAVES)
    ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, 4)
    ctx.set_tuning(tfmm.TUNE_P2P_WAVES, 0)
    try:
        ms = []
        for _ in range(3):
            ctx.evaluate(prm)
            ms.append(ctx.stats()["ms_p2p"])
    finally:
        ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, was[0])
        ctx.set_tuning(tfmm.TUNE_P2P_WAVES, was[1])
    t = sorted(ms)[1]
    pairs = st["n_p2p_calls"] - st["n_p2p_self"]
    tf = P2P_FLOPS_PER_PAIR * pairs / (t * 1e-3) / 1e12 if t > 0 else 0.0
    kernel = roof["kernel"].replace("_resident", "")
    return {"kernel": kernel, "ms_per_launch": t, "achieved": tf, "peak": roof["peak"],
            "unit": "TFLOP/s", "frac": tf / roof["peak"],
            "note": "the P2P kernel at full occupancy with nothing beside it (P2P schedule 4, "
                    "no resident-wave cap), median of 3 evaluations outside the timed steps"}


def errors(ctx, targets, direct):
    import numpy as np
    f = ctx.fields()[targets]

    def rel(a, b):
        return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    return rel(f[:, 0], direct[:, 0]), rel(f[:, 1:], direct[:, 1:])


def kernels(st, phases, precision, workload, basis="spherical", cart_direct=False,
            resident=False):
    """Achieved rates of the P2P kernel (20 flops per evaluated body pair)
    and the M2L tile kernel (9,680 flops per grouped pair; the Cartesian
    basis's direct kernel k_cart_m2l: CART_M2L_FLOPS_PER_PAIR per pair, FP64
    vector) over their HIP-event times on the streams they run on."""
    import tfmm
    mixed = precision == tfmm.MIXED
    pairs = st["n_p2p_calls"] - st["n_p2p_self"]
    p2p_ms = phases["ms_p2p"]
    p2p_tf = P2P_FLOPS_PER_PAIR * pairs / (p2p_ms * 1e-3) / 1e12 if p2p_ms > 0 else 0.0
    p2p_peak = PEAK_FP32_TFLOPS if mixed else PEAK_FP64_TFLOPS
    p2p_name = ("k_p2p_fp32_resident" if resident else "k_p2p_fp32") if mixed else (
        "k_p2p_fp64_resident" if resident else "k_p2p_fp64")
    grouped = st["n_m2l_exec"] - st["n_m2l_fallback"]
    tiles_ms = phases["ms_m2l_tiles"]
    m2l_tf = M2L_FLOPS_PER_PAIR * grouped / (tiles_ms * 1e-3) / 1e12 if tiles_ms > 0 else 0.0
    tile_name = "k_m2l_tiles<%s>" % ("float" if mixed else "double")
    tile_peak = PEAK_FP32_TFLOPS if mixed else PEAK_FP64_TFLOPS
    p2p_traffic = measured_traffic(workload, p2p_name)
    tile_traffic = measured_traffic(workload, tile_name)
    if basis == "cartesian" and cart_direct:
        pairs_m2l = st["n_m2l_exec"]
        cart_tf = (CART_M2L_FLOPS_PER_PAIR * pairs_m2l / (tiles_ms * 1e-3) / 1e12
                   if tiles_ms > 0 else 0.0)
        tile_traffic = measured_traffic(workload, "k_cart_m2l")
        m2l_entry = {"kernel": "k_cart_m2l", "bound": "valu", "achieved": cart_tf,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": cart_tf / PEAK_FP64_TFLOPS,
                     "traffic": tile_traffic["bytes"] if tile_traffic else None,
                     "traffic_source": tile_traffic["source"] if tile_traffic else None,
                     "flops_per_pair": CART_M2L_FLOPS_PER_PAIR, "pairs_per_launch": pairs_m2l,
                     "ms_per_launch": tiles_ms}
    else:
        m2l_entry = {"kernel": tile_name, "bound": "mfma", "achieved": m2l_tf,
                     "peak": tile_peak, "unit": "TFLOP/s", "frac": m2l_tf / tile_peak,
                     "traffic": tile_traffic["bytes"] if tile_traffic else None,
                     "traffic_source": tile_traffic["source"] if tile_traffic else None,
                     "traffic_knobs": tile_traffic["knobs"] if tile_traffic else None,
                     "flops_per_pair": M2L_FLOPS_PER_PAIR, "pairs_per_launch": grouped,
                     "ms_per_launch": tiles_ms}
    timing = ("HIP-event span of the launch in the timed evaluation, on the P2P's stream, "
              "beside the kernels that run with it")
    if resident:
        timing += ("; as resident waves at the default issue priority beside kernels that "
                   "raise theirs, the P2P takes the slots they leave, so the span is longer "
                   "than the kernel's own time (roofline.isolated)")
    return {
        "kernel": p2p_name, "bound": "valu", "achieved": p2p_tf, "peak": p2p_peak,
        "unit": "TFLOP/s", "frac": p2p_tf / p2p_peak, "timing": timing,
        "traffic": p2p_traffic["bytes"] if p2p_traffic else None,
        "traffic_source": p2p_traffic["source"] if p2p_traffic else None,
        "flops_per_pair": P2P_FLOPS_PER_PAIR, "body_pairs": pairs, "ms_per_launch": p2p_ms,
        "pct_fp32_peak": 100.0 * p2p_tf / PEAK_FP32_TFLOPS,
        "m2l_tiles": m2l_entry,
    }


def counts(st):
    return {"cells": st["ncells"], "leaves": st["nleaves"], "max_depth": st["max_depth"],
            "m2l_pairs": st["n_m2l_exec"], "m2l_groups": st["n_m2l_groups"],
            "m2l_items": st["n_m2l_items"], "m2l_fallback": st["n_m2l_fallback"],
            "p2p_leaf_pairs": st["n_p2p_leafpairs"],
            "p2p_body_pairs": st["n_p2p_calls"] - st["n_p2p_self"]}


def to_device(arrays):
    import torch
    dev = torch.device("cuda:0")
    out = [torch.from_numpy(a).to(dev) for a in arrays]
    torch.cuda.synchronize()
    return out


def set_knobs(ctx, args):
    import tfmm
    fixed = {}
    if getattr(args, "basis", "spherical") == "cartesian":
        ctx.set_basis(tfmm.BASIS_CARTESIAN)
        ctx.set_tuning(tfmm.TUNE_CART_M2L, getattr(args, "cart_m2l", 0))
    if args.graph:
        ctx.set_tuning(tfmm.TUNE_GRAPH, 2)
    if args.unit_factor is not None:
        ctx.set_tuning(tfmm.TUNE_M2L_UNIT, args.unit_factor)
        fixed["m2l_unit_factor"] = args.unit_factor
    if args.xcd_run is not None:
        ctx.set_tuning(tfmm.TUNE_M2L_XCD_RUN, args.xcd_run)
        fixed["m2l_xcd_run"] = args.xcd_run
    if args.p2p_schedule is not None:
        ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, args.p2p_schedule)
        fixed["p2p_schedule"] = args.p2p_schedule
    if getattr(args, "p2p_waves", None) is not None:
        ctx.set_tuning(tfmm.TUNE_P2P_WAVES, args.p2p_waves)
        fixed["p2p_waves"] = args.p2p_waves
    return fixed


def default_knobs(ctx):
    import tfmm
    ctx.set_tuning(tfmm.TUNE_M2L_UNIT, 0)
    ctx.set_tuning(tfmm.TUNE_M2L_XCD_RUN, 32)
    ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, 0)
    ctx.set_tuning(tfmm.TUNE_P2P_WAVES, 0)


def autotune(ctx, prm, fixed, enabled):
    import tfmm
    if fixed:
        return {"fixed": fixed}
    if not enabled:
        return None
    t0 = time.perf_counter()
    unit, run, sched, tms = ctx.autotune(prm, reps=5)
    return {"m2l_unit_factor": unit, "m2l_xcd_run": run, "p2p_schedule": sched,
            "p2p_waves": int(ctx.get_tuning(tfmm.TUNE_P2P_WAVES)),
            "ms_per_evaluation": tms, "seconds": time.perf_counter() - t0,
            "note": "16 tile-kernel settings, then up to 8 P2P schedule/resident-wave "
                    "settings, x 5 evaluations each, outside the timed steps"}


def learned(ctx, path, precision):
    """ncrit/theta from the learned picker (trial tree + statistics + FP32
    MLP inference), or None without weights."""
    import tfmm
    if precision != tfmm.MIXED or not os.path.exists(path):
        return None
    pk = tfmm.Picker.load(path)
    t0 = time.perf_counter()
    p = ctx.pick_params(pk, precision)
    pick_ms = (time.perf_counter() - t0) * 1e3
    pk.close()
    return {"ncrit": int(p.ncrit), "theta": float(p.theta), "ms": pick_ms,
            "note": "trial tree at ncrit 64 + tree statistics + FP32 MLP inference kernel, "
                    "timed outside the steps"}


def bench_single(args, dist, n, precision, use_picker, steps, warmup, fp64_steps, seed=0,
                 fallback_steps=0):
    """One problem of n bodies: the measured form (learned parameters when
    the picker applies), optionally the fallback parameters and the plain
    FP64 form of the same bodies."""
    import numpy as np
    import datasets
    import tfmm
    x, y, z, q = datasets.make(dist, n, seed=seed)
    dx = to_device((x, y, z, q))
    ctx = tfmm.Context(0)
    try:
        fixed = set_knobs(ctx, args)
        ctx.set_bodies_device(n, *[a.data_ptr() for a in dx])
        cart = args.basis == "cartesian"
        # the picker was trained on the spherical basis's timings
        picked = learned(ctx, args.picker, precision) if use_picker and not cart else None
        ncrit = picked["ncrit"] if picked else (args.ncrit or 64)
        theta = picked["theta"] if picked else args.theta
        prm = tfmm.make_params(ncrit, theta, precision)
        tuned = autotune(ctx, prm, fixed, not args.no_autotune)
        rng = np.random.default_rng(123)
        targets = np.sort(rng.choice(n, size=min(args.targets, n), replace=False))
        direct = ctx.direct(targets)
        ms, phases, st = time_config(ctx, prm, steps, warmup)
        ep, ea = errors(ctx, targets, direct)
        prec_name = "fp64" if precision == tfmm.FP64 else "mixed"
        workload = "%s N=%d P=10 ncrit=%d theta=%g %s%s%s" % (
            dist, n, ncrit, round(theta, 4), prec_name, " (learned)" if picked else "",
            (" cartesian basis" + (" direct M2L" if args.cart_m2l else "")) if cart else "")
        out = {"workload": workload, "value": n / (ms * 1e-3), "unit": "bodies/s",
               "ms_per_step": ms, "steps": steps, "n_bodies": n, "distribution": dist,
               "ncrit": ncrit, "theta": theta, "precision": prec_name,
               "error": {"rel_l2_pot": ep, "rel_l2_acc": ea, "targets": len(targets),
                         "against": "FP64 direct sum"},
               "picker": picked, "autotune": tuned, "phases_ms": phases, "counts": counts(st),
               "roofline": kernels(st, phases, precision, workload, args.basis,
                                   cart and args.cart_m2l == 1,
                                   ctx.get_tuning(tfmm.TUNE_P2P_WAVES) > 0),
               "basis": args.basis}
        if not args.graph and not cart:
            out["one_graph"] = one_graph(ctx, prm, steps, n)
        out["roofline"]["isolated"] = p2p_isolated(ctx, prm, st, out["roofline"])
        if fallback_steps > 0 and picked:
            prm_fb = tfmm.make_params(64, 0.4, precision)
            msf, phf, stf = time_config(ctx, prm_fb, fallback_steps, 1)
            ef = errors(ctx, targets, direct)
            out["fallback"] = {"value": n / (msf * 1e-3), "ms_per_step": msf, "ncrit": 64,
                               "theta": 0.4, "rel_l2_pot": ef[0], "rel_l2_acc": ef[1],
                               "learned_speedup": msf / ms}
        if fp64_steps > 0 and precision == tfmm.MIXED:
            prm64 = tfmm.make_params(args.ncrit or 64, args.theta, tfmm.FP64)
            default_knobs(ctx)  # the plain form keeps the default knobs
            ms64, ph64, st64 = time_config(ctx, prm64, fp64_steps, 1)
            e64 = errors(ctx, targets, direct)
            out["fp64_plain"] = {"value": n / (ms64 * 1e-3), "ms_per_step": ms64,
                                 "ncrit": args.ncrit or 64, "theta": args.theta,
                                 "rel_l2_pot": e64[0], "rel_l2_acc": e64[1],
                                 "phases_ms": ph64,
                                 "roofline": kernels(st64, ph64, tfmm.FP64, None),
                                 "speedup_of_headline": ms64 / ms}
        return out
    finally:
        ctx.close()
        del dx


def bench_batch(args, nprob, precision, steps, warmup):
    """nprob independent problems evaluated as one forest; theta fixed, the
    leaf size the fastest of 64/128/256 on these bodies (timed first)."""
    import numpy as np
    import datasets
    import tfmm
    probs, kinds = datasets.batch(nprob, seed=0)
    offsets = np.concatenate([[0], np.cumsum([len(p[0]) for p in probs])]).astype(np.int64)
    arrs = [np.concatenate([p[i] for p in probs]) for i in range(4)]
    n = int(offsets[-1])
    dx = to_device(arrs)
    ctx = tfmm.Context(0)
    try:
        fixed = set_knobs(ctx, args)
        ctx.set_batch_device(offsets, *[a.data_ptr() for a in dx])
        theta = args.theta
        if args.ncrit is None:
            sweep = {}
            for nc in (64, 128, 256):
                p_nc = tfmm.make_params(nc, theta, precision)
                ctx.evaluate(p_nc)
                ctx.sync()
                t0 = time.perf_counter()
                for _ in range(2):
                    ctx.evaluate(p_nc)
                ctx.sync()
                sweep[nc] = (time.perf_counter() - t0) / 2 * 1e3
            ncrit = min(sweep, key=sweep.get)
            picked = {"ncrit": ncrit, "theta": theta, "sweep_ms": sweep,
                      "note": "leaf size timed at 64/128/256 (theta fixed), outside the steps"}
        else:
            ncrit, picked = args.ncrit, None
        prm = tfmm.make_params(ncrit, theta, precision)
        tuned = autotune(ctx, prm, fixed, not args.no_autotune)
        rng = np.random.default_rng(123)
        targets = np.sort(rng.choice(n, size=min(args.targets, n), replace=False))
        direct = ctx.direct(targets)  # per problem: a target sees only its own problem
        ms, phases, st = time_config(ctx, prm, steps, warmup)
        ep, ea = errors(ctx, targets, direct)
        graph_run = None if args.graph else one_graph(ctx, prm, steps, n)
        prec_name = "fp64" if precision == tfmm.FP64 else "mixed"
        workload = ("batch of %d problems (cube/plummer/ellipsoid, N log-uniform in [2^10, 2^16])"
                    " total N=%d P=10 ncrit=%d theta=%g %s" % (nprob, n, ncrit, round(theta, 4),
                                                               prec_name))
        return {"workload": workload, "value": n / (ms * 1e-3), "unit": "bodies/s",
                "ms_per_step": ms, "steps": steps, "n_bodies": n, "problems": nprob,
                "distribution": {k: kinds.count(k) for k in datasets.BATCH_KINDS},
                "ncrit": ncrit, "theta": theta, "precision": prec_name, "graph": bool(args.graph),
                "error": {"rel_l2_pot": ep, "rel_l2_acc": ea, "targets": len(targets),
                          "against": "FP64 direct sum within each problem"},
                "picker": picked, "autotune": tuned, "phases_ms": phases, "counts": counts(st),
                "roofline": kernels(st, phases, precision, workload, resident=ctx.get_tuning(
                    tfmm.TUNE_P2P_WAVES) > 0),
                "one_graph": graph_run}, probs
    finally:
        ctx.close()
        del dx


def calibrate_copy():
    import torch
    x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")  # 256 MiB
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_per_copy": x.numel() * 4, "copies": 3}))


def line(metric, head, args, cpu):
    return {
        "metric": metric,
        "value": head["value"],
        "unit": "bodies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if head["precision"] == "fp64"
                  else "f32 P2P/M2L + f64 tree/P2M/M2M/L2L/L2P"),
        "data": "synthetic",
        "config": {"workload": head["workload"], "n_bodies": head["n_bodies"],
                   "distribution": head["distribution"], "expansion": 10,
                   "problems": head.get("problems", 1), "ncrit": head["ncrit"],
                   "theta": head["theta"], "precision": head["precision"]},
        "error": head["error"],
        "picker": head["picker"],
        "autotune": head["autotune"],
        "graph": bool(args.graph),
        "phases_ms": head["phases_ms"],
        "counts": head["counts"],
        "roofline": head["roofline"],
        "fp64_plain": head.get("fp64_plain"),
        "one_graph": head.get("one_graph"),
        "cpu_baseline": cpu,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU; this build has no multi-GPU path (--gpus %d refused)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    import tfmm

    precision = tfmm.FP64 if args.precision == "fp64" else tfmm.MIXED
    use_picker = not args.no_learned and args.ncrit is None
    if args.batch > 0:
        head, probs = bench_batch(args, args.batch, precision, args.steps, args.warmup)
        cpu = None if args.no_cpu else cpu_baseline_batch(probs, head["ncrit"], head["theta"],
                                                          args.cpu_sample)
        out = line("FMM bodies/s over %d batched small problems in one forest" % args.batch,
                   head, args, cpu)
        print(json.dumps(out))
        return
    head = bench_single(args, args.dist, args.n, precision, use_picker, args.steps, args.warmup,
                        args.fp64_steps)
    cpu = None if args.no_cpu else cpu_baseline(args.dist, args.cpu_sample, head["ncrit"],
                                                head["theta"])
    metric = HEADLINE_METRIC if (args.dist, args.n) == ("plummer", 4194304) else \
        "FMM bodies/s, %s N=%d" % (args.dist, args.n)
    out = line(metric, head, args, cpu)
    if not args.single and not args.only_headline:
        ks = args.config_steps
        c1_args = argparse.Namespace(**vars(args))
        c1_args.ncrit, c1_args.theta = 64, 0.4
        c1_args.basis = "spherical"
        c1_cart = argparse.Namespace(**vars(c1_args))
        c1_cart.basis = "cartesian"
        configs = {
            "c1": bench_single(c1_args, "cube", 1 << 20, tfmm.FP64, False, ks, 1, 0, seed=0),
            "c1_cartesian": bench_single(c1_cart, "cube", 1 << 20, tfmm.FP64, False, ks, 1, 0,
                                         seed=0),
            "c3": bench_batch(args, 256, tfmm.MIXED, ks, 1)[0],
            "c4": bench_single(args, "two_plummer", 1 << 23, tfmm.MIXED, not args.no_learned,
                               ks, 1, 2, fallback_steps=3),
        }
        out["configs"] = configs
    print(json.dumps(out))


if __name__ == "__main__":
    main()

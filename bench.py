"""Benchmark of the MI355X TapasFMM path.

One step = one full FMM evaluation (tree build + upward + traversal with M2L
and P2P + downward) of the workload's bodies, which are resident in HBM
before the timed region starts.  Prints one JSON line (rank 0).

Workload (default): 4,194,304 bodies from a Plummer sphere (a = 1, radii
truncated at 20a, isotropic), equal charges 1/N, expansion order P = 10,
ncrit = 64, theta = 0.4.  The error of the evaluated fields is measured
against the FP64 direct sum on 4096 random targets (outside the timed
region).  The CPU baseline is the reference's own pipeline (oracle/_ref,
1 MPI rank) timed on a bounded sample of the same distribution.

--batch K: K independent problems instead (N log-uniform in [2^10, 2^16],
each a uniform cube, a Plummer sphere or a 1:0.3:0.1 ellipsoid, charges
uniform in [-1, 1]) evaluated together as one forest, ncrit 64, theta 0.4;
the value is the total bodies of the batch per second.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tapas-reduce_amd"))

PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
P2P_FLOPS_PER_PAIR = 20
# complex multiply-adds of one M2L pair at P = 10 (sum_j (j+1)(P-j)^2 = 1210) x 8
M2L_FLOPS_PER_PAIR = 9680


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--dist", default="plummer")
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
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile-kernel knobs instead of autotuning them")
    ap.add_argument("--unit-factor", type=float, default=None,
                    help="M2L work-unit size in mean items per tile (0: one unit per tile)")
    ap.add_argument("--xcd-run", type=int, default=None,
                    help="M2L work units dealt to one XCD in a row")
    ap.add_argument("--p2p-schedule", type=int, default=None, choices=[0, 1],
                    help="0: P2P beside the M2L batch building, 1: beside the M2L tiles")
    ap.add_argument("--graph", action="store_true",
                    help="launch the numeric phase of every evaluation as one hipGraph")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 256 MiB device buffer three times (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    ap.add_argument("--batch", type=int, default=0,
                    help="evaluate this many independent small problems as one forest")
    return ap.parse_args()


def cpu_baseline(args, ncrit, theta):
    """Reference pipeline (oracle/_ref/tapas_ref) on a bounded sample, at the
    headline's ncrit and theta."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import datasets
    exe = oracle.ref_binary(count=False)
    if exe is None:
        return None
    n = args.cpu_sample
    x, y, z, q = datasets.make(args.dist, n, seed=1)
    t0 = time.time()
    r = oracle.run_ref(x, y, z, q, ncrit, theta, count=False, timeout=600)
    wall = time.time() - t0
    ph = r["times"]
    t = ph["tree"] + ph["upward"] + ph["traverse"] + ph["downward"]
    return {"value": n / t, "unit": "bodies/s", "cores": 1, "kind": "reference",
            "sample": "%d-body %s (same generator, ncrit %d, theta %.4g), reference TapasFMM "
                      "pipeline, 1 MPI rank; phases tree %.2fs upward %.2fs traverse %.2fs "
                      "downward %.2fs (upward includes the reference's double upward pass "
                      "on one rank); process wall %.1fs" % (n, args.dist, ncrit, theta, ph["tree"],
                                                             ph["upward"], ph["traverse"],
                                                             ph["downward"], wall)}


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


def errors(ctx, targets, direct):
    import numpy as np
    f = ctx.fields()[targets]

    def rel(a, b):
        return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    return rel(f[:, 0], direct[:, 0]), rel(f[:, 1:], direct[:, 1:])


def calibrate_copy():
    import torch
    x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")  # 256 MiB
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_per_copy": x.numel() * 4, "copies": 3}))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU; this build has no multi-GPU path (--gpus %d refused)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    import numpy as np
    import torch

    import datasets
    import tfmm

    precision = tfmm.FP64 if args.precision == "fp64" else tfmm.MIXED
    probs = None
    if args.batch > 0:
        probs, kinds = datasets.batch(args.batch, seed=0)
        offsets = np.concatenate([[0], np.cumsum([len(p[0]) for p in probs])]).astype(np.int64)
        x, y, z, q = [np.concatenate([p[i] for p in probs]) for i in range(4)]
        args.n = int(offsets[-1])
    else:
        x, y, z, q = datasets.make(args.dist, args.n, seed=0)
    dev = torch.device("cuda:0")
    tx = torch.from_numpy(x).to(dev)
    ty = torch.from_numpy(y).to(dev)
    tz = torch.from_numpy(z).to(dev)
    tq = torch.from_numpy(q).to(dev)
    torch.cuda.synchronize()

    ctx = tfmm.Context(0)
    if args.graph:
        ctx.set_tuning(tfmm.TUNE_GRAPH, 1)
    fixed_knobs = {}
    if args.unit_factor is not None:
        ctx.set_tuning(tfmm.TUNE_M2L_UNIT, args.unit_factor)
        fixed_knobs["m2l_unit_factor"] = args.unit_factor
    if args.xcd_run is not None:
        ctx.set_tuning(tfmm.TUNE_M2L_XCD_RUN, args.xcd_run)
        fixed_knobs["m2l_xcd_run"] = args.xcd_run
    if args.p2p_schedule is not None:
        ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, args.p2p_schedule)
        fixed_knobs["p2p_schedule"] = args.p2p_schedule
    if probs is not None:
        ctx.set_batch_device(offsets, tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), tq.data_ptr())
    else:
        ctx.set_bodies_device(args.n, tx.data_ptr(), ty.data_ptr(), tz.data_ptr(), tq.data_ptr())

    # parameters: the learned picker (mixed precision) or the fixed ncrit/theta
    picked = None
    ncrit, theta = (args.ncrit or 64), args.theta
    if probs is not None and args.ncrit is None:
        # a batch fixes theta only: the leaf size is the fastest candidate on
        # these bodies (the forest has no single-problem statistics for the picker)
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
    if (precision == tfmm.MIXED and not args.no_learned and probs is None
            and os.path.exists(args.picker)):
        pk = tfmm.Picker.load(args.picker)
        t0 = time.perf_counter()
        p = ctx.pick_params(pk, precision)
        pick_ms = (time.perf_counter() - t0) * 1e3
        ncrit, theta = int(p.ncrit), float(p.theta)
        picked = {"ncrit": ncrit, "theta": theta, "ms": pick_ms,
                  "note": "trial tree at ncrit 64 + tree statistics + FP32 MLP inference "
                          "kernel, timed outside the steps"}
        pk.close()
    prm = tfmm.make_params(ncrit, theta, precision)

    # tile-kernel knobs: the library's autotuner on these bodies (untimed)
    tuned = {"fixed": fixed_knobs} if fixed_knobs else None
    if not args.no_autotune and not fixed_knobs:
        t0 = time.perf_counter()
        unit, run, sched, tms = ctx.autotune(prm, reps=3)
        tuned = {"m2l_unit_factor": unit, "m2l_xcd_run": run, "p2p_schedule": sched,
                 "ms_per_evaluation": tms, "seconds": time.perf_counter() - t0,
                 "note": "16 candidate settings x 3 evaluations, outside the timed steps"}

    # accuracy reference: FP64 direct sum on random targets (untimed)
    rng = np.random.default_rng(123)
    targets = np.sort(rng.choice(args.n, size=min(args.targets, args.n), replace=False))
    direct = ctx.direct(targets)

    ms_step, phases, st = time_config(ctx, prm, args.steps, args.warmup)
    err_pot, err_acc = errors(ctx, targets, direct)

    pairs = st["n_p2p_calls"] - st["n_p2p_self"]
    p2p_ms = phases["ms_p2p"]
    p2p_tflops = P2P_FLOPS_PER_PAIR * pairs / (p2p_ms * 1e-3) / 1e12 if p2p_ms > 0 else 0.0
    peak = PEAK_FP32_TFLOPS if precision == tfmm.MIXED else PEAK_FP64_TFLOPS
    # dominant kernel: the batched M2L tile kernel (one launch per step),
    # timed by HIP events on the context stream around that launch
    grouped = st["n_m2l_exec"] - st["n_m2l_fallback"]
    tiles_ms = phases["ms_m2l_tiles"]
    m2l_tflops = M2L_FLOPS_PER_PAIR * grouped / (tiles_ms * 1e-3) / 1e12 if tiles_ms > 0 else 0.0

    fp64 = None
    if args.fp64_steps > 0 and precision == tfmm.MIXED:
        prm64 = tfmm.make_params(args.ncrit or 64, args.theta, tfmm.FP64)
        ctx.set_tuning(tfmm.TUNE_M2L_UNIT, 0)  # the plain form keeps the default knobs
        ctx.set_tuning(tfmm.TUNE_M2L_XCD_RUN, 32)
        ctx.set_tuning(tfmm.TUNE_P2P_SCHEDULE, 0)
        ms64, ph64, st64 = time_config(ctx, prm64, args.fp64_steps, 1)
        e64 = errors(ctx, targets, direct)
        fp64 = {"value": args.n / (ms64 * 1e-3), "ms_per_step": ms64, "ncrit": args.ncrit or 64,
                "theta": args.theta, "rel_l2_pot": e64[0], "rel_l2_acc": e64[1],
                "phases_ms": ph64, "speedup_of_headline": ms64 / ms_step}

    if args.no_cpu:
        cpu = None
    elif probs is not None:
        cpu = cpu_baseline_batch(probs, ncrit, theta, args.cpu_sample)
    else:
        cpu = cpu_baseline(args, ncrit, theta)
    if probs is not None:
        workload = ("batch of %d problems (cube/plummer/ellipsoid, N log-uniform in [2^10, 2^16])"
                    " total N=%d P=10 ncrit=%d theta=%g %s" % (len(probs), args.n, ncrit,
                                                               round(theta, 4), args.precision))
    else:
        workload = "%s N=%d P=10 ncrit=%d theta=%g %s%s" % (
            args.dist, args.n, ncrit, round(theta, 4), args.precision,
            " (learned)" if picked else "")
    tile_kernel = "k_m2l_tiles<%s>" % ("float" if precision == tfmm.MIXED else "double")
    traffic = measured_traffic(workload, tile_kernel)
    out = {
        "metric": ("FMM bodies/s at rel-L2 err <=1e-4, 4M Plummer; P2P % FP32 peak"
                   if probs is None else
                   "FMM bodies/s over %d batched small problems in one forest" % len(probs)),
        "value": args.n / (ms_step * 1e-3),
        "unit": "bodies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == tfmm.FP64 else "f32 P2P/M2L + f64 tree/P2M/M2M/L2L/L2P",
        "data": "synthetic",
        "config": {"workload": workload,
            "n_bodies": args.n, "distribution": args.dist if probs is None else
            {k: kinds.count(k) for k in datasets.BATCH_KINDS}, "expansion": 10,
            "problems": 1 if probs is None else len(probs),
            "ncrit": ncrit, "theta": theta, "precision": args.precision},
        "error": {"rel_l2_pot": err_pot, "rel_l2_acc": err_acc, "targets": len(targets),
                  "against": "FP64 direct sum"},
        "picker": picked,
        "autotune": tuned,
        "graph": bool(args.graph),
        "phases_ms": phases,
        "counts": {"cells": st["ncells"], "leaves": st["nleaves"], "max_depth": st["max_depth"],
                   "m2l_pairs": st["n_m2l_exec"], "m2l_groups": st["n_m2l_groups"],
                   "m2l_items": st["n_m2l_items"], "m2l_fallback": st["n_m2l_fallback"],
                   "p2p_leaf_pairs": st["n_p2p_leafpairs"], "p2p_body_pairs": pairs},
        "roofline": {"kernel": tile_kernel,
                     "bound": "mfma", "achieved": m2l_tflops, "peak": peak, "unit": "TFLOP/s",
                     "frac": m2l_tflops / peak,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "traffic_knobs": traffic["knobs"] if traffic else None,
                     "flops_per_pair": M2L_FLOPS_PER_PAIR, "pairs_per_launch": grouped,
                     "ms_per_launch": tiles_ms,
                     "p2p": {"tflops": p2p_tflops, "flops_per_pair": P2P_FLOPS_PER_PAIR,
                             "body_pairs": pairs, "ms": p2p_ms,
                             "pct_fp32_peak": 100.0 * p2p_tflops / PEAK_FP32_TFLOPS}},
        "fp64_plain": fp64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()

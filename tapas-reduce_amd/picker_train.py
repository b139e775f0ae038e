"""Generates the picker's training data from the method itself and trains it
on the GPU.

Every problem (distribution, N log-uniform in [2^16, 2^23], seed) is
evaluated in mixed precision at each (ncrit, theta) of the grid
{16, 32, 64, 128, 256} x {0.2, 0.3, 0.36, 0.4, 0.42, ..., 0.5, 0.6}
(finer where the error crosses 1e-4); its error is the larger of the
potential and acceleration rel-L2 errors against the FP64 direct sum on 1024
random targets, its cost the wall time of one evaluation (best of two).  The
label of a problem is the cheapest configuration whose error is <= the
target times a safety margin; the features are the 16 statistics of its
trial tree (ncrit = 64).  The network is then trained by the library's own
kernels (tfmm_picker_train) and its weights written to picker/weights.npz.

Measuring is split into slices that can run as separate processes:

    python tapas-reduce_amd/picker_train.py --first 0 --count 16 --measure-only
    ...
    python tapas-reduce_amd/picker_train.py --data picker/runs_000.npz picker/runs_016.npz ...
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import datasets  # noqa: E402
import tfmm  # noqa: E402

NCRITS = tfmm.NCRIT_CLASSES
THETAS = (0.2, 0.3, 0.36, 0.4, 0.42, 0.44, 0.46, 0.48, 0.5, 0.6)
DISTS = ("cube", "plummer", "ellipsoid", "two_plummer")


def problem(i, seed0):
    rng = np.random.default_rng(seed0 + i)
    dist = DISTS[i % len(DISTS)]
    n = int(round(2.0 ** rng.uniform(16, 23)))
    return dist, n, seed0 + i


def measure(ctx, x, y, z, q, targets, direct, precision):
    """errors (25,) and times (25,) over the grid, plus the features"""
    ctx.set_bodies(x, y, z, q)
    ctx.build_tree(tfmm.make_params(64, 0.4, precision))
    feat = ctx.tree_features()
    errs = np.zeros((len(NCRITS), len(THETAS)))
    times = np.zeros_like(errs)
    ctx.evaluate(tfmm.make_params(64, 0.4, precision))  # warm-up of this size
    for a, nc in enumerate(NCRITS):
        for b, th in enumerate(THETAS):
            best = 1e30
            try:
                for _ in range(2):
                    t0 = time.perf_counter()
                    ctx.evaluate(tfmm.make_params(nc, th, precision))
                    ctx.sync()
                    best = min(best, time.perf_counter() - t0)
            except tfmm.TfmmError as e:  # e.g. lists too large for the device
                print("  (ncrit %d, theta %g) not evaluated: %s" % (nc, th, e), flush=True)
                errs[a, b] = np.inf
                times[a, b] = np.inf
                continue
            f = ctx.fields()[targets]
            ep = np.sqrt(((f[:, 0] - direct[:, 0]) ** 2).sum() / (direct[:, 0] ** 2).sum())
            ea = np.sqrt(((f[:, 1:] - direct[:, 1:]) ** 2).sum() / (direct[:, 1:] ** 2).sum())
            errs[a, b] = max(ep, ea)
            times[a, b] = best
    return feat, errs, times


def labels(errs, times, target):
    """(class, theta) of the cheapest configuration meeting the target"""
    ok = errs <= target
    if ok.any():
        t = np.where(ok, times, np.inf)
        a, b = np.unravel_index(np.argmin(t), t.shape)
    else:
        a, b = np.unravel_index(np.argmin(errs), errs.shape)
    return int(a), float(THETAS[b])


def generate(first, count, seed0, ntarget, precision, log):
    ctx = tfmm.Context(0)
    F, E, T, meta = [], [], [], []
    for i in range(first, first + count):
        dist, n, seed = problem(i, seed0)
        x, y, z, q = datasets.make(dist, n, seed=seed)
        rng = np.random.default_rng(seed)
        targets = np.sort(rng.choice(n, size=min(ntarget, n), replace=False))
        ctx.set_bodies(x, y, z, q)
        direct = ctx.direct(targets)
        t0 = time.time()
        feat, errs, times = measure(ctx, x, y, z, q, targets, direct, precision)
        F.append(feat)
        E.append(errs)
        T.append(times)
        meta.append((dist, n, seed))
        log("problem %3d %-12s n=%-8d  %.1fs  best-time %.1f ms" % (
            i, dist, n, time.time() - t0, 1e3 * times.min()))
    ctx.close()
    return np.array(F), np.array(E), np.array(T), meta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--count", type=int, default=96)
    ap.add_argument("--measure-only", action="store_true")
    ap.add_argument("--margin", type=float, default=0.6,
                    help="labels use target-error * margin")
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--targets", type=int, default=1024)
    ap.add_argument("--target-error", type=float, default=1e-4)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--holdout", type=float, default=0.2)
    ap.add_argument("--out", default=os.path.join(HERE, "picker"))
    ap.add_argument("--data", nargs="*", default=None,
                    help="train on saved measurement slices instead of measuring")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)

    def log(msg):
        print(msg, flush=True)

    if args.data:
        F, E, T, meta = [], [], [], []
        for path in args.data:
            d = np.load(path, allow_pickle=False)
            F.append(d["features"])
            E.append(d["errors"])
            T.append(d["times"])
            meta += [tuple(m) for m in json.loads(str(d["meta"]))]
        F, E, T = np.concatenate(F), np.concatenate(E), np.concatenate(T)
    else:
        F, E, T, meta = generate(args.first, args.count, args.seed, args.targets, tfmm.MIXED,
                                 log)
        np.savez(os.path.join(args.out, "runs_%03d.npz" % args.first), features=F, errors=E,
                 times=T, meta=json.dumps(meta), ncrits=np.array(NCRITS),
                 thetas=np.array(THETAS))
        if args.measure_only:
            return
    lab = [labels(E[i], T[i], args.target_error * args.margin) for i in range(len(F))]
    cls = np.array([c for c, _ in lab], np.int32)
    th = np.array([t for _, t in lab], np.float32)
    n = len(F)
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(n)
    nh = int(round(args.holdout * n))
    test_idx, train_idx = perm[:nh], perm[nh:]
    pk = tfmm.Picker(0, seed=args.seed)
    loss = pk.train(F[train_idx], cls[train_idx], th[train_idx], steps=args.steps, lr=args.lr)
    theta_max = float(th[train_idx].max())  # the head must not extrapolate past its labels
    pk.set_theta_max(theta_max)
    np.savez(os.path.join(args.out, "weights.npz"), weights=pk.weights(),
             train_idx=train_idx, test_idx=test_idx, target_error=args.target_error,
             margin=args.margin, theta_max=theta_max)
    # the picked configurations on the held-out problems, evaluated at the
    # exact predicted (ncrit, theta) next to the fallback (64, 0.4)
    out = pk.predict(F)
    nc, pt = tfmm.decode(out, theta_max)
    report = []
    ctx = tfmm.Context(0)
    for i in test_idx:
        dist, nb, seed = meta[i]
        x, y, z, q = datasets.make(dist, int(nb), seed=int(seed))
        rng = np.random.default_rng(int(seed))
        targets = np.sort(rng.choice(int(nb), size=min(args.targets, int(nb)), replace=False))
        ctx.set_bodies(x, y, z, q)
        direct = ctx.direct(targets)
        res = {}
        for name, (a_nc, a_th) in (("picked", (int(nc[i]), float(pt[i]))), ("fallback", (64, 0.4))):
            prm = tfmm.make_params(a_nc, a_th, tfmm.MIXED)
            ctx.evaluate(prm)
            best = 1e30
            for _ in range(3):
                t0 = time.perf_counter()
                ctx.evaluate(prm)
                ctx.sync()
                best = min(best, time.perf_counter() - t0)
            f = ctx.fields()[targets]
            ep = np.sqrt(((f[:, 0] - direct[:, 0]) ** 2).sum() / (direct[:, 0] ** 2).sum())
            ea = np.sqrt(((f[:, 1:] - direct[:, 1:]) ** 2).sum() / (direct[:, 1:] ** 2).sum())
            res[name] = (a_nc, a_th, float(max(ep, ea)), best)
        report.append(dict(problem=meta[i], picked=res["picked"][:2], err=res["picked"][2],
                           time=res["picked"][3], fallback_err=res["fallback"][2],
                           fallback_time=res["fallback"][3]))
        log("held-out %-12s n=%-8d picked (%d, %.3f) err %.2e %.1f ms | fallback err %.2e "
            "%.1f ms" % (dist, int(nb), res["picked"][0], res["picked"][1], res["picked"][2],
                         1e3 * res["picked"][3], res["fallback"][2], 1e3 * res["fallback"][3]))
    ctx.close()
    ok = np.mean([r["err"] <= args.target_error for r in report]) if report else float("nan")
    gain = (np.exp(np.mean([np.log(r["fallback_time"] / r["time"]) for r in report]))
            if report else float("nan"))
    summary = dict(problems=n, runs=int(E.size), train=len(train_idx), heldout=len(test_idx),
                   loss_first=float(loss[0]), loss_last=float(loss[-1]),
                   heldout_meets_target=float(ok), heldout_geomean_speedup_vs_fallback=float(gain))
    with open(os.path.join(args.out, "summary.json"), "w") as f:
        json.dump(dict(summary=summary, heldout=report), f, indent=1)
    log(json.dumps(summary))


if __name__ == "__main__":
    main()

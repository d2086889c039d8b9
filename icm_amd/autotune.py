"""MI355X autotuner: picks the planner's block size B and schoolbook /
transform crossover per (n, k) bucket from timings on the device, and writes
the compiled-in tuning table (icm_amd/csrc/tuning_mi355x.h).

It replaces the reference's per-device dispatch calibration
(tools/calibrate_gpu.cu, tools/calibrate_gpu_best_b.cu ->
devices/b200/gpu_fft_config.h): instead of fitting a cost model of vendor FFT
tiers, it times the engine's own launch tiers (schoolbook sums for levels
whose children keep c <= sb_max coefficients, LDS transforms, four-step
transforms) end to end.  Acceptance rule (the reference's calibration
check): with the chosen configurations the pass time must not decrease as n
or k grows along the grid.

Usage (on a GPU host, after building):
    python -m icm_amd.autotune                 # full grid, writes the header
    python -m icm_amd.autotune --quick --dry-run
Rebuild (python -m icm_amd.build) afterwards to compile the new table in.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

import icm_amd

HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "tuning_mi355x.h")
SB_CANDIDATES = [0, 17, 33, 65, 129, 257, 512]


def workload(n, k, seed=3):
    rng = np.random.default_rng(seed)
    S = np.exp(rng.uniform(0.0, math.log(1e9), n))
    p = np.arange(k, 0, -1, dtype=np.float64)
    return S, p / p.sum()


def time_plan(plan, p, Q, reps):
    plan.run(p, Q)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.run(p, Q)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)) * 1e3


def k_values(n):
    # about one point per decade of k (the crossover moves with k as much as with n)
    ks = sorted({1, 9, 90, 900, max(1, n // 10), n})
    return [k for k in ks if k <= n]


def grid(quick):
    ns = [1024, 16384, 131072] if quick else [1024, 4096, 16384, 65536, 262144, 1048576]
    return [(n, k) for n in ns for k in k_values(n)]


def tune_shape(n, k, Q, reps, log):
    S, p = workload(n, k)
    best = None
    results = []
    for B in ([16, 32] if n <= 64 else [32, 64]):
        plan = icm_amd.Plan(S, k, block_size=B)
        for sb in SB_CANDIDATES:
            plan.set_schoolbook_max(sb)
            ms = time_plan(plan, p, Q, reps)
            results.append({"B": B, "sb_max": sb, "ms": ms})
            if best is None or ms < best["ms"]:
                best = {"B": B, "sb_max": sb, "ms": ms}
        plan.close()
    log(f"n={n:8d} k={k:8d}: best B={best['B']} sb_max={best['sb_max']} {best['ms']:.3f} ms "
        f"(transforms only, B=32: {next(r['ms'] for r in results if r['B'] == 32 and r['sb_max'] == 0):.3f} ms)")
    return best, results


def tune_batch(Q, reps, log, count=4096):
    """Block size and crossover of ragged batches, timed on the ragged
    benchmark mix (icm_amd.workloads.ragged_instances, 4096 instances)."""
    from icm_amd.workloads import ragged_instances
    stacks, pays = ragged_instances(count, seed=4)
    ks = [len(x) for x in pays]
    flat = np.concatenate(pays)
    best, results = None, []
    for B in (32, 64):
        bp = icm_amd.BatchPlan(stacks, ks, block_size=B)
        for sb in SB_CANDIDATES:
            bp.set_schoolbook_max(sb)
            bp.run_flat(flat, Q)
            ts = []
            for _ in range(reps):
                bp.run_flat(flat, Q)
                ts.append(bp.stats.total_ns / 1e6)
            ms = float(np.median(ts))
            results.append({"B": B, "sb_max": sb, "ms": ms})
            if best is None or ms < best["ms"]:
                best = {"B": B, "sb_max": sb, "ms": ms}
        bp.close()
    log(f"ragged batch ({count} instances): best B={best['B']} sb_max={best['sb_max']} {best['ms']:.3f} ms")
    return best, results


def small_work(n, k):
    """Schoolbook multiply-adds per quadrature node of the one-launch
    small-instance kernel (the planner's small_work, same geometry)."""
    nr, L = [n], 1
    while (1 << (L - 1)) < n:
        L += 1
    for _ in range(1, L):
        nr.append((nr[-1] + 1) // 2)
    c = [min((1 << l) + 1, k) for l in range(L)]
    g = [min(1 << l, k) for l in range(L)]
    w = sum(nr[l] * c[l] * min(c[l - 1], c[l]) for l in range(1, L - 1))
    return w + sum(nr[l - 1] * g[l - 1] * c[l - 1] for l in range(1, L))


def tune_small(Q, reps, log):
    """Crossover of the one-launch small-instance path against the
    multi-launch path, as a bound on small_work(n, k)."""
    rng = np.random.default_rng(5)
    wins, losses, rows = [], [], []
    for n in (64, 256, 1024, 2048, 4096):
        for k in (1, 3, 10, 30, 100, 300, 1000):
            if k > n:
                continue
            S = np.exp(rng.uniform(0.0, math.log(1e6), n))
            p = np.ones(k) / k
            plan = icm_amd.Plan(S, k)
            plan.set_small_path(1)
            if not plan.small_path:
                plan.close()
                continue
            t_small = time_plan(plan, p, Q, reps)
            plan.set_small_path(0)
            t_multi = time_plan(plan, p, Q, reps)
            plan.close()
            w = small_work(n, k)
            (wins if t_small < t_multi else losses).append(w)
            rows.append({"n": n, "k": k, "work": w, "small_ms": t_small, "multi_ms": t_multi})
            log(f"small path n={n} k={k} work={w}: {t_small:.3f} ms vs {t_multi:.3f} ms multi-launch")
    first_loss = min(losses) if losses else None
    below = [w for w in wins if first_loss is None or w < first_loss]
    if not below:
        return {"max_work": 0.0, "note": "small path never faster"}, rows
    hi = max(below)
    bound = math.sqrt(hi * first_loss) if first_loss else 2.0 * hi
    return {"max_work": bound, "note": f"fastest up to work {hi}, first loss at {first_loss}"}, rows


def check_monotone(table):
    """Pass time must not decrease with n (fixed k fraction) or with k (fixed n)."""
    bad = []
    by_n = {}
    for (n, k), v in table.items():
        by_n.setdefault(n, []).append((k, v["ms"]))
    for n, row in by_n.items():
        row.sort()
        for (k1, t1), (k2, t2) in zip(row, row[1:]):
            if t2 < t1 * 0.97:
                bad.append(f"n={n}: k={k2} ({t2:.3f} ms) faster than k={k1} ({t1:.3f} ms)")
    ns = sorted(by_n)
    for n1, n2 in zip(ns, ns[1:]):
        for frac in ("1", "n"):
            k1 = 1 if frac == "1" else n1
            k2 = 1 if frac == "1" else n2
            t1, t2 = table[(n1, k1)]["ms"], table[(n2, k2)]["ms"]
            if t2 < t1 * 0.97:
                bad.append(f"k={frac}: n={n2} ({t2:.3f} ms) faster than n={n1} ({t1:.3f} ms)")
    return bad


def write_header(table, path, batch=None, small=None):
    ns = sorted({n for n, _ in table})
    lines = ["// tuning_mi355x.h -- planner table for MI355X (gfx950), written by",
             "// icm_amd/autotune.py from timings on the device.  First row whose limits",
             "// cover (n, k) wins.", "#pragma once", "",
             "struct TuneRow {", "    int n_max, k_max;",
             "    int B;          // block size (players per leaf block)",
             "    int sb_max;     // schoolbook for levels with c_{l-1} <= sb_max", "};", "",
             "static const TuneRow kTuneTable[] = {"]
    # bucket limits at the geometric means of neighbouring grid points (the
    # grids are roughly log-spaced)
    for i, n in enumerate(ns):
        n_max = 0x7fffffff if i == len(ns) - 1 else math.isqrt(n * ns[i + 1])
        ks = sorted(k for nn, k in table if nn == n)
        for j, k in enumerate(ks):
            k_max = 0x7fffffff if j == len(ks) - 1 else math.isqrt(k * ks[j + 1])
            v = table[(n, k)]
            lines.append(f"    {{{n_max}, {k_max}, {v['B']}, {v['sb_max']}}},   // tuned at n={n} k={k}: "
                         f"{v['ms']:.3f} ms")
    lines += ["};", ""]
    b = batch or {"B": 32, "sb_max": 129, "ms": float("nan")}
    lines += ["// Ragged batches (one setting per batch), tuned on the 4096-instance ragged mix.",
              "struct BatchTune {", "    int B, sb_max;", "};",
              f"static const BatchTune kBatchTune = {{{b['B']}, {b['sb_max']}}};   // {b['ms']:.3f} ms", ""]
    sm = small or {"max_work": 262144.0, "note": "default"}
    lines += ["// Small instances in one launch (k_small): used when the tree fits one",
              "// workgroup's LDS and its schoolbook multiply-adds per quadrature node are at",
              "// most this many (crossover against the multi-launch path).",
              f"static const double kSmallMaxWork = {sm['max_work']:.1f};   // {sm['note']}", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--quick", action="store_true", help="small grid (three n values)")
    ap.add_argument("--Q", type=int, default=256)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--dry-run", action="store_true", help="print the table, do not write the header")
    ap.add_argument("--json", default="", help="also write all timings to this file")
    ap.add_argument("--from-json", default="", help="write the header from a timings file of an earlier run")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if args.from_json:
        with open(args.from_json) as f:
            data = json.load(f)
        table = {tuple(int(x) for x in key.split(",")): v for key, v in data["best"].items()}
        bad = check_monotone(table)
        log("acceptance: " + ("PASS" if not bad else f"{len(bad)} violations: {bad}"))
        write_header(table, HEADER, data.get("batch_best"), data.get("small_best"))
        log(f"wrote {HEADER}")
        return 0 if not bad else 1

    icm_amd.init()
    table, every = {}, {}
    for n, k in grid(args.quick):
        best, results = tune_shape(n, k, args.Q, args.reps, log)
        table[(n, k)] = best
        every[f"{n},{k}"] = results
    bad = check_monotone(table)
    for b in bad:
        log(f"monotonicity: {b}")
    log("acceptance: " + ("PASS (time monotone in n and k)" if not bad else f"{len(bad)} violations"))
    batch_best, batch_all = tune_batch(args.Q, args.reps, log)
    small_best, small_all = tune_small(args.Q, max(args.reps, 5), log)
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"best": {f"{n},{k}": v for (n, k), v in table.items()}, "all": every,
                       "monotone_violations": bad, "batch_best": batch_best, "batch_all": batch_all,
                       "small_best": small_best, "small_all": small_all}, f, indent=1)
    if not args.dry_run:
        write_header(table, HEADER, batch_best, small_best)
        log(f"wrote {HEADER}; rebuild with python -m icm_amd.build")
    return 0 if not bad else 1


if __name__ == "__main__":
    sys.exit(main())

"""Command-line benchmarks of the GPU engine (the reference's bench_gpu modes
that need no CPU reference, bench/bench_gpu.cu:194-278, and the 1-second
frontier search of tools/threshold_search_gpu.cu).

    python -m icm_amd.cli quick                 # grid n x k at Q = 64, one line per shape
    python -m icm_amd.cli bench <n> <k> [reps] [Q]
    python -m icm_amd.cli threshold [kn|k100|all]

Stacks and payouts follow the reference's generators in kind (uniform
stacks in [1, 100], payout n - r, unnormalised) with numpy's seeded PCG64.
The GPU-versus-CPU modes (verify, verify_ext) live in tests/verify_gpu.py
beside the CPU oracle they compare against.
"""
import sys
import time

import numpy as np

import icm_amd


def make_stacks(n, dist=0, seed=42):
    """Distributions of bench/bench_gpu.cu:25-44 (own generator)."""
    rng = np.random.default_rng(seed)
    if dist == 0:
        return 1.0 + 99.0 * rng.uniform(0.0, 1.0, n)
    if dist == 1:
        S = np.ones(n)
        S[0] = 10000.0
        return S
    if dist == 2:
        return 2.0 ** (np.arange(n) * 10.0 / n)
    if dist == 3:
        return np.full(n, 100.0)
    return 1.0 + rng.uniform(0.0, 1.0, n)


def make_payout(n, k):
    return (n - np.arange(k)).astype(np.float64)


def _plan_ms(plan, p, Q):
    plan.run(p, Q)          # warmup (first launches load the code objects)
    plan.run(p, Q)
    return plan.stats.total_ns / 1e6


def run_quick(Q=64):
    print(f"=== GPU QUICK GRID (Q={Q}) ===")
    for n in (256, 1024, 4096, 8192, 16384):
        S = make_stacks(n, 0)
        for k in (min(64, n), min(256, n), n):
            plan = icm_amd.Plan(S, k)
            ms = _plan_ms(plan, make_payout(n, k), Q)
            summ = plan.summary()
            plan.close()
            print(f"n={n:<7d} k={k:<7d} time={ms:8.2f} ms  B={summ['B']:<4d} "
                  f"peak={summ['planned_peak_vram_bytes'] / 2**20:8.1f} MB", flush=True)
    return 0


def run_bench(n, k, reps=3, Q=256):
    S = make_stacks(n, 0)
    p = make_payout(n, k)
    plan = icm_amd.Plan(S, k)
    summ = plan.summary()
    best = float("inf")
    for r in range(reps):
        ms = _plan_ms(plan, p, Q)
        best = min(best, ms)
        print(f"run={r}  total={ms:.2f} ms  B={summ['B']}  engine={summ['engine']}  "
              f"peak_vram={summ['planned_peak_vram_bytes'] / 2**20:.1f} MB", flush=True)
    plan.close()
    print(f"best={best:.2f} ms  (n={n} k={k} Q={Q})")
    return 0


def measure_ms(n, k, reps=5, Q=256, seed=None):
    """Median device time of `reps` passes after one warmup, plan reused
    (tools/threshold_search_gpu.cu:67-130); negative when the shape does not run."""
    rng = np.random.default_rng(123 + n if seed is None else seed)
    S = 1.0 + 99.0 * rng.uniform(0.0, 1.0, n)
    p = make_payout(n, k)
    try:
        plan = icm_amd.Plan(S, k)
    except icm_amd.ICMError as e:
        print(f"  [measure_ms] n={n} k={k} plan_create FAIL: {e}", file=sys.stderr, flush=True)
        return -1.0
    try:
        plan.run(p, Q)
        ts = []
        for _ in range(reps):
            plan.run(p, Q)
            ts.append(plan.stats.total_ns / 1e6)
    except icm_amd.ICMError as e:
        print(f"  [measure_ms] n={n} k={k} FAIL: {e}", file=sys.stderr, flush=True)
        return -1.0
    finally:
        plan.close()
    return float(sorted(ts)[len(ts) // 2])


def threshold(k_mode, limit_ms=1000.0):
    """Largest n with median time <= limit (binary search, k = n or k = 100)."""
    label = "k=n" if k_mode == 0 else "k=100"
    print(f"\n=== Binary search: {label} ===")
    print(f"{'n':<12s} {'k':<8s} {'median_ms':<12s} verdict")
    print("-" * 40)
    lo, hi = (262144, 4194304) if k_mode == 0 else (1048576, 16777216)
    kk = (lambda n: n) if k_mode == 0 else (lambda n: min(100, n))
    t_lo = measure_ms(lo, kk(lo))
    print(f"{lo:<12d} {kk(lo):<8d} {t_lo:<12.3f} {'UNDER' if 0 <= t_lo <= limit_ms else 'OVER'}", flush=True)
    if not (0 <= t_lo <= limit_ms):
        return None
    t_hi = measure_ms(hi, kk(hi))
    while 0 <= t_hi <= limit_ms:
        print(f"{hi:<12d} {kk(hi):<8d} {t_hi:<12.3f} UNDER", flush=True)
        lo, hi = hi, hi * 2
        t_hi = measure_ms(hi, kk(hi))
        if hi > (1 << 26):
            print(f"\n*** {label}: entire tested range under {limit_ms:.0f} ms (hi={lo}) ***")
            return lo
    print(f"{hi:<12d} {kk(hi):<8d} {t_hi:<12.3f} {'OVER' if t_hi > limit_ms else 'FAIL'}", flush=True)
    while hi - lo > max(1024, lo // 100):
        mid = (lo + hi) // 2
        t = measure_ms(mid, kk(mid))
        under = 0 <= t <= limit_ms
        print(f"{mid:<12d} {kk(mid):<8d} {t:<12.3f} {'UNDER' if under else 'OVER':<6s} [lo={lo} hi={hi}]",
              flush=True)
        if under:
            lo = mid
        else:
            hi = mid
    print(f"\n--- {label} threshold ---")
    print(f"Largest n <= {limit_ms:.0f}ms: {lo}")
    print(f"Smallest n > {limit_ms:.0f}ms: {hi}")
    print(f"Search bracket width: {hi - lo} ({100.0 * (hi - lo) / lo:.1f}% of lo)")
    return lo


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    icm_amd.init()
    if argv and argv[0] == "quick":
        return run_quick()
    if argv and argv[0] == "bench":
        if len(argv) < 3:
            print("Usage: python -m icm_amd.cli bench <n> <k> [reps] [Q]")
            return 1
        n, k = int(argv[1]), int(argv[2])
        reps = int(argv[3]) if len(argv) > 3 else 3
        Q = int(argv[4]) if len(argv) > 4 else 256
        return run_bench(n, k, reps, Q)
    if argv and argv[0] == "threshold":
        which = argv[1] if len(argv) > 1 else "all"
        if which not in ("kn", "k100", "all"):
            print("usage: python -m icm_amd.cli threshold [kn|k100|all]")
            return 1
        print("threshold search: binary search for the 1000 ms crossing")
        print("method: median of 5 reps per candidate n, Q=256, plan reused (icm_hip_plan_create + "
              "icm_hip_equity_with_plan)")
        res = {}
        if which in ("kn", "all"):
            res["k=n"] = threshold(0)
        if which in ("k100", "all"):
            res["k=100"] = threshold(1)
        print("\n========================================\nFINAL RESULTS:")
        for key, v in res.items():
            print(f"  {key:5s} threshold n = {v}" if v else f"  {key:5s} threshold: SEARCH FAILED")
        return 0
    print("Usage:\n  python -m icm_amd.cli quick\n  python -m icm_amd.cli bench <n> <k> [reps] [Q]\n"
          "  python -m icm_amd.cli threshold [kn|k100|all]\n  python tests/verify_gpu.py verify|verify_ext")
    return 1


if __name__ == "__main__":
    sys.exit(main())

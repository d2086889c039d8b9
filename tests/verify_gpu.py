"""GPU-versus-CPU verification grids of the reference's bench_gpu
(bench/bench_gpu.cu:64-192): every shape runs on the GPU through the C ABI
(icm_gpu_equity) and on the CPU oracle, per-player relative error with the
denominator 1 below 1e-14, pass below 1e-8.

    python tests/verify_gpu.py verify        # n 64 .. 65536, dists 0-2, k in {min(100,n), n}
    python tests/verify_gpu.py verify_ext    # n 64 .. 131072, dists 0-3, k in {16, 100, 512, n/2, n}
Both add k = 1024 and 2048 at n >= 65536.  ICM_GPU_VERIFY_EXT_MAX_N caps n
and ICM_GPU_VERIFY_EXT_LITE=1 trims the extended k grid, as in the
reference.  Exit status 0 when every shape passes.
"""
import ctypes
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import icm_amd  # noqa: E402
from icm_amd.cli import make_payout  # noqa: E402
from oracle import oracle as O  # noqa: E402


def reference_stacks(n, dist):
    """bench/bench_gpu.cu:25-47 with the same generator: glibc srand(42) /
    rand() / RAND_MAX, so the stacks are those the reference verifies."""
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(42)
    rand_max = 2147483647.0
    if dist == 0:
        return np.array([1.0 + 99.0 * (libc.rand() / rand_max) for _ in range(n)])
    if dist == 1:
        S = np.ones(n)
        S[0] = 10000.0
        return S
    if dist == 2:
        return 2.0 ** (np.arange(n) * 10.0 / n)
    if dist == 3:
        return np.full(n, 100.0)
    return np.array([1.0 + libc.rand() / rand_max for _ in range(n)])


def ext_ks(n, lite=False):
    """k list of the extended grid (bench/bench_gpu.cu:101-119)."""
    ks = [min(100, n), max(1, n // 2), n] if lite else [min(16, n), min(100, n), min(512, n), max(1, n // 2), n]
    if n >= 65536:
        ks += [1024, 2048]
    return ks


def rel_err_max(gpu, cpu):
    den = np.abs(cpu).copy()
    den[den < 1e-14] = 1.0
    r = np.abs(gpu - cpu) / den
    i = int(np.argmax(r))
    return float(r[i]), i


def run_verify(extended, Q=256, log=print):
    ns = [64, 128, 256, 1024, 4096, 16384, 65536, 131072] if extended else [64, 256, 1024, 4096, 16384, 65536]
    if extended:
        cap = int(os.environ.get("ICM_GPU_VERIFY_EXT_MAX_N", "0") or 0)
        if cap > 0 and any(n <= cap for n in ns):
            ns = [n for n in ns if n <= cap]
    dists = [0, 1, 2, 3] if extended else [0, 1, 2]
    lite = extended and os.environ.get("ICM_GPU_VERIFY_EXT_LITE", "0") not in ("", "0")
    all_pass = True
    worst = (0.0, 0, 0, 0, -1)
    icm_amd.init()
    for n in ns:
        for d in dists:
            S = reference_stacks(n, d)
            if extended:
                ks = ext_ks(n, lite)
            else:
                ks = [min(100, n), n] + ([1024, 2048] if n >= 65536 else [])
            for k in ks:
                p = make_payout(n, k)
                t0 = time.perf_counter()
                cpu = O.equity(S, p, Q=Q, B=32)
                t_cpu = (time.perf_counter() - t0) * 1e3
                plan = icm_amd.Plan(S, k)
                try:
                    gpu = plan.run(p, Q)
                except icm_amd.ICMError as e:
                    log(f"FAIL n={n} k={k} dist={d} gpu-error={e}")
                    all_pass = False
                    plan.close()
                    continue
                t_gpu = plan.stats.total_ns / 1e6
                B = plan.stats.B
                plan.close()
                err, idx = rel_err_max(gpu, cpu)
                ok = err < 1e-8
                all_pass &= ok
                if err > worst[0]:
                    worst = (err, n, k, d, idx)
                log(f"{'PASS' if ok else 'FAIL'} n={n:<7d} k={k:<7d} dist={d}  err={err:.3e}  "
                    f"cpu={t_cpu:.1f} ms gpu={t_gpu:.1f} ms  B={B}")
                if extended and not ok:
                    log(f"  diag: idx={idx} cpu={cpu[idx]:.17g} gpu={gpu[idx]:.17g} abs={abs(cpu[idx] - gpu[idx]):.3e}")
    if extended:
        log(f"verify_ext_worst: err={worst[0]:.3e} n={worst[1]} k={worst[2]} dist={worst[3]} idx={worst[4]}")
    return all_pass


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] not in ("verify", "verify_ext"):
        print("Usage:\n  python tests/verify_gpu.py verify\n  python tests/verify_gpu.py verify_ext")
        return 1

    def log(msg):
        print(msg, flush=True)

    return 0 if run_verify(argv[0] == "verify_ext", log=log) else 1


if __name__ == "__main__":
    sys.exit(main())

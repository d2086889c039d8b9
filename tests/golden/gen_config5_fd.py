"""Writes tests/golden/config5_fd.npz: central finite differences of the CPU
oracle's v . E at the adjoint workload (n = 131072, k = 1000, Q = 256,
stacks lognormal(0, 1.5), cotangent v ~ N(0, 1), linear payout k - j
normalised), on 256 random coordinates at h = 1e-6 S_i and the first 16 of
them at h = 1e-5 S_i.  The largest stack is excluded: the quadrature nodes
depend on max S, which the adjoint holds fixed.

Run from the repository root (about 40 minutes on 8 threads):
    python tests/golden/gen_config5_fd.py [threads]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

N, K, Q, SEED = 131072, 1000, 256, 5


def workload():
    rng = np.random.default_rng(SEED)
    S = rng.lognormal(0.0, 1.5, N)
    v = rng.normal(size=N)
    p = np.arange(K, 0, -1, dtype=np.float64)
    p /= p.sum()
    imax = int(np.argmax(S))
    coords = [int(i) for i in rng.choice(N, 264, replace=False) if i != imax][:256]
    return S, v, p, np.array(coords, dtype=np.int64)


def main():
    nt = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    S, v, p, coords = workload()

    def dot(SS):
        return float(np.dot(v, O.equity(SS, p, Q=Q, fft_min=64, nthreads=nt)))

    t0 = time.time()
    base = dot(S)
    fd6, fd5 = [], []
    for c, i in enumerate(coords):
        for rel, out in ((1e-6, fd6), (1e-5, fd5)):
            if rel == 1e-5 and c >= 16:
                continue
            h = rel * S[i]
            Sp = S.copy(); Sp[i] += h
            Sm = S.copy(); Sm[i] -= h
            out.append((dot(Sp) - dot(Sm)) / (2 * h))
        print(f"{c + 1}/{len(coords)} coordinate {i}: fd6 {fd6[-1]:.6e} ({time.time() - t0:.0f} s)", flush=True)
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "config5_fd.npz"),
             n=N, k=K, Q=Q, seed=SEED, coords=coords, fd6=np.array(fd6), fd5=np.array(fd5),
             vE=base, S_sum=S.sum(), S_head=S[:4])


if __name__ == "__main__":
    main()

"""Writes tests/golden/config3.npz: oracle values for the large forward
workload (n = 2^20 stacks log-uniform on [1, 1e9], seed 3, Q = 256; see
icm_amd.workloads.loguniform_field), at sampled players:

  * k = n/4, linear and quadratic payouts: the oracle's integrated equities
    (all n players computed; 1024 sampled players stored, plus sum(E) and
    sum(E * w) for a seeded normal w as checksums of the whole vector);
  * k = n/4, both payouts: Psi of the same players at 6 quadrature nodes
    spanning the rule (oracle hybrid engine, players in sorted order);
  * k = n, quadratic payout: V1 (direct sums) and V2 (integral form) of 64
    sampled players, giving the exact equity (2 V2 + 3 V1 - 2) / sum m^2.

Run from the repository root (about 9 minutes on 8 threads):
    python tests/golden/gen_config3.py [threads]
"""
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
import closed_forms as C  # noqa: E402
from icm_amd.workloads import loguniform_field  # noqa: E402
from oracle import oracle as O  # noqa: E402

N, Q, SEED = 1 << 20, 256, 3
NODES = [20, 70, 110, 150, 190, 235]


def samples():
    rng = np.random.default_rng(33)
    idx = np.sort(rng.choice(N, 1024, replace=False))
    w = np.random.default_rng(34).normal(size=N)
    idx64 = np.sort(np.random.default_rng(35).choice(N, 64, replace=False))
    return idx, w, idx64


def main():
    nt = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    idx, w, idx64 = samples()
    out = {"n": N, "Q": Q, "seed": SEED, "idx": idx, "idx64": idx64, "nodes": np.array(NODES)}
    t0 = time.time()
    k4 = N // 4
    for pay in ("linear", "quadratic"):
        S, p = loguniform_field(N, k4, SEED, pay)
        E = O.equity(S, p, Q=Q, fft_min=64, nthreads=nt)
        out[f"E_{pay}"] = E[idx]
        out[f"sumE_{pay}"] = E.sum()
        out[f"wE_{pay}"] = float(np.dot(w, E))
        print(f"k=n/4 {pay}: equities ({time.time() - t0:.0f} s)", flush=True)
        order = np.argsort(-S, kind="stable")          # stack-descending (ties by index)
        logv, _ = O.make_nodes(Q, S.max())
        a_nodes = [np.exp(S[order] * logv[q]) for q in NODES]

        def node(a):
            return O.inner(a, p, engine=O.ENGINE_HYBRID, B=32, fft_min=64)

        with ThreadPoolExecutor(nt) as ex:
            psi = list(ex.map(node, a_nodes))
        inv = np.empty(N, dtype=np.int64)
        inv[order] = np.arange(N)
        out[f"psi_{pay}"] = np.array([ps[inv[idx]] for ps in psi])   # [node][sample], original player ids
        print(f"k=n/4 {pay}: node Psi ({time.time() - t0:.0f} s)", flush=True)
    S, _ = loguniform_field(N, N, SEED, "quadratic")
    out["V1"] = C.v1_direct(S, idx64)
    out["V1m_int"], out["V2"] = C.v_integrals(S, idx64)
    out["S_sum"] = S.sum()
    out["S_head"] = S[:4]
    print(f"k=n closed forms ({time.time() - t0:.0f} s)", flush=True)
    np.savez(os.path.join(HERE, "config3.npz"), **out)


if __name__ == "__main__":
    main()

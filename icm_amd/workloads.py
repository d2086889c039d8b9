"""Synthetic inputs of the benchmark shapes (numpy PCG64, seeded).

loguniform_field: the large forward workload: n stacks log-uniform on
[1, 1e9] (the 1e9:1 extreme ratio), linear payout k - j or quadratic
(k - j)^2, normalised to sum 1.

ragged_instances: independent tournaments of uneven size for ragged batches:
n = floor(exp(U[ln 2, ln n_max])), k = min(n, choice{1, 3, 9, ceil(n/10), n}),
stacks Dirichlet(1) x 1e7 chips, payouts geometric with ratio U[0.5, 0.95],
normalised to sum 1.
"""
import math

import numpy as np


def loguniform_field(n, k, seed=3, payout="linear"):
    rng = np.random.default_rng(seed)
    S = np.exp(rng.uniform(0.0, math.log(1e9), n))
    p = np.arange(k, 0, -1, dtype=np.float64)
    if payout == "quadratic":
        p = p * p
    elif payout != "linear":
        raise ValueError(f"payout must be 'linear' or 'quadratic', got {payout!r}")
    return S, p / p.sum()


def ragged_instances(count, seed=4, n_max=4096, n_min=2):
    rng = np.random.default_rng(seed)
    stacks, payouts = [], []
    for _ in range(count):
        n = int(math.floor(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        n = max(n_min, min(n, n_max))
        k = min(n, [1, 3, 9, math.ceil(n / 10), n][int(rng.integers(5))])
        S = rng.dirichlet(np.ones(n)) * 1e7
        S = np.maximum(S, 1e-300)          # Dirichlet draws can underflow to 0 for large n
        r = rng.uniform(0.5, 0.95)
        p = r ** np.arange(k)
        stacks.append(S)
        payouts.append(p / p.sum())
    return stacks, payouts

"""GPU schoolbook tier (direct sums for levels with few coefficients).

Pins: the same plan with every level as transforms (schoolbook_max = 0) and
with every eligible level as schoolbook sums (schoolbook_max = 512) agree to
rounding, and both match the oracle; subsets (pruned parent lists), tiny k,
a missing right child at every level and the root level are covered.
"""
import math
import sys

import numpy as np
import pytest

import icm_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(gpu, cpu):
    den = np.abs(cpu).copy()
    den[den < 1e-14] = 1.0
    return float(np.max(np.abs(gpu - cpu) / den))


@pytest.mark.parametrize("n,k", [(40, 40), (97, 3), (333, 1), (1000, 9), (2049, 205), (5000, 5000),
                                 (20000, 450), (65536, 65536)])
def test_schoolbook_vs_transforms_and_oracle(n, k):
    rng = np.random.default_rng(n + 7 * k)
    S = np.exp(rng.uniform(0.0, math.log(1e6), n))
    p = 0.85 ** np.arange(k)
    p /= p.sum()
    say(f"schoolbook n={n} k={k}")
    plan = icm_amd.Plan(S, k)
    plan.set_schoolbook_max(0)
    e_fft = plan.run(p, Q=256)
    plan.set_schoolbook_max(512)
    e_sb = plan.run(p, Q=256)
    plan.close()
    d = float(np.max(np.abs(e_sb - e_fft)) / np.max(np.abs(e_fft)))
    assert d <= 1e-12, f"schoolbook vs transforms {d:.2e}"
    if n <= 20000:
        ref = O.equity(S, p, Q=256, B=32)
        assert rel_err(e_sb, ref) < 1e-8
        assert np.max(np.abs(e_sb - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_schoolbook_subset_pruned():
    rng = np.random.default_rng(3)
    n, k = 30000, 300
    S = rng.uniform(1.0, 1000.0, n)
    p = np.linspace(1.0, 0.1, k)
    targets = rng.choice(n, 40, replace=False)
    plan = icm_amd.Plan(S, k, schoolbook_max=512)
    full = plan.run(p, Q=128)
    plan.set_targets(targets)
    sub = plan.run(p, Q=128)
    plan.close()
    mask = np.zeros(n, dtype=bool)
    mask[targets] = True
    assert np.all(sub[~mask] == 0.0)
    assert np.max(np.abs(sub[mask] - full[mask])) <= 1e-13 * np.max(np.abs(full))


def test_schoolbook_v1_closed_form():
    """k = n linear payout: every level with c <= 512 schoolbook, vs V1."""
    rng = np.random.default_rng(11)
    n = 3000
    S = rng.uniform(1.0, 100.0, n)
    p = np.arange(n, 0, -1, dtype=float)
    plan = icm_amd.Plan(S, n, schoolbook_max=512)
    e = plan.run(p, Q=256)
    plan.close()
    v1 = O.v1_exact(S)
    assert np.max(np.abs(e - v1) / v1) < 1e-9

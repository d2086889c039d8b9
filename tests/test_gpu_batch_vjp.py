"""GPU adjoint of ragged batches (icm_hip_batch_vjp) through the C ABI.

Pins:
  * every instance's (E, dS) against the single-instance adjoint
    (icm_hip_equity_vjp, itself pinned by finite differences of the oracle and
    the analytic V1 gradient), instance by instance;
  * full 64 x 64 Jacobians of a batch of 512 n = 64 instances (one batched
    adjoint over 32768 replicas with unit cotangents) against the analytic
    Jacobian of the V1 closed form (k = n, payout n - r):
    dV1_i/dS_i = sum_{j!=i} S_j/(S_i+S_j)^2, dV1_i/dS_j = -S_i/(S_i+S_j)^2;
  * HIP-graph replays of the adjoint against the eager run, bit for bit.
"""
import sys

import numpy as np
import pytest

import icm_amd
from icm_amd.workloads import ragged_instances

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_batch_vjp_matches_single_instance():
    stacks, payouts = ragged_instances(160, seed=41, n_max=2048)
    rng = np.random.default_rng(2)
    vs = [rng.normal(size=len(s)) for s in stacks]
    plan = icm_amd.BatchPlan(stacks, [len(p) for p in payouts])
    eqs, dss = plan.vjp(payouts, vs, Q=256)
    fwd = plan.run(payouts, Q=256)
    plan.close()
    worst = 0.0
    for i in rng.choice(len(stacks), 40, replace=False):
        e1, d1 = icm_amd.equity_vjp(stacks[i], payouts[i], vs[i], Q=256)
        de = np.max(np.abs(eqs[i] - e1)) / np.max(np.abs(e1))
        dd = np.max(np.abs(dss[i] - d1)) / max(np.max(np.abs(d1)), 1e-300)
        df = np.max(np.abs(eqs[i] - fwd[i])) / np.max(np.abs(fwd[i]))
        worst = max(worst, dd)
        assert de <= 1e-12 and df <= 1e-12, f"instance {i}: equity {de:.2e} / {df:.2e}"
        assert dd <= 1e-10, f"instance {i} n={len(stacks[i])} k={len(payouts[i])}: dS {dd:.2e}"
    say(f"  worst dS difference vs single-instance adjoint: {worst:.2e}")


def test_batch_jacobians_512_x_64_v1():
    rng = np.random.default_rng(512)
    n, count = 64, 512
    stacks = [rng.uniform(1.0, 100.0, n) for _ in range(count)]
    payouts = [np.arange(n, 0, -1, dtype=float)] * count
    say(f"jacobians: {count} instances x {n} replicas")
    jacs, eqs = icm_amd.jacobian_batch(stacks, payouts, Q=256)
    worst = 0.0
    off = ~np.eye(n, dtype=bool)
    for S, J, E in zip(stacks, jacs, eqs):
        Si, Sj = S[:, None], S[None, :]
        den = (Si + Sj) ** 2
        exact = np.where(off, -Si / den, 0.0)
        np.fill_diagonal(exact, np.where(off, Sj / den, 0.0).sum(axis=1))
        v1 = (S[:, None] / (S[:, None] + S[None, :]))[off].reshape(n, n - 1).sum(axis=1) + 1.0
        assert np.max(np.abs(E - v1) / v1) < 1e-9
        worst = max(worst, np.max(np.abs(J - exact)) / np.max(np.abs(exact)))
    say(f"  worst |J - J_V1| / max|J_V1| = {worst:.2e}")
    assert worst < 1e-8


def test_batch_vjp_graph_replay_matches_eager():
    stacks, payouts = ragged_instances(100, seed=43, n_max=1500)
    ks = [len(p) for p in payouts]
    eager = icm_amd.BatchPlan(stacks, ks, graphs=False)
    graph = icm_amd.BatchPlan(stacks, ks, graphs=True)
    rng = np.random.default_rng(4)
    for rep, Q in enumerate([256, 256, 128, 256]):
        vs = [rng.normal(size=len(s)) for s in stacks]
        e1, d1 = eager.vjp(payouts, vs, Q=Q)
        _ = graph.run(payouts, Q=Q)                  # forward graph between adjoint replays
        e2, d2 = graph.vjp(payouts, vs, Q=Q)
        for a, b in zip(d1, d2):
            assert np.array_equal(a, b), f"replay {rep} Q={Q}"
        for a, b in zip(e1, e2):
            assert np.array_equal(a, b)
        say(f"  adjoint replay {rep} Q={Q}: identical")
    eager.close()
    graph.close()

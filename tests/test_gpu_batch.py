"""GPU ragged batches through the C ABI (icm_hip_batch_*).

Pins:
  * every instance of a ragged batch against the oracle's equities (same
    quadrature, per-instance nodes), per player rel. error < 1e-8 and
    |gpu - cpu| <= 1e-11 max|E|;
  * k = 1 instances against the closed form E_i = payout_0 S_i / sum S
    (python/tests/test_icm.py:76-101 semantics);
  * the batch against the single-instance GPU path, instance by instance;
  * HIP-graph replays (new payouts, a changed Q, eager work between replays)
    against the eager run, bit for bit;
  * q-passes and block sizes do not change the result beyond rounding.
"""
import math
import sys

import numpy as np
import pytest

import icm_amd
from icm_amd.workloads import ragged_instances
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(gpu, cpu):
    den = np.abs(cpu).copy()
    den[den < 1e-14] = 1.0
    return float(np.max(np.abs(gpu - cpu) / den))


def test_batch_vs_oracle():
    stacks, payouts = ragged_instances(300, seed=11, n_max=1024)
    say(f"batch of {len(stacks)} instances, {sum(len(s) for s in stacks)} players")
    out = icm_amd.equity_batch(stacks, payouts, Q=256)
    worst = 0.0
    for i, (S, p, e) in enumerate(zip(stacks, payouts, out)):
        ref = O.equity(S, p, Q=256, B=32)
        r = rel_err(e, ref)
        a = float(np.max(np.abs(e - ref)) / np.max(np.abs(ref)))
        worst = max(worst, a)
        assert r < 1e-8 and a <= 1e-11, f"instance {i} n={len(S)} k={len(p)}: rel {r:.2e} abs {a:.2e}"
    say(f"  worst |gpu - cpu| / max|E| = {worst:.2e}")


def test_batch_large_instances_vs_oracle():
    """Instances at the top of the size range (n up to 4096, k = n and k = n/10)."""
    rng = np.random.default_rng(12)
    stacks, payouts = [], []
    for n, k in [(4096, 4096), (4000, 400), (3001, 3001), (2048, 2048), (2049, 205), (1500, 9)]:
        S = rng.dirichlet(np.ones(n)) * 1e7
        p = 0.8 ** np.arange(k)
        stacks.append(S)
        payouts.append(p / p.sum())
    out = icm_amd.equity_batch(stacks, payouts, Q=256)
    for S, p, e in zip(stacks, payouts, out):
        ref = O.equity(S, p, Q=256, B=32)
        a = float(np.max(np.abs(e - ref)) / np.max(np.abs(ref)))
        say(f"  n={len(S)} k={len(p)}: {a:.2e}")
        assert rel_err(e, ref) < 1e-8 and a <= 1e-11


def test_batch_full_config_k1_closed_form_and_single_path():
    """4096 instances of the benchmark batch: k = 1 closed form for every k = 1
    instance, and a sample against the single-instance GPU path."""
    stacks, payouts = ragged_instances(4096, seed=4)
    say(f"full batch: {sum(len(s) for s in stacks)} players")
    plan = icm_amd.BatchPlan(stacks, [len(p) for p in payouts])
    out = plan.run(payouts, Q=256)
    say(f"  plan {plan.summary()}")
    n_k1 = 0
    for S, p, e in zip(stacks, payouts, out):
        assert np.all(np.isfinite(e))
        if len(p) == 1:
            n_k1 += 1
            exact = p[0] * S / S.sum()
            assert np.max(np.abs(e - exact)) <= 1e-10 * p[0], f"k=1 n={len(S)}"
    assert n_k1 > 100
    rng = np.random.default_rng(0)
    for i in rng.choice(len(stacks), 24, replace=False):
        single = icm_amd.equity(stacks[i], payouts[i], Q=256)
        d = float(np.max(np.abs(out[i] - single)) / np.max(np.abs(single)))
        assert d <= 1e-12, f"instance {i} n={len(stacks[i])} k={len(payouts[i])}: {d:.2e}"
    plan.close()


def test_batch_graph_replay_matches_eager():
    stacks, payouts = ragged_instances(200, seed=21, n_max=2048)
    ks = [len(p) for p in payouts]
    eager = icm_amd.BatchPlan(stacks, ks, graphs=False)
    graph = icm_amd.BatchPlan(stacks, ks, graphs=True)
    rng = np.random.default_rng(3)
    for rep, Q in enumerate([256, 256, 64, 256, 64]):
        pays = [np.sort(rng.uniform(0.1, 1.0, k))[::-1] for k in ks]
        e = eager.run(pays, Q=Q)
        _ = icm_amd.equity(stacks[0], pays[0], Q=Q)      # eager work between replays
        g = graph.run(pays, Q=Q)
        for a, b in zip(e, g):
            assert np.array_equal(a, b), f"replay {rep} Q={Q}"
        say(f"  replay {rep} Q={Q}: identical")
    eager.close()
    graph.close()


@pytest.mark.parametrize("q_batch,B", [(48, 32), (0, 16), (0, 64)])
def test_batch_passes_and_block_sizes(q_batch, B):
    stacks, payouts = ragged_instances(120, seed=31, n_max=3000)
    ks = [len(p) for p in payouts]
    base = icm_amd.equity_batch(stacks, payouts, Q=256)
    plan = icm_amd.BatchPlan(stacks, ks, block_size=B, q_batch=q_batch)
    out = plan.run(payouts, Q=256)
    for a, b in zip(base, out):
        assert np.max(np.abs(a - b)) <= 1e-12 * max(np.max(np.abs(a)), 1e-300)
    plan.close()


def test_batch_edge_instances():
    """n = 1 and 2, k = n, one instance, equal stacks, extreme ratios."""
    stacks = [np.array([5.0]), np.array([1.0, 3.0]), np.array([2.0, 2.0, 2.0]),
              np.array([1.0, 1e9, 10.0, 1e-3]), np.full(33, 7.0), np.geomspace(1e-4, 1e5, 65)]
    payouts = [np.array([1.0]), np.array([0.7, 0.3]), np.array([0.5, 0.3, 0.2]),
               np.array([0.6, 0.4]), np.linspace(1, 0.1, 33), np.array([1.0])]
    out = icm_amd.equity_batch(stacks, payouts, Q=256)
    for S, p, e in zip(stacks, payouts, out):
        ref = O.equity(S, p, Q=256, B=32)
        assert rel_err(e, ref) < 1e-8
        assert np.max(np.abs(e - ref)) <= 1e-11 * max(np.max(np.abs(ref)), 1e-300)
    assert out[0][0] == pytest.approx(1.0, abs=1e-12)
    one = icm_amd.equity_batch([stacks[3]], [payouts[3]], Q=256)[0]
    assert np.max(np.abs(one - out[3])) <= 1e-15


def test_batch_with_four_step_instances_vs_oracle():
    """Instances with a level above one LDS transform (k > 2048 and n > 4096)
    run inside the batch through single-instance plans (four-step levels):
    a mixed batch holding an n = 65536, k = n instance and an n = 8192,
    k = 3000 one beside batched instances, every instance against the
    oracle (rel < 1e-8, abs <= 1e-11 max|E|); a second run equals the first;
    the batch adjoint matches each instance's own adjoint."""
    rng = np.random.default_rng(65536)
    sizes = [(300, 20), (65536, 65536), (1000, 1000), (8192, 3000), (17, 5), (4096, 4096)]
    stacks = [rng.lognormal(0.0, 1.0, n) * 100.0 for n, _ in sizes]
    payouts = []
    for _, k in sizes:
        p = 0.8 ** np.arange(k)
        payouts.append(p / p.sum())
    bp = icm_amd.BatchPlan(stacks, [len(p) for p in payouts])
    try:
        info = bp.summary()
        say(f"mixed batch: {info}")
        out = bp.run(payouts, 256)
        again = bp.run(payouts, 256)
        v = [rng.normal(size=len(S)) for S in stacks]
        eq_v, ds_v = bp.vjp(payouts, v, 256)
    finally:
        bp.close()
    for i, (S, p) in enumerate(zip(stacks, payouts)):
        ref = O.equity(S, p, Q=256, B=32)
        r = rel_err(out[i], ref)
        a = float(np.max(np.abs(out[i] - ref)) / np.max(np.abs(ref)))
        say(f"  instance {i} n={len(S)} k={len(p)}: rel {r:.2e}, abs {a:.2e}")
        assert r < 1e-8 and a <= 1e-11
        assert np.array_equal(out[i], again[i])
        assert np.max(np.abs(eq_v[i] - out[i])) <= 1e-13 * np.max(np.abs(out[i]))
        e1, d1 = icm_amd.equity_vjp(S, p, v[i], Q=256)
        d_err = float(np.max(np.abs(ds_v[i] - d1)) / np.max(np.abs(d1)))
        say(f"    adjoint vs single instance: {d_err:.2e} of max|dS|")
        assert d_err < 1e-11


def test_batch_leaf_extract_players_per_lane_bit_identical(monkeypatch):
    """The batch leaf extract with 2 players of a block per lane gives the
    same bits as one player per lane (a fresh plan per setting: a captured
    graph keeps the kernel it was recorded with)."""
    stacks, payouts = ragged_instances(400, seed=21, n_max=3000)
    out = {}
    for ppl in ("1", "2"):
        monkeypatch.setenv("ICM_HIP_LEAF_PPL", ppl)
        bp = icm_amd.BatchPlan(stacks, [len(p) for p in payouts])
        try:
            out[ppl] = np.concatenate(bp.run(payouts, 256))
        finally:
            bp.close()
    say(f"batch leaf extract per-lane players: {len(stacks)} instances")
    assert np.array_equal(out["1"], out["2"])


def test_batch_leaf_build_eight_lanes_vs_one_lane(monkeypatch):
    """The batch leaf build with eight lanes per block (B = 64) against one
    lane per block (ICM_HIP_LEAF_COOP=0): the same recurrence; equal to
    rounding (the two kernels may form their multiply-adds differently),
    and both runs of each kernel deterministic."""
    stacks, payouts = ragged_instances(400, seed=23, n_max=4096)
    out = {}
    for coop in ("0", "1"):
        monkeypatch.setenv("ICM_HIP_LEAF_COOP", coop)
        bp = icm_amd.BatchPlan(stacks, [len(p) for p in payouts])
        try:
            out[coop] = np.concatenate(bp.run(payouts, 256))
            assert np.array_equal(out[coop], np.concatenate(bp.run(payouts, 256)))
        finally:
            bp.close()
    d = float(np.max(np.abs(out["1"] - out["0"])))
    say(f"batch leaf build, eight lanes vs one lane per block: max |diff| {d:.2e}, "
        f"identical: {np.array_equal(out['0'], out['1'])}")
    assert d <= 1e-14 * float(np.max(np.abs(out["0"])))

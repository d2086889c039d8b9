"""Write-before-read audit of the multi-launch path on the GPU.

With icm_hip_plan_set_poison on, every device buffer that a q-batch writes
before it reads (block and level products, g ping-pong, four-step scratch
bufA / bufB, the spectrum caches, the per-chunk partials, and the adjoint's
tangent and B arrays) is filled with NaN bytes before each q-batch.  A kernel
that read a position its own batch had not written -- a value left by the
previous q-batch or an earlier run -- would then put a NaN into the result.
Each case checks that the poisoned run is finite and bit-identical to the
unpoisoned one, and that two poisoned runs agree (no run-to-run variation).

Shapes span several q-batches on four-step levels (n = 20000 and 65536,
q_batch 1 and 2, k = n and k = n/4), with column fusion off and on, the
spectrum cache on and off, pruned subsets and the adjoint; one case runs the
full n = 2^20, k = n pass in one q-batch.
"""
import math

import numpy as np
import pytest

import icm_amd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _inputs(n, k, seed):
    rng = np.random.default_rng(seed)
    S = np.exp(rng.uniform(0.0, math.log(1e6), n))
    r = np.arange(k, dtype=float)
    p = (k - r) ** 2
    p /= p.sum()
    return rng, S, p


def _same(a, b, what):
    assert np.all(np.isfinite(a)), f"{what}: NaN or inf in the poisoned run"
    assert np.array_equal(a, b), f"{what}: poisoned run differs, max |diff| {np.max(np.abs(a - b)):.3e}"


def _poisoned(plan, Q, q_batch):
    """The fills ran: at least the block products of every q-batch."""
    nbatch = -(-Q // q_batch)
    blocks = 8 * q_batch * plan.summary()["B"] * (plan.n // plan.summary()["B"])
    say(f"  poisoned {plan.poisoned_bytes / 1e9:.3f} GB over {nbatch} q-batches")
    assert plan.poisoned_bytes >= nbatch * blocks


@pytest.mark.parametrize("fusion", [0, 15])
@pytest.mark.parametrize("n,k,q_batch,Q", [(20000, 20000, 1, 3), (20000, 20000, 2, 5), (65536, 65536, 1, 3),
                                           (65536, 65536, 2, 5), (65536, 16384, 2, 5)])
def test_forward_reads_only_what_its_batch_wrote(n, k, q_batch, Q, fusion):
    rng, S, p = _inputs(n, k, n + k + q_batch + fusion)
    plan = icm_amd.Plan(S, k, q_batch=q_batch)
    try:
        plan.set_column_fusion(fusion)
        say(f"poison n={n} k={k} q_batch={q_batch} Q={Q} fusion={fusion} cache={plan.spectrum_cache}")
        ref = plan.run(p, Q=Q)
        assert plan.poisoned_bytes == 0
        plan.set_poison(True)
        got = plan.run(p, Q=Q)
        _poisoned(plan, Q, q_batch)
        _same(got, ref, "full run")
        _same(plan.run(p, Q=Q), got, "second poisoned run")
        targets = rng.choice(n, 64, replace=False)
        plan.set_targets(targets)
        sub = plan.run(p, Q=Q)
        plan.set_poison(False)
        sub_ref = plan.run(p, Q=Q)
        _same(sub, sub_ref, "pruned subset")
        plan.set_targets(None)
        plan.set_spectrum_cache(False)
        fresh = plan.run(p, Q=Q)
        plan.set_poison(True)
        _same(plan.run(p, Q=Q), fresh, "spectrum cache off")
    finally:
        plan.close()


@pytest.mark.parametrize("ckpt", [0, 3])
@pytest.mark.parametrize("n,q_batch,Q", [(20000, 2, 5), (65536, 1, 3)])
def test_adjoint_reads_only_what_its_batch_wrote(n, q_batch, Q, ckpt):
    rng, S, p = _inputs(n, n, n + 77)
    v = rng.normal(size=n)
    plan = icm_amd.Plan(S, n, q_batch=q_batch)
    try:
        plan.set_adjoint_checkpoint(ckpt)
        say(f"poison adjoint n={n} q_batch={q_batch} Q={Q} checkpoint interval {ckpt} cache={plan.spectrum_cache}")
        e_ref, ds_ref = plan.vjp(p, v, Q=Q)
        plan.set_poison(True)
        e, ds = plan.vjp(p, v, Q=Q)
        _poisoned(plan, Q, q_batch)
        _same(e, e_ref, "adjoint equity")
        _same(ds, ds_ref, "adjoint dS")
        e2, ds2 = plan.vjp(p, v, Q=Q)
        _same(ds2, ds, "second poisoned adjoint")
    finally:
        plan.close()


def test_full_size_pass_reads_only_what_it_wrote():
    n = 1 << 20
    _, S, p = _inputs(n, n, 20)
    plan = icm_amd.Plan(S, n)
    try:
        say(f"poison n={n} k=n Q=256 q_batch={plan.summary()['q_batch']}")
        ref = plan.run(p, Q=256)
        plan.set_poison(True)
        got = plan.run(p, Q=256)
        _poisoned(plan, 256, plan.summary()["q_batch"])
        _same(got, ref, "n = 2^20 pass")
    finally:
        plan.close()

"""GPU: the one-launch small-instance path (one workgroup per quadrature node,
whole tree in LDS; the counterpart of the reference's single-kernel path,
src/gpu/gpu_kernels.cu:1547-1668) against the oracle and against the
multi-launch path of the same plan, including config 1 (n = 1024, k = 10,
stacks uniform integers in [1, 1e6], payout 1/(j+1) normalised), more
quadrature nodes than workgroups, edge sizes and subsets."""
import numpy as np
import pytest

import icm_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def rel_err(gpu, cpu):
    den = np.abs(cpu).copy()
    den[den < 1e-14] = 1.0
    return float(np.max(np.abs(gpu - cpu) / den))


def config1():
    rng = np.random.default_rng(1)
    S = rng.integers(1, 10 ** 6, 1024, endpoint=True).astype(np.float64)
    p = 1.0 / np.arange(1, 11)
    return S, p / p.sum()


def test_config1_takes_the_small_path_and_matches_the_oracle():
    S, p = config1()
    pl = icm_amd.Plan(S, len(p))
    try:
        assert pl.small_path, "config 1 should run as one launch"
        e_small = pl.run(p, 256)
        pl.set_small_path(0)
        assert not pl.small_path
        e_multi = pl.run(p, 256)
    finally:
        pl.close()
    ref = O.equity(S, p, Q=256, fft_min=0)
    e1, e2 = rel_err(e_small, ref), rel_err(e_small, e_multi)
    say(f"config1: small path vs oracle {e1:.2e}, vs multi-launch {e2:.2e}")
    assert e1 < 1e-8 and e2 < 1e-12


@pytest.mark.parametrize("n,k", [(1, 1), (2, 1), (2, 2), (3, 2), (5, 5), (10, 4), (33, 7), (257, 257),
                                 (1000, 1000), (1500, 3), (4096, 40), (6000, 2)])
def test_small_path_shapes(n, k):
    rng = np.random.default_rng(n * 7 + k)
    S = np.exp(rng.uniform(0.0, np.log(1e6), n))
    p = np.sort(rng.uniform(0.1, 1.0, k))[::-1]
    pl = icm_amd.Plan(S, k)
    try:
        pl.set_small_path(1)
        if not pl.small_path:
            pytest.skip(f"n={n} k={k} does not fit one workgroup's LDS")
        outs = {Q: pl.run(p, Q) for Q in (1, 2, 17, 256)}
        pl.set_small_path(0)
        multi = pl.run(p, 256)
    finally:
        pl.close()
    worst = 0.0
    for Q, e in outs.items():
        ref = O.equity(S, p, Q=Q, fft_min=0 if n <= 2000 else 64)
        worst = max(worst, rel_err(e, ref))
    say(f"n={n} k={k}: worst rel err vs oracle over Q in (1, 2, 17, 256) {worst:.2e}")
    assert worst < 1e-8
    assert rel_err(outs[256], multi) < 1e-12


def test_more_nodes_than_workgroups_and_node_cache():
    S, p = config1()
    pl = icm_amd.Plan(S, len(p))
    try:
        assert pl.small_path
        a = pl.run(p, 2000)          # workgroups loop over several nodes each
        b = pl.run(p, 256)           # nodes uploaded again for a new Q
        c = pl.run(p[::-1].copy(), 256)
        d = pl.run(p, 2000)
    finally:
        pl.close()
    assert np.array_equal(a, d)
    assert rel_err(a, O.equity(S, p, Q=2000, fft_min=0)) < 1e-8
    assert rel_err(b, O.equity(S, p, Q=256, fft_min=0)) < 1e-8
    assert rel_err(c, O.equity(S, p[::-1].copy(), Q=256, fft_min=0)) < 1e-8


def test_small_path_subset():
    S, p = config1()
    t = np.array([0, 77, 512, 1023], dtype=np.int32)
    pl = icm_amd.Plan(S, len(p))
    try:
        pl.set_targets(t)
        assert pl.small_path
        e = pl.run(p, 256)
    finally:
        pl.close()
    ref = O.equity(S, p, Q=256, fft_min=0)
    mask = np.zeros(len(S), dtype=bool)
    mask[t] = True
    assert rel_err(e[mask], ref[mask]) < 1e-8 and np.all(e[~mask] == 0.0)


def test_small_path_adjoint_unchanged():
    """The adjoint keeps the multi-launch kernels: equities from the VJP pass
    equal the small path's forward."""
    S, p = config1()
    v = np.random.default_rng(3).normal(size=len(S))
    pl = icm_amd.Plan(S, len(p))
    try:
        e = pl.run(p, 256)
        eq, ds = pl.vjp(p, v, 256)
    finally:
        pl.close()
    assert rel_err(eq, e) < 1e-12 and np.all(np.isfinite(ds))

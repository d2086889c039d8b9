"""GPU adjoint: dS = d(v . E)/dS through the C ABI.

Pins:
  * per node, H_j = sum_i c_i dPsi_i/da_j against central differences of the
    oracle's Phi(a) = sum_i c_i Psi_i(a): Psi_i is multilinear in a, so Phi is
    linear in each a_j and the central difference is exact up to rounding;
  * the analytic gradient of the V1 closed form (k = n, payout n - r):
    dV1_i/dS_i = sum_{j!=i} S_j/(S_i+S_j)^2, dV1_i/dS_j = -S_i/(S_i+S_j)^2;
  * central finite differences (h = 1e-6 S_i) of the oracle's FP64 equities;
  * the equities returned by the adjoint pass equal the forward pass.
Quadrature nodes are held fixed (they depend on max S through y_hi), so the
finite-difference coordinates exclude the largest stack.
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


def conditioned_a(n, k, seed):
    """a_j with about k/2 expected 'x' factors, so Psi is O(payout) for any k
    (with k << n and a ~ U(0,1) every Psi is ~1e-50 and differences are noise)."""
    rng = np.random.default_rng(seed)
    a = 1.0 - (k / (2.0 * n)) * rng.uniform(0.0, 2.0, n)
    return np.clip(a, 0.02, 0.999), rng


@pytest.mark.parametrize("n,k", [(10, 4), (40, 13), (300, 300), (2000, 150), (5000, 5000), (20000, 20000)])
def test_node_adjoint_vs_exact_differences(n, k):
    a, rng = conditioned_a(n, k, n + k)
    c = rng.normal(size=n)
    p = np.sort(rng.uniform(0, 1, k))[::-1]
    say(f"node vjp n={n} k={k}")
    psi, h = icm_amd.node_vjp(a, c, p)
    fwd = icm_amd.node_inner(a, p)
    assert np.max(np.abs(psi - fwd)) <= 1e-13 * p.max()
    eng = O.ENGINE_NAIVE if n <= 2000 else O.ENGINE_HYBRID

    def phi(aa):
        return float(np.dot(c, O.inner(aa, p, engine=eng, B=32, fft_min=64)))

    coords = range(n) if n <= 40 else rng.choice(n, 12, replace=False)
    scale = 0.0
    errs = []
    for j in coords:
        hh = 0.5 * min(a[j], 1.0 - a[j], 0.1)
        if hh < 1e-6:
            continue
        ap = a.copy(); ap[j] += hh
        am = a.copy(); am[j] -= hh
        fd = (phi(ap) - phi(am)) / (2 * hh)
        scale = max(scale, abs(fd))
        errs.append(abs(h[j] - fd))
    err = max(errs) / max(scale, 1e-300)
    say(f"  max |H - FD| / max|FD| = {err:.2e} over {len(errs)} coordinates")
    assert err < 1e-8


@pytest.mark.parametrize("cache", ["1", "3"])
def test_node_adjoint_saturated_top_levels(cache, monkeypatch):
    """a_j ~ 1e-5: the top child coefficients (products of 1 - a_j over
    thousands of players) are O(1), so the wrapped-term fixes and the
    missing-right-child parents of the four-step levels carry weight.
    Exact central differences of the oracle; spectrum cache on (memory
    strategy 1) and off (3)."""
    monkeypatch.setenv("ICM_HIP_MEMORY_STRATEGY", cache)
    n = k = 20000
    rng = np.random.default_rng(20000)
    a = 1e-5 * rng.uniform(0.2, 2.0, n)
    a[rng.choice(n, 40, replace=False)] = rng.uniform(0.05, 0.95, 40)
    c = rng.normal(size=n)
    p = np.sort(rng.uniform(0, 1, k))[::-1]
    say(f"node vjp saturated n={n} memory_strategy={cache}")
    psi, h = icm_amd.node_vjp(a, c, p)
    ref_psi = O.inner(a, p, engine=O.ENGINE_HYBRID, B=32, fft_min=64)
    e_psi = float(np.max(np.abs(psi - ref_psi))) / p.max()

    def phi(aa):
        return float(np.dot(c, O.inner(aa, p, engine=O.ENGINE_HYBRID, B=32, fft_min=64)))

    # Phi is linear in each a_j, so a secant over a wide interval is exact
    # (a step of order a_j ~ 1e-5 would leave only rounding noise)
    errs, scale = [], 0.0
    for j in rng.choice(n, 12, replace=False):
        ap = a.copy(); ap[j] = 0.75
        am = a.copy(); am[j] = 0.25
        fd = (phi(ap) - phi(am)) / 0.5
        scale = max(scale, abs(fd))
        errs.append(abs(h[j] - fd))
    err = max(errs) / max(scale, 1e-300)
    say(f"  Psi vs oracle {e_psi:.2e}; max |H - FD| / max|FD| = {err:.2e}")
    assert e_psi < 1e-12 and err < 1e-8


@pytest.mark.parametrize("n", [200, 2000])
def test_vjp_v1_analytic_gradient(n):
    rng = np.random.default_rng(n)
    S = rng.uniform(1.0, 100.0, n)
    p = np.arange(n, 0, -1, dtype=float)
    v = rng.normal(size=n)
    eq, ds = icm_amd.equity_vjp(S, p, v, Q=256)
    Si, Sj = S[:, None], S[None, :]
    den = (Si + Sj) ** 2
    off = ~np.eye(n, dtype=bool)
    dii = np.where(off, Sj / den, 0.0).sum(axis=1)          # dV1_i/dS_i
    dij = np.where(off, -Si / den, 0.0)                      # dV1_i/dS_j
    exact = v * dii + (v[:, None] * dij).sum(axis=0)
    v1 = O.v1_exact(S)
    e_eq = np.max(np.abs(eq - v1) / v1)
    e_ds = np.max(np.abs(ds - exact)) / np.max(np.abs(exact))
    say(f"V1 n={n}: equity {e_eq:.2e}, gradient {e_ds:.2e}")
    assert e_eq < 1e-9 and e_ds < 1e-8


@pytest.mark.parametrize("n,k", [(300, 50), (1500, 1500)])
def test_vjp_vs_oracle_finite_differences(n, k):
    rng = np.random.default_rng(7 * n + k)
    S = np.exp(rng.uniform(0.0, math.log(1e4), n))
    p = 0.85 ** np.arange(k)
    p /= p.sum()
    v = rng.normal(size=n)
    eq, ds = icm_amd.equity_vjp(S, p, v, Q=256)
    fwd = icm_amd.equity(S, p, Q=256)
    assert np.max(np.abs(eq - fwd)) <= 1e-12 * np.max(np.abs(fwd))
    imax = int(np.argmax(S))
    coords = [i for i in rng.choice(n, 17, replace=False) if i != imax][:16]
    scale = np.max(np.abs(ds))

    def fd(i, rel):
        h = rel * S[i]
        Sp = S.copy(); Sp[i] += h
        Sm = S.copy(); Sm[i] -= h
        # schoolbook products (fft_min=0): the least noisy oracle engine
        return (np.dot(v, O.equity(Sp, p, Q=256, fft_min=0)) - np.dot(v, O.equity(Sm, p, Q=256, fft_min=0))) / (2 * h)

    # The difference quotient's rounding noise scales like 1/h (about 1e-11
    # absolute at h = 1e-6 S_i here, ~1e-5 x max|dS| at n = 1500), hence 3e-5 at h = 1e-6 S_i;
    # h = 1e-5 S_i has 10x less noise and O(h^2) ~ 1e-10 bias.
    e6 = max(abs(ds[i] - fd(i, 1e-6)) for i in coords) / scale
    e5 = max(abs(ds[i] - fd(i, 1e-5)) for i in coords) / scale
    say(f"FD n={n} k={k}: max |dS - FD| / max|dS| = {e6:.2e} (h=1e-6 S), {e5:.2e} (h=1e-5 S)")
    assert e6 < 3e-5 and e5 < 1e-6


def test_vjp_large_shape_self_consistency():
    """n = 131072, k = 1000, lognormal(sigma 1.5) stacks: adjoint vs differences of the GPU forward."""
    rng = np.random.default_rng(5)
    n, k = 131072, 1000
    S = rng.lognormal(0.0, 1.5, n)
    p = np.arange(k, 0, -1, dtype=float)
    p /= p.sum()
    v = rng.normal(size=n)
    plan = icm_amd.Plan(S, k)
    eq, ds = plan.vjp(p, v, Q=256)
    fwd = plan.run(p, Q=256)
    plan.close()
    assert np.max(np.abs(eq - fwd)) <= 1e-12 * np.max(np.abs(fwd))
    imax = int(np.argmax(S))
    coords = [i for i in rng.choice(n, 5, replace=False) if i != imax][:4]
    scale = np.max(np.abs(ds))
    for i in coords:
        h = 1e-5 * S[i]
        Sp = S.copy(); Sp[i] += h
        Sm = S.copy(); Sm[i] -= h
        fd = (np.dot(v, icm_amd.equity(Sp, p, Q=256)) - np.dot(v, icm_amd.equity(Sm, p, Q=256))) / (2 * h)
        say(f"large n={n} coord {i}: dS={ds[i]:.6e} FD={fd:.6e}")
        assert abs(ds[i] - fd) <= 1e-5 * scale + 1e-6 * abs(fd)


@pytest.mark.parametrize("n,k,q_batch,Q", [(65536, 65536, 0, 64), (20000, 20000, 1, 3)])
def test_vjp_spectrum_cache_matches_recompute(n, k, q_batch, Q):
    """The adjoint reads the forward products' cached child transforms in its
    A and B propagates and forms the tangent products dP_X = dP_L P_R + P_L dP_R
    in one spectral pass from the cached P and dP transforms; without the
    cache it recomputes the transforms and builds the two tangent terms in
    separate passes.  Both agree to rounding (the tangent terms are summed
    before instead of after the inverse transform; the fused LDS levels
    transform A + iB together, so B's rounding reaches A and the equities),
    within one q-batch and across several."""
    rng = np.random.default_rng(n + Q)
    S = rng.uniform(1.0, 100.0, n)
    r = np.arange(k, dtype=float)
    p = (k - r) ** 2
    p /= p.sum()
    v = rng.normal(size=n)
    plan = icm_amd.Plan(S, k, q_batch=q_batch)
    say(f"vjp spectrum cache n={n} k={k} q_batch={q_batch} Q={Q} on={plan.spectrum_cache}")
    assert plan.spectrum_cache
    e1, d1 = plan.vjp(p, v, Q=Q)
    plan.set_spectrum_cache(False)
    e0, d0 = plan.vjp(p, v, Q=Q)
    plan.close()
    erre = np.max(np.abs(e1 - e0)) / np.max(np.abs(e0))
    say(f"  E cache vs recompute: {erre:.2e} of max|E|")
    assert erre < 1e-12
    scale = np.max(np.abs(d0))
    err = np.max(np.abs(d1 - d0)) / scale
    say(f"  dS cache vs recompute: {err:.2e} of max|dS|")
    assert err < 1e-11


@pytest.mark.parametrize("n,k,Q", [(262144, 262144, 16), (65536, 65536, 32), (300001, 300001, 8)])
def test_vjp_column_fusion_matches_unfused(n, k, Q):
    """The adjoint's column fusions (up-sweep P and dP builds handing the level
    above its pair transforms, down-sweep B then A propagates handing the
    level below its g-pair transforms, with the end coefficients the wrapped
    and exact terms read) against the same plan with every fusion off and the
    spectrum caches on: at k = n the wrapped terms apply at every four-step
    level; n = 2^18 has three fused level pairs each way, n = 300001 odd node
    counts (19, 5, 3: parents without a right child) at fused levels.  The fused kernels
    compute the same sums as the separate column kernels, except the F = 8192
    build's hand-over (a different rounding, as in the forward): dS within
    1e-13 of max|dS|, equities within 1e-13 of max|E|."""
    rng = np.random.default_rng(n + 5)
    S = np.exp(rng.uniform(0.0, math.log(1e6), n))
    r = np.arange(k, dtype=float)
    p = (k - r) ** 2
    p /= p.sum()
    v = rng.normal(size=n)
    plan = icm_amd.Plan(S, k)
    try:
        assert plan.spectrum_cache
        e1, d1 = plan.vjp(p, v, Q=Q)
        plan.set_column_fusion(False)
        e0, d0 = plan.vjp(p, v, Q=Q)
    finally:
        plan.close()
    de = float(np.max(np.abs(e1 - e0)) / np.max(np.abs(e0)))
    dd = float(np.max(np.abs(d1 - d0)) / np.max(np.abs(d0)))
    say(f"vjp column fusion n={n} k={k} Q={Q}: E {de:.1e}, dS {dd:.1e} of max "
        f"(identical: E {np.array_equal(e1, e0)}, dS {np.array_equal(d1, d0)})")
    assert np.all(np.isfinite(d1))
    assert de <= 1e-13 and dd <= 1e-13


@pytest.mark.parametrize("n,k,q_batch,Q", [(65536, 65536, 0, 16), (20000, 20000, 2, 5), (131072, 1000, 0, 16)])
def test_vjp_checkpointing_matches_keeping_every_level(n, k, q_batch, Q):
    """Checkpointing keeps the tangent products of every m-th tree level and
    rebuilds a segment from its lower checkpoint when the reverse sweep needs
    its coefficients (levels that read the tangent pairs' cached spectra need
    none).  The equities are identical; dS agrees with the default adjoint
    that keeps every level to rounding (<= 1e-14 max|dS|: most coordinates
    are identical, a few differ in the last bits -- measured up to 3e-17 of
    max|dS|); the plan without spectrum caches agrees to 1e-11."""
    rng = np.random.default_rng(n + k + 11)
    S = np.exp(rng.uniform(0.0, math.log(1e4), n))
    r = np.arange(k, dtype=float)
    p = (k - r) ** 2
    p /= p.sum()
    v = rng.normal(size=n)
    plan = icm_amd.Plan(S, k, q_batch=q_batch)
    L = len(icm_amd.plan_geometry(n, k, plan.summary()["B"])["F"])
    say(f"vjp checkpoint n={n} k={k} q_batch={q_batch} Q={Q} levels={L}")
    e_def, d_def = plan.vjp(p, v, Q=Q)
    assert plan.adjoint_layout[0] == 0
    runs = {}
    for m in (2, 3, 4, L):
        plan.set_adjoint_checkpoint(m)
        runs[m] = plan.vjp(p, v, Q=Q)
        say(f"  interval {m}: layout {plan.adjoint_layout}")
        assert plan.adjoint_layout[0] == m
    plan.set_adjoint_checkpoint(0)
    plan.set_spectrum_cache(False)
    e0, d0 = plan.vjp(p, v, Q=Q)
    plan.close()
    scale = float(np.max(np.abs(d_def)))
    for m, (e, d) in runs.items():
        diff = float(np.max(np.abs(d - d_def)))
        say(f"  interval {m}: max |dS - dS(every level)| = {diff / scale:.1e} of max|dS|, "
            f"{int(np.sum(d != d_def))} of {len(d)} coordinates differ")
        assert np.array_equal(e, e_def) and diff <= 1e-14 * scale, f"interval {m}: max |dS diff| {diff:.3e}"
    err = np.max(np.abs(d0 - d_def)) / np.max(np.abs(d_def))
    say(f"  without spectrum caches vs the cached adjoint: {err:.2e} of max|dS|")
    assert err < 1e-11 and np.max(np.abs(e0 - e_def)) <= 1e-12 * np.max(np.abs(e_def))


def test_vjp_checkpointing_chosen_under_a_memory_budget():
    """With a device budget that leaves fewer than 32 q-planes per adjoint
    batch when every tangent level is kept, the auto mode checkpoints and
    gets more planes; the results match the unconstrained plan to rounding."""
    n = k = 65536
    rng = np.random.default_rng(99)
    S = rng.uniform(1.0, 100.0, n)
    p = np.arange(k, 0, -1, dtype=float)
    p /= p.sum()
    v = rng.normal(size=n)
    ref = icm_amd.Plan(S, k, q_batch=64)
    e_ref, d_ref = ref.vjp(p, v, Q=64)
    ref.close()
    # the forward's 64 planes fit; the 150 MB left hold about 23 adjoint
    # planes with every tangent level (6.3 MB each) and about 35 with
    # checkpoints (4.2 MB each)
    fwd, _ = icm_amd.plan_dry_run(n, k, 10 ** 12, q_batch=64)
    budget = fwd["planned_peak_vram_bytes"] + 150_000_000
    plan = icm_amd.Plan(S, k, q_batch=64, mem_budget_bytes=budget, mem_fraction=1.0)
    try:
        e, d = plan.vjp(p, v, Q=64)
        m, qv = plan.adjoint_layout
        say(f"budget {budget / 1e9:.2f} GB: forward q_batch {plan.summary()['q_batch']}, adjoint interval {m}, "
            f"{qv} planes")
        assert m >= 2
    finally:
        plan.close()
    assert np.max(np.abs(d - d_ref)) <= 1e-11 * np.max(np.abs(d_ref))
    assert np.max(np.abs(e - e_ref)) <= 1e-12 * np.max(np.abs(e_ref))


@pytest.mark.parametrize("n,k", [(20000, 20000), (65536, 1000), (5000, 37), (4133, 64)])
def test_adjoint_leaves_players_per_lane_bit_identical(n, k, monkeypatch):
    """The B = 64 adjoint leaves with 2 players of a block per lane (staged
    differences of A and of B's adjoint, double-buffered planes) give the same
    bits as one player per lane: equities and dS, and per node Psi and H
    with a covering all three division branches."""
    rng = np.random.default_rng(n + 3 * k)
    S = np.exp(rng.uniform(0.0, math.log(1e6), n))
    p = np.arange(k, 0, -1, dtype=float) ** 2
    p /= p.sum()
    v = rng.normal(size=n)
    a = rng.uniform(0.0, 1.0, n)
    a[::5] = 1e-17
    c = rng.normal(size=n)
    out = {}
    for ppl in ("1", "2"):
        monkeypatch.setenv("ICM_HIP_LEAF_PPL", ppl)
        plan = icm_amd.Plan(S, k, block_size=64)
        e, d = plan.vjp(p, v, Q=32)
        plan.close()
        out[ppl] = (e, d) + tuple(icm_amd.node_vjp(a, c, p))
    say(f"adjoint leaves per-lane players n={n} k={k}")
    for i, what in enumerate(("E", "dS", "Psi", "H")):
        assert np.array_equal(out["1"][i], out["2"][i]), \
            f"{what}: max |diff| {np.max(np.abs(out['1'][i] - out['2'][i])):.3e}"

"""GPU parity at the workloads' full sizes.

  * large forward (n = 2^20, stacks log-uniform on [1, 1e9], Q = 256):
    - k = n, quadratic payout (k - j)^2 against the exact equities
      (2 V2 + 3 V1 - 2) / sum m^2 of 64 sampled players (V1 direct, V2 in
      its integral form; tests/closed_forms.py), relative error < 1e-9;
    - k = n/4, linear and quadratic payouts: the integrated equities against
      the CPU oracle at 1024 sampled players (< 1e-8 relative, denominator 1
      below 1e-14 as bench/bench_gpu.cu:18-23) and two checksums of the
      whole vector (sum E and sum w E for a seeded normal w);
    - k = n/4, both payouts: Psi at 6 quadrature nodes spanning the rule
      against the oracle (< 1e-12 x max payout);
    the oracle values come from tests/golden/gen_config3.py (config3.npz);
  * ragged batch of 4096 instances (icm_amd.workloads.ragged_instances,
    seed 4): every instance against the oracle run here;
  * adjoint at n = 131072, k = 1000 (stacks lognormal(0, 1.5), v ~ N(0, 1)):
    dS against central differences of the oracle's v . E at h = 1e-6 S_i on
    256 coordinates and h = 1e-5 S_i on 16 (tests/golden/gen_config5_fd.py,
    config5_fd.npz), each coordinate within 2e-7 max|dS| plus the difference
    quotient's rounding noise eps sum|v E| / h_i;
  * the analytic gradient of the quadratic closed form at n = 64 and 96.
"""
import math
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import icm_amd
from icm_amd.workloads import loguniform_field, ragged_instances
from oracle import oracle as O

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import closed_forms as C  # noqa: E402

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)


def rel_err(gpu, cpu):
    den = np.abs(cpu).copy()
    den[den < 1e-14] = 1.0
    return float(np.max(np.abs(gpu - cpu) / den))


def _fixture(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _check_inputs(f, S):
    assert abs(S.sum() - float(f["S_sum"])) <= 1e-12 * S.sum()
    assert np.array_equal(S[:4], f["S_head"]), "workload generator differs from the one the fixture used"


# ------------------------------------------------------------ large forward
def test_config3_quadratic_closed_form_k_eq_n():
    f = _fixture("config3.npz")
    n = int(f["n"])
    S, p = loguniform_field(n, n, int(f["seed"]), "quadratic")
    _check_inputs(f, S)
    idx = f["idx64"]
    norm = n * (n + 1.0) * (2.0 * n + 1.0) / 6.0
    exact = (2.0 * f["V2"] + 3.0 * f["V1"] - 2.0) / norm
    # the integral rule reproduces V1 - 1 (pins the shared t grid)
    assert np.max(np.abs(f["V1m_int"] + 1.0 - f["V1"]) / f["V1"]) < 1e-12
    say(f"quadratic k=n n={n}")
    eq = icm_amd.equity(S, p, Q=int(f["Q"]))
    err = float(np.max(np.abs(eq[idx] - exact) / exact))
    say(f"  max rel err vs 2 V2 + 3 V1 - 2 over {len(idx)} players: {err:.2e}; sum rule {abs(eq.sum() - 1):.1e}")
    assert err < 1e-9
    assert abs(eq.sum() - 1.0) < 1e-9


@pytest.mark.parametrize("payout", ["linear", "quadratic"])
def test_config3_quarter_k_vs_oracle(payout):
    f = _fixture("config3.npz")
    n = int(f["n"])
    k = n // 4
    S, p = loguniform_field(n, k, int(f["seed"]), payout)
    _check_inputs(f, S)
    say(f"k=n/4 {payout} n={n}")
    eq = icm_amd.equity(S, p, Q=int(f["Q"]))
    idx = f["idx"]
    err = rel_err(eq[idx], f[f"E_{payout}"])
    w = np.random.default_rng(34).normal(size=n)
    scale = float(np.dot(np.abs(w), np.abs(eq)))
    d_sum = abs(eq.sum() - float(f[f"sumE_{payout}"]))
    d_w = abs(float(np.dot(w, eq)) - float(f[f"wE_{payout}"])) / scale
    say(f"  sampled rel err {err:.2e}; |sum E diff| {d_sum:.1e}; |w.E diff|/|w||E| {d_w:.1e}")
    assert err < 1e-8 and d_sum < 1e-11 and d_w < 1e-11


@pytest.mark.parametrize("payout", ["linear", "quadratic"])
def test_config3_quarter_k_nodes_vs_oracle(payout):
    f = _fixture("config3.npz")
    n = int(f["n"])
    k = n // 4
    S, p = loguniform_field(n, k, int(f["seed"]), payout)
    order = np.argsort(-S, kind="stable")
    inv = np.empty(n, dtype=np.int64)
    inv[order] = np.arange(n)
    logv, _ = O.make_nodes(int(f["Q"]), S.max())
    idx = f["idx"]
    for r, q in enumerate(f["nodes"]):
        psi = icm_amd.node_inner(np.exp(S[order] * logv[q]), p)
        err = float(np.max(np.abs(psi[inv[idx]] - f[f"psi_{payout}"][r]))) / p.max()
        say(f"  node q={q}: max |Psi - oracle| / max payout = {err:.2e}")
        assert err < 1e-12


# ------------------------------------------------------------ ragged batch
def test_config4_every_instance_vs_oracle():
    stacks, pays = ragged_instances(4096, seed=4)
    bp = icm_amd.BatchPlan(stacks, [len(p) for p in pays])
    try:
        out = bp.run(pays, 256)
    finally:
        bp.close()
    say(f"batch: {len(stacks)} instances, {sum(len(s) for s in stacks)} players; oracle per instance")

    def ref(i):
        return O.equity(stacks[i], pays[i], Q=256, fft_min=64, nthreads=1)

    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        refs = list(ex.map(ref, range(len(stacks))))
    worst_rel, worst_abs, bad = 0.0, 0.0, []
    for i, (e, c) in enumerate(zip(out, refs)):
        r = rel_err(e, c)
        a = float(np.max(np.abs(e - c)) / np.max(np.abs(c)))
        worst_rel, worst_abs = max(worst_rel, r), max(worst_abs, a)
        if not (r < 1e-8 and a <= 1e-11):
            bad.append((i, len(e), len(pays[i]), r, a))
    say(f"  worst rel {worst_rel:.2e}, worst |gpu - cpu| / max|E| {worst_abs:.2e}, failing {len(bad)}")
    assert not bad, f"instances failing: {bad[:5]}"


# ------------------------------------------------------------ adjoint
def test_config5_vjp_vs_oracle_finite_differences():
    f = _fixture("config5_fd.npz")
    n, k, Q = int(f["n"]), int(f["k"]), int(f["Q"])
    rng = np.random.default_rng(int(f["seed"]))
    S = rng.lognormal(0.0, 1.5, n)
    v = rng.normal(size=n)
    p = np.arange(k, 0, -1, dtype=np.float64)
    p /= p.sum()
    _check_inputs(f, S)
    eq, ds = icm_amd.equity_vjp(S, p, v, Q=Q)
    assert abs(float(np.dot(v, eq)) - float(f["vE"])) <= 1e-11 * float(np.dot(np.abs(v), eq))
    c = f["coords"]
    scale = float(np.max(np.abs(ds)))
    # Per coordinate: 2e-7 max|dS| (the adjoint's accuracy and the O(h^2)
    # truncation) plus the rounding noise of a central difference of the
    # oracle's v . E, eps sum|v E| / h_i (h_i = rel S_i: small stacks have
    # tiny steps; observed noise at most 0.1 of this term).
    noise = np.finfo(np.float64).eps * float(np.sum(np.abs(v * eq)))
    worst = 0.0
    for rel, fd, cc in ((1e-6, f["fd6"], c), (1e-5, f["fd5"], c[:len(f["fd5"])])):
        err = np.abs(ds[cc] - fd)
        bound = 2e-7 * scale + noise / (rel * S[cc])
        worst = max(worst, float(np.max(err / bound)))
        say(f"config5 adjoint, h={rel:g} S on {len(cc)} coords: max |dS - FD| / max|dS| = "
            f"{float(np.max(err)) / scale:.2e}, max error / bound = {float(np.max(err / bound)):.3f}")
    assert worst <= 1.0


def _quadratic_gradient(S, v):
    """d(v . E)/dS for E_i = (2 V2_i + 3 V1_i - 2) / sum m^2 (exact)."""
    n = len(S)
    Si, Sj = S[:, None], S[None, :]
    off = ~np.eye(n, dtype=bool)
    d1 = (Si + Sj) ** 2
    dV1_ii = np.where(off, Sj / d1, 0.0).sum(axis=1)
    dV1_ij = np.where(off, -Si / d1, 0.0)                          # [i, m], m != i
    T = S[:, None, None] + S[None, :, None] + S[None, None, :]    # [i, j, l]
    ii = np.arange(n)
    valid = (ii[None, :, None] != ii[:, None, None]) & (ii[None, None, :] != ii[:, None, None]) \
        & (ii[None, :, None] != ii[None, None, :])
    T2 = np.where(valid, 1.0 / (T * T), 0.0)
    # dV2_i/dS_i = sum_{j<l} (S_j + S_l) / T^2 = (1/2) sum_{j != l} (T - S_i) / T^2
    dV2_ii = 0.5 * np.sum(T2 * (T - S[:, None, None]), axis=(1, 2))
    # dV2_i/dS_m = - sum_{l != i, m} S_i / T_{iml}^2
    dV2_im = -S[:, None] * np.sum(T2, axis=2)
    norm = n * (n + 1.0) * (2.0 * n + 1.0) / 6.0
    J = (2.0 * dV2_im + 3.0 * dV1_ij) / norm
    J[ii, ii] = (2.0 * dV2_ii + 3.0 * dV1_ii) / norm
    return v @ J


@pytest.mark.parametrize("n", [64, 96])
def test_vjp_quadratic_closed_form_gradient(n):
    rng = np.random.default_rng(100 + n)
    S = rng.uniform(1.0, 100.0, n)
    p = np.arange(n, 0, -1, dtype=np.float64) ** 2
    p /= p.sum()
    v = rng.normal(size=n)
    eq, ds = icm_amd.equity_vjp(S, p, v, Q=256)
    norm = n * (n + 1.0) * (2.0 * n + 1.0) / 6.0
    exact_e = (2.0 * O.v2_exact(S) + 3.0 * O.v1_exact(S) - 2.0) / norm
    exact_g = _quadratic_gradient(S, v)
    e_eq = float(np.max(np.abs(eq - exact_e) / exact_e))
    e_ds = float(np.max(np.abs(ds - exact_g)) / np.max(np.abs(exact_g)))
    say(f"quadratic n={n}: equity {e_eq:.2e}, gradient {e_ds:.2e}")
    assert e_eq < 1e-9 and e_ds < 1e-8

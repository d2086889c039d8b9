"""CPU suite: pins the oracle (oracle/icm_oracle.c) before it is trusted.

Checks, each chosen so that a plausible slip in the restatement fails one:
  * the reference's own golden vectors (web/verify/reference.json, 16 cases,
    tolerance 5e-12 x max|E| as web/verify/verify_node.mjs:10,36-45);
  * exact ICM by enumerating finishing orders (tiny n, Malmuth-Harville);
  * closed forms V1 / V2 / k=1 / n=2 (src/cpu/icm.c:145-170,
    python/tests/test_icm.py:76-101, web/verify/gen_reference.c:139-150);
  * per-node closed forms that hold for every quadrature node at k = n;
  * node rule against scipy's log_ndtr (independent log Phi);
  * engine cross-checks (naive / linear / hybrid schoolbook / hybrid FFT),
    as bench/bench.c:733-767;
  * subset semantics (python/tests/test_icm.py:104-140).
"""
import itertools
import json
import math
import os

import numpy as np
import pytest
import scipy.special

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")


def _cases():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c["name"])
def test_golden_vectors(case):
    S = np.array(case["stacks"])
    p = np.array(case["payouts"])
    ref = np.array(case["equity"])
    fft_min = 64 if case["n"] >= 3000 else 0
    eq = O.equity(S, p, Q=case["Q"], engine=O.ENGINE_HYBRID, B=32, fft_min=fft_min)
    err = np.max(np.abs(eq - ref)) / np.max(np.abs(ref))
    assert err <= 5e-12, err


def _mh_exact(S, payout):
    """Exact ICM: enumerate finishing orders with Malmuth-Harville probabilities."""
    n = len(S)
    eq = np.zeros(n)
    for order in itertools.permutations(range(n)):
        prob = 1.0
        rem = float(sum(S))
        for i in order:
            prob *= S[i] / rem
            rem -= S[i]
        for pos, i in enumerate(order):
            if pos < len(payout):
                eq[i] += prob * payout[pos]
    return eq


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_exact_enumeration_small_n(seed):
    rng = np.random.default_rng(seed)
    n = 6
    S = rng.uniform(1.0, 50.0, n)
    payout = np.sort(rng.uniform(0, 1, 4))[::-1]
    exact = _mh_exact(S, payout)
    for eng in (O.ENGINE_NAIVE, O.ENGINE_LINEAR, O.ENGINE_HYBRID):
        eq = O.equity(S, payout, Q=256, engine=eng, B=2 if eng == O.ENGINE_HYBRID else 32)
        np.testing.assert_allclose(eq, exact, rtol=1e-10, atol=0)


def test_v1_closed_form_k_equals_n():
    rng = np.random.default_rng(5)
    n = 300
    S = rng.uniform(1, 100, n)
    payout = np.arange(n, 0, -1, dtype=float)      # pi_r = n - r
    eq = O.equity(S, payout, Q=256, engine=O.ENGINE_HYBRID, B=32, fft_min=0)
    v1 = O.v1_exact(S)
    np.testing.assert_allclose(eq, v1, rtol=5e-12)


def test_v2_closed_form():
    rng = np.random.default_rng(6)
    n = 40
    S = rng.uniform(1, 100, n)
    # pi_r = C(n-1-r, 2) gives V2 (src/cpu/icm.c:154-157)
    payout = np.array([(n - 1 - r) * (n - 2 - r) / 2.0 for r in range(n)])
    eq = O.equity(S, payout, Q=256, engine=O.ENGINE_HYBRID, B=8, fft_min=0)
    v2 = O.v2_exact(S)
    np.testing.assert_allclose(eq, v2, rtol=5e-11)


def test_k1_closed_form():
    S = np.array([100.0 * (i + 1) for i in range(10)])
    eq = O.equity(S, [1.0], Q=256)
    np.testing.assert_allclose(eq, S / S.sum(), rtol=1e-10)


def test_n2_closed_form():
    eq = O.equity(np.array([7500.0, 2500.0]), [1.0], Q=256)
    assert abs(eq[0] - 0.75) < 1e-12


@pytest.mark.parametrize("n,k", [(5, 3), (40, 40), (257, 20), (1000, 150)])
def test_sum_rule(n, k):
    # Q = 1024: at Q = 256 the erfc-trap rule itself misses the sum rule by
    # ~1e-7 for stacks spread over 1e4 (every engine agrees to 1e-16 there).
    rng = np.random.default_rng(n)
    S = rng.uniform(1, 1e4, n)
    p = np.sort(rng.uniform(0, 1, k))[::-1]
    eq = O.equity(S, p, Q=1024, fft_min=64)
    assert abs(eq.sum() - p.sum()) <= 1e-10 * p.sum()


def test_node_rule_against_scipy():
    for Smax in (1.0, 100.0, 1e6, 1e9):
        Q = 256
        logv, w = O.make_nodes(Q, Smax)
        y_lo = -7.7
        y_hi = max(6.5, math.sqrt(2.0) * math.sqrt(math.log(Smax) + 25.0))
        y = y_lo + np.arange(Q) * ((y_hi - y_lo) / (Q - 1))
        np.testing.assert_allclose(logv, scipy.special.log_ndtr(y), rtol=1e-13, atol=1e-300)
        h = (y_hi - y_lo) / (Q - 1)
        np.testing.assert_allclose(w, h * np.exp(-0.5 * y * y) / math.sqrt(2 * math.pi), rtol=1e-14)


def _per_node_forms(a):
    A = a.sum() - a
    B2 = (a * a).sum() - a * a
    return 1.0 + A, 1.0 + 3.0 * A + A * A - B2


@pytest.mark.parametrize("engine,B,fft_min", [(O.ENGINE_NAIVE, 32, 0), (O.ENGINE_LINEAR, 32, 0),
                                              (O.ENGINE_HYBRID, 16, 0), (O.ENGINE_HYBRID, 32, 16)])
def test_per_node_closed_forms(engine, B, fft_min):
    rng = np.random.default_rng(11)
    n = 700
    a = rng.uniform(0, 1, n)
    a[:5] = [0.0, 1.0, 0.5, 1e-16, 1 - 1e-12]
    lin, quad = _per_node_forms(a)
    r = np.arange(n, dtype=float)
    ones = O.inner(a, np.ones(n), engine=engine, B=B, fft_min=fft_min)
    np.testing.assert_allclose(ones, 1.0, rtol=1e-12)
    psi1 = O.inner(a, n - r, engine=engine, B=B, fft_min=fft_min)
    np.testing.assert_allclose(psi1, lin, rtol=1e-12)
    psi2 = O.inner(a, (n - r) ** 2, engine=engine, B=B, fft_min=fft_min)
    np.testing.assert_allclose(psi2, quad, rtol=1e-11)


@pytest.mark.parametrize("n,k", [(64, 64), (500, 37), (2048, 2048), (3000, 450)])
def test_engine_cross_check(n, k):
    rng = np.random.default_rng(n + k)
    S = np.exp(rng.uniform(0, math.log(1e6), n))
    p = 0.9 ** np.arange(k)
    p /= p.sum()
    ref = O.equity(S, p, Q=64, engine=O.ENGINE_LINEAR)
    for B, fm in ((16, 0), (32, 0), (32, 32), (64, 64)):
        eq = O.equity(S, p, Q=64, engine=O.ENGINE_HYBRID, B=B, fft_min=fm)
        np.testing.assert_allclose(eq, ref, rtol=1e-11, atol=1e-15)


def test_subset_matches_full_and_zeros():
    rng = np.random.default_rng(9)
    n, k = 1024, 100
    S = rng.uniform(1, 100, n)
    p = np.arange(n, n - k, -1, dtype=float)
    targets = rng.choice(n, 10, replace=False)
    full = O.equity(S, p, Q=256, fft_min=64)
    sub = O.equity(S, p, Q=256, targets=targets, fft_min=64)
    np.testing.assert_allclose(sub[targets], full[targets], rtol=1e-11)
    mask = np.ones(n, bool)
    mask[targets] = False
    assert np.all(sub[mask] == 0.0)


def test_planted_mistake_is_caught():
    """A dropped term in the per-node closed form must be detected."""
    rng = np.random.default_rng(3)
    a = rng.uniform(0, 1, 50)
    psi = O.inner(a, 50 - np.arange(50.0), engine=O.ENGINE_HYBRID, B=8)
    wrong = 1.0 + a.sum()          # forgot to leave player i out
    assert np.max(np.abs(psi - wrong)) > 1e-3

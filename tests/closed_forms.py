"""Exact ICM equities for the linear and quadratic payouts at k = n, for
sampled players of large fields (test infrastructure).

With c_i the number of players eliminated before player i (Malmuth-Harville
finishing order; pairwise P(i outlasts j) = S_i / (S_i + S_j)):

  * payout n - r = 1 + c_i gives V1_i = 1 + sum_{j != i} S_i / (S_i + S_j)
    (src/cpu/icm.c:145-152);
  * payout C(c_i, 2) gives V2_i = sum_{j < l, j, l != i} S_i / (S_i + S_j + S_l)
    (src/cpu/icm.c:154-170), O(n^2) per player directly;
  * payout (n - r)^2 = 1 + 3 c_i + 2 C(c_i, 2) gives E_i = 2 V2_i + 3 V1_i - 2.

For large n, V2 uses its integral form: S_i / (S_i + S_j + S_l) =
int_0^inf S_i e^{-(S_i + S_j + S_l) t} dt, so
  V2_i = int S_i e^{-S_i t} (A_i(t)^2 - B2_i(t)) / 2 dt,
  A_i = sum_{j != i} e^{-S_j t},  B2_i = sum_{j != i} e^{-2 S_j t},
evaluated with the trapezoidal rule in u = log t.  The integrand is analytic
in the strip |Im u| < pi/2 and decays at both ends, so the rule converges
like exp(-pi^2 / h): h = 0.05 is far below the 1e-16 level.  The same rule
gives V1 - 1 (tested against the direct sum).
"""
import math

import numpy as np


def v1_direct(S, idx):
    S = np.asarray(S, dtype=np.float64)
    out = []
    for i in idx:
        r = S[i] / (S[i] + S)
        out.append(1.0 + (np.sum(r) - 0.5))       # drop the j = i term (1/2)
    return np.array(out)


def v_integrals(S, idx, h=0.05, t_lo=None, t_hi=None, chunk=1 << 13):
    """(V1 - 1, V2) for players idx by the log-t trapezoidal rule."""
    S = np.asarray(S, dtype=np.float64)
    smin, smax = float(S.min()), float(S.max())
    if t_lo is None:
        t_lo = 1e-22 / (smax * len(S) ** 2)
    if t_hi is None:
        t_hi = 80.0 / smin
    u = np.arange(math.log(t_lo), math.log(t_hi) + h, h)
    t = np.exp(u)
    A = np.zeros_like(t)
    B2 = np.zeros_like(t)
    for c0 in range(0, len(S), chunk):
        e = np.exp(-np.outer(S[c0:c0 + chunk], t))
        A += e.sum(axis=0)
        B2 += (e * e).sum(axis=0)
    v1m, v2 = [], []
    for i in idx:
        ei = np.exp(-S[i] * t)
        Ai = A - ei
        B2i = B2 - ei * ei
        base = S[i] * ei * t * h                  # S_i e^{-S_i t} dt, dt = t du
        v1m.append(float(np.sum(base * Ai)))
        v2.append(float(np.sum(base * 0.5 * (Ai * Ai - B2i))))
    return np.array(v1m), np.array(v2)


def quadratic_equity(S, idx):
    """Normalised equities of players idx for payout (n - r)^2 / sum_m m^2."""
    n = len(S)
    v1 = v1_direct(S, idx)
    _, v2 = v_integrals(S, idx)
    norm = n * (n + 1.0) * (2.0 * n + 1.0) / 6.0
    return (2.0 * v2 + 3.0 * v1 - 2.0) / norm

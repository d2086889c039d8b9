"""Host-side tools (no GPU): autotuner table writer and acceptance check,
CLI input generators, ragged workload generator."""
import math
import re

import numpy as np
import pytest

from icm_amd import autotune
from icm_amd.cli import make_payout, make_stacks
from icm_amd.workloads import ragged_instances


def parse_header(path):
    text = open(path).read()
    rows = [tuple(int(x) for x in m) for m in re.findall(r"\{(\d+), (\d+), (\d+), (\d+)\},", text)]
    batch = re.search(r"kBatchTune = \{(\d+), (\d+)\}", text)
    return rows, (int(batch.group(1)), int(batch.group(2))) if batch else None


def lookup(rows, n, k):
    """The planner's rule (icm_common.hpp tune_row): first row covering (n, k)."""
    for r in rows:
        if n <= r[0] and k <= r[1]:
            return r
    return rows[-1]


def test_header_rows_cover_grid_and_lookup_matches(tmp_path):
    table = {}
    for n in (1024, 16384, 1048576):
        for k in (1, 9, n // 10, n):
            table[(n, k)] = {"B": 64 if k > 9 else 32, "sb_max": 129 if k > 9 else 512, "ms": math.log(n * k + 2)}
    out = tmp_path / "tune.h"
    autotune.write_header(table, str(out), {"B": 64, "sb_max": 129, "ms": 1.0})
    rows, batch = parse_header(out)
    assert batch == (64, 129)
    assert len(rows) == len(table)
    assert rows[-1][0] == 0x7FFFFFFF and rows[-1][1] == 0x7FFFFFFF
    # every tuned shape maps to its own row
    for (n, k), v in table.items():
        r = lookup(rows, n, k)
        assert (r[2], r[3]) == (v["B"], v["sb_max"]), (n, k, r)
    # shapes between grid points take a neighbour's row, never nothing
    for n, k in [(5000, 3), (300000, 40000), (1 << 22, 1 << 22), (2, 1)]:
        r = lookup(rows, n, k)
        assert r[2] in (32, 64)


def test_monotonicity_check_flags_a_faster_larger_shape():
    table = {(1024, 1): {"ms": 1.0}, (1024, 1024): {"ms": 2.0},
             (4096, 1): {"ms": 1.5}, (4096, 4096): {"ms": 3.0}}
    assert autotune.check_monotone(table) == []
    table[(4096, 4096)] = {"ms": 1.0}          # larger n, larger k, faster: a violation
    bad = autotune.check_monotone(table)
    assert bad and any("4096" in b for b in bad)


def test_committed_table_is_consistent():
    rows, batch = parse_header(autotune.HEADER)
    assert rows and batch
    assert all(r[2] in (16, 32, 64) for r in rows)
    assert all(0 <= r[3] <= 512 for r in rows)
    assert rows[-1][0] == 0x7FFFFFFF and rows[-1][1] == 0x7FFFFFFF
    assert batch[0] in (16, 32, 64)


@pytest.mark.parametrize("dist", [0, 1, 2, 3])
def test_cli_stack_distributions(dist):
    S = make_stacks(1000, dist)
    assert S.shape == (1000,) and np.all(S > 0)
    if dist == 0:
        assert 1.0 <= S.min() and S.max() <= 100.0
    if dist == 1:
        assert S[0] == 10000.0 and np.all(S[1:] == 1.0)
    if dist == 2:
        assert S[0] == 1.0 and abs(S[-1] - 2.0 ** (999 * 10.0 / 1000)) < 1e-9
    if dist == 3:
        assert np.all(S == 100.0)
    assert np.array_equal(make_payout(10, 4), [10.0, 9.0, 8.0, 7.0])


def test_ragged_instances_follow_the_benchmark_rules():
    stacks, payouts = ragged_instances(500, seed=4)
    assert len(stacks) == len(payouts) == 500
    ks = set()
    for S, p in zip(stacks, payouts):
        n, k = len(S), len(p)
        assert 2 <= n <= 4096 and 1 <= k <= n
        assert k in {1, 3, 9, math.ceil(n / 10), n} or k == n
        assert abs(S.sum() - 1e7) < 1e-3 and np.all(S > 0)
        assert abs(p.sum() - 1.0) < 1e-12
        if k > 1:
            r = p[1] / p[0]
            ok = p[:-1] > 1e-250            # r^j underflows to 0 for very long payouts
            assert 0.5 <= r <= 0.95 and np.allclose(p[1:][ok] / p[:-1][ok], r)
        ks.add(k == n)
    assert ks == {True, False}
    # same seed, same batch
    s2, p2 = ragged_instances(500, seed=4)
    assert all(np.array_equal(a, b) for a, b in zip(stacks, s2))

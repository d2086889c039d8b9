"""CPU suite: the planner's memory policy and option handling, run host-only
through icm_hip_plan_dry_run (the planner and the allocation sequence against
a simulated device; no GPU call).

Reference policy: on a failed allocation the q-batch is halved and the plan
allocated again, at most 4 times (src/gpu/gpu_memory.cu:267-279); the
q-batch override ICM_GPU_Q_BATCH (src/gpu/gpu_api.cu:139-144).
"""
import os

import pytest

import icm_amd

GB = 1 << 30


def test_headline_shape_fits_all_nodes_in_one_batch():
    s, r = icm_amd.plan_dry_run(1 << 20, 1 << 20, 288 * GB)
    assert r == 0 and s["q_batch"] == 256
    assert s["planned_peak_vram_bytes"] <= 0.6 * 288 * GB
    assert s["n_tier3"] > 0


def test_bytes_are_linear_in_q_batch():
    n, k = 65536, 65536
    b = [icm_amd.plan_dry_run(n, k, 64 * GB, q_batch=q)[0]["planned_peak_vram_bytes"] for q in (1, 2, 3, 5)]
    per = b[1] - b[0]
    assert per > 0
    assert b[2] - b[1] == per and b[3] - b[2] == 2 * per


def test_forced_q_batch_halves_until_it_fits():
    n, k = 65536, 65536
    one = icm_amd.plan_dry_run(n, k, 64 * GB, q_batch=1)[0]["planned_peak_vram_bytes"]
    two = icm_amd.plan_dry_run(n, k, 64 * GB, q_batch=2)[0]["planned_peak_vram_bytes"]
    per = two - one
    fixed = one - per
    cap = fixed + 10 * per            # room for 10 q-planes
    s, r = icm_amd.plan_dry_run(n, k, cap, q_batch=64)
    assert (s["q_batch"], r) == (8, 3)           # 64 -> 32 -> 16 -> 8
    assert s["planned_peak_vram_bytes"] <= cap


def test_gives_up_after_four_halvings():
    n, k = 65536, 65536
    one = icm_amd.plan_dry_run(n, k, 64 * GB, q_batch=1)[0]["planned_peak_vram_bytes"]
    two = icm_amd.plan_dry_run(n, k, 64 * GB, q_batch=2)[0]["planned_peak_vram_bytes"]
    cap = one + (two - one) // 2      # fits 1 q-plane, not 2
    with pytest.raises(icm_amd.ICMError, match="after 4 halvings"):
        icm_amd.plan_dry_run(n, k, cap, q_batch=64)      # 64 .. 4 all fail
    s, r = icm_amd.plan_dry_run(n, k, cap, q_batch=16)   # 16 -> 8 -> 4 -> 2 -> 1
    assert (s["q_batch"], r) == (1, 4)
    with pytest.raises(icm_amd.ICMError, match="budget exceeded"):
        icm_amd.plan_dry_run(n, k, one // 2, q_batch=1)


def test_budget_share_sets_q_batch():
    n, k = 65536, 65536
    a = icm_amd.plan_dry_run(n, k, 8 * GB)[0]
    b = icm_amd.plan_dry_run(n, k, 8 * GB, mem_fraction=0.3)[0]
    assert b["q_batch"] < a["q_batch"]
    assert a["planned_peak_vram_bytes"] <= 0.6 * 8 * GB
    assert b["planned_peak_vram_bytes"] <= 0.3 * 8 * GB


def test_memory_strategy_recompute_drops_the_spectrum_cache():
    n, k = 1 << 18, 1 << 18
    a = icm_amd.plan_dry_run(n, k, 64 * GB, q_batch=4)[0]
    b = icm_amd.plan_dry_run(n, k, 64 * GB, q_batch=4, memory_strategy=3)[0]
    assert b["planned_peak_vram_bytes"] < a["planned_peak_vram_bytes"]


def test_q_batch_environment_override(monkeypatch):
    n, k = 65536, 65536
    monkeypatch.setenv("ICM_GPU_Q_BATCH", "3")
    assert icm_amd.plan_dry_run(n, k, 64 * GB)[0]["q_batch"] == 3
    monkeypatch.setenv("ICM_GPU_Q_BATCH", "999")          # out of range: ignored, as the reference
    assert icm_amd.plan_dry_run(n, k, 64 * GB)[0]["q_batch"] > 3
    monkeypatch.delenv("ICM_GPU_Q_BATCH")
    assert icm_amd.plan_dry_run(n, k, 64 * GB, q_batch=5)[0]["q_batch"] == 5


def test_block_size_parameter():
    assert icm_amd.plan_dry_run(100000, 1000, 8 * GB, block_size=16)[0]["B"] == 16
    assert icm_amd.plan_dry_run(100000, 1000, 8 * GB, block_size=64)[0]["B"] == 64


def _dry_with_options(**fields):
    import ctypes
    lib = icm_amd.load_library()
    o = icm_amd._options()
    for k_, v in fields.items():
        setattr(o, k_, v)
    prm = icm_amd._params(mem_budget_bytes=8 * GB)
    sm = icm_amd.PlanSummary()
    r = ctypes.c_int(0)
    rc = lib.icm_hip_plan_dry_run(4096, 4096, ctypes.byref(o), ctypes.byref(prm), ctypes.byref(sm),
                                  ctypes.byref(r))
    return rc, lib.icm_hip_last_error().decode()


def test_reference_option_fields_keep_their_meaning():
    # -1 (the reference's auto) and 0 are accepted; counts this engine
    # cannot honour are rejected with a message instead of reinterpreted
    assert _dry_with_options(force_uncached_fused_levels=-1, force_uncached_cufft_levels=-1)[0] == 0
    assert _dry_with_options(force_uncached_fused_levels=0, force_uncached_cufft_levels=0)[0] == 0
    rc, msg = _dry_with_options(force_uncached_fused_levels=2)
    assert rc == -1 and "no per-level cache count" in msg
    rc, msg = _dry_with_options(force_uncached_cufft_levels=32)
    assert rc == -1 and "no per-level cache count" in msg
    for ms in (-1, 0, 1, 2, 3, 4):   # 3 = recompute; other values mean auto, as the reference
        assert _dry_with_options(memory_strategy=ms)[0] == 0


def test_invalid_parameters_are_rejected():
    with pytest.raises(icm_amd.ICMError, match="block_size"):
        icm_amd.plan_dry_run(4096, 4096, 8 * GB, block_size=24)
    with pytest.raises(icm_amd.ICMError, match="q_batch"):
        icm_amd.plan_dry_run(4096, 4096, 8 * GB, q_batch=300)
    with pytest.raises(icm_amd.ICMError, match="mem_fraction"):
        icm_amd.plan_dry_run(4096, 4096, 8 * GB, mem_fraction=1.5)
    with pytest.raises(icm_amd.ICMError, match="mem_budget_bytes"):
        icm_amd.plan_dry_run(4096, 4096, 0)

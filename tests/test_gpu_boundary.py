"""GPU: the drop-in boundary exactly as the reference's callers bind it.

  * python/icm/__init__.py:86-119 loads the library with ctypes and sets
    argtypes on icm_init / icm_equity / icm_equity_subset only: the same
    binding, on a fresh handle, against the oracle;
  * the GPU API of src/gpu/icm_gpu.h (icm_gpu_init, icm_gpu_equity,
    icm_gpu_equity_subset, the plan calls) with options as bench/bench_gpu.cu
    sets them and with every field -1;
  * option values the engine cannot honour are rejected with a message;
  * icm_gpu_release_pooled_memory (src/gpu/icm_gpu.h:70-74) and the probes;
  * the out-of-memory policy on the device: a q-batch halved under a memory
    cap gives the same equities, and a failed adjoint allocation leaves the
    plan (and new plans) working.
"""
import ctypes

import numpy as np
import pytest

import icm_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def say(msg):
    print(msg, flush=True)


def rel_err(gpu, cpu):
    den = np.abs(cpu).copy()
    den[den < 1e-14] = 1.0
    return float(np.max(np.abs(gpu - cpu) / den))


def _reference_binding():
    """ctypes setup of python/icm/__init__.py:86-119, verbatim in effect."""
    lib = ctypes.CDLL(icm_amd.library_path())
    lib.icm_init.argtypes = [ctypes.c_char_p]
    lib.icm_init.restype = None
    lib.icm_equity.argtypes = [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D]
    lib.icm_equity.restype = None
    lib.icm_equity_subset.argtypes = [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, _I, ctypes.c_int]
    lib.icm_equity_subset.restype = None
    return lib


def _case(n=700, k=90, seed=1):
    rng = np.random.default_rng(seed)
    S = np.ascontiguousarray(rng.uniform(1.0, 1e4, n))
    p = np.ascontiguousarray(0.85 ** np.arange(k))
    return S, p / p.sum()


def test_reference_python_loader_binding():
    lib = _reference_binding()
    lib.icm_init(None)
    S, p = _case()
    n, k = len(S), len(p)
    out = np.zeros(n)
    lib.icm_equity(n, S.ctypes.data_as(_D), 256, p.ctypes.data_as(_D), k, out.ctypes.data_as(_D))
    ref = O.equity(S, p, Q=256)
    e1 = rel_err(out, ref)
    t = np.ascontiguousarray([3, 17, 400, 699], dtype=np.int32)
    sub = np.full(n, 7.0)
    lib.icm_equity_subset(n, S.ctypes.data_as(_D), 256, p.ctypes.data_as(_D), k, sub.ctypes.data_as(_D),
                          t.ctypes.data_as(_I), len(t))
    ref_sub = O.equity(S, p, Q=256, targets=t)
    mask = np.zeros(n, dtype=bool)
    mask[t] = True
    e2 = rel_err(sub[mask], ref_sub[mask])
    say(f"reference binding: icm_equity {e1:.2e}, icm_equity_subset {e2:.2e}")
    assert e1 < 1e-8 and e2 < 1e-8
    assert np.all(sub[~mask] == 0.0)           # non-targets zeroed (src/gpu/gpu_api.cu:662-670)


class GpuOptions(ctypes.Structure):
    """IcmGpuOptions as declared by src/gpu/icm_gpu.h:12-34."""
    _fields_ = [(f, ctypes.c_int) for f in ("device_id", "use_cufftdx", "enable_graphs", "enable_q_pipeline",
                                            "memory_strategy", "force_uncached_fused_levels",
                                            "force_uncached_cufft_levels", "fast_mode", "verbose")]


def _bench_options():
    o = GpuOptions()                       # bench/bench_gpu.cu:134-141
    o.use_cufftdx, o.enable_graphs, o.enable_q_pipeline, o.memory_strategy = 1, 0, 1, 0
    o.force_uncached_fused_levels = o.force_uncached_cufft_levels = -1
    return o


def _all_minus_one():
    o = GpuOptions()
    for f, _ in GpuOptions._fields_:
        setattr(o, f, -1)
    return o


@pytest.mark.parametrize("opts", ["bench", "minus_one", "null"])
def test_reference_gpu_api(opts):
    lib = ctypes.CDLL(icm_amd.library_path())
    assert lib.icm_gpu_init(0) == 1
    lib.icm_gpu_equity.argtypes = [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_void_p,
                                   ctypes.POINTER(icm_amd.RunStats)]
    lib.icm_gpu_equity_subset.argtypes = [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, _I, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.POINTER(icm_amd.RunStats)]
    o = {"bench": _bench_options(), "minus_one": _all_minus_one(), "null": None}[opts]
    optp = ctypes.cast(ctypes.byref(o), ctypes.c_void_p) if o is not None else None
    S, p = _case(1500, 1500, 2)
    n, k = len(S), len(p)
    out = np.zeros(n)
    st = icm_amd.RunStats()
    rc = lib.icm_gpu_equity(n, S.ctypes.data_as(_D), 256, p.ctypes.data_as(_D), k, out.ctypes.data_as(_D), optp,
                            ctypes.byref(st))
    assert rc == 0
    ref = O.equity(S, p, Q=256)
    e1 = rel_err(out, ref)
    t = np.ascontiguousarray([0, 5, 1499, 2000, -3], dtype=np.int32)   # out-of-range targets ignored
    sub = np.zeros(n)
    rc = lib.icm_gpu_equity_subset(n, S.ctypes.data_as(_D), 256, p.ctypes.data_as(_D), k,
                                   sub.ctypes.data_as(_D), t.ctypes.data_as(_I), len(t), optp, None)
    assert rc == 0
    e2 = rel_err(sub[[0, 5, 1499]], ref[[0, 5, 1499]])
    say(f"icm_gpu_equity ({opts}): {e1:.2e}; subset {e2:.2e}; total {st.total_ns / 1e6:.2f} ms B={st.B}")
    assert e1 < 1e-8 and e2 < 1e-8 and st.total_ns > 0 and st.B in (16, 32, 64)
    assert np.count_nonzero(sub) == 3


def test_reference_plan_api():
    lib = ctypes.CDLL(icm_amd.library_path())
    assert lib.icm_gpu_init(0) == 1
    lib.icm_gpu_plan_create.restype = ctypes.c_void_p
    lib.icm_gpu_plan_create.argtypes = [ctypes.c_int, _D, ctypes.c_int, ctypes.c_void_p]
    lib.icm_gpu_equity_with_plan.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D,
                                             ctypes.POINTER(icm_amd.RunStats)]
    lib.icm_gpu_plan_summary.argtypes = [ctypes.c_void_p, ctypes.POINTER(icm_amd.PlanSummary)]
    lib.icm_gpu_plan_destroy.argtypes = [ctypes.c_void_p]
    lib.icm_gpu_last_error.restype = ctypes.c_char_p
    S, p = _case(5000, 5000, 3)
    o = _bench_options()
    h = lib.icm_gpu_plan_create(len(S), S.ctypes.data_as(_D), len(p), ctypes.cast(ctypes.byref(o), ctypes.c_void_p))
    assert h
    try:
        sm = icm_amd.PlanSummary()
        assert lib.icm_gpu_plan_summary(h, ctypes.byref(sm)) == 0 and sm.n == 5000 and sm.q_batch >= 1
        ref = O.equity(S, p, Q=128)
        for _ in range(2):                     # plan reuse
            out = np.zeros(len(S))
            assert lib.icm_gpu_equity_with_plan(h, 128, p.ctypes.data_as(_D), out.ctypes.data_as(_D), None) == 0
            assert rel_err(out, ref) < 1e-8
    finally:
        lib.icm_gpu_plan_destroy(h)
    # counts this engine cannot honour: NULL plan and a message, not a silent reinterpretation
    o.force_uncached_fused_levels = 3
    h = lib.icm_gpu_plan_create(len(S), S.ctypes.data_as(_D), len(p), ctypes.cast(ctypes.byref(o), ctypes.c_void_p))
    assert not h and b"per-level cache count" in lib.icm_gpu_last_error()


def test_release_pooled_memory_and_probes():
    lib = icm_amd.load_library()
    icm_amd.init()
    assert lib.icm_gpu_release_pooled_memory() == 0
    copy = icm_amd.measure_hbm_bandwidth_gbps()
    r, w, c = icm_amd.measure_stream_gbps(1 << 30)
    b, p = icm_amd.measure_fused_pair_ns(1024, 4096, quick=True)
    say(f"probes: device copy {copy:.0f} GB/s (bytes copied), stream read {r:.0f} write {w:.0f} copy {c:.0f} GB/s; "
        f"fused 1024: build {b:.1f} ns/parent, propagate {p:.1f} ns/parent")
    assert 500 < copy < 8000 and 1000 < r < 8500 and 1000 < w < 8500 and 1000 < c < 8500
    assert b > 0 and p > 0
    with pytest.raises(icm_amd.ICMError):
        icm_amd.measure_fused_pair_ns(1000, 16)


def test_halved_q_batch_under_a_memory_cap_gives_the_same_equities():
    S, p = _case(65536, 65536, 4)
    Q = 64
    full = icm_amd.Plan(S, len(p))
    ref = full.run(p, Q)
    full.close()
    # without the spectrum cache (memory_strategy 3) the per-plane bytes are
    # the same at every q-batch: room for 10 planes, so 64 -> 32 -> 16 -> 8
    one = icm_amd.plan_dry_run(len(S), len(p), 1 << 40, q_batch=1, memory_strategy=3)[0]
    two = icm_amd.plan_dry_run(len(S), len(p), 1 << 40, q_batch=2, memory_strategy=3)[0]
    one, two = one["planned_peak_vram_bytes"], two["planned_peak_vram_bytes"]
    cap = one + 10 * (two - one)
    dry, retries = icm_amd.plan_dry_run(len(S), len(p), cap, q_batch=64, memory_strategy=3)
    assert (dry["q_batch"], retries) == (8, 3)
    pl = icm_amd.Plan(S, len(p), q_batch=64, mem_budget_bytes=cap, memory_strategy=3)
    try:
        sm = pl.summary()
        out = pl.run(p, Q)
    finally:
        pl.close()
    err = float(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
    say(f"capped plan: q_batch {sm['q_batch']} ({sm['planned_peak_vram_bytes'] / 1e9:.2f} GB <= {cap / 1e9:.2f} GB), "
        f"max diff {err:.1e}")
    assert sm["q_batch"] == 8 and sm["planned_peak_vram_bytes"] == dry["planned_peak_vram_bytes"] <= cap
    assert err < 1e-13


def test_failed_adjoint_allocation_leaves_plans_working():
    S, p = _case(20000, 20000, 5)
    v = np.random.default_rng(6).normal(size=len(S))
    base = icm_amd.Plan(S, len(p), q_batch=4)
    ref = base.run(p, 16)
    need = base.summary()["planned_peak_vram_bytes"]
    base.close()
    # cap just above the forward's bytes: the adjoint buffers cannot fit
    pl = icm_amd.Plan(S, len(p), q_batch=4, mem_budget_bytes=need + (1 << 20))
    try:
        held = pl.summary()["planned_peak_vram_bytes"]
        with pytest.raises(icm_amd.ICMError, match="adjoint buffers"):
            pl.vjp(p, v, 16)
        assert pl.summary()["planned_peak_vram_bytes"] == held     # nothing left allocated
        again = pl.run(p, 16)                                       # no stale error reported
    finally:
        pl.close()
    # the capped plan may run without the spectrum cache (and so without the
    # column hand-overs that need it): equal to rounding, not bit for bit
    assert np.max(np.abs(again - ref)) <= 1e-13 * np.max(np.abs(ref))
    eq, ds = icm_amd.Plan(S, len(p), q_batch=4).vjp(p, v, 16)
    assert np.max(np.abs(eq - ref)) <= 1e-13 * np.max(np.abs(ref)) and np.all(np.isfinite(ds))


def test_batch_plan_under_a_memory_cap():
    rng = np.random.default_rng(8)
    stacks = [rng.uniform(1, 100, int(n)) for n in rng.integers(2, 3000, 64)]
    pays = [0.9 ** np.arange(max(1, len(s) // 7)) for s in stacks]
    pays = [x / x.sum() for x in pays]
    ref = icm_amd.equity_batch(stacks, pays, Q=64)
    bp = icm_amd.BatchPlan(stacks, [len(x) for x in pays], q_batch=64)
    need64 = bp.summary()["device_bytes"]
    bp.close()
    bp = icm_amd.BatchPlan(stacks, [len(x) for x in pays], q_batch=64, mem_budget_bytes=need64 // 3)
    try:
        sm = bp.summary()
        out = bp.run(pays, 64)
    finally:
        bp.close()
    say(f"capped batch: q-planes per pass {sm['q_batch']}, {sm['device_bytes'] / 1e9:.3f} GB")
    assert sm["q_batch"] < 64 and sm["device_bytes"] <= need64 // 3
    for a, b in zip(out, ref):
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))

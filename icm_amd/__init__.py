"""icm_amd -- MI355X-native ICM equity engine (ctypes front end of libicm_hip.so).

Mirrors the reference Python package (python/icm/__init__.py): ``init``,
``equity`` and ``equity_subset`` with the same argument meaning, validation
and error messages, plus the reusable plan of the reference GPU header
(src/gpu/icm_gpu.h:79-111) as :class:`Plan`.  Every computation runs in the
HIP library; there is no CPU fallback: if the library is missing, loading
raises :class:`ICMError`.
"""
import ctypes
import os

import numpy as np

__all__ = ["ICMError", "init", "equity", "equity_subset", "equity_vjp", "node_inner", "node_vjp", "Plan",
           "BatchPlan", "plan_geometry", "plan_dry_run", "make_nodes", "library_path", "load_library",
           "measure_hbm_bandwidth_gbps", "measure_stream_gbps", "measure_lds_gbps", "measure_fused_pair_ns"]

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libicm_hip.so"

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


class ICMError(Exception):
    """Error raised by the ICM library (mirrors python/icm/__init__.py)."""


class Options(ctypes.Structure):
    """Layout of IcmHipOptions (= IcmGpuOptions, src/gpu/icm_gpu.h:12-34)."""
    _fields_ = [("device_id", ctypes.c_int), ("use_cufftdx", ctypes.c_int),
                ("enable_graphs", ctypes.c_int), ("enable_q_pipeline", ctypes.c_int),
                ("memory_strategy", ctypes.c_int),
                ("force_uncached_fused_levels", ctypes.c_int),
                ("force_uncached_cufft_levels", ctypes.c_int),
                ("fast_mode", ctypes.c_int), ("verbose", ctypes.c_int)]


class PlanParams(ctypes.Structure):
    """Layout of IcmHipPlanParams (engine parameters with no IcmGpuOptions field)."""
    _fields_ = [("block_size", ctypes.c_int), ("q_batch", ctypes.c_int),
                ("mem_fraction", ctypes.c_double), ("mem_budget_bytes", ctypes.c_size_t)]


class RunStats(ctypes.Structure):
    """Layout of IcmHipRunStats (= IcmGpuRunStats, src/gpu/icm_gpu.h:36-50)."""
    _fields_ = [("total_ns", ctypes.c_double), ("sort_ns", ctypes.c_double),
                ("quadrature_overhead_ns", ctypes.c_double),
                ("block_build_ns", ctypes.c_double), ("tree_build_ns", ctypes.c_double),
                ("tree_propagate_cached_ns", ctypes.c_double),
                ("tree_propagate_recomputed_ns", ctypes.c_double),
                ("leaf_extract_ns", ctypes.c_double), ("peak_vram_bytes", ctypes.c_size_t),
                ("engine", ctypes.c_int), ("B", ctypes.c_int),
                ("uncached_fused_levels", ctypes.c_int),
                ("uncached_cufft_levels", ctypes.c_int)]


class PlanSummary(ctypes.Structure):
    """Layout of IcmHipPlanSummary (= IcmGpuPlanSummary, src/gpu/icm_gpu.h:52-63)."""
    _fields_ = [("n", ctypes.c_int), ("k", ctypes.c_int), ("B", ctypes.c_int),
                ("engine", ctypes.c_int), ("n_levels", ctypes.c_int),
                ("n_tier1", ctypes.c_int), ("n_tier2", ctypes.c_int),
                ("n_tier3", ctypes.c_int), ("q_batch", ctypes.c_int),
                ("planned_peak_vram_bytes", ctypes.c_size_t)]


# (name, restype, argtypes) of every symbol declared in include/icm_hip.h
_SIGS = [
    ("icm_hip_init", ctypes.c_int, [ctypes.c_int]),
    ("icm_hip_shutdown", None, []),
    ("icm_hip_last_error", ctypes.c_char_p, []),
    ("icm_hip_release_pooled_memory", ctypes.c_int, []),
    ("icm_hip_plan_create", ctypes.c_void_p, [ctypes.c_int, _D, ctypes.c_int, ctypes.POINTER(Options)]),
    ("icm_hip_plan_create_ex", ctypes.c_void_p,
     [ctypes.c_int, _D, ctypes.c_int, ctypes.POINTER(Options), ctypes.POINTER(PlanParams)]),
    ("icm_hip_plan_dry_run", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int, ctypes.POINTER(Options), ctypes.POINTER(PlanParams),
      ctypes.POINTER(PlanSummary), _I]),
    ("icm_hip_plan_destroy", None, [ctypes.c_void_p]),
    ("icm_hip_plan_summary", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PlanSummary)]),
    ("icm_hip_equity_with_plan", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, ctypes.POINTER(RunStats)]),
    ("icm_hip_equity_with_plan_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(RunStats)]),
    ("icm_hip_equity", ctypes.c_int, [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D]),
    ("icm_hip_equity_subset", ctypes.c_int,
     [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, _I, ctypes.c_int]),
    ("icm_hip_node_inner", ctypes.c_int, [ctypes.c_int, _D, _D, ctypes.c_int, _D]),
    ("icm_gpu_init", ctypes.c_int, [ctypes.c_int]),
    ("icm_gpu_shutdown", None, []),
    ("icm_gpu_last_error", ctypes.c_char_p, []),
    ("icm_gpu_release_pooled_memory", ctypes.c_int, []),
    ("icm_gpu_plan_create", ctypes.c_void_p, [ctypes.c_int, _D, ctypes.c_int, ctypes.POINTER(Options)]),
    ("icm_gpu_plan_destroy", None, [ctypes.c_void_p]),
    ("icm_gpu_plan_summary", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PlanSummary)]),
    ("icm_gpu_equity_with_plan", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, ctypes.POINTER(RunStats)]),
    ("icm_gpu_equity", ctypes.c_int,
     [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.POINTER(Options), ctypes.POINTER(RunStats)]),
    ("icm_gpu_equity_subset", ctypes.c_int,
     [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, _I, ctypes.c_int,
      ctypes.POINTER(Options), ctypes.POINTER(RunStats)]),
    ("icm_gpu_measure_hbm_bandwidth_gbps", ctypes.c_int, [_D]),
    ("icm_gpu_measure_fused_pair_ns", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D]),
    ("icm_gpu_measure_fused_r2c_pair_ns", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D]),
    ("icm_hip_measure_stream_gbps", ctypes.c_int, [ctypes.c_size_t, _D, _D, _D]),
    ("icm_hip_measure_lds_gbps", ctypes.c_int, [_D]),
    ("icm_hip_equity_noalloc", ctypes.c_int, [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D]),
    ("icm_init", None, [ctypes.c_char_p]),
    ("icm_equity", None, [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D]),
    ("icm_equity_subset", None, [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, _I, ctypes.c_int]),
    ("icm_hip_plan_geometry", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, _I, _I, _I]),
    ("icm_hip_make_nodes", None, [ctypes.c_int, ctypes.c_double, _D, _D]),
    ("icm_hip_plan_set_targets", ctypes.c_int, [ctypes.c_void_p, _I, ctypes.c_int]),
    ("icm_hip_vjp_with_plan", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D, ctypes.POINTER(RunStats)]),
    ("icm_hip_vjp_with_plan_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.POINTER(RunStats)]),
    ("icm_hip_equity_vjp", ctypes.c_int, [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, _D, _D]),
    ("icm_hip_node_vjp", ctypes.c_int, [ctypes.c_int, _D, _D, _D, ctypes.c_int, _D, _D]),
    ("icm_hip_plan_set_profiling", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_plan_set_schoolbook_max", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_plan_get_schoolbook_max", ctypes.c_int, [ctypes.c_void_p]),
    ("icm_hip_plan_set_small_path", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_plan_get_small_path", ctypes.c_int, [ctypes.c_void_p]),
    ("icm_hip_plan_set_column_fusion", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_plan_set_poison", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_plan_set_adjoint_checkpoint", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_plan_adjoint_layout", ctypes.c_int, [ctypes.c_void_p, _I, _I]),
    ("icm_hip_selfcheck_row_layouts", ctypes.c_int, []),
    ("icm_hip_plan_poisoned_bytes", ctypes.c_longlong, [ctypes.c_void_p]),
    ("icm_hip_plan_set_spectrum_cache", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_plan_get_spectrum_cache", ctypes.c_int, [ctypes.c_void_p]),
    ("icm_hip_plan_launch_count", ctypes.c_int, [ctypes.c_void_p]),
    ("icm_hip_plan_launch_stat", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _I, _D, _D]),
    ("icm_hip_plan_kernel_count", ctypes.c_int, [ctypes.c_void_p]),
    ("icm_hip_plan_kernel_stat", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _I, _D, _D, _D]),
    ("icm_hip_batch_plan_create", ctypes.c_void_p, [ctypes.c_int, _I, _D, _I, ctypes.POINTER(Options)]),
    ("icm_hip_batch_plan_create_ex", ctypes.c_void_p,
     [ctypes.c_int, _I, _D, _I, ctypes.POINTER(Options), ctypes.POINTER(PlanParams)]),
    ("icm_hip_batch_plan_destroy", None, [ctypes.c_void_p]),
    ("icm_hip_batch_equity", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, ctypes.POINTER(RunStats)]),
    ("icm_hip_batch_equity_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(RunStats)]),
    ("icm_hip_equity_batch", ctypes.c_int, [ctypes.c_int, _I, _D, _I, ctypes.c_int, _D, _D]),
    ("icm_hip_batch_vjp", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D, ctypes.POINTER(RunStats)]),
    ("icm_hip_batch_vjp_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.POINTER(RunStats)]),
    ("icm_hip_batch_summary", ctypes.c_int,
     [ctypes.c_void_p, _I, _I, _I, _I, _I, ctypes.POINTER(ctypes.c_size_t)]),
    ("icm_hip_batch_set_profiling", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_batch_set_schoolbook_max", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_batch_set_persistent", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_batch_is_persistent", ctypes.c_int, [ctypes.c_void_p]),
    ("icm_hip_batch_set_spin_limit", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("icm_hip_batch_kernel_count", ctypes.c_int, [ctypes.c_void_p]),
    ("icm_hip_batch_kernel_stat", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _I, _D, _D, _D]),
]

_lib = None


def library_path():
    """Path of the in-tree library ($ICM_HIP_LIB overrides)."""
    return os.environ.get("ICM_HIP_LIB", os.path.join(HERE, _LIB_NAME))


def load_library():
    """Load libicm_hip.so and bind every C-ABI symbol (raises ICMError)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.isfile(path):
        raise ICMError(
            f"Cannot find {_LIB_NAME} at {path}. Build it with 'python -m icm_amd.build'.")
    try:
        lib = ctypes.CDLL(path)
    except OSError as e:
        raise ICMError(f"Failed to load ICM library from {path}: {e}") from e
    for name, res, args in _SIGS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _err(lib, what):
    msg = lib.icm_hip_last_error()
    return ICMError(f"{what} failed: {msg.decode() if msg else 'unknown error'}")


def _to_double_array(data, name):
    arr = np.ascontiguousarray(data, dtype=np.float64)
    if arr.ndim != 1:
        raise ValueError(f"{name} must be a 1-D array, got shape {arr.shape}")
    if arr.size == 0:
        raise ValueError(f"{name} must not be empty")
    return arr


def _to_int_array(data, name):
    arr = np.ascontiguousarray(data, dtype=np.int32)
    if arr.ndim != 1:
        raise ValueError(f"{name} must be a 1-D array, got shape {arr.shape}")
    if arr.size == 0:
        raise ValueError(f"{name} must not be empty")
    return arr


def _ptr(arr, ctype=ctypes.c_double):
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


_initialized = False


def init(wisdom_path=None, device_id=0):
    """Select the GPU (python/icm/__init__.py:155-178; wisdom_path is accepted
    for signature compatibility and ignored: no FFT plans are loaded)."""
    global _initialized
    lib = load_library()
    if lib.icm_hip_init(int(device_id)) != 1:
        raise _err(lib, "icm_hip_init")
    _initialized = True


def _auto_init():
    if not _initialized:
        init()


def _options(verbose=False, graphs=True, memory_strategy=0):
    """IcmHipOptions with the reference's defaults (-1 = auto for the level counts)."""
    o = Options()
    o.enable_graphs = int(bool(graphs))
    o.memory_strategy = int(memory_strategy)
    o.force_uncached_fused_levels = -1
    o.force_uncached_cufft_levels = -1
    o.verbose = int(bool(verbose))
    return o


def _params(block_size=0, q_batch=0, mem_fraction=0.0, mem_budget_bytes=0):
    p = PlanParams()
    p.block_size = int(block_size)
    p.q_batch = int(q_batch)
    p.mem_fraction = float(mem_fraction)
    p.mem_budget_bytes = int(mem_budget_bytes)
    return p


def _validate(n, k, Q):
    if k < 1 or k > n:
        raise ValueError(f"payouts length k={k} must satisfy 1 <= k <= n={n}")
    if Q < 1:
        raise ValueError(f"Q must be positive, got {Q}")


def equity(stacks, payouts, Q=256):
    """ICM equities for all players (python/icm/__init__.py:191-235)."""
    _auto_init()
    s = _to_double_array(stacks, "stacks")
    p = _to_double_array(payouts, "payouts")
    n, k = len(s), len(p)
    _validate(n, k, Q)
    out = np.zeros(n, dtype=np.float64)
    lib = load_library()
    if lib.icm_hip_equity(n, _ptr(s), int(Q), _ptr(p), k, _ptr(out)) != 0:
        raise _err(lib, "icm_hip_equity")
    return out


def equity_subset(stacks, payouts, targets, Q=256):
    """Equities for target players only; others are 0 (python/icm/__init__.py:238-292)."""
    _auto_init()
    s = _to_double_array(stacks, "stacks")
    p = _to_double_array(payouts, "payouts")
    t = _to_int_array(targets, "targets")
    n, k = len(s), len(p)
    _validate(n, k, Q)
    if len(t) < 1:
        raise ValueError("targets must not be empty")
    if np.any(t < 0) or np.any(t >= n):
        raise ValueError(f"target indices must be in [0, {n-1}]")
    out = np.zeros(n, dtype=np.float64)
    lib = load_library()
    if lib.icm_hip_equity_subset(n, _ptr(s), int(Q), _ptr(p), k, _ptr(out),
                                 _ptr(t, ctypes.c_int), len(t)) != 0:
        raise _err(lib, "icm_hip_equity_subset")
    return out


def node_inner(a, payouts):
    """Psi_i of one quadrature node for given a_i (GPU tree, given order)."""
    _auto_init()
    a = _to_double_array(a, "a")
    p = _to_double_array(payouts, "payouts")
    n, k = len(a), len(p)
    _validate(n, k, 1)
    out = np.zeros(n)
    lib = load_library()
    if lib.icm_hip_node_inner(n, _ptr(a), _ptr(p), k, _ptr(out)) != 0:
        raise _err(lib, "icm_hip_node_inner")
    return out


def equity_vjp(stacks, payouts, v, Q=256):
    """(equity, dS) with dS = d(v . equity)/dS, quadrature nodes held fixed."""
    _auto_init()
    s = _to_double_array(stacks, "stacks")
    p = _to_double_array(payouts, "payouts")
    vv = _to_double_array(v, "v")
    n, k = len(s), len(p)
    _validate(n, k, Q)
    if len(vv) != n:
        raise ValueError(f"cotangent length {len(vv)} != n={n}")
    eq = np.zeros(n)
    ds = np.zeros(n)
    lib = load_library()
    if lib.icm_hip_equity_vjp(n, _ptr(s), int(Q), _ptr(p), k, _ptr(vv), _ptr(eq), _ptr(ds)) != 0:
        raise _err(lib, "icm_hip_equity_vjp")
    return eq, ds


def node_vjp(a, c, payouts):
    """One node: (Psi_i, H_i = sum_j c_j dPsi_j/da_i) for given a, c (given order)."""
    _auto_init()
    a = _to_double_array(a, "a")
    c = _to_double_array(c, "c")
    p = _to_double_array(payouts, "payouts")
    n, k = len(a), len(p)
    _validate(n, k, 1)
    psi = np.zeros(n)
    h = np.zeros(n)
    lib = load_library()
    if lib.icm_hip_node_vjp(n, _ptr(a), _ptr(c), _ptr(p), k, _ptr(psi), _ptr(h)) != 0:
        raise _err(lib, "icm_hip_node_vjp")
    return psi, h


class Plan:
    """Reusable plan (IcmGpuPlan, src/gpu/icm_gpu.h:79-111): sort, geometry and
    device memory are set up once for (stacks, k); ``run`` takes Q and payout."""

    def __init__(self, stacks, k, block_size=0, q_batch=0, verbose=False, schoolbook_max=None,
                 mem_fraction=0.0, mem_budget_bytes=0, memory_strategy=0):
        _auto_init()
        self._lib = load_library()
        s = _to_double_array(stacks, "stacks")
        self.n, self.k = len(s), int(k)
        _validate(self.n, self.k, 1)
        opts = _options(verbose=verbose, memory_strategy=memory_strategy)
        prm = _params(block_size, q_batch, mem_fraction, mem_budget_bytes)
        self._h = self._lib.icm_hip_plan_create_ex(self.n, _ptr(s), self.k, ctypes.byref(opts), ctypes.byref(prm))
        if not self._h:
            raise _err(self._lib, "icm_hip_plan_create")
        self.stats = RunStats()
        if schoolbook_max is not None:
            self.set_schoolbook_max(schoolbook_max)

    def set_schoolbook_max(self, cmax):
        """Levels whose children keep <= cmax coefficients run as direct sums."""
        if self._lib.icm_hip_plan_set_schoolbook_max(self._h, int(cmax)) != 0:
            raise _err(self._lib, "icm_hip_plan_set_schoolbook_max")

    @property
    def schoolbook_max(self):
        return self._lib.icm_hip_plan_get_schoolbook_max(self._h)

    def set_small_path(self, mode):
        """-1: tuned choice; 0: never; 1: whenever the tree fits one workgroup."""
        if self._lib.icm_hip_plan_set_small_path(self._h, int(mode)) != 0:
            raise _err(self._lib, "icm_hip_plan_set_small_path")

    @property
    def small_path(self):
        """True when forward runs go through the one-launch small-instance kernel."""
        return self._lib.icm_hip_plan_get_small_path(self._h) == 1

    def set_column_fusion(self, on):
        """Column fusion of the four-step levels: True / False, or a mask (1
        propagate, 2 build, 4 top build into the root, 8 the F = 8192 build).
        Bits 1, 2 and 4 give results bit-identical to no fusion; bit 8 changes
        the rounding of the F = 8192 level's hand-over (1e-13 relative)."""
        mode = -1 if on is True else (0 if on is False else int(on))
        if self._lib.icm_hip_plan_set_column_fusion(self._h, mode) != 0:
            raise _err(self._lib, "icm_hip_plan_set_column_fusion")

    def set_poison(self, on):
        """Diagnostic: NaN-fill every buffer a q-batch writes before it reads,
        before each q-batch, so that a read of a position the batch has not
        written shows up as NaN in the result."""
        if self._lib.icm_hip_plan_set_poison(self._h, int(bool(on))) != 0:
            raise _err(self._lib, "icm_hip_plan_set_poison")

    def set_adjoint_checkpoint(self, interval):
        """Adjoint memory: keep the tangent products of every ``interval``-th
        tree level and recompute the rest in the reverse sweep (>= 2), keep
        every level (0), or choose from free memory (-1)."""
        if self._lib.icm_hip_plan_set_adjoint_checkpoint(self._h, int(interval)) != 0:
            raise _err(self._lib, "icm_hip_plan_set_adjoint_checkpoint")

    @property
    def adjoint_layout(self):
        """(checkpoint interval in use, q-planes per adjoint batch), or None
        before the first adjoint call."""
        m, q = ctypes.c_int(0), ctypes.c_int(0)
        if self._lib.icm_hip_plan_adjoint_layout(self._h, ctypes.byref(m), ctypes.byref(q)) != 0:
            return None
        return m.value, q.value

    @property
    def poisoned_bytes(self):
        """Bytes NaN-filled by the last run (0 with poisoning off)."""
        return self._lib.icm_hip_plan_poisoned_bytes(self._h)

    def set_spectrum_cache(self, on):
        """Keep (True) or recompute (False) the build's column transforms of
        the four-step levels' children for the propagate sweep."""
        if self._lib.icm_hip_plan_set_spectrum_cache(self._h, int(bool(on))) != 0:
            raise _err(self._lib, "icm_hip_plan_set_spectrum_cache")

    @property
    def spectrum_cache(self):
        return bool(self._lib.icm_hip_plan_get_spectrum_cache(self._h) == 1)

    def summary(self):
        sm = PlanSummary()
        if self._lib.icm_hip_plan_summary(self._h, ctypes.byref(sm)) != 0:
            raise _err(self._lib, "icm_hip_plan_summary")
        return {f: getattr(sm, f) for f, _ in PlanSummary._fields_}

    def run(self, payouts, Q=256):
        p = _to_double_array(payouts, "payouts")
        if len(p) != self.k:
            raise ValueError(f"payouts length {len(p)} != plan k={self.k}")
        if Q < 1:
            raise ValueError(f"Q must be positive, got {Q}")
        out = np.zeros(self.n)
        if self._lib.icm_hip_equity_with_plan(self._h, int(Q), _ptr(p), _ptr(out),
                                              ctypes.byref(self.stats)) != 0:
            raise _err(self._lib, "icm_hip_equity_with_plan")
        return out

    def run_device(self, d_payout_ptr, d_equity_ptr, Q=256):
        """Device-resident run: raw device pointers (e.g. torch .data_ptr())."""
        if self._lib.icm_hip_equity_with_plan_device(self._h, int(Q), ctypes.c_void_p(d_payout_ptr),
                                                     ctypes.c_void_p(d_equity_ptr),
                                                     ctypes.byref(self.stats)) != 0:
            raise _err(self._lib, "icm_hip_equity_with_plan_device")
        return self.stats.total_ns

    def vjp(self, payouts, v, Q=256):
        """(equity, dS) for cotangent v (host arrays)."""
        p = _to_double_array(payouts, "payouts")
        vv = _to_double_array(v, "v")
        if len(p) != self.k or len(vv) != self.n:
            raise ValueError("payouts / cotangent length mismatch")
        eq = np.zeros(self.n)
        ds = np.zeros(self.n)
        if self._lib.icm_hip_vjp_with_plan(self._h, int(Q), _ptr(p), _ptr(vv), _ptr(eq), _ptr(ds),
                                           ctypes.byref(self.stats)) != 0:
            raise _err(self._lib, "icm_hip_vjp_with_plan")
        return eq, ds

    def vjp_device(self, d_payout_ptr, d_v_ptr, d_equity_ptr, d_ds_ptr, Q=256):
        """Device-resident adjoint pass (raw device pointers; outputs may be 0)."""
        if self._lib.icm_hip_vjp_with_plan_device(
                self._h, int(Q), ctypes.c_void_p(d_payout_ptr), ctypes.c_void_p(d_v_ptr),
                ctypes.c_void_p(d_equity_ptr or None), ctypes.c_void_p(d_ds_ptr or None),
                ctypes.byref(self.stats)) != 0:
            raise _err(self._lib, "icm_hip_vjp_with_plan_device")
        return self.stats.total_ns

    def set_targets(self, targets=None):
        """Restrict later runs to target players (None: all players)."""
        if targets is None:
            rc = self._lib.icm_hip_plan_set_targets(self._h, None, 0)
        else:
            t = _to_int_array(targets, "targets")
            if np.any(t < 0) or np.any(t >= self.n):
                raise ValueError(f"target indices must be in [0, {self.n-1}]")
            self._keep_targets = t
            rc = self._lib.icm_hip_plan_set_targets(self._h, _ptr(t, ctypes.c_int), len(t))
        if rc != 0:
            raise _err(self._lib, "icm_hip_plan_set_targets")

    def set_profiling(self, on=True):
        """Per-kernel event timing for later runs (clears the aggregates)."""
        if self._lib.icm_hip_plan_set_profiling(self._h, int(bool(on))) != 0:
            raise _err(self._lib, "icm_hip_plan_set_profiling")

    def kernel_stats(self):
        """{name: (launches, total_ns, alg_bytes, alg_flops)} since set_profiling."""
        out = {}
        buf = ctypes.create_string_buffer(64)
        for i in range(self._lib.icm_hip_plan_kernel_count(self._h)):
            ln = ctypes.c_int(0)
            ns, by, fl = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
            self._lib.icm_hip_plan_kernel_stat(self._h, i, buf, 64, ctypes.byref(ln), ctypes.byref(ns),
                                               ctypes.byref(by), ctypes.byref(fl))
            out[buf.value.decode()] = (ln.value, ns.value, by.value, fl.value)
        return out

    def launch_log(self):
        """[(name, level, ns, alg_bytes)] of every profiled launch (level > 0:
        build of that tree level, < 0: its propagate, 0: leaves)."""
        out = []
        buf = ctypes.create_string_buffer(64)
        lv = ctypes.c_int(0)
        ns, by = ctypes.c_double(0), ctypes.c_double(0)
        for i in range(self._lib.icm_hip_plan_launch_count(self._h)):
            self._lib.icm_hip_plan_launch_stat(self._h, i, buf, 64, ctypes.byref(lv), ctypes.byref(ns),
                                               ctypes.byref(by))
            out.append((buf.value.decode(), lv.value, ns.value, by.value))
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.icm_hip_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ragged(stacks_list, payouts_list):
    if len(stacks_list) != len(payouts_list) or len(stacks_list) == 0:
        raise ValueError("need one payout vector per instance and at least one instance")
    ss = [_to_double_array(s, "stacks") for s in stacks_list]
    ps = [_to_double_array(p, "payouts") for p in payouts_list]
    for s, p in zip(ss, ps):
        _validate(len(s), len(p), 1)
    n_off = np.zeros(len(ss) + 1, dtype=np.int32)
    k_off = np.zeros(len(ss) + 1, dtype=np.int32)
    n_off[1:] = np.cumsum([len(s) for s in ss])
    k_off[1:] = np.cumsum([len(p) for p in ps])
    return np.concatenate(ss), np.concatenate(ps), n_off, k_off


class BatchPlan:
    """Ragged batch of independent instances (instance i: stacks_list[i] and
    k_list[i] payouts).  ``run(payouts_list, Q)`` returns one equity array per
    instance; the first run for a Q is recorded as a HIP graph and replayed
    by later runs."""

    def __init__(self, stacks_list, k_list, block_size=0, q_batch=0, graphs=True, verbose=False,
                 mem_fraction=0.0, mem_budget_bytes=0):
        _auto_init()
        self._lib = load_library()
        if len(stacks_list) != len(k_list) or len(stacks_list) == 0:
            raise ValueError("need one k per instance and at least one instance")
        ss = [_to_double_array(s, "stacks") for s in stacks_list]
        for s, k in zip(ss, k_list):
            _validate(len(s), int(k), 1)
        self.n_off = np.zeros(len(ss) + 1, dtype=np.int32)
        self.k_off = np.zeros(len(ss) + 1, dtype=np.int32)
        self.n_off[1:] = np.cumsum([len(s) for s in ss])
        self.k_off[1:] = np.cumsum([int(k) for k in k_list])
        S = np.concatenate(ss)
        opts = _options(verbose=verbose, graphs=graphs)
        prm = _params(block_size, q_batch, mem_fraction, mem_budget_bytes)
        self._h = self._lib.icm_hip_batch_plan_create_ex(len(ss), _ptr(self.n_off, ctypes.c_int), _ptr(S),
                                                         _ptr(self.k_off, ctypes.c_int), ctypes.byref(opts),
                                                         ctypes.byref(prm))
        if not self._h:
            raise _err(self._lib, "icm_hip_batch_plan_create")
        self.n_inst = len(ss)
        self.stats = RunStats()

    def split(self, flat):
        return [flat[self.n_off[i]:self.n_off[i + 1]] for i in range(self.n_inst)]

    def run_flat(self, payout_flat, Q=256):
        p = _to_double_array(payout_flat, "payouts")
        if len(p) != self.k_off[-1]:
            raise ValueError(f"flat payout length {len(p)} != {self.k_off[-1]}")
        if Q < 1:
            raise ValueError(f"Q must be positive, got {Q}")
        out = np.zeros(int(self.n_off[-1]))
        if self._lib.icm_hip_batch_equity(self._h, int(Q), _ptr(p), _ptr(out), ctypes.byref(self.stats)) != 0:
            raise _err(self._lib, "icm_hip_batch_equity")
        return out

    def run(self, payouts_list, Q=256):
        if len(payouts_list) != self.n_inst:
            raise ValueError("need one payout vector per instance")
        ps = [_to_double_array(p, "payouts") for p in payouts_list]
        for i, p in enumerate(ps):
            if len(p) != self.k_off[i + 1] - self.k_off[i]:
                raise ValueError(f"instance {i}: payout length {len(p)} != k")
        return self.split(self.run_flat(np.concatenate(ps), Q))

    def vjp_flat(self, payout_flat, v_flat, Q=256):
        """(equity, dS) flat arrays for a flat cotangent v (same order as the stacks)."""
        p = _to_double_array(payout_flat, "payouts")
        v = _to_double_array(v_flat, "v")
        if len(p) != self.k_off[-1] or len(v) != self.n_off[-1]:
            raise ValueError("flat payout / cotangent length mismatch")
        if Q < 1:
            raise ValueError(f"Q must be positive, got {Q}")
        eq = np.zeros(int(self.n_off[-1]))
        ds = np.zeros(int(self.n_off[-1]))
        if self._lib.icm_hip_batch_vjp(self._h, int(Q), _ptr(p), _ptr(v), _ptr(eq), _ptr(ds),
                                       ctypes.byref(self.stats)) != 0:
            raise _err(self._lib, "icm_hip_batch_vjp")
        return eq, ds

    def vjp(self, payouts_list, v_list, Q=256):
        """Per-instance (equity, dS) lists for per-instance cotangents."""
        if len(payouts_list) != self.n_inst or len(v_list) != self.n_inst:
            raise ValueError("need one payout vector and one cotangent per instance")
        eq, ds = self.vjp_flat(np.concatenate([_to_double_array(p, "payouts") for p in payouts_list]),
                               np.concatenate([_to_double_array(v, "v") for v in v_list]), Q)
        return self.split(eq), self.split(ds)

    def vjp_device(self, d_payout_ptr, d_v_ptr, d_equity_ptr, d_ds_ptr, Q=256):
        if self._lib.icm_hip_batch_vjp_device(self._h, int(Q), ctypes.c_void_p(d_payout_ptr),
                                              ctypes.c_void_p(d_v_ptr), ctypes.c_void_p(d_equity_ptr or None),
                                              ctypes.c_void_p(d_ds_ptr or None), ctypes.byref(self.stats)) != 0:
            raise _err(self._lib, "icm_hip_batch_vjp_device")
        return self.stats.total_ns

    def run_device(self, d_payout_ptr, d_equity_ptr, Q=256):
        if self._lib.icm_hip_batch_equity_device(self._h, int(Q), ctypes.c_void_p(d_payout_ptr),
                                                 ctypes.c_void_p(d_equity_ptr), ctypes.byref(self.stats)) != 0:
            raise _err(self._lib, "icm_hip_batch_equity_device")
        return self.stats.total_ns

    def summary(self):
        vals = [ctypes.c_int(0) for _ in range(5)]
        nbytes = ctypes.c_size_t(0)
        if self._lib.icm_hip_batch_summary(self._h, *[ctypes.byref(v) for v in vals], ctypes.byref(nbytes)) != 0:
            raise _err(self._lib, "icm_hip_batch_summary")
        keys = ("n_inst", "n_players", "n_blocks", "max_levels", "q_batch")
        out = {k: v.value for k, v in zip(keys, vals)}
        out["device_bytes"] = nbytes.value
        return out

    def set_schoolbook_max(self, cmax):
        """Schoolbook / transform crossover of the forward levels."""
        if self._lib.icm_hip_batch_set_schoolbook_max(self._h, int(cmax)) != 0:
            raise _err(self._lib, "icm_hip_batch_set_schoolbook_max")

    def set_persistent(self, on=True):
        """Run the forward batch as one persistent launch per q-pass (block
        size 64) instead of one launch per phase and tree level."""
        if self._lib.icm_hip_batch_set_persistent(self._h, int(bool(on))) != 0:
            raise _err(self._lib, "icm_hip_batch_set_persistent")

    @property
    def persistent(self):
        return bool(self._lib.icm_hip_batch_is_persistent(self._h))

    def set_spin_limit(self, spins):
        """Bound (polls) of a persistent work item's wait for its inputs."""
        if self._lib.icm_hip_batch_set_spin_limit(self._h, int(spins)) != 0:
            raise _err(self._lib, "icm_hip_batch_set_spin_limit")

    def set_profiling(self, on=True):
        if self._lib.icm_hip_batch_set_profiling(self._h, int(bool(on))) != 0:
            raise _err(self._lib, "icm_hip_batch_set_profiling")

    def kernel_stats(self):
        out = {}
        buf = ctypes.create_string_buffer(64)
        for i in range(self._lib.icm_hip_batch_kernel_count(self._h)):
            ln = ctypes.c_int(0)
            ns, by, fl = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
            self._lib.icm_hip_batch_kernel_stat(self._h, i, buf, 64, ctypes.byref(ln), ctypes.byref(ns),
                                                ctypes.byref(by), ctypes.byref(fl))
            out[buf.value.decode()] = (ln.value, ns.value, by.value, fl.value)
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.icm_hip_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def equity_batch(stacks_list, payouts_list, Q=256):
    """One-shot ragged batch: list of equity arrays, one per instance."""
    _auto_init()
    S, P, n_off, k_off = _ragged(stacks_list, payouts_list)
    if Q < 1:
        raise ValueError(f"Q must be positive, got {Q}")
    out = np.zeros(len(S))
    lib = load_library()
    if lib.icm_hip_equity_batch(len(n_off) - 1, _ptr(n_off, ctypes.c_int), _ptr(S), _ptr(k_off, ctypes.c_int),
                                int(Q), _ptr(P), _ptr(out)) != 0:
        raise _err(lib, "icm_hip_equity_batch")
    return [out[n_off[i]:n_off[i + 1]] for i in range(len(n_off) - 1)]


def jacobian_batch(stacks_list, payouts_list, Q=256, graphs=False):
    """Full Jacobians dE/dS (n_i x n_i, row j = dE_j/dS) of every instance,
    from one batched adjoint pass: instance i is replicated n_i times with
    the unit cotangents e_0 .. e_{n_i - 1}."""
    ss = [_to_double_array(s, "stacks") for s in stacks_list]
    ps = [_to_double_array(p, "payouts") for p in payouts_list]
    if len(ss) != len(ps):
        raise ValueError("need one payout vector per instance")
    rep_s, rep_k, rep_p, rep_v = [], [], [], []
    for s, p in zip(ss, ps):
        n = len(s)
        for j in range(n):
            rep_s.append(s)
            rep_k.append(len(p))
            rep_p.append(p)
            e = np.zeros(n)
            e[j] = 1.0
            rep_v.append(e)
    bp = BatchPlan(rep_s, rep_k, graphs=graphs)
    try:
        eq, ds = bp.vjp_flat(np.concatenate(rep_p), np.concatenate(rep_v), Q)
    finally:
        bp.close()
    out, eqs, off = [], [], 0
    for s in ss:
        n = len(s)
        rows = ds[off:off + n * n].reshape(n, n)
        eqs.append(eq[off:off + n].copy())
        out.append(rows.copy())
        off += n * n
    return out, eqs


def plan_geometry(n, k, B=32, max_levels=64):
    """Host-only planner query: per-level (n_real, c, G, F)."""
    lib = load_library()
    arrs = [np.zeros(max_levels, dtype=np.int32) for _ in range(4)]
    L = lib.icm_hip_plan_geometry(int(n), int(k), int(B), max_levels,
                                  *[_ptr(a, ctypes.c_int) for a in arrs])
    if L < 0:
        raise ValueError("invalid geometry arguments")
    return {name: a[:L].copy() for name, a in zip(("n_real", "c", "G", "F"), arrs)}


def make_nodes(Q, Smax):
    """(logv, w) of the erfc-trap rule computed by the library (host code)."""
    lib = load_library()
    logv = np.zeros(Q)
    w = np.zeros(Q)
    lib.icm_hip_make_nodes(int(Q), float(Smax), _ptr(logv), _ptr(w))
    return logv, w


def plan_dry_run(n, k, mem_budget_bytes, block_size=0, q_batch=0, mem_fraction=0.0, memory_strategy=0):
    """Host-only planning against a simulated device with mem_budget_bytes
    free: (summary dict, q-batch halvings).  Raises ICMError when one q-plane
    does not fit."""
    lib = load_library()
    sm = PlanSummary()
    retries = ctypes.c_int(0)
    if lib.icm_hip_plan_dry_run(int(n), int(k), ctypes.byref(_options(memory_strategy=memory_strategy)),
                                ctypes.byref(_params(block_size, q_batch, mem_fraction, mem_budget_bytes)),
                                ctypes.byref(sm), ctypes.byref(retries)) != 0:
        raise _err(lib, "icm_hip_plan_dry_run")
    return {f: getattr(sm, f) for f, _ in PlanSummary._fields_}, retries.value


def measure_hbm_bandwidth_gbps():
    """Device-copy rate of the reference probe (bytes copied per second)."""
    _auto_init()
    lib = load_library()
    v = ctypes.c_double(0)
    if lib.icm_gpu_measure_hbm_bandwidth_gbps(ctypes.byref(v)) != 1:
        raise _err(lib, "icm_gpu_measure_hbm_bandwidth_gbps")
    return v.value


def measure_stream_gbps(nbytes=1 << 30):
    """(read, write, copy) HBM rates of this library's streaming kernels."""
    _auto_init()
    lib = load_library()
    r, w, c = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
    if lib.icm_hip_measure_stream_gbps(int(nbytes), ctypes.byref(r), ctypes.byref(w), ctypes.byref(c)) != 1:
        raise _err(lib, "icm_hip_measure_stream_gbps")
    return r.value, w.value, c.value


def measure_lds_gbps():
    """LDS read rate (bytes delivered to lanes per second) of conflict-free 16-byte reads."""
    _auto_init()
    lib = load_library()
    v = ctypes.c_double(0)
    if lib.icm_hip_measure_lds_gbps(ctypes.byref(v)) != 1:
        raise _err(lib, "icm_hip_measure_lds_gbps")
    return v.value


def measure_fused_pair_ns(fft_n, batch, quick=False):
    """(build, propagate) device ns per parent of one LDS-level launch."""
    _auto_init()
    lib = load_library()
    b, c = ctypes.c_double(0), ctypes.c_double(0)
    if lib.icm_gpu_measure_fused_pair_ns(int(fft_n), int(batch), int(bool(quick)), ctypes.byref(b),
                                         ctypes.byref(c)) != 1:
        raise _err(lib, "icm_gpu_measure_fused_pair_ns")
    return b.value, c.value

"""GPU: icm_hip_equity_noalloc (all device memory from the library's static
scratch array) against icm_hip_equity on the same inputs, two calls in a row
(the second smaller, so stale scratch contents from the first would show),
and its size limits."""
import ctypes

import numpy as np
import pytest

import icm_amd

pytestmark = pytest.mark.gpu
_D = ctypes.POINTER(ctypes.c_double)


def _lib():
    lib = icm_amd.load_library()
    icm_amd.init()
    f = lib.icm_hip_equity_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D]
    return lib


def _both(lib, S, p, Q):
    S = np.ascontiguousarray(S, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    a = np.full(len(S), np.nan)
    b = np.full(len(S), np.nan)
    assert lib.icm_hip_equity_noalloc(len(S), S.ctypes.data_as(_D), Q, p.ctypes.data_as(_D), len(p),
                                      a.ctypes.data_as(_D)) == 0, lib.icm_hip_last_error()
    assert lib.icm_hip_equity(len(S), S.ctypes.data_as(_D), Q, p.ctypes.data_as(_D), len(p),
                              b.ctypes.data_as(_D)) == 0, lib.icm_hip_last_error()
    return a, b


def _rel(a, b):
    den = np.abs(b).copy()
    den[den < 1e-14] = 1.0
    return float(np.max(np.abs(a - b) / den))


def test_noalloc_matches_the_plan_path_twice_shrinking():
    lib = _lib()
    rng = np.random.default_rng(77)
    S1 = rng.integers(1, 10 ** 6, 1024, endpoint=True).astype(np.float64)     # config 1 shape
    p1 = 1.0 / np.arange(1, 11)
    p1 /= p1.sum()
    a1, b1 = _both(lib, S1, p1, 256)
    S2 = rng.uniform(1.0, 100.0, 300)                                          # smaller, more payouts
    p2 = 0.9 ** np.arange(40)
    a2, b2 = _both(lib, S2, p2 / p2.sum(), 64)
    e1, e2 = _rel(a1, b1), _rel(a2, b2)
    print(f"noalloc vs icm_hip_equity: {e1:.2e} (n=1024 k=10 Q=256), {e2:.2e} (n=300 k=40 Q=64)", flush=True)
    assert e1 < 1e-8 and e2 < 1e-8
    assert np.all(np.isfinite(a1)) and np.all(np.isfinite(a2))


def test_noalloc_rejects_sizes_beyond_its_limits():
    lib = _lib()
    out = np.zeros(5000)
    S = np.ones(5000)
    p = np.ones(5)
    assert lib.icm_hip_equity_noalloc(5000, S.ctypes.data_as(_D), 256, p.ctypes.data_as(_D), 5,
                                      out.ctypes.data_as(_D)) == -1
    assert b"n <= 4096" in lib.icm_hip_last_error()
    S = np.ones(4096)
    p = np.ones(4096)
    assert lib.icm_hip_equity_noalloc(4096, S.ctypes.data_as(_D), 256, p.ctypes.data_as(_D), 4096,
                                      out.ctypes.data_as(_D)) == -1
    assert b"LDS" in lib.icm_hip_last_error()
    assert lib.icm_hip_equity_noalloc(100, S.ctypes.data_as(_D), 4096, p.ctypes.data_as(_D), 5,
                                      out.ctypes.data_as(_D)) == -1

"""ctypes loader for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / CPU baseline.  The product
package (icm_amd) never imports this module.

The C functions restate the reference's CPU library (src/cpu/icm.c); each one
cites the lines it follows in oracle/icm_oracle.c.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ENGINE_NAIVE = 0
ENGINE_LINEAR = 1
ENGINE_HYBRID = 2

_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_U8 = ctypes.POINTER(ctypes.c_ubyte)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.orc_make_nodes.argtypes = [ctypes.c_int, ctypes.c_double, _D, _D]
    lib.orc_make_nodes.restype = None
    lib.orc_v1_exact.argtypes = [ctypes.c_int, _D, _D]
    lib.orc_v1_exact.restype = None
    lib.orc_v2_exact.argtypes = [ctypes.c_int, _D, _D]
    lib.orc_v2_exact.restype = None
    lib.orc_inner.argtypes = [ctypes.c_int, _D, _D, ctypes.c_int, _D,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, _U8]
    lib.orc_inner.restype = ctypes.c_int
    lib.orc_equity.argtypes = [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int,
                               _D, _I, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int]
    lib.orc_equity.restype = ctypes.c_int
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(_D)


def make_nodes(Q, Smax):
    """(logv[Q], w[Q]) of the erfc-trap rule (src/cpu/icm.c:117-132)."""
    lib = _load()
    logv = np.zeros(Q)
    w = np.zeros(Q)
    lib.orc_make_nodes(int(Q), float(Smax), _dp(logv), _dp(w))
    return logv, w


def v1_exact(S):
    lib = _load()
    S = np.ascontiguousarray(S, dtype=np.float64)
    out = np.zeros_like(S)
    lib.orc_v1_exact(len(S), _dp(S), _dp(out))
    return out


def v2_exact(S):
    lib = _load()
    S = np.ascontiguousarray(S, dtype=np.float64)
    out = np.zeros_like(S)
    lib.orc_v2_exact(len(S), _dp(S), _dp(out))
    return out


def inner(a, payout, engine=ENGINE_HYBRID, B=32, fft_min=0, active=None):
    """Per-node Psi_i for given a_j (players in the given order)."""
    lib = _load()
    a = np.ascontiguousarray(a, dtype=np.float64)
    payout = np.ascontiguousarray(payout, dtype=np.float64)
    out = np.zeros_like(a)
    act = None
    if active is not None:
        act = np.ascontiguousarray(active, dtype=np.uint8)
    rc = lib.orc_inner(len(a), _dp(a), _dp(payout), len(payout), _dp(out),
                       int(engine), int(B), int(fft_min),
                       act.ctypes.data_as(_U8) if act is not None else None)
    if rc != 0:
        raise ValueError("orc_inner: invalid arguments")
    return out


def equity(S, payout, Q=256, targets=None, engine=ENGINE_HYBRID, B=32,
           fft_min=64, q_lo=0, q_hi=None, nthreads=0):
    """Equities via the restated integration loop (src/cpu/icm.c:2523-2611)."""
    lib = _load()
    S = np.ascontiguousarray(S, dtype=np.float64)
    payout = np.ascontiguousarray(payout, dtype=np.float64)
    n = len(S)
    out = np.zeros(n)
    tg = None
    nt = 0
    if targets is not None:
        tg = np.ascontiguousarray(targets, dtype=np.int32)
        nt = len(tg)
    if q_hi is None:
        q_hi = Q
    rc = lib.orc_equity(n, _dp(S), int(Q), _dp(payout), len(payout), _dp(out),
                        tg.ctypes.data_as(_I) if tg is not None else None, nt,
                        int(engine), int(min(B, n)) if engine == ENGINE_HYBRID else int(B),
                        int(fft_min), int(q_lo), int(q_hi), int(nthreads))
    if rc != 0:
        raise ValueError("orc_equity: invalid arguments")
    return out

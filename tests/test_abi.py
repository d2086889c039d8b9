"""CPU suite: the C-ABI library loads and exports every declared symbol, and
the host-only parts (planner geometry, node rule) behave.  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

import icm_amd
from oracle import oracle as O

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "icm_hip.h")


def _declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", txt)
    return sorted(set(n for n in names if n.startswith("icm_")))


def test_library_exports_every_declared_symbol():
    lib = icm_amd.load_library()
    names = _declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), f"missing export {name}"


def test_python_binding_covers_header():
    bound = {n for n, _, _ in icm_amd._SIGS}
    assert set(_declared_functions()) <= bound


def test_row_pair_index_maps():
    """The four-step row-pair kernels' index maps, checked on the host by the
    library itself: layouts one to one, every row owned by one workgroup,
    representatives covering both tiles once in mirror-bin pairs (F1 = 2 ..
    4096).  A mirror off by one (ib at K2 - r instead of K2 - 1 - r) fails it."""
    lib = icm_amd.load_library()
    assert lib.icm_hip_selfcheck_row_layouts() == 0, lib.icm_hip_last_error()


def test_reference_loader_contract():
    """python/icm/__init__.py binds icm_init/icm_equity/icm_equity_subset by
    name with these argtypes; they must resolve in the HIP library."""
    lib = ctypes.CDLL(icm_amd.library_path())
    lib.icm_init.argtypes = [ctypes.c_char_p]
    lib.icm_equity.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_double)]
    assert lib.icm_equity_subset is not None


def test_node_rule_matches_oracle_bitwise():
    for Smax in (1.0, 37.5, 1e6, 1e9):
        for Q in (2, 17, 256):
            a = icm_amd.make_nodes(Q, Smax)
            b = O.make_nodes(Q, Smax)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("n,k,B", [(1, 1, 16), (10, 3, 16), (1000, 10, 32), (1000, 1000, 32),
                                   (65536, 65536, 64), (1 << 20, 1 << 20, 32), (1 << 20, 1 << 18, 32),
                                   (131072, 1000, 32), (4097, 4097, 32), (300, 299, 16)])
def test_geometry_invariants(n, k, B):
    g = icm_amd.plan_geometry(n, k, B)
    L = len(g["F"])
    nb = (n + B - 1) // B
    assert g["n_real"][0] == nb and g["n_real"][-1] == 1
    assert 2 ** (L - 1) >= nb > 2 ** (L - 2) if L > 1 else nb == 1
    for l in range(L):
        D = B << l
        assert g["c"][l] == min(D + 1, k) and g["G"][l] == min(D, k)
    assert g["G"][-1] == k            # the root g is the payout itself
    for l in range(1, L):
        F, c, G = g["F"][l], g["c"][l - 1], g["G"][l - 1]
        assert F & (F - 1) == 0 and F >= 16
        assert F >= G + c - 1          # propagate correlation never wraps
        assert F >= 2 * c - 2          # build aliases at most coefficient 0
        assert g["c"][l] <= F + 1      # parent fits the transform (+ exact top term)
        if l >= 2:
            assert g["n_real"][l] == (g["n_real"][l - 1] + 1) // 2


def test_geometry_rejects_bad_input():
    for args in ((0, 1, 32), (5, 6, 32), (5, 0, 32), (100, 10, 24)):
        with pytest.raises(ValueError):
            icm_amd.plan_geometry(*args)


def test_python_validation_matches_reference():
    """Validation happens before any GPU call (python/tests/test_icm.py:183-208)."""
    with pytest.raises(ValueError, match="must not be empty"):
        icm_amd._to_double_array([], "stacks")
    with pytest.raises(ValueError, match="k=3 must satisfy 1 <= k <= n=2"):
        icm_amd._validate(2, 3, 256)
    with pytest.raises(ValueError, match="Q must be positive"):
        icm_amd._validate(3, 3, 0)

"""The reference's extended GPU-versus-CPU grid (bench/bench_gpu.cu:65-192,
`bench_gpu verify_ext`) under pytest: n in {64, 128, 256, 1024, 4096, 16384,
65536, 131072}, stack distributions 0-3 with the reference's own generator
(glibc srand(42) / rand()), k in {16, 100, 512, n/2, n} plus 1024 and 2048
at n >= 65536, payout n - m, Q = 256.  Pass: per-player relative error
below 1e-8 with denominator 1 below 1e-14 (bench/bench_gpu.cu:18-23,157).
ICM_GPU_VERIFY_EXT_MAX_N caps n and ICM_GPU_VERIFY_EXT_LITE=1 trims the k
list, as in the reference.  Each case runs through icm_gpu_equity."""
import ctypes
import os
import sys

import numpy as np
import pytest

import icm_amd
from oracle import oracle as O

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from verify_gpu import ext_ks, reference_stacks, rel_err_max  # noqa: E402

pytestmark = pytest.mark.gpu

NS = [64, 128, 256, 1024, 4096, 16384, 65536, 131072]
_cap = int(os.environ.get("ICM_GPU_VERIFY_EXT_MAX_N", "0") or 0)
if _cap > 0 and any(n <= _cap for n in NS):
    NS = [n for n in NS if n <= _cap]
LITE = os.environ.get("ICM_GPU_VERIFY_EXT_LITE", "0") not in ("", "0")
_D = ctypes.POINTER(ctypes.c_double)


@pytest.mark.parametrize("dist", [0, 1, 2, 3])
@pytest.mark.parametrize("n", NS)
def test_verify_ext(n, dist):
    lib = icm_amd.load_library()
    icm_amd.init()
    S = reference_stacks(n, dist)
    worst = 0.0
    for k in sorted(set(ext_ks(n, LITE))):
        p = (n - np.arange(k)).astype(np.float64)
        print(f"verify_ext n={n} dist={dist} k={k}", flush=True)
        cpu = O.equity(S, p, Q=256, fft_min=64)
        gpu = np.zeros(n)
        st = icm_amd.RunStats()
        rc = lib.icm_gpu_equity(n, S.ctypes.data_as(_D), 256, p.ctypes.data_as(_D), k, gpu.ctypes.data_as(_D),
                                None, ctypes.byref(st))
        assert rc == 0, lib.icm_gpu_last_error()
        err, idx = rel_err_max(gpu, cpu)
        worst = max(worst, err)
        print(f"  err={err:.3e} idx={idx} gpu={st.total_ns / 1e6:.2f} ms B={st.B}", flush=True)
        assert err < 1e-8, f"n={n} k={k} dist={dist} idx={idx} cpu={cpu[idx]!r} gpu={gpu[idx]!r}"
    print(f"verify_ext_worst n={n} dist={dist}: {worst:.3e}", flush=True)

"""Build libicm_hip.so (HIP kernels for gfx950 + host C++) in-tree.

Usage: python -m icm_amd.build   (or call build())
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libicm_hip.so")
SOURCES = ["icm_kernels.hip", "icm_host.cpp", "icm_batch.cpp", "icm_probe.cpp", "icm_noalloc.hip"]
ARCH = os.environ.get("ICM_HIP_ARCH", "gfx950")


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.isfile(c):
            return c
    raise RuntimeError("hipcc not found")


def _stale(obj, src):
    """True when obj is older than src or any header beside the sources."""
    if not os.path.isfile(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith((".hpp", ".h"))]
    deps.append(os.path.join(HERE, "..", "include", "icm_hip.h"))
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=False, incremental=False):
    """Compile every source into icm_amd/libicm_hip.so (always recompiles
    unless incremental, which skips objects newer than their inputs)."""
    hipcc = _hipcc()
    objs = []
    tmp = os.path.join(HERE, "_build")
    os.makedirs(tmp, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
        if incremental and not _stale(obj, os.path.join(CSRC, src)):
            objs.append(obj)
            continue
        cmd = [hipcc, "-c", "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
               "-I", os.path.join(HERE, "..", "include"),
               "-x", "hip" if src.endswith(".hip") else "c++",
               os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [hipcc, "-c", "-O3", "-fPIC", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
                   "-I", os.path.join(HERE, "..", "include"),
                   os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    out_tmp = LIB + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", out_tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out_tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))

#!/usr/bin/env python3
"""Benchmark: ICM forward equities at n = 2^20, k = n, Q = 256, FP64, one MI355X.

A step is one full forward pass: the erfc-trap quadrature (256 nodes) over
the product tree, propagate and leaf division for all n players, with the
inputs (stacks inside the plan, payout) already resident in HBM.  The timed
region is exactly --steps passes bracketed by device synchronisation.

Prints ONE JSON line on stdout.  Extra fields:
  roofline       dominant kernel (by summed device time in the timed steps):
                 algorithmic bytes of the reference-path byte model (block
                 size 32 geometry) for the tree levels its launches serve /
                 its mean launch time (HIP events on the plan's stream), per
                 level the same over all of that level's launches, its own
                 compulsory bytes as a second fraction, and the device's
                 measured streaming rates beside the nominal 8 TB/s.
  path_roofline  the whole pass against the same byte model.
  configs_1_2    config 1 (n=1024, k=10) and config 2 (n=65536, k=n), each
                 with its CPU leg (the oracle) timed in the same run.
  cpu_baseline   the CPU oracle (oracle/, a port of the reference CPU engine)
                 on a bounded sample of quadrature nodes of the same workload,
                 extrapolated to Q = 256, on all host threads and on one.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix)
B200_REF_S = 0.50790           # reference fwd n=2^20 k=n Q=256 on B200 (RESULTS.md:552)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(n, k, seed=3):
    rng = np.random.default_rng(seed)
    S = np.exp(rng.uniform(0.0, math.log(1e9), n))     # log-uniform [1, 1e9]
    p = np.arange(k, 0, -1, dtype=np.float64)            # linear payout k - j
    p /= p.sum()
    return S, p


def next_pow2(v):
    p = 1
    while p < v:
        p <<= 1
    return p


def path_model(n, k, B=32):
    """Per-call-node HBM bytes of the reference path's level-materialising
    model: reference geometry (src/cpu/icm.c:1205-1263) with block size 32,
    every level read and written once.  Returns {"total", "fixed",
    "build": {l: bytes}, "prop": {l: bytes}} (levels of that geometry)."""
    nb = (n + B - 1) // B
    N = next_pow2(nb)
    L = int(math.log2(N)) + 1
    r = [nb]
    for _ in range(1, L):
        r.append((r[-1] + 1) // 2)
    psz = [min(B * (1 << (l + 1)), k) for l in range(L)]
    below = [0] + [1 if psz[l] == 2 * psz[l - 1] and psz[l - 1] >= 2 else 0 for l in range(1, L)]
    g = [min(B, psz[0])]
    for l in range(1, L):
        g.append(min(g[-1] + psz[l - 1] - 1, psz[l]))
    peff = [0] * L
    geff = [0] * L
    for l in range(1, L):
        cps = psz[l - 1]
        peff[l] = cps // 2 + 1 if below[l] else cps
        mx = (psz[l - 1] + cps // 2) if below[l] else psz[l]
        geff[l] = min(g[l - 1] + peff[l] - 1, mx)
    fixed = 8.0 * (3 * n + 2 * nb * 33 + nb * g[0])
    build = {l: 8.0 * (r[l - 1] * psz[l - 1] + r[l] * psz[l]) for l in range(1, L - 1)}
    prop = {l: 8.0 * (r[l] * geff[l] + r[l - 1] * (peff[l] + g[l - 1])) for l in range(1, L)}
    return {"total": fixed + sum(build.values()) + sum(prop.values()), "fixed": fixed, "build": build, "prop": prop}


def path_model_bytes(n, k, B=32):
    return path_model(n, k, B)["total"]


def host_info():
    """lscpu model, sockets, cores and threads of the host running the CPU leg."""
    info = {"model": None, "sockets": None, "cores_per_socket": None, "threads_per_core": None,
            "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core"}
        for line in out.splitlines():
            k_, _, v = line.partition(":")
            if k_.strip() in keys:
                v = v.strip()
                info[keys[k_.strip()]] = int(v) if v.isdigit() else v
    except Exception as e:  # reported, not fatal
        info["error"] = str(e)
    return info


def fft_backend():
    """The oracle's products use its own radix-2 FFT; FFTW is named when the
    host has it (it is not linked: the oracle never loads it)."""
    import ctypes.util
    lib = ctypes.util.find_library("fftw3")
    return "own radix-2 FFT (oracle/icm_oracle.c)" + (f"; {lib} present, unused" if lib else "; no FFTW on host")


def cpu_baseline(n, k, Q, S, p, nthreads):
    """The oracle on a bounded sample of quadrature nodes of the workload, on
    nthreads host threads (one node per thread) and on one thread (one node),
    each extrapolated to Q nodes."""
    from oracle import oracle as O
    q_lo = Q // 2 - nthreads // 2
    q_hi = q_lo + nthreads
    t0 = time.perf_counter()
    O.equity(S, p, Q=Q, engine=O.ENGINE_HYBRID, B=32, fft_min=64, q_lo=q_lo, q_hi=q_hi, nthreads=nthreads)
    dt = time.perf_counter() - t0
    t_full = dt * Q / (q_hi - q_lo)
    t1 = time.perf_counter()
    O.equity(S, p, Q=Q, engine=O.ENGINE_HYBRID, B=32, fft_min=64, q_lo=Q // 2, q_hi=Q // 2 + 1, nthreads=1)
    dt1 = time.perf_counter() - t1
    return {"value": n / t_full, "unit": "player-equities/s", "cores": nthreads, "kind": "port",
            "sample": f"quadrature nodes {q_lo}..{q_hi - 1} of {Q} at n={n} k={k} "
                      f"({q_hi - q_lo} nodes, one per thread, {dt:.1f} s), extrapolated to Q={Q}",
            "single_thread": {"value": n / (dt1 * Q), "unit": "player-equities/s",
                              "sample": f"quadrature node {Q // 2} on 1 thread ({dt1:.1f} s), extrapolated to Q={Q}"},
            "engine": "hybrid block + subproduct tree (B=32), products with >= 64 terms by FFT",
            "fft_backend": fft_backend(), "host": host_info()}


def roofline_fields(launches, steps, n, k, B, Q, peak_gbs):
    """Roofline of the dominant kernel (by summed device time over the timed
    steps) from the per-launch log [(name, level, ns, alg_bytes)]:
      frac            algorithmic bytes of the reference-path model for the
                      levels the kernel's launches serve / their time / peak;
      frac_compulsory the kernel's own compulsory bytes / time / peak;
      levels          per propagate level served: model bytes over the time
                      of all of that level's launches (column + row + column
                      passes of the four-step levels)."""
    tot = {}
    for name, lv, ns, by in launches:
        t = tot.setdefault(name, [0, 0.0, 0.0])
        t[0] += 1
        t[1] += ns
        t[2] += by
    name, (cnt, ns, by) = max(tot.items(), key=lambda kv: kv[1][1])
    model = path_model(n, k, 32)
    shift = int(round(math.log2(B / 32)))     # this engine's level l = the model's level l + shift
    def level_model(lv):              # model bytes of one level for all Q nodes (one pass)
        lm = abs(lv) + shift
        return Q * (model["prop"].get(lm, 0.0) if lv < 0 else model["build"].get(lm, 0.0))

    served = sorted({lv for nm, lv, _, _ in launches if nm == name})
    # every timed step runs each served level once (in one launch per q-batch)
    per_launch_model = steps * sum(level_model(lv) for lv in served) / cnt
    # per level: every launch with that level tag, summed over the timed steps
    levels = {}
    for lv in served:
        t_lv = sum(t for nm, l2, t, b in launches if l2 == lv) / steps
        levels[str(lv)] = {"model_GB": level_model(lv) / 1e9, "ms": t_lv / 1e6,
                           "frac": level_model(lv) / (t_lv * 1e-9) / 1e9 / peak_gbs,
                           "kernels": sorted({nm for nm, l2, _, _ in launches if l2 == lv})}
    avg_s = ns / cnt * 1e-9
    # every kernel: launches per step, mean duration, compulsory bytes per launch
    kernels = {nm: {"launches_per_step": c // steps, "mean_us": t / c / 1e3, "compulsory_GB": b / c / 1e9,
                    "frac_compulsory": b / t / peak_gbs, "share": t / sum(v[1] for v in tot.values())}
               for nm, (c, t, b) in sorted(tot.items(), key=lambda kv: -kv[1][1])}
    return name, {
        "bound": "hbm", "achieved": per_launch_model / avg_s / 1e9, "peak": peak_gbs, "unit": "GB/s",
        "frac": per_launch_model / avg_s / 1e9 / peak_gbs, "kernel": name,
        "launches_per_step": cnt // steps, "mean_launch_us": avg_s * 1e6,
        "alg_bytes_per_launch": per_launch_model,
        "alg_bytes_basis": f"reference-path model, block size 32 (model level = level + {shift})",
        "compulsory_bytes_per_launch": by / cnt, "frac_compulsory": by / cnt / avg_s / 1e9 / peak_gbs,
        "levels": levels, "kernels": kernels}


def config1_field(icm_amd, reps=50, cpu_threads=16):
    """Config 1 (n=1024, k=10, Q=256, stacks uniform integers in [1, 1e6],
    payout 1/(j+1) normalised): device and wall time per call with the plan
    reused, and the oracle's linear engine (the reference CPU's choice at
    k=10) on all threads and on one, in the same run; parity in the same run."""
    from oracle import oracle as O
    rng = np.random.default_rng(1)
    S = rng.integers(1, 10 ** 6, 1024, endpoint=True).astype(np.float64)
    p = 1.0 / np.arange(1, 11)
    p /= p.sum()
    pl = icm_amd.Plan(S, len(p))
    small = pl.small_path
    for _ in range(5):
        pl.run(p, 256)
    dev, wall = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        e = pl.run(p, 256)
        wall.append(time.perf_counter() - t0)
        dev.append(pl.stats.total_ns / 1e6)
    pl.close()

    def cpu_ms(nt, r=5):
        ts = []
        for _ in range(r):
            t0 = time.perf_counter()
            ref = O.equity(S, p, Q=256, engine=O.ENGINE_LINEAR, nthreads=nt)
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3, ref

    cpu_all, ref = cpu_ms(cpu_threads)
    cpu_one, _ = cpu_ms(1)
    den = np.abs(ref)
    den[den < 1e-14] = 1.0
    return {"workload": "n=1024 k=10 Q=256", "small_path": bool(small),
            "gpu_device_ms": float(np.median(dev)), "gpu_wall_ms": float(np.median(wall)) * 1e3,
            "cpu_ms": cpu_all, "cpu_threads": cpu_threads, "cpu_single_thread_ms": cpu_one,
            "cpu_kind": "port (oracle linear engine)", "max_rel_err_vs_cpu": float(np.max(np.abs(e - ref) / den)),
            "reference_published": {"cpu_16t_ms": 0.129, "cpu_1t_ms": 1.33, "b200_k64_ms": 0.153}}


def config2_field(icm_amd, reps=10, cpu_threads=16):
    """Config 2 (n=65536, k=n, lognormal(0,1) stacks scaled to 1e9 chips,
    payout 0.8^j normalised, Q=256): GPU (plan reused) and the oracle on the
    host threads in the same run, with their agreement."""
    from oracle import oracle as O
    rng = np.random.default_rng(2)
    n = 65536
    S = rng.lognormal(0.0, 1.0, n)
    S *= 1e9 / S.sum()
    p = 0.8 ** np.arange(n)
    p /= p.sum()
    pl = icm_amd.Plan(S, n)
    pl.run(p, 256)
    dev, wall = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        e = pl.run(p, 256)
        wall.append(time.perf_counter() - t0)
        dev.append(pl.stats.total_ns / 1e6)
    pl.close()
    t0 = time.perf_counter()
    ref = O.equity(S, p, Q=256, fft_min=64, nthreads=cpu_threads)
    cpu_s = time.perf_counter() - t0
    den = np.abs(ref)
    den[den < 1e-14] = 1.0
    gpu_ms = float(np.median(dev))
    return {"workload": "n=65536 k=65536 Q=256", "gpu_device_ms": gpu_ms,
            "gpu_wall_ms": float(np.median(wall)) * 1e3, "player_equities_per_s": n / (gpu_ms * 1e-3),
            "cpu_ms": cpu_s * 1e3, "cpu_threads": cpu_threads, "cpu_kind": "port (oracle hybrid engine)",
            "gpu_over_cpu": cpu_s * 1e3 / gpu_ms, "max_rel_err_vs_cpu": float(np.max(np.abs(e - ref) / den)),
            "reference_published": {"b200_ms": 20.26, "cpu_16t_ms": 862.0}}


# ragged batch: 4096 independent instances (n log-uniform [2, 4096],
# k in {1, 3, 9, ceil(n/10), n}), one graph launch per batch, against
# looping the plain drop-in (icm_hip_equity per instance) on a sample
def batched_field(icm_amd, torch, dev, Q, reps=5, sample=256):
    from icm_amd.workloads import ragged_instances
    stacks, pays = ragged_instances(4096, seed=4)
    ks = [len(x) for x in pays]
    t_c0 = time.perf_counter()
    bp = icm_amd.BatchPlan(stacks, ks)
    create_ms = (time.perf_counter() - t_c0) * 1e3
    dP = torch.tensor(np.concatenate(pays), dtype=torch.float64, device=dev)
    dE = torch.empty(int(bp.n_off[-1]), dtype=torch.float64, device=dev)
    bp.run_device(dP.data_ptr(), dE.data_ptr(), Q)          # records the graph
    torch.cuda.synchronize()
    t_a = time.perf_counter()
    for _ in range(reps):
        bp.run_device(dP.data_ptr(), dE.data_ptr(), Q)
    torch.cuda.synchronize()
    batch_ms = (time.perf_counter() - t_a) * 1e3 / reps
    # the same batch as one persistent launch per q-pass (opt-in mode)
    persist = None
    try:
        bp.set_persistent(True)
        bp.run_device(dP.data_ptr(), dE.data_ptr(), Q)      # records its graph
        torch.cuda.synchronize()
        t_p = time.perf_counter()
        for _ in range(reps):
            bp.run_device(dP.data_ptr(), dE.data_ptr(), Q)
        torch.cuda.synchronize()
        pms = (time.perf_counter() - t_p) * 1e3 / reps
        persist = {"ms_per_batch": pms, "instances_per_s": len(stacks) / (pms * 1e-3),
                   "vs_per_level_graph": batch_ms / pms,
                   "launches_per_batch": 1, "queues": "one per XCD (8)"}
        bp.set_persistent(False)
    except icm_amd.ICMError as e:
        persist = {"error": str(e)}
    bp.set_profiling(True)
    bp.run_device(dP.data_ptr(), dE.data_ptr(), Q)
    bst = bp.kernel_stats()
    bp.set_profiling(False)
    summ_b = bp.summary()
    bp.close()
    rng = np.random.default_rng(9)
    idx = rng.choice(len(stacks), sample, replace=False)
    icm_amd.equity(stacks[idx[0]], pays[idx[0]], Q)
    t_l = time.perf_counter()
    for i in idx:
        icm_amd.equity(stacks[i], pays[i], Q)
    loop_s = time.perf_counter() - t_l
    loop_inst_s = sample / loop_s
    # the same sample with one plan per instance created beforehand (plan reuse)
    plans = [icm_amd.Plan(stacks[i], len(pays[i])) for i in idx]
    for pl, i in zip(plans, idx):
        pl.run(pays[i], Q)
    t_r = time.perf_counter()
    for pl, i in zip(plans, idx):
        pl.run(pays[i], Q)
    reuse_s = time.perf_counter() - t_r
    for pl in plans:
        pl.close()
    reuse_inst_s = sample / reuse_s
    inst_s = len(stacks) / (batch_ms * 1e-3)
    kern = {k_: {"launches": v[0], "ms": v[1] / 1e6,
                 "GBps": (v[2] / (v[1] * 1e-9) / 1e9) if v[1] > 0 else 0.0}
            for k_, v in sorted(bst.items(), key=lambda kv: -kv[1][1])}
    return {"instances": len(stacks), "players": summ_b["n_players"], "Q": Q,
            "ms_per_batch": batch_ms, "instances_per_s": inst_s,
            "loop_dropin_instances_per_s": loop_inst_s,
            "loop_sample": f"icm_hip_equity on {sample} random instances of the batch",
            "speedup_vs_loop": inst_s / loop_inst_s,
            "loop_plan_reused_instances_per_s": reuse_inst_s, "speedup_vs_plan_reused_loop": inst_s / reuse_inst_s,
            "plan_create_ms": create_ms,
            "device_bytes": summ_b["device_bytes"], "kernels_eager": kern,
            "mode": "graph of per-level launches (default)", "persistent": persist}


# full 64 x 64 Jacobians of 512 n = 64 instances: one batched adjoint over
# 512 x 64 replicas with unit cotangents (graph replay timed)
def jacobian_field(icm_amd, torch, dev, Q, reps=3):
    rng = np.random.default_rng(512)
    n, count = 64, 512
    stacks = [rng.uniform(1.0, 100.0, n) for _ in range(count)]
    rep_s = [s for s in stacks for _ in range(n)]
    bp = icm_amd.BatchPlan(rep_s, [n] * len(rep_s))
    dP = torch.tensor(np.tile(np.arange(n, 0, -1, dtype=np.float64), len(rep_s)), device=dev)
    dV = torch.tensor(np.tile(np.eye(n).ravel(), count), dtype=torch.float64, device=dev)
    dE = torch.empty(len(rep_s) * n, dtype=torch.float64, device=dev)
    dD = torch.empty_like(dE)
    bp.run_device(dP.data_ptr(), dE.data_ptr(), Q)
    bp.vjp_device(dP.data_ptr(), dV.data_ptr(), dE.data_ptr(), dD.data_ptr(), Q)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        bp.run_device(dP.data_ptr(), dE.data_ptr(), Q)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        bp.vjp_device(dP.data_ptr(), dV.data_ptr(), dE.data_ptr(), dD.data_ptr(), Q)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    bp.close()
    fwd_ms = (t1 - t0) * 1e3 / reps
    jac_ms = (t2 - t1) * 1e3 / reps
    return {"instances": count, "n": n, "replicas": len(rep_s), "jacobians_ms": jac_ms,
            "forward_same_batch_ms": fwd_ms, "ratio": jac_ms / fwd_ms,
            "jacobians_per_s": count / (jac_ms * 1e-3)}


# forward + adjoint: one pass returns E and dS = d(v.E)/dS; ratio to the forward
def fwd_vjp_ratio(icm_amd, torch, dev, Q, Sx, px, reps, checkpoint=None):
    pl = icm_amd.Plan(Sx, len(px))
    if checkpoint is not None:
        pl.set_adjoint_checkpoint(checkpoint)
    dp = torch.tensor(px, dtype=torch.float64, device=dev)
    de = torch.empty(len(Sx), dtype=torch.float64, device=dev)
    dd = torch.empty(len(Sx), dtype=torch.float64, device=dev)
    dv = torch.tensor(np.random.default_rng(1).normal(size=len(Sx)), dtype=torch.float64, device=dev)
    pl.run_device(dp.data_ptr(), de.data_ptr(), Q)
    pl.vjp_device(dp.data_ptr(), dv.data_ptr(), de.data_ptr(), dd.data_ptr(), Q)
    torch.cuda.synchronize()
    t_a = time.perf_counter()
    for _ in range(reps):
        pl.run_device(dp.data_ptr(), de.data_ptr(), Q)
    torch.cuda.synchronize()
    t_b = time.perf_counter()
    for _ in range(reps):
        pl.vjp_device(dp.data_ptr(), dv.data_ptr(), de.data_ptr(), dd.data_ptr(), Q)
    torch.cuda.synchronize()
    t_c = time.perf_counter()
    interval, planes = pl.adjoint_layout
    pl.close()
    fwd_ms = (t_b - t_a) * 1e3 / reps
    vjp_ms = (t_c - t_b) * 1e3 / reps
    return {"fwd_ms": fwd_ms, "fwd_plus_vjp_ms": vjp_ms, "ratio": vjp_ms / fwd_ms,
            "checkpoint_interval": interval, "adjoint_q_planes": planes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--k", type=int, default=0, help="0 = n")
    ap.add_argument("--Q", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--only", choices=["all", "forward", "batched", "small"], default="all",
                    help="forward: the headline pass only (for profiler runs); batched: the ragged batch only; "
                         "small: configs 1 and 2 only")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (a plain copy kernel of known size, "
                         "the scale of profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the path shards over quadrature nodes, see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import torch
    import icm_amd

    n = args.n
    k = args.k or n
    Q = args.Q
    S, p = workload(n, k)
    icm_amd.init()
    dev = torch.device("cuda:0")
    if args.calib_copy_mib > 0:
        cnt = args.calib_copy_mib * (1 << 20) // 8
        src = torch.ones(cnt, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {cnt * 8} bytes read, {cnt * 8} bytes written")
        del src, dst
    if args.only == "small":
        nt = args.cpu_threads or max(1, min(16, os.cpu_count() or 1))
        c1 = config1_field(icm_amd, cpu_threads=nt)
        c2 = config2_field(icm_amd, cpu_threads=nt)
        print(json.dumps({"metric": "config 1 latency (ms) and config 2 throughput", "value": c1["gpu_device_ms"],
                          "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "f64",
                          "configs_1_2": {"config1": c1, "config2": c2}}), flush=True)
        return
    if args.only == "batched":
        batched = batched_field(icm_amd, torch, dev, Q)
        log(f"batched: {batched}")
        print(json.dumps({"metric": "instances/s, ragged batch of 4096 (config 4)",
                          "value": batched["instances_per_s"], "unit": "instances/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f64", "batched": batched}), flush=True)
        return
    plan = icm_amd.Plan(S, k)
    summ = plan.summary()
    log(f"plan: {summ}")
    d_p = torch.tensor(p, dtype=torch.float64, device=dev)
    d_e = torch.empty(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    for i in range(args.warmup):
        plan.run_device(d_p.data_ptr(), d_e.data_ptr(), Q)
        log(f"warmup {i}: {plan.stats.total_ns / 1e6:.2f} ms")
    plan.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        plan.run_device(d_p.data_ptr(), d_e.data_ptr(), Q)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kst = plan.kernel_stats()
    llog = plan.launch_log()
    plan.set_profiling(False)
    eq = d_e.cpu().numpy()
    log(f"sum(E) - sum(payout) = {eq.sum() - p.sum():.3e}")
    ms = (t1 - t0) * 1e3 / args.steps
    value = n * args.steps / (t1 - t0)

    # dominant kernel (summed device time over the timed steps), its roofline
    # against the reference-path byte model and its own compulsory bytes
    tot_kernel_ns = sum(v[1] for v in kst.values())
    name, roof = roofline_fields(llog, args.steps, n, k, summ["B"], Q, HBM_PEAK_GBS)
    traffic = None
    # HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
    # this workload (scaled by a 1 GiB copy), committed as a profile summary
    tfile = os.path.join(ROOT, "profiles", f"traffic_fwd_n{n}_k{k}.md")
    if os.path.isfile(tfile) and Q == 256:
        for line in open(tfile):
            if line.startswith("per_launch_bytes_json:"):
                traffic = json.loads(line.split(":", 1)[1]).get(name)
    launches = kst[name][0]
    roof.update({"traffic": traffic, "share_of_device_time": kst[name][1] / tot_kernel_ns,
                 "flops_per_launch": kst[name][3] / launches,
                 "fp64_frac": kst[name][3] / launches / (roof["mean_launch_us"] * 1e-6) / (FP64_PEAK_TFLOPS * 1e12)})
    # measured streaming rates on this device beside the nominal 8 TB/s
    try:
        r_gbs, w_gbs, c_gbs = icm_amd.measure_stream_gbps(1 << 30)
        copy_gbs = icm_amd.measure_hbm_bandwidth_gbps()
        roof["measured_hbm"] = {"read_GBps": r_gbs, "write_GBps": w_gbs, "copy_GBps": c_gbs,
                                "device_copy_bytes_copied_GBps": copy_gbs,
                                "frac_of_measured_copy": roof["achieved"] / c_gbs}
    except Exception as e:  # reported, not fatal
        roof["measured_hbm"] = {"error": str(e)}
    model_bytes = path_model_bytes(n, k) * Q
    path_roof = {"bound": "hbm", "model_bytes_per_step": model_bytes,
                 "achieved": model_bytes / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": model_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    kernels = {k_: {"launches": v[0], "ms_total": v[1] / 1e6 / args.steps,
                    "GBps": (v[2] / (v[1] * 1e-9) / 1e9) if v[1] > 0 else 0.0}
               for k_, v in sorted(kst.items(), key=lambda kv: -kv[1][1])}
    log("kernels (per step): " + json.dumps(kernels, indent=1))

    subset = adjoint = batched = pcie = None
    if args.only == "all":
        # the same pass through the host-buffer entry (payout upload and equity
        # download over PCIe inside the call)
        plan.run(p, Q)
        t_h = time.perf_counter()
        for _ in range(2):
            plan.run(p, Q)
        host_ms = (time.perf_counter() - t_h) * 1e3 / 2
        pcie = {"entry": "icm_hip_equity_with_plan (host payout and equity)", "ms_per_step": host_ms,
                "value": n / (host_ms * 1e-3)}
        log(f"pcie-inclusive: {pcie}")

        # subset variant of the same workload: 64 random target players, pruned tree
        rng = np.random.default_rng(11)
        targets = rng.choice(n, 64, replace=False)
        plan.set_targets(targets)
        plan.run_device(d_p.data_ptr(), d_e.data_ptr(), Q)
        torch.cuda.synchronize()
        ts0 = time.perf_counter()
        for i in range(args.steps):
            plan.run_device(d_p.data_ptr(), d_e.data_ptr(), Q)
        torch.cuda.synchronize()
        sub_ms = (time.perf_counter() - ts0) * 1e3 / args.steps
        plan.set_targets(None)
        subset = {"targets": 64, "ms_per_step": sub_ms, "speedup_vs_full": ms / sub_ms}
        log(f"subset (64 targets): {sub_ms:.2f} ms/step")

        r5 = np.random.default_rng(5)
        S5 = r5.lognormal(0.0, 1.5, 131072)
        p5 = np.arange(1000, 0, -1, dtype=np.float64)
        p5 /= p5.sum()
        adjoint = {"n131072_k1000": fwd_vjp_ratio(icm_amd, torch, dev, Q, S5, p5, 3)}
    plan.close()
    if args.only == "all":
        adjoint["n1048576_k1048576"] = fwd_vjp_ratio(icm_amd, torch, dev, Q, S, p, 2)
        # the same with the tangent products of every 4th tree level kept and
        # the rest recomputed in the reverse sweep (adjoint checkpointing)
        adjoint["n1048576_k1048576_checkpoint4"] = fwd_vjp_ratio(icm_amd, torch, dev, Q, S, p, 2, checkpoint=4)
        adjoint["jacobians_512_n64"] = jacobian_field(icm_amd, torch, dev, Q)
        log(f"adjoint: {adjoint}")
        batched = batched_field(icm_amd, torch, dev, Q)
        log(f"batched: {batched}")

    nt = args.cpu_threads or max(1, min(16, os.cpu_count() or 1))
    small = None
    if args.only == "all":
        small = {"config1": config1_field(icm_amd, cpu_threads=nt)}
        log(f"config 1: {small['config1']}")
        small["config2"] = config2_field(icm_amd, cpu_threads=nt)
        log(f"config 2: {small['config2']}")

    cpu = None
    if not args.no_cpu_baseline:
        log(f"cpu baseline: {nt} threads ...")
        try:
            cpu = cpu_baseline(n, k, Q, S, p, nt)
        except Exception as e:  # reported, not fatal
            cpu = {"error": str(e)}
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "player-equities/s at n=2^20 k=n Q=256; batched inst/s; fwd+VJP vs fwd time",
        "value": value, "unit": "player-equities/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": (value / (n / B200_REF_S)) if (n == 1 << 20 and k == n and Q == 256) else None,
        "dtype": "f64",
        "data": "synthetic: stacks log-uniform [1,1e9] (seed 3), payout k-j normalized",
        "config": {"workload": f"icm_forward n={n} k={k} Q={Q}", "n": n, "k": k, "Q": Q,
                   "block_size": summ["B"], "q_batch": summ["q_batch"], "parallelism": "single-gpu"},
        "roofline": roof,
        "path_roofline": path_roof,
        "subset": subset,
        "adjoint": adjoint,
        "batched": batched,
        "configs_1_2": small,
        "pcie_inclusive": pcie,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

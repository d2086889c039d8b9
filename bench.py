#!/usr/bin/env python3
"""Benchmark: ICM forward equities at n = 2^20, k = n, Q = 256, FP64, one MI355X.

A step is one full forward pass: the erfc-trap quadrature (256 nodes) over
the product tree, propagate and leaf division for all n players, with the
inputs (stacks inside the plan, payout) already resident in HBM.  The timed
region is exactly --steps passes bracketed by device synchronisation.

Prints ONE JSON line on stdout.  Extra fields:
  roofline       dominant kernel (by summed device time in the timed steps):
                 compulsory bytes (or FLOPs) per launch / mean launch time,
                 both measured live with HIP events on the plan's stream.
  path_roofline  the whole pass against the implementation-independent byte
                 model of the reference path (block size 32 geometry).
  cpu_baseline   the CPU oracle (oracle/, a port of the reference CPU engine)
                 on a bounded sample of quadrature nodes of the same workload,
                 extrapolated to Q = 256.
"""
import argparse
import json
import math
import os
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


def path_model_bytes(n, k, B=32):
    """HBM bytes per call-node of the reference path's level-materialising
    model: reference geometry (src/cpu/icm.c:1205-1263) with block size 32,
    every level read and written once."""
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
    by = 3 * n + 2 * nb * 33 + nb * g[0]
    for l in range(1, L - 1):
        by += r[l - 1] * psz[l - 1] + r[l] * psz[l]
    for l in range(1, L):
        by += r[l] * geff[l] + r[l - 1] * (peff[l] + g[l - 1])
    return 8.0 * by


def cpu_baseline(n, k, Q, S, p, nthreads):
    from oracle import oracle as O
    q_lo = Q // 2 - nthreads // 2
    q_hi = q_lo + nthreads
    t0 = time.perf_counter()
    O.equity(S, p, Q=Q, engine=O.ENGINE_HYBRID, B=32, fft_min=64, q_lo=q_lo, q_hi=q_hi, nthreads=nthreads)
    dt = time.perf_counter() - t0
    t_full = dt * Q / (q_hi - q_lo)
    return {"value": n / t_full, "unit": "player-equities/s", "cores": nthreads, "kind": "port",
            "sample": f"quadrature nodes {q_lo}..{q_hi - 1} of {Q} at n={n} k={k} "
                      f"({q_hi - q_lo} nodes, one per thread, {dt:.1f} s), extrapolated to Q={Q}"}


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
    inst_s = len(stacks) / (batch_ms * 1e-3)
    kern = {k_: {"launches": v[0], "ms": v[1] / 1e6,
                 "GBps": (v[2] / (v[1] * 1e-9) / 1e9) if v[1] > 0 else 0.0}
            for k_, v in sorted(bst.items(), key=lambda kv: -kv[1][1])}
    return {"instances": len(stacks), "players": summ_b["n_players"], "Q": Q,
            "ms_per_batch": batch_ms, "instances_per_s": inst_s,
            "loop_dropin_instances_per_s": loop_inst_s,
            "loop_sample": f"icm_hip_equity on {sample} random instances of the batch",
            "speedup_vs_loop": inst_s / loop_inst_s, "plan_create_ms": create_ms,
            "device_bytes": summ_b["device_bytes"], "kernels_eager": kern}


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
def fwd_vjp_ratio(icm_amd, torch, dev, Q, Sx, px, reps):
    pl = icm_amd.Plan(Sx, len(px))
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
    pl.close()
    fwd_ms = (t_b - t_a) * 1e3 / reps
    vjp_ms = (t_c - t_b) * 1e3 / reps
    return {"fwd_ms": fwd_ms, "fwd_plus_vjp_ms": vjp_ms, "ratio": vjp_ms / fwd_ms}


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
    ap.add_argument("--only", choices=["all", "forward", "batched"], default="all",
                    help="forward: the headline pass only (for profiler runs); batched: the ragged batch only")
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
    plan.set_profiling(False)
    eq = d_e.cpu().numpy()
    log(f"sum(E) - sum(payout) = {eq.sum() - p.sum():.3e}")
    ms = (t1 - t0) * 1e3 / args.steps
    value = n * args.steps / (t1 - t0)

    # dominant kernel (summed device time over the timed steps)
    tot_kernel_ns = sum(v[1] for v in kst.values())
    name, (launches, ns, by, fl) = max(kst.items(), key=lambda kv: kv[1][1])
    avg_s = ns / launches * 1e-9
    bytes_per_launch = by / launches
    flops_per_launch = fl / launches
    hbm_frac = bytes_per_launch / avg_s / (HBM_PEAK_GBS * 1e9)
    fp_frac = flops_per_launch / avg_s / (FP64_PEAK_TFLOPS * 1e12)
    traffic = None
    # HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
    # this workload (scaled by a 1 GiB copy), committed as a profile summary
    tfile = os.path.join(ROOT, "profiles", f"traffic_fwd_n{n}_k{k}.md")
    if os.path.isfile(tfile) and Q == 256:
        for line in open(tfile):
            if line.startswith("per_launch_bytes_json:"):
                traffic = json.loads(line.split(":", 1)[1]).get(name)
    if hbm_frac >= fp_frac:
        roof = {"bound": "hbm", "achieved": bytes_per_launch / avg_s / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": hbm_frac}
    else:
        roof = {"bound": "fp64", "achieved": flops_per_launch / avg_s / 1e12, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": fp_frac}
    roof.update({"traffic": traffic, "kernel": name, "launches_per_step": launches // args.steps,
                 "mean_launch_us": avg_s * 1e6, "alg_bytes_per_launch": bytes_per_launch,
                 "share_of_device_time": ns / tot_kernel_ns})
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
        adjoint["jacobians_512_n64"] = jacobian_field(icm_amd, torch, dev, Q)
        log(f"adjoint: {adjoint}")
        batched = batched_field(icm_amd, torch, dev, Q)
        log(f"batched: {batched}")

    cpu = None
    if not args.no_cpu_baseline:
        nt = args.cpu_threads or max(1, min(16, os.cpu_count() or 1))
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
        "pcie_inclusive": pcie,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the FedNL round engine on one MI355X.

Workload (default): plain FedNL, FP64 throughout, L2-logistic (lambda=1e-3),
142 clients x 350 samples, d=300 (299 binary features at ~4% density +
intercept, ~3% positives, planted logistic model), TopK k=8d on the upper
triangle of each client's Hessian difference, alpha=1 (option 2), FedNL
option b. A step is one FedNL round over all 142 clients (client gradients,
Hessians, compressors, learned-Hessian updates, master aggregation, Cholesky
Newton solve). Inputs are resident in HBM before timing starts.

Prints one JSON line. `value` is rounds/s. Also reported: time to
||grad f||^2 <= 1e-17 from x0 = 0, a roofline record of the dominant kernel
(measured with HIP events on the engine's stream during the timed rounds) and
the CPU baseline: the reference's own building blocks (oracle/_ref) run with
the host's cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-‖∇f‖²≤1e-17 & rounds/s vs plain FP64 GPU; solves/s + p99 latency"
PEAK_FP64_MFMA_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
TOL = 1e-17


def _sync_all(eng):
    eng.sync()
    try:
        import torch  # noqa: F401
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def kernel_roofline(eng, times, prob):
    """Roofline record of the kernel with the largest device time."""
    n, m, d, k, S = eng.n, eng.m, eng.d, eng.k, eng.S
    name, (ms, launches) = max(times.items(), key=lambda kv: kv[1][0])
    avg_s = ms * 1e-3 / max(launches, 1)
    if name == "k_hessian_f64":
        flops = float(n) * m * d * (d + 1)  # upper triangle incl. diagonal, per round
        ach = flops / avg_s / 1e12
        return name, dict(bound="mfma", achieved=ach, peak=PEAK_FP64_MFMA_TFLOPS, unit="TFLOP/s",
                          frac=ach / PEAK_FP64_MFMA_TFLOPS, traffic=None), avg_s
    if name in ("k_select", "k_emit"):
        by = float(n) * (8 * S + 12 * k + 16 * k)
    elif name == "k_margin":
        by = float(n) * (m * d * 8 + m * 8 * 3 + 8 * d)
    elif name == "k_master_h":
        by = float(n) * k * 12 + 16.0 * n * k
    else:
        by = float(S) * 8
    ach = by / avg_s / 1e9
    return name, dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS,
                      traffic=None), avg_s


def cpu_baseline(prob, rounds, cores):
    import oracle as O
    kind = "reference" if O.ref_available() else "port"
    t0 = time.perf_counter()
    if kind == "reference":
        out = O.ref_fednl_run(prob, "topk", 8.0, rounds=rounds, workers=cores, tracking=False)
        sec = out["seconds"]
        per_round = (sec[-1] - sec[0]) / max(rounds - 1, 1) if rounds > 1 else sec[0]
    else:
        f = O.FedNL(prob, "topk", k_mult=8.0)
        ts = []
        for _ in range(rounds):
            a = time.perf_counter()
            f.round(tracking=False)
            ts.append(time.perf_counter() - a)
        per_round = float(np.mean(ts[1:] if len(ts) > 1 else ts))
        cores = 1
    total = time.perf_counter() - t0
    return dict(value=1.0 / per_round, unit="rounds/s", cores=cores, kind=kind,
                sample=f"{rounds} rounds of the same workload (setup included in {total:.1f}s wall), "
                       f"rate from the round timestamps")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--tol-rounds", type=int, default=400, help="round budget of the time-to-tolerance solve")
    ap.add_argument("--cpu-rounds", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration) and exit")
    args = ap.parse_args()

    if args.copy_probe_mb:
        import torch
        n = args.copy_probe_mb * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 3}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the FedNL simulation runs every client "
              "on one device); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    from ufednl_amd import datagen
    from ufednl_amd.engine import Engine

    prob = datagen.config2()
    n, m, d = prob.features.shape

    # time to tolerance from x0 = 0 (device timestamps of each round)
    eng_tol = Engine.from_problem(prob, compressor="topk", k_mult=8.0, tracking=False)
    eng_tol.run(args.tol_rounds, check=False)
    tr = eng_tol.trace()
    g2 = tr["grad_norm"] ** 2
    hit = np.nonzero(g2 <= TOL)[0]
    r_tol = int(hit[0]) if hit.size else None
    t_tol = float(tr["seconds"][r_tol]) if r_tol is not None else None
    final_g2 = float(g2[-1])
    eng_tol.close()

    # throughput
    eng = Engine.from_problem(prob, compressor="topk", k_mult=8.0, tracking=False)
    eng.run(args.warmup, check=False)
    _sync_all(eng)
    eng.kernel_timing(True)
    _sync_all(eng)
    t0 = time.perf_counter()
    eng.enqueue(args.steps)
    _sync_all(eng)
    t1 = time.perf_counter()
    times = eng.kernel_times()
    eng.kernel_timing(False)
    elapsed = t1 - t0
    rps = args.steps / elapsed
    kname, roof, kavg = kernel_roofline(eng, times, prob)

    cpu = None
    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline(prob, args.cpu_rounds, cores)
        except Exception as ex:  # report, do not fail the GPU measurement
            cpu = dict(value=None, unit="rounds/s", cores=cores, kind="unavailable", sample=str(ex))

    line = {
        "metric": METRIC,
        "value": rps,
        "unit": "rounds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded planted-logistic binary features, generated in-process)",
        "config": {"workload": "fednl_topk8d_fp64_n142_m350_d300", "clients": n, "samples_per_client": m,
                   "d": d, "k": eng.k, "compressor": "topk", "lambda": prob.lam, "option": "b", "alpha": eng.alpha,
                   "parallelism": "all clients on one GPU"},
        "time_to_tol_s": t_tol,
        "rounds_to_tol": r_tol,
        "tol": TOL,
        "final_grad_norm_sq": final_g2,
        "kernel_ms_per_round": {k_: v[0] / max(v[1], 1) for k_, v in times.items()},
        "roofline_kernel": kname,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

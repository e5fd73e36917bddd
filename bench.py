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
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_BF16_MFMA_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
TOL = 1e-17
# HBM bytes per launch of each kernel on this workload (rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes scaled by a plain-copy calibration)
TRAFFIC_FILE = "fednl_round_topk_n142_m350_d300_hbm_traffic.txt"


def _sync_all(eng):
    eng.sync()
    try:
        import torch  # noqa: F401
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def algorithmic_work(name, n, m, d, k, S, nnz=None):
    """Algorithmic work of one launch of each round kernel: ("mfma", flops)
    for the FP64 arithmetic kernels, ("hbm", bytes) for the data-movement
    ones. Bytes count each input read once and each output written once.
    nnz: the design's nonzero count (the sparse forms read only those)."""
    if name == "k_hessian_f64":
        # upper triangle incl. diagonal of A_c^T D_c A_c: m FMAs per entry
        return "mfma", float(n) * m * d * (d + 1)
    if name == "k_hessian_sparse":
        # binary designs, sums over co-occurring samples: learned read and
        # Delta written per upper entry, the sample weights D and g read, the
        # client gradient written (the per-entry sample lists, a fixed
        # encoding of the design, are the kernel's own overhead)
        return "hbm", float(n) * (16.0 * S + 16.0 * m + 8.0 * d)
    if name == "k_hessian_f32":
        # the same products in FP32
        return "mfma", float(n) * m * d * (d + 1)
    if name == "k_hessian_bf16":
        # three BF16 products per term (hi*hi, hi*lo, lo*hi)
        return "mfma", 3.0 * n * m * d * (d + 1)
    if name == "k_grad_f64":
        # features once, gradient weights, client gradients written
        return "hbm", float(n) * m * (8.0 * d + 8.0) + float(n) * d * 8.0
    if name in ("k_newton", "k_newton_mc"):
        # Cholesky d^3/3, two triangular solves 2 d^2
        return "mfma", d ** 3 / 3.0 + 2.0 * d * d
    if name == "k_select":
        # |Delta_c| read once, k positions + count written; then in the same
        # workgroup the messages (as k_emit below)
        return "hbm", float(n) * (8.0 * S + 4.0 * k + 4.0) + float(n) * (k * 40.0 + S / 8.0)
    if name == "k_margin":
        # features once, labels, x; sigmoid weights, gradient weights, partials
        return "hbm", float(n) * m * (8.0 * d + 8.0 + 16.0) + 8.0 * d
    if name == "k_margin_sparse":
        # binary designs: the samples' nonzero column ids (2 B each), row
        # offsets and label signs, x; D, g and the -log sigma partials written
        z = float(nnz) if nnz is not None else float(n) * m * d
        return "hbm", 2.0 * z + float(n) * m * (4.0 + 8.0 + 16.0 + 0.5) + 8.0 * d
    if name == "k_emit":
        # Delta and learned at the k positions, learned written back, message
        # (value, flat index, position) and the selection bitmap written
        return "hbm", float(n) * (k * (8.0 + 16.0 + 8.0 + 4.0 + 4.0) + S / 8.0)
    if name == "k_master":
        # selection bitmaps, the selected Delta values, H read and written
        # (W written), client gradients read, the average written
        return "hbm", float(n) * (S / 8.0 + 8.0 * k) + 24.0 * S + float(n) * d * 8.0 + 8.0 * d
    return "hbm", 8.0 * S


def peak_of(name, bound):
    """Published MI355X peak for a kernel's bound, in TFLOP/s or GB/s."""
    if bound == "hbm":
        return PEAK_HBM_GBS
    return {"k_hessian_f32": PEAK_FP32_MFMA_TFLOPS,
            "k_hessian_bf16": PEAK_BF16_MFMA_TFLOPS}.get(name, PEAK_FP64_MFMA_TFLOPS)


def load_traffic(path):
    """HBM bytes per launch from the committed counter summary: lines
    `kernel fetch_kib write_kib fetch_factor write_factor`."""
    out = {}
    try:
        with open(path) as fh:
            for ln in fh:
                f = ln.split()
                if len(f) == 5 and not ln.startswith("#"):
                    kib_r, kib_w, fr, fw = (float(x) for x in f[1:])
                    out[f[0]] = 1024.0 * (kib_r * fr + kib_w * fw)
    except OSError:
        pass
    return out


def kernel_roofline(eng, times, traffic, nnz=None):
    """Roofline record of the kernel with the largest device time."""
    n, m, d, k, S = eng.n, eng.m, eng.d, eng.k, eng.S
    name, (ms, launches) = max(times.items(), key=lambda kv: kv[1][0])
    avg_s = ms * 1e-3 / max(launches, 1)
    bound, work = algorithmic_work(name, n, m, d, k, S, nnz)
    if bound == "mfma":
        ach, unit = work / avg_s / 1e12, "TFLOP/s"
    else:
        ach, unit = work / avg_s / 1e9, "GB/s"
    peak = peak_of(name, bound)
    return name, dict(bound=bound, achieved=ach, peak=peak, unit=unit, frac=ach / peak,
                      traffic=traffic.get(name)), avg_s


def time_to_tol(eng, rounds):
    """Runs `rounds` rounds; returns (round index, seconds) of the first
    ||grad f||^2 <= TOL (None if not reached) and the round rate."""
    eng.run(rounds, check=False)
    tr = eng.trace()
    g2 = tr["grad_norm"] ** 2
    hit = np.nonzero(g2 <= TOL)[0]
    r = int(hit[0]) if hit.size else None
    sec = tr["seconds"]
    rate = (len(sec) - 1) / (sec[-1] - sec[0]) if len(sec) > 1 and sec[-1] > sec[0] else None
    return r, (float(sec[r]) if r is not None else None), rate


def hessian_rel_error(eng_mode, eng_ref, x):
    """Max over clients of ||H_mode(x) - H_fp64(x)||_F / ||H_fp64(x)||_F."""
    a = eng_mode.hessians_at(x)
    b = eng_ref.hessians_at(x)
    d = eng_ref.d
    diag = np.array([j * (j + 3) // 2 for j in range(d)])

    def fro(v):
        return np.sqrt(2.0 * np.sum(v * v, axis=1) - np.sum(v[:, diag] ** 2, axis=1))

    return float(np.max(fro(a - b) / fro(b)))


def _warm(eng, rounds=300):
    """Sustained rounds on an existing engine right before a timed section,
    so that it starts at the GPU's working clocks."""
    if eng is not None:
        eng.run(rounds, check=False)


def small_configs_leg(rounds=1000, warm=None):
    """Round rates of the small workloads, where launch overhead weighs most:
    50 x 221, d=68, TopK k=8d and 80 x 407, d=123 with RandSeqK and TopLEK,
    `rounds` rounds from x0 = 0 enqueued at once, with plain launches and with
    each round replayed from a captured HIP graph (FEDNL_ROUND_GRAPH=1)."""
    from ufednl_amd import datagen
    from ufednl_amd.engine import Engine
    out = {}
    for wname, prob, comp in (("n50_m221_d68_topk", datagen.config1(), "topk"),
                              ("n80_m407_d123_seqk", datagen.config3(), "seqk"),
                              ("n80_m407_d123_toplek", datagen.config3(), "toplek")):
        res = {}
        for form, env in (("plain", "0"), ("graph", "1")):
            old = os.environ.get("FEDNL_ROUND_GRAPH")
            os.environ["FEDNL_ROUND_GRAPH"] = env
            try:
                eng = Engine.from_problem(prob, compressor=comp, k_mult=8.0, tracking=False)
            finally:
                if old is None:
                    del os.environ["FEDNL_ROUND_GRAPH"]
                else:
                    os.environ["FEDNL_ROUND_GRAPH"] = old
            eng.run(5, check=False)  # graph capture and first replays
            _warm(warm)
            _sync_all(eng)
            t0 = time.perf_counter()
            eng.enqueue(rounds)
            _sync_all(eng)
            res[f"rounds_per_s_{form}"] = rounds / (time.perf_counter() - t0)
            res["hessian_form"] = eng.hessian_form
            eng.close()
        res["rounds"] = rounds
        out[wname] = res
        print(f"small configs {wname}: {res}", file=sys.stderr, flush=True)
    return out


def _hessian_stage_us(kt):
    """Per-launch time of the Hessian stage, whichever kernel runs it."""
    name = next(k for k in kt if k.startswith("k_hessian"))
    return 1e3 * kt[name][0] / max(kt[name][1], 1)


def mixed_precision_leg(rounds_c2, rounds_c3, warm=None):
    """FP64 / FP32 / BF16-split A^T D A on the 142 x 350, d=300 TopK workload
    and on the 80 x 407, d=123 RandSeqK / TopLEK workload. "fp64" is the
    engine's default FP64 form (on these binary designs the sparse one),
    "fp64_dense" the FP64 MFMA contraction; the FP32 and BF16 modes run the
    dense contraction."""
    from ufednl_amd import datagen
    from ufednl_amd.engine import Engine
    out = {}
    for wname, prob, comp, rounds in (("n142_m350_d300_topk", datagen.config2(), "topk", rounds_c2),
                                      ("n80_m407_d123_seqk", datagen.config3(), "seqk", rounds_c3),
                                      ("n80_m407_d123_toplek", datagen.config3(), "toplek", rounds_c3)):
        res = {}
        ref = None
        for mode in ("fp64", "fp64_dense", "fp32", "bf16x2"):
            kw = dict(hessian="fp64", hessian_form="dense") if mode == "fp64_dense" else dict(hessian=mode)
            eng = Engine.from_problem(prob, compressor=comp, k_mult=8.0, tracking=False, **kw)
            _warm(warm)
            r, t, rate = time_to_tol(eng, rounds)
            eng.kernel_timing(True)
            eng.run(5, check=False)
            kt = eng.kernel_times()
            eng.kernel_timing(False)
            # the Hessian stage: the contraction of the mode (below FP64 with the sums of the margin
            # pass's gradient partials on the diagonal tiles)
            entry = dict(rounds_to_tol=r, time_to_tol_s=t, rounds_per_s=rate, hessian_form=eng.hessian_form,
                         hessian_stage_us=_hessian_stage_us(kt))
            if mode == "fp64":
                ref = eng
                entry["x_final"] = eng.x()
            else:
                xs = [np.zeros(prob.features.shape[2]), res["fp64"]["x_final"]]
                entry["hessian_rel_err_max"] = max(hessian_rel_error(eng, ref, x) for x in xs)
                entry["final_x_max_abs_diff_vs_fp64"] = float(np.max(np.abs(eng.x() - res["fp64"]["x_final"])))
                eng.close()
            res[mode] = entry
        ref.close()
        for v in res.values():
            v.pop("x_final", None)
        base = res["fp64"]["time_to_tol_s"]
        base_d = res["fp64_dense"]["time_to_tol_s"]
        for mode in ("fp64_dense", "fp32", "bf16x2"):
            t = res[mode]["time_to_tol_s"]
            res[mode]["time_to_tol_speedup_vs_fp64"] = (base / t) if (base and t) else None
            if mode != "fp64_dense":
                res[mode]["time_to_tol_speedup_vs_fp64_dense"] = (base_d / t) if (base_d and t) else None
        out[wname] = res
    return out


def _converged_round(eng):
    """First round index with ||grad f||^2 <= TOL (None if not reached)."""
    g2 = eng.trace()["grad_norm"] ** 2
    hit = np.nonzero(g2 <= TOL)[0]
    return int(hit[0]) if hit.size else None


def batched_leg(count, max_rounds=300, chunk=10, modes=("fp64", "fp32", "bf16x2"), warm=None):
    """count independent dense-Gaussian problems (d~U[32,512], n~U[16,256],
    m~U[100,600], TopK k=8d), per Hessian mode: stepped as one batch (one
    launch per kernel per round, graph replay) in chunks of rounds, members
    leaving once they reach ||grad f||^2 <= 1e-17, until all have or
    max_rounds; solves/s counts every problem over the whole wall time.
    Beside it, the same problems on their own streams: every engine's rounds
    to tolerance enqueued first, then one wait."""
    from ufednl_amd import datagen
    from ufednl_amd.engine import Batch, Engine
    shapes = datagen.batched_shapes(count)
    pool = datagen.gaussian_pool(max(n * m * d for d, n, m in shapes), seed=40)
    probs = [datagen.pooled_problem(f"b{i}", pool, 0, n, m, d, seed=100 + i) for i, (d, n, m) in enumerate(shapes)]
    out = dict(problems=count, shapes="d~U[32,512], n~U[16,256], m~U[100,600], N(0,1) features (float32-exact), "
                                      "TopK k=8d, lambda 1e-3, option b, to ||grad f||^2 <= 1e-17")
    for mode in modes:
        engs = [Engine.from_problem(p, compressor="topk", k_mult=8.0, hessian=mode, tracking=False) for p in probs]
        _warm(warm)
        # members that reached the tolerance leave the batch after each chunk
        # of rounds (the batch is rebuilt, and its round graph captured again,
        # over the rest)
        live = list(engs)
        t0 = time.perf_counter()
        done = 0
        member_rounds = 0
        while live and done < max_rounds:
            b = Batch(live, use_graph=True)
            b.run(chunk)
            b.close()
            done += chunk
            member_rounds += chunk * len(live)
            live = [e for e in live if _converged_round(e) is None]
        wall = time.perf_counter() - t0
        r_tol = [_converged_round(e) for e in engs]
        conv = [r for r in r_tol if r is not None]
        ent = dict(batch_rounds=done, member_rounds=member_rounds, wall_s=wall,
                   member_round_us=1e6 * wall / max(member_rounds, 1), converged=len(conv),
                   unconverged=count - len(conv), solves_per_s=len(conv) / wall,
                   rounds_to_tol_p50=float(np.percentile(conv, 50)) if conv else None,
                   rounds_to_tol_max=max(conv) if conv else None)
        for e in engs:
            e.close()
        if mode == "fp64" and conv:
            # the same problems, each on its own stream, rounds to tolerance
            # enqueued for all before one wait
            engs = [Engine.from_problem(p, compressor="topk", k_mult=8.0, hessian=mode, tracking=False)
                    for p in probs]
            _warm(warm)
            t0 = time.perf_counter()
            for e, r in zip(engs, r_tol):
                e.enqueue((r if r is not None else done) + 1)
            for e in engs:
                e.sync()
            ws = time.perf_counter() - t0
            ent.update(streams_wall_s=ws, streams_solves_per_s=len(conv) / ws, batch_speedup_vs_streams=ws / wall)
            for e in engs:
                e.close()
        out[mode] = ent
    return out


def serving_leg(count, rate, budget_gb, max_rounds, modes=("fp64", "fp32", "bf16x2"), seed=5, warm=None):
    """Poisson arrivals (rate per second, fixed seed) of FedNL solves with
    dense N(0,1) features (float32-exact), n in [32,512], m in [200,1000],
    d log-uniform in [64,2048], TopK k=8d, each run to ||grad f||^2 <= 1e-17;
    continuous batching under a device-memory budget, per Hessian mode.
    Every request's features are a window of one shared pool of draws."""
    from ufednl_amd import datagen
    from ufednl_amd import engine as E
    rng = np.random.default_rng(seed)
    shapes = datagen.serving_shapes(count, seed=seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    t0 = time.perf_counter()
    size = max(n * m * d for d, n, m in shapes) + (1 << 24)
    pool = datagen.gaussian_pool(size, seed=seed)
    probs = []
    for i, (d, n, m) in enumerate(shapes):
        off = int(rng.integers(0, size - n * m * d + 1))
        probs.append(datagen.pooled_problem(f"r{i}", pool, off, n, m, d, seed=1000 + i))
    gen_s = time.perf_counter() - t0
    out = dict(requests=count, arrival_rate_per_s=rate, budget_gb=budget_gb, data_gen_s=gen_s,
               shapes="n~U[32,512], m~U[200,1000], d log-U[64,2048], N(0,1) (float32-exact), TopK k=8d, "
                      "lambda 1e-3, to ||grad f||^2 <= 1e-17",
               d_max=max(d for d, _, _ in shapes))
    for mode in modes:
        reqs = [dict(arrival_s=float(t), problem=p, max_rounds=max_rounds, tol=TOL,
                     opts=dict(compressor="topk", k_mult=8.0, hessian=mode)) for p, t in zip(probs, arrivals)]
        _warm(warm)
        res, _ = E.serve(reqs, hbm_budget_bytes=budget_gb * 1e9, rounds_per_step=4)
        lat = np.array([r["finished_s"] - rq["arrival_s"] for r, rq in zip(res, reqs)])
        ok = [r["status"] == 0 for r in res]
        span = max(r["finished_s"] for r in res) - float(arrivals[0])
        out[mode] = dict(converged=int(sum(ok)), unconverged_d=sorted(p.d for p, o in zip(probs, ok) if not o),
                         solves_per_s=sum(ok) / span if span > 0 else None, span_s=span,
                         latency_p50_s=float(np.percentile(lat, 50)), latency_p99_s=float(np.percentile(lat, 99)),
                         rounds_to_tol_p50=float(np.median([r["rounds_to_tol"] for r in res if r["status"] == 0]))
                         if any(ok) else None)
        print(f"serving {mode}: {out[mode]}", file=sys.stderr, flush=True)
    base = out.get("fp64", {}).get("solves_per_s")
    for mode in modes:
        if mode != "fp64" and base:
            out[mode]["solves_per_s_vs_fp64"] = out[mode]["solves_per_s"] / base
    return out


CPU_BUILD = ("oracle/_ref: the reference's own sources compiled by oracle/Makefile.ref with g++ -O2 -mavx2 -mfma "
             "-ffp-contract=off (the reference's Release build adds -O3 -march=native -flto); client work on "
             "persistent worker threads woken by condition variables (dopt's workers spin and yield); master "
             "consumes clients in order (--debug forcesequpdate)")


def cpu_baseline(prob, rounds, cores, tol_rounds=None):
    """The reference's building blocks (oracle/_ref) on the host's cores: the
    round rate over `rounds` rounds and, with tol_rounds, the time to
    ||grad f||^2 <= 1e-17 from x0 = 0 read from the round timestamps."""
    import oracle as O
    kind = "reference" if O.ref_available() else "port"
    t0 = time.perf_counter()
    extra = {}
    if kind == "reference":
        out = O.ref_fednl_run(prob, "topk", 8.0, rounds=rounds, workers=cores, tracking=False)
        sec = out["seconds"]
        per_round = (sec[-1] - sec[0]) / max(rounds - 1, 1) if rounds > 1 else sec[0]
        if tol_rounds:
            run = O.ref_fednl_run(prob, "topk", 8.0, rounds=tol_rounds, workers=cores, tracking=False)
            hit = np.nonzero(run["grad_norm"] ** 2 <= TOL)[0]
            extra = dict(time_to_tol_s=float(run["seconds"][hit[0]]) if hit.size else None,
                         rounds_to_tol=int(hit[0]) if hit.size else None, tol_rounds_run=tol_rounds)
        extra["build"] = CPU_BUILD
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
    res = dict(value=1.0 / per_round, unit="rounds/s", cores=cores, kind=kind,
               sample=f"{rounds} rounds of the same workload for the rate"
                      + (f", then {tol_rounds} rounds from x0 for the time to tolerance" if tol_rounds else "")
                      + f" ({total:.1f}s wall with setup); rates and times from the round timestamps")
    res.update(extra)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=320,
                    help="rounds timed from x0 = 0 (320: this workload reaches ||grad f||^2 <= 1e-17 "
                         "after about 305, so the rate is that of the solve's own rounds)")
    ap.add_argument("--warmup", type=int, default=500,
                    help="untimed rounds first: the GPU clocks up under sustained load")
    ap.add_argument("--calib-rounds", type=int, default=0,
                    help="rounds from x0 = 0 with every kernel timed (per-kernel table); 0: --steps")
    ap.add_argument("--tol-rounds", type=int, default=400, help="round budget of the time-to-tolerance solve")
    ap.add_argument("--cpu-rounds", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="only the headline, time-to-tolerance and CPU legs")
    ap.add_argument("--legs", default="mixed,small,batched,serving",
                    help="extra legs to run: any of mixed, small, batched, serving")
    ap.add_argument("--batch-problems", type=int, default=64)
    ap.add_argument("--batch-max-rounds", type=int, default=300)
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=4.0, help="requests per second (Poisson)")
    ap.add_argument("--serve-budget-gb", type=float, default=160.0)
    ap.add_argument("--serve-max-rounds", type=int, default=1000)
    ap.add_argument("--hessian", default="fp64", choices=("fp64", "fp32", "bf16x2"),
                    help="Hessian form of the headline engines (fp64: the metric's form; the others for profiling)")
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

    # warmup: sustained rounds on a second engine bring the GPU to its working
    # clocks; that engine stays the warm-up source of the other legs. Then the
    # per-kernel table: the timed region's rounds (from x0 = 0) on a fresh
    # engine with every kernel timed by HIP events on its stream. Order of
    # the round launches: warmup, calibration, 300 warm-up rounds, timed region.
    eng_w = Engine.from_problem(prob, compressor="topk", k_mult=8.0, tracking=False, hessian=args.hessian)
    eng_w.run(args.warmup, check=False)
    calib_rounds = args.calib_rounds or args.steps
    eng_c = Engine.from_problem(prob, compressor="topk", k_mult=8.0, tracking=False, hessian=args.hessian)
    eng_c.kernel_timing(True)
    eng_c.enqueue(calib_rounds)
    _sync_all(eng_c)
    calib = eng_c.kernel_times()
    eng_c.close()
    dominant = max(calib.items(), key=lambda kv: kv[1][0])[0]
    eng = Engine.from_problem(prob, compressor="topk", k_mult=8.0, tracking=False, hessian=args.hessian)
    _warm(eng_w)

    # the timed region: K rounds from x0 = 0 on a fresh engine, HIP events
    # around the dominant kernel only; the time to tolerance comes from the
    # same run's per-round device timestamps
    eng.kernel_timing(True, only=dominant)
    _sync_all(eng)
    t0 = time.perf_counter()
    eng.enqueue(args.steps)
    _sync_all(eng)
    t1 = time.perf_counter()
    times = eng.kernel_times()
    eng.kernel_timing(False)
    elapsed = t1 - t0
    rps = args.steps / elapsed
    tr = eng.trace()
    g2 = tr["grad_norm"] ** 2
    hit = np.nonzero(g2 <= TOL)[0]
    r_tol = int(hit[0]) if hit.size else None
    t_tol = float(tr["seconds"][r_tol]) if r_tol is not None else None
    final_g2 = float(g2[-1])
    traffic = load_traffic(os.path.join(ROOT, "profiles", TRAFFIC_FILE))
    nnz = int(np.count_nonzero(prob.features))
    kname, roof, kavg = kernel_roofline(eng, times, traffic, nnz)
    all_roofs = {}
    for name_, (ms_, nl_) in calib.items():
        b_, w_ = algorithmic_work(name_, eng.n, eng.m, eng.d, eng.k, eng.S, nnz)
        sec_ = ms_ * 1e-3 / max(nl_, 1)
        pk_ = peak_of(name_, b_) * (1e12 if b_ == "mfma" else 1e9)
        all_roofs[name_] = dict(bound=b_, us=sec_ * 1e6, frac=w_ / sec_ / pk_, traffic=traffic.get(name_))

    # the metric's time to tolerance, whatever --steps is: a fresh engine from
    # x0 = 0, rounds enqueued at once, the first round with ||grad f||^2 <=
    # 1e-17 read from the device timestamps
    _warm(eng_w)
    eng_t = Engine.from_problem(prob, compressor="topk", k_mult=8.0, tracking=False)
    r_tol2, t_tol2, rate_tol = time_to_tol(eng_t, args.tol_rounds)
    eng_t.close()
    if t_tol is None:
        r_tol, t_tol = r_tol2, t_tol2

    legs = set() if args.no_extra else {x.strip() for x in args.legs.split(",") if x.strip()}
    extra = {}
    if "mixed" in legs:
        extra["mixed_precision"] = mixed_precision_leg(args.tol_rounds, 500, warm=eng_w)
    if "small" in legs:
        extra["small_configs"] = small_configs_leg(1000, warm=eng_w)
    if "batched" in legs:
        extra["batched"] = batched_leg(args.batch_problems, args.batch_max_rounds, warm=eng_w)
    if "serving" in legs:
        extra["serving"] = serving_leg(args.serve_requests, args.serve_rate, args.serve_budget_gb,
                                       args.serve_max_rounds, warm=eng_w)

    cpu = None
    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline(prob, args.cpu_rounds, cores, (r_tol + 5) if r_tol is not None else None)
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
        "dtype": "f64" if args.hessian == "fp64" else f"f64 ({args.hessian} Hessian contraction)",
        "data": "synthetic (seeded planted-logistic binary features, generated in-process)",
        "config": {"workload": "fednl_topk8d_fp64_n142_m350_d300", "clients": n, "samples_per_client": m,
                   "d": d, "k": eng.k, "compressor": "topk", "lambda": prob.lam, "option": "b", "alpha": eng.alpha,
                   "hessian_form": eng.hessian_form, "parallelism": "all clients on one GPU"},
        "time_to_tol_s": t_tol,
        "rounds_to_tol": r_tol,
        "time_to_tol_leg": {"rounds_to_tol": r_tol2, "time_to_tol_s": t_tol2, "rounds_per_s": rate_tol,
                            "rounds": args.tol_rounds},
        "tol": TOL,
        "final_grad_norm_sq": final_g2,
        "kernel_ms_per_round": {k_: v[0] / max(v[1], 1) for k_, v in calib.items()},
        "calib_rounds": calib_rounds,
        "timed_region": "rounds 0..steps-1 from x0 = 0 after warmup rounds on a second engine, HIP events "
                        "around the roofline kernel only; kernel_ms_per_round: the same rounds from x0 = 0 "
                        "on a third engine with every kernel timed",
        "kernel_roofline_all": all_roofs,
        "roofline_kernel": kname,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own code.

The reference here is dopt's C++ building blocks compiled by
oracle/Makefile.ref into oracle/_ref/libdopt_ref.so (selectors, RNGs,
logistic oracle, Cholesky + substitutions) with the FedNL round
orchestrated as TrainLoop.h does under --debug forcesequpdate
(oracle/ref_harness.cpp). Inputs are seeded; inputs and outputs are both
stored, so the fixtures stand on their own.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from ufednl_amd import datagen  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def ref_topk_packed(vals, d, k):
    L = O.lda(d)
    full = np.zeros(d * L)
    full[O.upper_list(d)] = vals
    out = np.zeros(max(k, len(vals)), dtype=np.uint32)
    c = O.ref().ref_topk(O.dp(full), d, k, O.up(out))
    return out[:c].copy()


def deltas(rng, kind, S):
    if kind == "gauss":
        return rng.standard_normal(S)
    if kind == "ties":
        return rng.integers(-3, 4, S).astype(np.float64)
    if kind == "zeros":
        return np.where(rng.random(S) < 0.5, 0.0, -0.0)
    return np.round(rng.standard_normal(S), 1)


def selectors():
    rng = np.random.default_rng(20240601)
    cases = {}
    n = 0
    for d, k in [(5, 1), (5, 3), (9, 12), (17, 40), (30, 100), (40, 320)]:
        for kind in ("gauss", "ties", "zeros", "round1"):
            S = d * (d + 1) // 2
            v = deltas(rng, kind, S)
            cases[f"topk_{n}_d"] = np.array([d, k])
            cases[f"topk_{n}_in"] = v
            cases[f"topk_{n}_out"] = ref_topk_packed(v, d, k)
            n += 1
    cases["topk_count"] = np.array([n])
    # TopLEK: one worker stream of 6 calls per case (stream starts at seed 1234)
    m = 0
    for d, k in [(6, 5), (12, 30), (20, 60), (30, 150)]:
        for kind in ("gauss", "ties"):
            S = d * (d + 1) // 2
            L = O.lda(d)
            vals = np.stack([deltas(rng, kind, S) for _ in range(6)])
            mats = np.stack([O.unpack_full(v, d) for v in vals])
            oi = np.zeros(6 * k, dtype=np.uint32)
            oc = np.zeros(6, dtype=np.int32)
            O.ref().ref_toplek_stream(6, O.dp(np.ascontiguousarray(mats.reshape(-1))), d, k, k * (1.0 / S),
                                      O.up(oi), O.ip(oc))
            cases[f"toplek_{m}_d"] = np.array([d, k])
            cases[f"toplek_{m}_in"] = vals
            cases[f"toplek_{m}_cnt"] = oc
            cases[f"toplek_{m}_out"] = oi.reshape(6, k)
            m += 1
    cases["toplek_count"] = np.array([m])
    # RandK / RandSeqK: 4 rounds of one client generator
    q = 0
    for d, k in [(5, 3), (10, 20), (23, 100), (40, 320)]:
        for seed in (123, 130, 977):
            for name, fn in (("randk", O.ref().ref_randk), ("seqk", O.ref().ref_randseqk)):
                out = np.zeros(4 * k, dtype=np.uint32)
                fn(seed, 4, d, k, O.up(out))
                cases[f"{name}_{q}_dks"] = np.array([d, k, seed])
                cases[f"{name}_{q}_out"] = out.reshape(4, k)
            q += 1
    cases["rand_count"] = np.array([q])
    # LCG streams
    for seed in (0, 1, 123, 1234, 4294967295):
        oi = np.zeros(64, dtype=np.uint32)
        orl = np.zeros(64)
        O.ref().ref_lcg_integer(seed, 64, O.up(oi), O.dp(orl))
        rr = np.zeros(64)
        O.ref().ref_lcg_real(seed, 64, O.dp(rr))
        cases[f"lcg_{seed}_int"] = oi
        cases[f"lcg_{seed}_intreal"] = orl
        cases[f"lcg_{seed}_real"] = rr
    np.savez_compressed(os.path.join(OUT, "selectors.npz"), **cases)


def logistic():
    prob = datagen.small_problem(n=2, m=37, d_feat=9, density=0.3, seed=7)
    d, m = prob.d, prob.m
    L = O.lda(d)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(d) * 0.5
    t = np.zeros(m); s = np.zeros(m); g = np.zeros(d); h = np.zeros(d * L); f = np.zeros(1)
    O.ref().ref_client_eval(O.dp(np.ascontiguousarray(prob.features[0])), O.dp(np.ascontiguousarray(prob.labels[0])),
                            d, m, prob.lam, O.dp(x), O.dp(t), O.dp(s), O.dp(g), O.dp(h), O.dp(f))
    A = rng.standard_normal((d, d))
    Hs = A @ A.T / d + 0.1 * np.eye(d)
    Hup = np.array([Hs[i, j] for j in range(d) for i in range(j + 1)])
    b = rng.standard_normal(d)
    dx = np.zeros(d); ltr = np.zeros(d * L)
    ok = O.ref().ref_newton_solve(O.dp(O.unpack_full(Hup, d)), d, 0.25, O.dp(b), O.dp(dx), O.dp(ltr))
    np.savez_compressed(os.path.join(OUT, "logistic.npz"), feat=prob.features[0], labels=prob.labels[0],
                        lam=prob.lam, x=x, margin=t, sigmoid=s, grad=g, hess_up=O.pack_upper(h, d), f=f,
                        newton_H_up=Hup, newton_b=b, newton_shift=0.25, newton_dx=dx, newton_ok=ok)


def trajectories():
    out = {}
    for comp, kmult, seed in [("topk", 2.0, 11), ("toplek", 2.0, 12), ("randk", 2.0, 13), ("seqk", 2.0, 14)]:
        prob = datagen.small_problem(n=6, m=40, d_feat=11, density=0.35, seed=seed)
        R = 25
        r = O.ref_fednl_run(prob, comp, kmult, rounds=R, workers=2, tracking=True, dump_rounds=3)
        out[f"{comp}_grad_norm"] = r["grad_norm"]
        out[f"{comp}_f"] = r["f"]
        out[f"{comp}_x_norm"] = r["x_norm"]
        out[f"{comp}_x"] = r["x"]
        for key in ("grad", "lk", "cnt", "idx", "dx"):
            out[f"{comp}_dump_{key}"] = r["dump"][key]
        out[f"{comp}_meta"] = np.array([6, 40, 11, 35, seed, R])
        out[f"{comp}_features"] = prob.features.astype(np.uint8)  # binary design incl. intercept
        out[f"{comp}_labels"] = prob.labels.astype(np.int8)
    np.savez_compressed(os.path.join(OUT, "trajectories.npz"), **out)


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref/libdopt_ref.so is missing: build it with make -C oracle -f Makefile.ref")
    selectors()
    logistic()
    trajectories()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))

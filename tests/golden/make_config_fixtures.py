#!/usr/bin/env python3
"""Writes tests/golden/configs.npz: whole FedNL runs of the measured workloads
computed by the reference's own code.

The runs use dopt's building blocks compiled by oracle/Makefile.ref into
oracle/_ref/libdopt_ref.so, with the round orchestrated as TrainLoop.h does
under --debug forcesequpdate (oracle/ref_harness.cpp), 8 worker threads
(TopK, RandK and RandSeqK runs do not depend on the worker count; TopLEK's
random stream does, TrainLoop.h:664-686, so its engine runs use 8 workers too).

Stored per case: the seeded problem's SHA-256 (features then labels, float64
bytes), ||grad f(x_r)|| for every round, the first round with
||grad f||^2 <= 1e-17 (-1 if none) and the iterate after the last round.
The workloads (ufednl_amd/datagen.py):
  config1  50 x 221, d=68,  Bernoulli(0.5) features, TopK k=8d, alpha=1
  config2  142 x 350, d=300, binary ~4% features, ~3% positives, TopK k=8d, alpha=1;
           RandK and RandSeqK k=8d with alpha = 1/(w+1)
  config3  80 x 407, d=123, binary ~11% features, TopLEK k=8d alpha=1 (8 workers),
           RandSeqK k=8d alpha = 1/(w+1)
lambda = 1e-3, option b, x0 = 0 throughout.

Run from the repository root (needs oracle/_ref, about 2 minutes on 8 cores):
    python tests/golden/make_config_fixtures.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from ufednl_amd import datagen  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs.npz")
TOL = 1e-17

# (case, workload, compressor, rounds, workers)
CASES = [
    ("c1_topk", "config1", "topk", 200, 8),
    ("c2_topk", "config2", "topk", 360, 8),
    ("c2_randk", "config2", "randk", 60, 8),
    ("c2_seqk", "config2", "seqk", 60, 8),
    ("c3_toplek", "config3", "toplek", 700, 8),
    ("c3_seqk", "config3", "seqk", 80, 8),
]


def problem_sha(prob):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(prob.features, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(prob.labels, dtype=np.float64).tobytes())
    return h.hexdigest()


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libdopt_ref.so is missing (build it with oracle/Makefile.ref)")
    probs = {}
    out = {}
    for case, work, comp, rounds, workers in CASES:
        if work not in probs:
            probs[work] = getattr(datagen, work)()
        prob = probs[work]
        res = O.ref_fednl_run(prob, comp, k_mult=8.0, rounds=rounds, workers=workers, tracking=False)
        if res["status"] != 0:
            sys.exit(f"{case}: reference run failed ({res['status']})")
        g = res["grad_norm"]
        hit = np.nonzero(g * g <= TOL)[0]
        r_star = int(hit[0]) if len(hit) else -1
        out[f"{case}_sha"] = np.array(problem_sha(prob))
        out[f"{case}_grad_norm"] = g
        out[f"{case}_x"] = res["x"]
        out[f"{case}_r_star"] = np.array(r_star)
        out[f"{case}_workers"] = np.array(workers)
        print(f"{case}: {rounds} rounds, first ||grad f||^2 <= {TOL:g} at round {r_star}, "
              f"last {g[-1] ** 2:.3e}", flush=True)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()

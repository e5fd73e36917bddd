#!/usr/bin/env python3
"""Writes the real-data fixtures, with the reference's own code compiled into
oracle/_ref and the datasets the reference ships (dopt/datasets/):

  tests/golden/datasets.json  every dataset file there: its zlib CRC-32, and,
      as dopt's loader reads it for train() (TrainLoop.h:123-170) with and
      without --add_intercept, N, d and the two digests bin_fednl_local prints
      under --debug crc32fordata (Reports.h:25-33).
  tests/golden/realdata.npz   phishing, a9a and w8a as dopt loads them
      (without the intercept): the attribute values as codes into a value
      table (bit-packed when there are two values) and the +-1 labels; and
      ||grad f(x_r)|| of 20 FedNL rounds of the reference building blocks on
      the --add_intercept --reshuffle_train split (the samples permuted by
      dopt::shuffle with seed 123, TrainLoop.h:227-232; client c takes
      samples [c m, (c+1) m) of the permuted order, the rest dropped,
      TrainLoop.h:212-250), lambda 1e-3, option b, RandK and TopK
      k = 8d, alpha as --theoretical_alpha --use_theoretical_alpha_option_2.

Run from the repository root: python tests/golden/make_realdata_fixtures.py
"""
import json
import os
import sys
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from ufednl_amd.datagen import Problem  # noqa: E402

DATA = "/root/reference/dopt/datasets"
HERE = os.path.dirname(os.path.abspath(__file__))
CLIENTS = {"phishing": 50, "a9a": 80, "w8a": 142}
ROUNDS = 20
SEED = 123


def split(X, y, order, n):
    """The --add_intercept --reshuffle_train client split."""
    N = X.shape[0]
    m = N // n
    Xi = np.hstack([X, np.ones((N, 1))])
    idx = np.asarray(order[: m * n])
    return Problem("real", np.ascontiguousarray(Xi[idx].reshape(n, m, -1)),
                   np.ascontiguousarray(y[idx].reshape(n, m)), 1e-3, np.zeros(Xi.shape[1]))


def main():
    if not O.ref_available() or not os.path.isdir(DATA):
        raise SystemExit("needs oracle/_ref and the reference datasets")
    digests = {}
    for name in sorted(os.listdir(DATA)):
        path = os.path.join(DATA, name)
        raw = open(path, "rb").read()
        ent = {"file_crc32": zlib.crc32(raw) & 0xFFFFFFFF, "file_bytes": len(raw)}
        for ic in (0, 1):
            X, y, cx, cy = O.ref_load_libsvm(path, ic)
            ent[f"intercept{ic}"] = {"N": int(X.shape[0]), "d": int(X.shape[1]), "crc_samples": cx, "crc_labels": cy}
        digests[name] = ent
        print(name, ent["intercept1"], flush=True)
    with open(os.path.join(HERE, "datasets.json"), "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)

    arrays = {}
    for name, n in CLIENTS.items():
        X, y, _, _ = O.ref_load_libsvm(os.path.join(DATA, name), 0)
        vals = np.unique(X)
        codes = np.searchsorted(vals, X).astype(np.uint8)
        arrays[f"{name}_vals"] = vals
        arrays[f"{name}_shape"] = np.array(X.shape, dtype=np.int64)
        if len(vals) == 2:
            arrays[f"{name}_bits"] = np.packbits(codes.astype(bool), axis=1)
        else:
            arrays[f"{name}_codes"] = codes
        arrays[f"{name}_pos"] = y > 0
        order = np.zeros(X.shape[0], dtype=np.uint32)
        O.ref().ref_shuffle(SEED, X.shape[0], X.shape[0], O.up(order))
        prob = split(X, y, order, n)
        for comp in ("randk", "topk"):
            out = O.ref_fednl_run(prob, comp, 8.0, rounds=ROUNDS, workers=1, tracking=False)
            assert out["status"] == 0
            arrays[f"{name}_{comp}_grad_norm"] = out["grad_norm"]
            print(name, comp, out["grad_norm"][[0, 5, 10, -1]], flush=True)
    np.savez_compressed(os.path.join(HERE, "realdata.npz"), **arrays)
    print("wrote", sorted(arrays))


if __name__ == "__main__":
    main()

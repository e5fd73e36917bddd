#!/usr/bin/env python3
"""Writes tests/golden/reshuffle.npz: the sample permutations of dopt's
--reshuffle_train (TrainLoop.h:227-232: dopt::shuffle over all samples with
RandomGenIntegerLinear seeded by --reshuffle_train_seed, Shuffle.h), computed
by the reference's own code compiled into oracle/_ref.

Run from the repository root: python tests/golden/make_cli_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reshuffle.npz")


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built (needs the reference sources)")
    arrays = {}
    for n in (273, 1000):
        for seed in (123, 7):
            out = np.zeros(n, dtype=np.uint32)
            O.ref().ref_shuffle(seed, n, n, O.up(out))
            arrays[f"perm_{seed}_{n}"] = out
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, sorted(arrays))


if __name__ == "__main__":
    main()

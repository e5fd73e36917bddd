"""Phase breakdown of the TopK selection kernel (k_select_heap) from the
device-clock stamps that a diagnostic build of the library records.

The diagnostic build compiles the same sources with -DFEDNL_PROBE into
ufednl_amd/libfednl_hip_probe.so (python -m ufednl_amd.build --probe); the
product library has no stamps. Each client's workgroup adds, per round:
  build  heap construction over the first k positions (MinHeap.h:147-191),
  scan   positions k .. S-1: the streamed |Delta| ring and the admissions
         (MinHeap.h:78-85), with the admission count and their shader cycles,
  tail   heap read-back, (TopLEK) the energy walk's extractions
         (SpecialMathRoutinesForMatrix.h:441-499), the sorted index list and,
         for TopK, the message values and the learned update.
build/scan/tail are wall_clock64 ticks (100 MHz); the scan's, the
admissions' and the energy walk's shader cycles come from clock64.

Run: python -m ufednl_amd.phases [--rounds 0,10,40,150,300] [--workload c2|c3]
     [--compressor topk|toplek]
(c2: 142 clients x 350 samples, d = 300; c3: 80 x 407, d = 123; k = 8d, FP64.)
"""
from __future__ import annotations

import argparse
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE_LIB = os.path.join(HERE, "libfednl_hip_probe.so")
SLOTS = 16384
BASE = 1000  # per-client records: 8 slots each from here


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rounds", default="0,10,40,150,300")
    ap.add_argument("--workload", default="c2", choices=("c2", "c3"))
    ap.add_argument("--compressor", default="topk", choices=("topk", "toplek"))
    args = ap.parse_args(argv)
    if not os.path.exists(PROBE_LIB):
        raise SystemExit(f"{PROBE_LIB} is missing: python -m ufednl_amd.build --probe")
    from ufednl_amd import datagen
    from ufednl_amd import engine as E
    E.LIB_PATH = PROBE_LIB
    L = E.lib()
    L.fednl_probe_read.argtypes = [ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int, ctypes.c_int]
    buf = (ctypes.c_ulonglong * SLOTS)()
    prob = datagen.config2() if args.workload == "c2" else datagen.config3()
    eng = E.Engine.from_problem(prob, compressor=args.compressor, k_mult=8.0, tracking=False)
    done = 0
    print(f"k_select_heap phases, {args.compressor} k={eng.k}, {eng.n} clients x {eng.m} samples, d={eng.d}",
          flush=True)
    lek = args.compressor == "toplek"
    print("round | build max us | scan max us | tail max us | admissions mean / max | "
          "cycles per admission | slowest client: admissions, admission cycles, scan cycles"
          + (" | extractions mean, cycles per extraction" if lek else ""))
    for r in sorted(int(x) for x in args.rounds.split(",")):
        if r > done:
            eng.run(r - done, check=False)
            done = r
        L.fednl_probe_read(buf, SLOTS, 1)
        eng.run(1, check=False)
        done += 1
        L.fednl_probe_read(buf, SLOTS, 0)
        a = np.frombuffer(buf, dtype=np.uint64).astype(np.int64)
        sel = a[BASE:BASE + 8 * eng.n].reshape(eng.n, 8)
        adm, cyc, scan = sel[:, 3], sel[:, 4], sel[:, 5]
        s = int(np.argmax(sel[:, 1]))
        print(f"{r:5d} | {sel[:, 0].max() / 100:12.1f} | {sel[:, 1].max() / 100:11.1f} | {sel[:, 2].max() / 100:11.1f} | "
              f"{adm.mean():10.0f} / {adm.max():4d} | {cyc.sum() / max(adm.sum(), 1):20.0f} | "
              f"{s}: {adm[s]}, {cyc[s]}, {scan[s]}"
              + (f" | {sel[:, 6].mean():.0f}, {sel[:, 7].sum() / max(sel[:, 6].sum(), 1):.0f}" if lek else ""),
              flush=True)
        if lek:
            w = a[2000:2000 + 4 * eng.n].reshape(eng.n, 4)
            ne = max(sel[:, 6].sum(), 1)
            print(f"      walk cycles per extraction: last element out {w[:, 0].sum() / ne:.0f}, "
                  f"sift {w[:, 1].sum() / ne:.0f}, energy test {w[:, 2].sum() / ne:.0f}", flush=True)
    eng.close()


if __name__ == "__main__":
    main()

"""Builds the native parts in-tree.

  ufednl_amd/libfednl_hip.so   HIP kernels + host engine + C ABI (hipcc, gfx950)
  ufednl_amd/bin_fednl_local   drop-in CLI linked against it
  ufednl_amd/bin_fednl_local_pp  the FedNL-PP CLI, linked the same way
  ufednl_amd/libfednl_hip_probe.so  (--probe only) the library with device-clock
      phase stamps, read by python -m ufednl_amd.phases

Run: python -m ufednl_amd.build  (or __graft_entry__.build()).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libfednl_hip.so")
CLI = os.path.join(HERE, "bin_fednl_local")
CLI_PP = os.path.join(HERE, "bin_fednl_local_pp")
PROBE_LIB = os.path.join(HERE, "libfednl_hip_probe.so")

ARCH = os.environ.get("FEDNL_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    # exact-arithmetic kernels (message values, learned update, master sums)
    # must not be contracted into FMA; kernels use fma() where they want one
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
    f"-I{INC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib():
    srcs = [os.path.join(CSRC, "fednl_kernels.hip"), os.path.join(CSRC, "fednl_cli.cpp")]
    srcs = [s for s in srcs if os.path.exists(s)]
    tmp = LIB + ".tmp"
    run([hipcc(), *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)


def build_cli():
    for src, out in (("bin_fednl_local_main.cpp", CLI), ("bin_fednl_local_pp_main.cpp", CLI_PP)):
        main = os.path.join(CSRC, src)
        run(["g++", "-O2", "-std=c++17", f"-I{INC}", "-o", out, main, f"-L{HERE}", "-lfednl_hip",
             f"-Wl,-rpath,$ORIGIN"])


def build_oracle():
    od = os.path.join(ROOT, "oracle")
    run(["make", "-s", "-C", od, "-f", "Makefile"])
    if os.path.isdir("/root/reference/dopt"):
        run(["make", "-s", "-C", od, "-f", "Makefile.ref"])


def build_probe():
    """The diagnostic library with device-clock phase stamps (-DFEDNL_PROBE),
    read by ufednl_amd.phases; never loaded by the product path."""
    srcs = [os.path.join(CSRC, "fednl_kernels.hip"), os.path.join(CSRC, "fednl_cli.cpp")]
    run([hipcc(), *HIP_FLAGS, "-DFEDNL_PROBE", "-shared", "-o", PROBE_LIB, *srcs])


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if "--probe" in argv:
        build_probe()
        return
    build_lib()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    sys.exit(main())

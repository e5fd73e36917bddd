"""MI355X-native FedNL round engine (drop-in for dopt's bin_fednl_local path).

The compute path is libfednl_hip.so (HIP kernels for gfx950 behind a C ABI,
include/fednl_hip.h). This package only binds it; there is no CPU fallback:
importing ``ufednl_amd.engine`` raises if the library is missing.
"""
from .engine import (COMPRESSORS, Engine, lib, lib_path, newton_solve, select)  # noqa: F401

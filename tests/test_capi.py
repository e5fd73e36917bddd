"""The C-ABI library loads on a host without a GPU and exports every
function include/fednl_hip.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fednl_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([a-zA-Z_][a-zA-Z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n not in ("typedef",)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("fednl_create", "fednl_run", "fednl_select", "fednl_newton_solve", "doptExecuteLocal"):
        assert must in names, names


def test_library_exports_all_declared():
    from ufednl_amd.engine import lib, lib_path
    lib()  # loads; raises if missing
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path()], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_error_path_without_compute():
    from ufednl_amd.engine import lib
    L = lib()
    L.fednl_last_error.restype = ctypes.c_char_p
    assert L.fednl_create(None, 1, 1, 1, None, None, None) == -1
    assert b"null" in L.fednl_last_error()

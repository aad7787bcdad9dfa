"""CPU-side checks of the C ABI library: it loads and exports every entry
point that include/scihpc_hip.h declares (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "scihpc_hip.h")
LIB = os.path.join(ROOT, "scihpc_amd", "libscihpc_hip.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"SCIHPC_API\s+(?:int|const char \*)\s*(scihpc_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "scihpc_mg_solve" in names and "scihpc_ls_tvd_rk3" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the HIP library first (python -m scihpc_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s+(scihpc_\w+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    # nothing else leaks out of the library
    extra = sorted(s for s in exported if s not in declared())
    assert not extra, extra


def test_library_loads_and_reports_errors_without_gpu_work():
    lib = ctypes.CDLL(LIB)
    lib.scihpc_last_error.restype = ctypes.c_char_p
    assert lib.scihpc_last_error() == b""
    # a bad argument is rejected before any device call
    h = ctypes.c_void_p()
    rc = lib.scihpc_mg_create(4, 8, 8, 8, ctypes.c_double(1.0), ctypes.c_double(1.0),
                              ctypes.c_double(1.0), 0, ctypes.byref(h))
    assert rc == 2
    assert b"ndim" in lib.scihpc_last_error()

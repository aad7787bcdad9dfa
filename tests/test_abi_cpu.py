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


def test_ctypes_structs_match_header(tmp_path):
    """The ctypes mirrors of the option and statistics structs have the
    header's size and field offsets (compiled with the host C compiler)."""
    from scihpc_amd import _capi

    pairs = [("scihpc_solve_opts", _capi.SolveOpts), ("scihpc_solve_stats", _capi.SolveStats),
             ("scihpc_fpc_params", _capi.FpcParams), ("scihpc_flow_params", _capi.FlowParams)]
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "scihpc_hip.h"', "int main(void) {"]
    for cname, py in pairs:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in pairs:
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)

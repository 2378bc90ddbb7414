"""The C-ABI library: loads, exports every symbol of include/apss.h, and
rejects bad arguments before touching the device.  No GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import apss
from helpers import tiny_matrix

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "apss.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)
    skip = {"if", "while", "for", "return", "sizeof", "typedef", "void", "defined"}
    return sorted({n for n in names if n.startswith(("apss_", "similarity_")) and n not in skip
                   and not n.endswith("_fn")})


def test_library_loads_and_exports_header_symbols():
    L = apss.lib()
    declared = header_functions()
    assert "apss_run" in declared and "similarity_join_device" in declared
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"
    assert set(apss.ABI_SYMBOLS) == set(declared)


def test_version_and_strerror():
    L = apss.lib()
    assert b"gfx950" in L.apss_version()
    for code in (0, -1, -2, -3, -4, -5, -6):
        assert L.apss_strerror(code)
    assert L.apss_strerror(-99) == b"unknown error"


def _call(rp, col, val, d, t, p):
    L = apss.lib()
    X = apss.make_csr(rp, col, val, d)
    o = apss.make_opts()
    r = apss.Result()
    return L.apss_run(ctypes.byref(X), float(t), int(p), ctypes.byref(o), ctypes.byref(r))


def test_argument_errors_without_device():
    rp, col, val, d = tiny_matrix([{1: 1.0}, {1: 1.0}], 4)
    assert _call(rp, col, val, d, float("nan"), 0) == apss.APSS_E_INVALID
    assert _call(rp, col, val, d, 0.5, 5) == apss.APSS_E_INVALID     # p > n_cols
    bad = rp.copy()
    bad[0] = 1
    assert _call(bad, col, val, d, 0.5, 0) == apss.APSS_E_INVALID    # row_ptr[0] != 0
    bad = rp.copy()
    bad[-1] = 7
    assert _call(bad, col, val, d, 0.5, 0) == apss.APSS_E_INVALID    # row_ptr[n] != nnz
    L = apss.lib()
    o = apss.make_opts()
    r = apss.Result()
    assert L.apss_run(None, 0.5, 0, ctypes.byref(o), ctypes.byref(r)) == apss.APSS_E_INVALID
    assert b"finite" in L.apss_last_error() or L.apss_last_error()


def test_struct_layouts():
    # the Python mirror must match the C layout (sizes fixed by the header)
    assert ctypes.sizeof(apss.CSR) == 56
    assert ctypes.sizeof(apss.Opts) == 72
    assert apss.PAIR_DTYPE.itemsize == 12
    assert ctypes.sizeof(apss.Result) == 256


def _nm(path, dynamic):
    args = ["nm", "-D", "--defined-only", path] if dynamic else ["nm", "-u", path]
    return subprocess.run(args, check=True, capture_output=True, text=True).stdout


def test_cxx_linkage_caller_links(tmp_path):
    # A caller compiled against the reference's declaration (C++ linkage,
    # similarity_search.h:3-14) references the mangled symbol; libapss.so
    # defines it beside the C one, so the link succeeds unchanged.
    src = os.path.join(ROOT, "tests", "caller", "ref_caller.cpp")
    exe = str(tmp_path / "ref_caller")
    pkg = os.path.join(ROOT, "apss-cuda_amd")
    subprocess.run(["g++", "-std=c++17", src, "-L" + pkg, "-lapss", "-Wl,-rpath," + pkg, "-o", exe],
                   check=True, capture_output=True)
    undefined = [ln.split()[-1] for ln in _nm(exe, False).splitlines() if "similarity_join_device" in ln]
    assert undefined and all(u.startswith("_Z22similarity_join_device") for u in undefined)
    exported = _nm(os.path.join(pkg, "libapss.so"), True)
    for sym in undefined + ["similarity_join_device", "apss_compat_last_count"]:
        assert re.search(r"\bT " + re.escape(sym) + r"\b", exported), f"{sym} not exported"

"""The C-ABI library loads and exports every entry point include/*.h declares (no GPU needed)."""

import ctypes
import os
import re
import subprocess

from lucid_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(lucid_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared_symbols():
    assert os.path.exists(_lib.LIB_PATH), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = declared_functions()
    assert declared, "no declarations parsed"
    assert declared == set(_lib.EXPORTED), declared ^ set(_lib.EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (lucid_\w+)", out))
    missing = declared - exported
    assert not missing, f"missing exports: {missing}"
    for name in declared:
        assert hasattr(lib, name)


def test_error_strings_and_invalid_args():
    lib = _lib.load()
    assert lib.lucid_error_string(0) == b"ok"
    assert b"capacity" in lib.lucid_error_string(_lib.LUCID_ERR_CAPACITY)
    # invalid arguments are rejected before any device call
    h = ctypes.c_void_p()
    bad = _lib.CreateInfo(0, 10, 0, 0, 0)
    assert lib.lucid_create(ctypes.byref(bad), ctypes.byref(h)) == _lib.LUCID_ERR_INVALID_ARG
    too_big = _lib.CreateInfo(4096, 4096, 0, 0, 0)
    assert lib.lucid_create(ctypes.byref(too_big), ctypes.byref(h)) == _lib.LUCID_ERR_INVALID_ARG
    assert lib.lucid_render(None, None, None, None) == _lib.LUCID_ERR_INVALID_ARG
    assert lib.lucid_finish(None) == _lib.LUCID_ERR_INVALID_ARG
    assert lib.lucid_bin_count(None) == 0

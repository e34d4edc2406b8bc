"""The C ABI library loads without a GPU and exports every entry point the
header declares (no compute calls here)."""

import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "petrbf.h")
LIB = os.path.join(ROOT, "petrbf_amd", "libpetrbf.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prbf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "prbf_matvec" in names and "prbf_solve" in names and "prbf_rasm_apply" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libpetrbf.so not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (prbf_[a-z0-9_]+)", out))
    assert set(declared()) <= exported


def test_python_binding_lists_every_export():
    from petrbf_amd import _lib
    assert sorted(_lib.EXPORTS) == declared()


def test_library_carries_gfx950_code():
    data = open(LIB, "rb").read()
    assert b"gfx950" in data

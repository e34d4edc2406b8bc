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


def test_static_archive_exports_every_symbol_and_links_a_c_caller(tmp_path):
    """libpetrbf.a (same objects) defines every declared entry point, and a C
    program that includes the header links against it with hipcc (no run:
    there may be no GPU)."""
    ar = os.path.join(ROOT, "petrbf_amd", "libpetrbf.a")
    if not os.path.exists(ar):
        pytest.fail("libpetrbf.a not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    out = subprocess.run(["nm", "--defined-only", ar], capture_output=True, text=True).stdout
    defined = set(re.findall(r"\bT (prbf_[a-z0-9_]+)", out))
    assert set(declared()) <= defined
    src = tmp_path / "caller.c"
    src.write_text(
        '#include "petrbf.h"\n'
        "int main(void) {\n"
        "  prbf_params p = {2, 0.1, 3.2, 16.0, 13.2};\n"
        "  prbf_ctx* c = 0;\n"
        "  prbf_solve_opts o;\n"
        "  prbf_solve_opts_default(&o);\n"
        "  if (prbf_create(&c, &p) != PRBF_OK) return 1;\n"
        "  prbf_destroy(c);\n"
        "  return o.restart == 30 ? 0 : 2;\n"
        "}\n")
    obj = tmp_path / "caller.o"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                    "-o", str(obj)], check=True)
    exe = tmp_path / "caller"
    from petrbf_amd.build import ARCH, HIPCC
    r = subprocess.run([HIPCC, f"--offload-arch={ARCH}", str(obj), ar, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert exe.exists()

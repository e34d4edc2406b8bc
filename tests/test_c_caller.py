"""A C program drives the library through include/petrbf.h alone (no Python
on the compute path): the N = 10^4 plain solve with host buffers must match
the committed oracle fixture (same iteration count, gamma to 1e-10) and its
residual, taken through prbf_evaluate, must meet rtol 1e-10."""

import os
import subprocess
import sys

import numpy as np
import pytest

from petrbf_amd import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


def _build(tmp_path):
    exe = tmp_path / "c1_caller"
    libdir = os.path.join(ROOT, "petrbf_amd")
    subprocess.run(["gcc", "-std=c99", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "c", "c1_caller.c"), "-o", str(exe), "-L", libdir, "-lpetrbf",
                    "-Wl,-rpath," + libdir], check=True)
    return exe


def test_c_caller_compiles_against_the_header(tmp_path):
    if not os.path.exists(os.path.join(ROOT, "petrbf_amd", "libpetrbf.so")):
        pytest.fail("libpetrbf.so not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    assert _build(tmp_path).exists()


@pytest.mark.gpu
def test_c_caller_solve_matches_fixture(tmp_path):
    print("[gpu-test] C caller, N=10^4 plain solve", flush=True)
    sys.stdout.flush()
    exe = _build(tmp_path)
    p = problems.lattice_2d_problem(100, seed=0x5EED0001)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([p.n], np.int64).tofile(f)
        np.array([p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc], np.float64).tofile(f)
        for a in (p.x, p.y, p.w[0]):
            np.ascontiguousarray(a, np.float64).tofile(f)
    r = subprocess.run([str(exe), str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(fout, "rb").read()
    st, its, reason = np.frombuffer(raw[:12], np.int32)
    g = np.frombuffer(raw[12:12 + 8 * p.n], np.float64)
    ag = np.frombuffer(raw[12 + 8 * p.n:], np.float64)
    fx = np.load(os.path.join(HERE, "golden", "c1_oracle.npz"))
    assert (st, reason) == (0, 2)
    assert its == int(fx["its"])
    assert np.linalg.norm(g - fx["gamma"]) <= 1e-10 * np.linalg.norm(fx["gamma"])
    assert np.linalg.norm(ag - p.w[0]) <= 1e-9 * np.linalg.norm(p.w[0])

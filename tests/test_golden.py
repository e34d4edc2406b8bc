"""The N = 10^4 CPU workload against the committed fixture
tests/golden/c1_oracle.npz (written by tests/golden/make_c1.py from the
oracle): the oracle must keep reproducing it (CPU), and the library must match
it (GPU): lists bit for bit, matvec to 1e-12, the plain solve to 1e-10 with
the same iteration count."""

import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from petrbf_amd import problems
from tests.golden.make_c1 import c1_digests

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture():
    f = np.load(os.path.join(HERE, "golden", "c1_oracle.npz"))
    return {k: f[k] for k in f.files}


def _problem():
    p = problems.lattice_2d_problem(100, seed=0x5EED0001)
    x = np.random.default_rng(101).standard_normal(p.n)
    return p, x


def _digests(perm, cells, bcsr, tcsr):
    return c1_digests(np.asarray(perm, np.int32), [np.asarray(c, np.int32) for c in cells],
                      [np.asarray(bcsr[0], np.int64), np.asarray(bcsr[1], np.int32)],
                      [np.asarray(tcsr[0], np.int64), np.asarray(tcsr[1], np.int32)])


def test_oracle_reproduces_c1_fixture():
    f = _fixture()
    p, x = _problem()
    o = Oracle.from_problem(p)
    pi, _ = o.perm()
    d = _digests(pi, o.cells(), o.buffer_csr(), o.trunc_csr())
    for k, v in d.items():
        assert v == str(f[f"sha_{k}"]), k
    y = o.matvec(x)
    assert np.linalg.norm(y - f["y"]) <= 1e-14 * np.linalg.norm(f["y"])
    o.rasm_setup()
    r = o.solve_sorted(p.w[0][pi], rtol=1e-10, restart=30, max_it=1000)
    assert r["reason"] == 2 and r["its"] == int(f["its"])
    g = np.empty(p.n)
    g[pi] = r["gamma"]
    assert np.linalg.norm(g - f["gamma"]) <= 1e-12 * np.linalg.norm(f["gamma"])
    assert np.allclose(r["hist"], f["hist"], rtol=1e-10, atol=0)


@pytest.mark.gpu
def test_library_matches_c1_fixture():
    from petrbf_amd.rbf import Context
    print("[gpu-test] c1 fixture", flush=True)
    f = _fixture()
    p, x = _problem()
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y)
    pi, _ = ctx.permutation()
    d = _digests(pi, ctx.cells(), ctx.buffer_csr(), ctx.trunc_csr())
    for k, v in d.items():
        assert v == str(f[f"sha_{k}"]), k
    y = ctx.matvec(x)
    assert np.linalg.norm(y - f["y"]) <= 1e-12 * np.linalg.norm(f["y"])
    g, info = ctx.solve(p.w[0], ksp="gmres", pc="rasm", rtol=1e-10, restart=30, max_it=1000)
    assert info.reason == 2 and info.its == int(f["its"])
    assert np.linalg.norm(g - f["gamma"]) <= 1e-10 * np.linalg.norm(f["gamma"])
    # the block solves differ from the oracle's LU in rounding (~kappa*eps)
    assert np.allclose(info.hist, f["hist"], rtol=1e-9, atol=1e-10 * f["hist"][0])
    ctx.close()

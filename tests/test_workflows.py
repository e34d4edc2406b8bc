"""End-to-end workflows of the reference's harnesses:

* the cylinder input fixture (2d/cdata read by 2d/cylinder.cxx:82-90, run line
  `cylinder .9 5 1.9`: sigma 0.007, cell 5 sigma, buffer (int)(5*1.9) = 9
  sigma, truncation 5 + 4 = 9 sigma, 2d/cylinder.cxx:45-60);
* solve on the lattice, evaluate on the h/2 grid (2d/main2.cxx:53,126-137);
* the FP32 library entry RBFInterpolation with remeshing (3d/wrapper.cxx:93-101).
"""

import os
import sys

import numpy as np
import pytest

from oracle.oracle import Oracle
from petrbf_amd import problems
from tests.cases import lattice

HERE = os.path.dirname(os.path.abspath(__file__))
CYL = dict(sigma=0.007, nsigma_box=5.0, sigma_buffer=float(int(5 * 1.9)), sigma_trunc=9.0)


def _cdata():
    d = np.load(os.path.join(HERE, "golden", "cylinder_cdata.npy"))
    return d[:, 0], d[:, 1], d[:, 2], d[:, 3]


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def test_cylinder_fixture_oracle_solve():
    x, y, g, e = _cdata()
    o = Oracle(2, CYL["sigma"], CYL["nsigma_box"], CYL["sigma_buffer"], CYL["sigma_trunc"], x, y)
    pi, _ = o.perm()
    r = o.solve_sorted(e[pi], rtol=1e-10)
    assert r["reason"] == 2
    w = o.matvec(r["gamma"])
    assert np.linalg.norm(w - e[pi]) <= 1e-9 * np.linalg.norm(e)


@pytest.mark.gpu
def test_cylinder_fixture_gpu_parity():
    _announce("cylinder cdata")
    from petrbf_amd.rbf import rbf_interpolation
    x, y, g, e = _cdata()
    o = Oracle(2, CYL["sigma"], CYL["nsigma_box"], CYL["sigma_buffer"], CYL["sigma_trunc"], x, y)
    pi, _ = o.perm()
    r = o.solve_sorted(e[pi], rtol=1e-10)
    g_cpu = np.empty_like(e)
    g_cpu[pi] = r["gamma"]
    g_gpu, its = rbf_interpolation(x, y, e, **CYL)
    assert its == r["its"]
    assert np.linalg.norm(g_gpu - g_cpu) <= 1e-10 * np.linalg.norm(g_cpu)


@pytest.mark.gpu
def test_cylinder_fixture_uneven_blocks_both_orders():
    """The cylinder fixture's blocks hold 15 to 87 unknowns (1 to 6 tile
    rows: single-tile blocks, odd tile counts under 2-column panels).  The
    FP64 preconditioner of every factorisation order and panel width agrees
    with the oracle's LU of each block, and FGMRES on the FP32 preconditioner
    of either order reaches rtol 1e-10 with the true residual to match."""
    _announce("cylinder uneven blocks")
    from petrbf_amd.rbf import Context
    x, y, g, e = _cdata()
    o = Oracle(2, CYL["sigma"], CYL["nsigma_box"], CYL["sigma_buffer"], CYL["sigma_trunc"], x, y)
    o.rasm_setup()
    pi, _ = o.perm()
    rng = np.random.default_rng(8)
    r = rng.standard_normal(x.shape[0])
    want = o.rasm_apply(r)
    for order, nw, npc in (("right", 4, 1), ("right", 8, 2), ("left", 4, 1), ("left", 4, 2), ("left", 4, 4)):
        ctx = Context(2, CYL["sigma"], CYL["nsigma_box"], CYL["sigma_buffer"], CYL["sigma_trunc"])
        ctx.set_tuning(chol_waves=nw, chol_panel=npc, chol_order=order)
        ctx.set_points(x, y)
        ctx.rasm_setup("rasm")
        got = ctx.rasm_apply(r)
        assert np.linalg.norm(got - want) <= 1e-8 * np.linalg.norm(want), (order, nw, npc)
        ctx.close()
    for order in ("right", "left"):
        ctx = Context(2, CYL["sigma"], CYL["nsigma_box"], CYL["sigma_buffer"], CYL["sigma_trunc"])
        ctx.set_tuning(chol_waves=4, chol_panel=2, chol_order=order)
        ctx.set_points(x, y)
        gam, inf = ctx.solve(e, ksp="fgmres", pc="rasm32", rtol=1e-10)
        assert inf.reason == 2, order
        w = np.empty_like(e)
        w[pi] = o.matvec(gam[pi])
        assert np.linalg.norm(w - e) <= 1e-9 * np.linalg.norm(e), order
        ctx.close()


@pytest.mark.gpu
def test_h2_grid_evaluation_matches_oracle():
    """Solve on the 100x100 lattice, evaluate on the 199x199 h/2 grid."""
    _announce("h/2 grid evaluation")
    from petrbf_amd.rbf import Context, vorticity_evaluation
    p = lattice(100, seed=0x5EED0001)
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y)
    gamma, info = ctx.solve(p.w[0])
    ctx.close()
    xt, yt, _ = problems.jittered_lattice_2d(199, seed=3, jitter=0.0)
    w_gpu = vorticity_evaluation(xt, yt, p.x, p.y, gamma, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc)
    oe = Oracle(2, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, xt, yt, None, p.x, p.y)
    pit, pjt = oe.perm()
    ws = oe.matvec(gamma[pjt])
    w_cpu = np.empty_like(ws)
    w_cpu[pit] = ws
    assert np.linalg.norm(w_gpu - w_cpu) <= 1e-12 * np.linalg.norm(w_cpu)
    exact = problems.lamb_oseen(xt, yt)
    err = np.linalg.norm(w_gpu - exact) / np.linalg.norm(exact)
    _announce(f"h/2 grid interpolation error {err:.3e}")
    assert err < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_rbf_interpolation_library_entry_remesh(dim):
    """Particles carrying strengths are remeshed onto a grid (evaluation), then
    the grid strengths are solved, all through the FP32 entry point."""
    _announce(f"RBFInterpolation f32 dim {dim}")
    from petrbf_amd.rbf import RBFInterpolation
    if dim == 2:
        p = lattice(40, seed=21)
        xg, yg, _ = problems.jittered_lattice_2d(40, seed=22, jitter=0.05)
        zg = None
        pts = [p.x, p.y]
        grid = [xg, yg]
    else:
        p = problems.lattice_3d_problem(12, seed=23)
        xg, yg, zg, _ = problems.jittered_lattice_3d(12, seed=24, jitter=0.05)
        pts = [p.x, p.y, p.z]
        grid = [xg, yg, zg]
    gp = (p.w[0] * p.h ** dim).astype(np.float32)
    wg0 = np.zeros(xg.shape[0], np.float32)
    out, its = RBFInterpolation(xg, yg, zg, wg0, p.x, p.y, p.z, gp, p.sigma, p.nsigma_box, p.sigma_buffer,
                                p.sigma_trunc)
    # the same pipeline on the oracle, from the same FP32 inputs
    f = lambda a: np.asarray(a, np.float32).astype(np.float64)
    G = [f(a) for a in grid]
    P = [f(a) for a in pts]
    zpar = (G[2] if dim == 3 else None, P[2] if dim == 3 else None)
    oe = Oracle(dim, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, G[0], G[1], zpar[0], P[0], P[1], zpar[1])
    pit, pjt = oe.perm()
    ws = oe.matvec(f(gp)[pjt])
    w = np.empty_like(ws)
    w[pit] = ws
    o = Oracle(dim, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, G[0], G[1], zpar[0])
    pi, _ = o.perm()
    r = o.solve_sorted(w[pi], rtol=1e-10)
    g_ref = np.empty_like(w)
    g_ref[pi] = r["gamma"]
    assert its == r["its"]
    assert np.linalg.norm(out - g_ref) <= 1e-6 * np.linalg.norm(g_ref)

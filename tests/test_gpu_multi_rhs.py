"""Batched multi-right-hand-side Krylov solves (prbf_solve_multi).

The reference solves the three components of the 3D vorticity field with one
RBFInterpolation call each (3d/wrapper.cxx:93-101; RHS 3d/main.cxx:104).  The
batched solve must give each right-hand side exactly what a single solve
gives (same kernels and reduction order per vector), while sharing each
matvec and preconditioner apply: checked bit for bit against single solves,
and against the oracle's GMRES/FGMRES per component.
"""

import sys

import numpy as np
import pytest

from oracle.oracle import Oracle
from petrbf_amd import problems
from petrbf_amd.rbf import Context

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _three_rhs_2d(n_side):
    p = problems.lattice_2d_problem(n_side, seed=0x5EED0D01)
    w = np.stack([problems.lamb_oseen(p.x, p.y, 0.01), problems.lamb_oseen(p.x - 0.2, p.y + 0.1, 0.02),
                  np.cos(2 * p.x) * np.sin(3 * p.y)])
    return p, w


@pytest.mark.parametrize("ksp,pc", [("gmres", "rasm"), ("fgmres", "rasm32"), ("fgmres", "none")])
def test_batched_equals_single_solves_2d(ksp, pc):
    """Three 2D right-hand sides (two Lamb-Oseen fields, one smooth trig
    field): the batched solve equals three single solves bit for bit (gamma,
    iterations, residual history), kernel by kernel and graph-captured."""
    _announce(f"batched 2D {ksp}/{pc}")
    p, w = _three_rhs_2d(60)
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y)
    max_it = 60 if pc == "none" else 1000
    g3, infos = ctx.solve_multi(w, ksp=ksp, pc=pc, rtol=1e-10, restart=30, max_it=max_it)
    for r in range(3):
        for graph in (False, True):
            g1, i1 = ctx.solve(w[r], ksp=ksp, pc=pc, rtol=1e-10, restart=30, max_it=max_it, graph=graph)
            assert (infos[r].its, infos[r].reason) == (i1.its, i1.reason), (r, graph)
            assert np.array_equal(infos[r].hist, i1.hist), (r, graph)
            assert np.array_equal(g3[r], g1), (r, graph)
        if pc != "none":
            assert infos[r].reason == 2
    _announce(f"its {[i.its for i in infos]}")
    ctx.close()


def test_batched_plain_solve_matches_oracle_per_component():
    """Plain GMRES(30) + FP64 RASM on three right-hand sides at once: every
    component matches the oracle's single solve (same iterations, gamma within
    1e-10, the parity bar of the plain form)."""
    _announce("batched plain vs oracle")
    p, w = _three_rhs_2d(100)
    o = Oracle.from_problem(p)
    pi, _ = o.perm()
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y)
    g3, infos = ctx.solve_multi(w, ksp="gmres", pc="rasm", rtol=1e-10, restart=30, max_it=1000)
    for r in range(3):
        ref = o.solve_sorted(w[r][pi], rtol=1e-10, restart=30, max_it=1000)
        g_cpu = np.empty(p.n)
        g_cpu[pi] = ref["gamma"]
        assert infos[r].reason == 2 and ref["reason"] == 2
        assert infos[r].its == ref["its"], (r, infos[r].its, ref["its"])
        assert np.linalg.norm(g3[r] - g_cpu) <= 1e-10 * np.linalg.norm(g_cpu), r
    ctx.close()


def test_batched_3d_vortex_ring():
    """The 3D vector-field workload's structure at 12^3 (3x3x3-spacing cells,
    one-cell buffer, 4-sigma cutoff, vortex-ring vorticity as three
    right-hand sides, the z component zero): the batched FGMRES equals the
    single solves bit for bit, the zero component returns at once, and the
    first cycle's residual history follows the oracle's FGMRES with exact FP64
    blocks."""
    _announce("batched 3D vortex ring")
    p = problems.lattice_3d_problem(12, seed=0x5EED0C05)
    o = Oracle.from_problem(p)
    o.rasm_setup()
    pi, _ = o.perm()
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y, p.z)
    g3, infos = ctx.solve_multi(p.w, ksp="fgmres", pc="rasm", rtol=1e-10, restart=30, max_it=60)
    assert infos[2].reason == 3 and infos[2].its == 0 and not g3[2].any()
    for r in range(3):
        g1, i1 = ctx.solve(p.w[r], ksp="fgmres", pc="rasm", rtol=1e-10, restart=30, max_it=60, graph=False)
        assert np.array_equal(g3[r], g1) and np.array_equal(infos[r].hist, i1.hist), r
    for r in range(2):
        ref = o.solve_sorted(p.w[r][pi], rtol=1e-10, restart=30, max_it=30, flexible=True)
        h_gpu, h_cpu = infos[r].hist[:31], np.asarray(ref["hist"][:31])
        rel = np.abs(h_gpu - h_cpu).max() / h_cpu[0]
        _announce(f"component {r}: first-cycle residual history vs oracle max rel diff {rel:.2e}; "
                  f"after 60 its {infos[r].hist[-1] / infos[r].hist[0]:.3f} of the initial residual")
        assert rel <= 1e-6, (r, rel)
    ctx.close()

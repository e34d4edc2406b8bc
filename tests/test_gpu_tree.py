"""GPU parity of the adaptive cluster structure (prbf_set_points_adaptive)
against the oracle's restatement (oracle/petrbf_oracle.cpp tree_build,
itself pinned against its brute-force definition in test_tree_oracle.py):
leaves, sorted order and both lists bit for bit, the matvec within 1e-12,
the FP64 RASM blocks of uneven leaves against the oracle's LU, the plain
solve's gamma within 1e-10, the mixed solve's true residual; and on the
non-uniform mixture of the C3 workload, whose blocks are numerically
singular, the regularised setup and the solve's behaviour."""

import hashlib
import sys

import numpy as np
import pytest

from oracle.oracle import Oracle
from petrbf_amd import problems
from petrbf_amd.rbf import Context
from tests.test_tree_oracle import CASES as TREE_CASES

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _pair(dim, sigma, nb, sb, st, pts, cap, dmax):
    o = Oracle(dim, sigma, nb, sb, st, *pts, tree=(cap, dmax))
    ctx = Context(dim, sigma, nb, sb, st)
    ctx.set_points_adaptive(*pts, cap=cap, max_depth=dmax)
    return o, ctx


def _compare_structure(o, ctx, dim):
    assert ctx.ncell == o.ncell
    for a, b in zip(ctx.permutation(), o.perm()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(ctx.cells(), o.cells()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(ctx.buffer_csr(), o.buffer_csr()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(ctx.trunc_csr(), o.trunc_csr()):
        np.testing.assert_array_equal(a, b)
    lg, lo = ctx.leaves(), o.leaves()
    for k in ("xc", "yc", "depth", "cell0", "blen", "tlen") + (("zc",) if dim == 3 else ()):
        np.testing.assert_array_equal(lg[k], lo[k], err_msg=k)


@pytest.mark.parametrize("name", sorted(TREE_CASES))
def test_tree_structure_bit_exact(name):
    dim, sigma, nb, sb, st, pts, cap, dmax = TREE_CASES[name]
    _announce(f"tree structure {name}")
    o, ctx = _pair(dim, sigma, nb, sb, st, pts, cap, dmax)
    _compare_structure(o, ctx, dim)
    rng = np.random.default_rng(29)
    for nrhs in (1, 3):
        g = rng.standard_normal((nrhs, o.ni))
        want = o.matvec(g)
        assert np.linalg.norm(ctx.matvec(g) - want) <= 1e-12 * np.linalg.norm(want)
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_tree_without_splits_equals_uniform(dim):
    """A cap above every cell's occupancy gives prbf_set_points' structure."""
    _announce(f"tree without splits {dim}D")
    p = problems.lattice_2d_problem(30, seed=5) if dim == 2 else problems.lattice_3d_problem(10, seed=5)
    pts = [p.x, p.y] + ([p.z] if dim == 3 else [])
    u = Context.from_problem(p)
    u.set_points(*pts)
    t = Context.from_problem(p)
    t.set_points_adaptive(*pts, cap=10 ** 6, max_depth=4)
    assert t.ncell == u.ncell
    for f in ("permutation", "cells", "buffer_csr", "trunc_csr"):
        for a, b in zip(getattr(t, f)(), getattr(u, f)()):
            np.testing.assert_array_equal(a, b)
    u.close()
    t.close()


def _uneven_lattice():
    """48x48 jittered lattice, a 12-spacing buffer, cells of more than 15
    particles split once: 34 leaves at depth 0 and 440 at depth 1, blocks of
    63 to 784 unknowns, well conditioned (the lattice's h/sigma = 0.8)."""
    return problems.lattice_2d_problem(48, seed=0x5EED0001, buffer_cells=12), (15, 3)


def test_tree_rasm_and_plain_solve_match_oracle():
    p, tree = _uneven_lattice()
    _announce("tree rasm + plain solve")
    o = Oracle.from_problem(p, tree=tree)
    o.rasm_setup()
    r = np.random.default_rng(37).standard_normal(p.n)
    want = o.rasm_apply(r)
    for order in (None, ("right", 4, 1), ("left", 4, 2)):
        ctx = Context.from_problem(p)
        if order:
            ctx.set_tuning(chol_order=order[0], chol_waves=order[1], chol_panel=order[2])
        ctx.set_points_adaptive(p.x, p.y, cap=tree[0], max_depth=tree[1])
        depth = ctx.leaves()["depth"]
        assert (depth == 0).any() and (depth == 1).any()
        ctx.rasm_setup("rasm")
        got = ctx.rasm_apply(r)
        assert np.linalg.norm(got - want) <= 1e-8 * np.linalg.norm(want), order
        ctx.close()
    pi, _ = o.perm()
    ref = o.solve_sorted(p.w[0][pi], rtol=1e-10, restart=30, max_it=1000)
    g_cpu = np.empty(p.n)
    g_cpu[pi] = ref["gamma"]
    ctx = Context.from_problem(p)
    ctx.set_points_adaptive(p.x, p.y, cap=tree[0], max_depth=tree[1])
    g, inf = ctx.solve(p.w[0], ksp="gmres", pc="rasm", rtol=1e-10, restart=30, max_it=1000)
    _announce(f"tree plain solve: its gpu {inf.its} cpu {ref['its']}")
    assert inf.reason == 2 and ref["reason"] == 2 and inf.its == ref["its"]
    assert np.linalg.norm(g - g_cpu) <= 1e-10 * np.linalg.norm(g_cpu)
    assert inf.n_regularised == 0
    g32, inf32 = ctx.solve(p.w[0], ksp="fgmres", pc="rasm32", rtol=1e-10, restart=30, max_it=1000)
    w = np.empty(p.n)
    w[pi] = o.matvec(g32[pi])
    res = np.linalg.norm(w - p.w[0]) / np.linalg.norm(p.w[0])
    _announce(f"tree mixed solve: its {inf32.its}, true residual {res:.2e}")
    assert inf32.reason == 2 and res <= 1.05e-10
    ctx.close()


def test_mixture_blocks_are_regularised_and_solve_stays_finite():
    """The C3 mixture at N = 2000: its points cluster (nearest-neighbour
    distances down to 1% of the mean spacing), so its blocks are numerically
    singular even in FP64.  The mixed and plain setups regularise those
    pivots, report how many blocks needed it, and FGMRES's residual estimate
    stays finite and never grows."""
    p = problems.mixture_2d_problem(2000, seed=0x5EED0E03)
    _announce("mixture regularised setup")
    ctx = Context.from_problem(p)
    ctx.set_points_adaptive(p.x, p.y, cap=12, max_depth=8)
    for pc in ("rasm32", "rasm"):
        g, inf = ctx.solve(p.w[0], ksp="fgmres", pc=pc, rtol=1e-10, restart=30, max_it=60)
        h = inf.hist
        _announce(f"{pc}: regularised blocks {inf.n_regularised} of {ctx.ncell}, its {inf.its}, "
                  f"residual {h[-1] / h[0]:.3e}")
        assert inf.n_regularised > 0
        assert np.isfinite(g).all() and np.isfinite(h).all()
        assert (np.diff(h[:31]) <= 1e-12 * h[0]).all()  # the estimate never grows within the first cycle
    ctx.close()


def _sha(*arrays):
    hh = hashlib.sha256()
    for a in arrays:
        hh.update(np.ascontiguousarray(a).tobytes())
    return hh.hexdigest()


def test_full_size_mixture_structure_against_oracle():
    """The C3 workload's 4*10^6-point mixture: the adaptive structure
    (cap 12, depth 8) bit-identical to the oracle's (SHA-256 of the
    permutation, leaf ranges and both lists), the matvec within 1e-12."""
    p = problems.mixture_2d_problem(4_000_000)
    _announce("full size mixture: oracle structure")
    o = Oracle.from_problem(p, tree=(12, 8))
    pi, _ = o.perm()
    ref = _sha(pi, *o.cells()[:2], *o.buffer_csr(), *o.trunc_csr())
    ctx = Context.from_problem(p)
    ctx.set_points_adaptive(p.x, p.y, cap=12, max_depth=8)
    assert ctx.ncell == o.ncell
    got = _sha(ctx.permutation()[0], *ctx.cells()[:2], *ctx.buffer_csr(), *ctx.trunc_csr())
    assert got == ref
    _announce("full size mixture: matvec")
    g = np.random.default_rng(8).standard_normal(p.n)
    want = o.matvec(g)
    assert np.linalg.norm(ctx.matvec(g) - want) <= 1e-12 * np.linalg.norm(want)
    ctx.close()

"""GPU parity at the edges the lattice tests do not reach, and at full size.

* clusters of more than 128 and more than 256 targets (the matvec's second
  and third 128-target chunks; the reference GPU path handles at most 128,
  2d/matmultgpu.cxx:23-25): lists bit-exact, matvec within 1e-12;
* clusters of more than 32 core particles (three and four 16-row chunks of
  core rows): the FP64 preconditioner of both factorisation orders against
  the oracle's LU (2d/matmult.cxx:83-93 blocks, PCASM restrict apply);
* the dense block export (mysubmat, 2d/matmult.cxx:53-104) against the
  oracle's block and buffer list;
* the mixed preconditioner's precision rule (blocks under 32 unknowns in
  FP64) and its FP32 blocks against an independent FP32 factorisation, for
  both forms of the FP32 backward substitution (two tile rows per step or
  one), odd and even tile-row counts and two or three right-hand-side
  chunks;
* the N = 4*10^6 lattice workload: lists, matvec and the true residual of
  the plain and mixed gamma, all against the oracle.
"""

import hashlib
import sys

import numpy as np
import pytest

from oracle.oracle import Oracle
from petrbf_amd import problems
from petrbf_amd.rbf import Context
from tests._fp32ref import fp32_apply

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _crowded():
    """A 16x16 background lattice plus 300 points in one cell (a crowded
    cluster of more than 256 targets next to ordinary ones)."""
    xs, ys, h = problems.jittered_lattice_2d(16, seed=9)
    sigma = h / 0.8
    box = 3.2 * sigma + 1e-6  # cell side (2d/get_cluster.h:29)
    u = problems.uniform01(0x5EED00CC, 600).reshape(-1, 2)
    # strictly inside cell (2, 1) of the grid anchored at the lattice's minimum
    x = np.concatenate([xs, xs.min() - 1e-6 + (2.05 + 0.9 * u[:, 0]) * box])
    y = np.concatenate([ys, ys.min() - 1e-6 + (1.05 + 0.9 * u[:, 1]) * box])
    return x, y, sigma, 3.2, 16.0, 13.2


def _cases_big_clusters():
    out = {}
    for cpb in (12, 17):  # 144 and 289 targets per interior cluster
        p = problems.lattice_2d_problem(3 * cpb + 6, seed=0x5EED0100 + cpb, cells_per_box=cpb, buffer_cells=2)
        out[f"lattice_cells{cpb}"] = (p.x, p.y, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc)
    out["crowded_cell"] = _crowded()
    return out


BIG = _cases_big_clusters()


@pytest.mark.parametrize("name", sorted(BIG))
def test_big_clusters_lists_and_matvec(name):
    x, y, sigma, nb, sb, st = BIG[name]
    _announce(f"big clusters {name}")
    o = Oracle(2, sigma, nb, sb, st, x, y, None)
    ctx = Context(2, sigma, nb, sb, st)
    ctx.set_points(x, y)
    ista, iend, _, _ = o.cells()
    tmax = int((iend - ista + 1).max())
    assert tmax > 128, tmax
    for a, b in zip(ctx.permutation(), o.perm()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(ctx.cells(), o.cells()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(ctx.buffer_csr(), o.buffer_csr()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(ctx.trunc_csr(), o.trunc_csr()):
        np.testing.assert_array_equal(a, b)
    rng = np.random.default_rng(23)
    for nrhs in (1, 3):
        g = rng.standard_normal((nrhs, x.shape[0]))
        want = o.matvec(g)
        got = ctx.matvec(g)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), nrhs
        # and cluster by cluster for the crowded ones (every 128-target chunk)
        for c in np.nonzero(iend - ista + 1 > 128)[0]:
            sl = slice(ista[c], iend[c] + 1)
            assert np.linalg.norm(got[..., sl] - want[..., sl]) <= 1e-12 * np.linalg.norm(want[..., sl]), c
    _announce(f"{name}: largest cluster {tmax} targets")
    ctx.close()


@pytest.mark.parametrize("cpb", [6, 8])
def test_many_core_rows_both_orders(cpb):
    """Cells of 6x6 and 8x8 lattice spacings: 36 and 64 core particles per
    block (three and four 16-row chunks of core rows, the last one cut to T
    rounded up to 4), blocks of 100 and 144 unknowns.  FP64 preconditioner
    of every order against the oracle's LU."""
    _announce(f"core rows T>32 cells {cpb}")
    p = problems.lattice_2d_problem(48, seed=0x5EED0200 + cpb, cells_per_box=cpb, buffer_cells=2)
    o = Oracle.from_problem(p)
    o.rasm_setup()
    ista, iend, _, _ = o.cells()
    assert int((iend - ista + 1).max()) > 32
    rng = np.random.default_rng(31)
    r = rng.standard_normal(p.n)
    want = o.rasm_apply(r)
    for order, nw, npc in ((None, 0, 0), ("right", 4, 1), ("right", 8, 2), ("left", 4, 1), ("left", 4, 2)):
        ctx = Context.from_problem(p)
        if order:
            ctx.set_tuning(chol_waves=nw, chol_panel=npc, chol_order=order)
        ctx.set_points(p.x, p.y)
        ctx.rasm_setup("rasm")
        got = ctx.rasm_apply(r)
        assert np.linalg.norm(got - want) <= 1e-8 * np.linalg.norm(want), (order, nw, npc)
        # core rows are rows of A_c^{-1}: M_c A_c = I on the core
        if order is None:
            for c in range(0, ctx.ncell, 7):
                A, _ = ctx.block(c)
                M = ctx.core_rows(c)
                if A.size:
                    assert np.abs(M @ A - np.eye(A.shape[0])[:M.shape[0]]).max() < 1e-6, c
        ctx.close()


@pytest.mark.parametrize("name", ["lattice24", "thin_buffer", "crowded_cell", "lattice3d"])
def test_block_export_matches_oracle(name):
    """prbf_rasm_get_block: the dense block and the subdomain indices of
    every cluster against the oracle's (2d/matmult.cxx:83-93)."""
    _announce(f"block export {name}")
    if name == "lattice24":
        p = problems.lattice_2d_problem(24, seed=11)
        args = (2, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, p.x, p.y, None)
    elif name == "thin_buffer":
        p = problems.lattice_2d_problem(20, seed=12, buffer_cells=2)
        args = (2, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, p.x, p.y, None)
    elif name == "crowded_cell":
        x, y, sigma, nb, sb, st = _crowded()
        args = (2, sigma, nb, sb, st, x, y, None)
    else:
        p = problems.lattice_3d_problem(9, seed=5)
        args = (3, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, p.x, p.y, p.z)
    dim, sigma, nb, sb, st, x, y, z = args
    o = Oracle(dim, sigma, nb, sb, st, x, y, z)
    ctx = Context(dim, sigma, nb, sb, st)
    ctx.set_points(x, y, z)
    bptr, bidx = o.buffer_csr()
    checked = 0
    for c in range(ctx.ncell):
        A, idx = ctx.block(c)
        np.testing.assert_array_equal(idx, bidx[bptr[c]:bptr[c + 1]])
        if idx.size == 0:
            continue
        B = o.block(c)
        assert A.shape == B.shape
        assert np.abs(A - B).max() <= 1e-14 * np.abs(B).max(), c
        np.testing.assert_array_equal(A, A.T)
        checked += 1
    assert checked > 0
    ctx.close()


def _fp32_bound_check(ctx, p, r, z32, z64, label, slack=10.0):
    """||z32 - z64|| within `slack` times what LAPACK's FP32 factorisation of
    the same blocks gives, plus a floor for blocks that FP32 solves exactly."""
    x, y, zc = ctx.sorted_points()
    bptr, bidx = ctx.buffer_csr()
    ista, iend, _, _ = ctx.cells()
    coords = [x, y, zc] if p.dim == 3 else [x, y, None]
    zref, _ = fp32_apply(coords, bptr, bidx, ista, iend, p.sigma, r)
    e_gpu = np.linalg.norm(z32 - z64)
    e_ref = np.linalg.norm(zref - z64)
    _announce(f"{label}: FP32 apply error {e_gpu / np.linalg.norm(z64):.2e}, "
              f"LAPACK FP32 reference {e_ref / np.linalg.norm(z64):.2e}")
    assert e_gpu <= slack * e_ref + 1e-6 * np.linalg.norm(z64), (e_gpu, e_ref)
    return e_gpu, e_ref


def test_fp32_blocks_match_an_independent_fp32_factorisation():
    """The FP32 preconditioner's distance from the exact one is of the order
    of FP32 rounding times the blocks' conditioning, measured by LAPACK's own
    FP32 Cholesky on the same blocks (in place of a fixed loose bound)."""
    _announce("fp32 apply vs LAPACK fp32 reference")
    p = problems.lattice_2d_problem(40, seed=11)
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y)
    r = np.random.default_rng(11).standard_normal(p.n)
    ctx.rasm_setup("rasm")
    z64 = ctx.rasm_apply(r)
    ctx.rasm_setup("rasm32")
    z32 = ctx.rasm_apply(r)
    _fp32_bound_check(ctx, p, r, z32, z64, "lattice 40x40")
    ctx.close()


@pytest.mark.parametrize("cpb,buf", [(5, 5), (6, 5)])
def test_fp32_backward_substitution_forms(cpb, buf):
    """Cells of 5x5 and 6x6 spacings with a 5-spacing buffer: interior blocks
    of 225 and 256 unknowns (15 and 16 tile rows, so the two-row backward
    substitution of the FP32 2-column kernels ends on a single row or on a
    pair; edge blocks give other counts) and 25 or 36 core particles (two
    and three 16-column chunks of right-hand sides).  Both FP32 forms, two
    rows per step (left-looking, 2-column panels) and one row per step
    (left-looking 1 and 4 columns, right-looking), are as close to the exact
    FP64 preconditioner as LAPACK's FP32 factorisation of the same blocks,
    and the FP64 one matches the oracle's LU."""
    _announce(f"fp32 substitution forms cells {cpb} buffer {buf}")
    p = problems.lattice_2d_problem(40, seed=0x5EED0300 + cpb, cells_per_box=cpb, buffer_cells=buf)
    o = Oracle.from_problem(p)
    o.rasm_setup()
    ista, iend, _, _ = o.cells()
    bptr, _ = o.buffer_csr()
    kt = set(((np.diff(bptr) + 15) // 16).tolist())
    assert {15, 16} & kt and int((iend - ista + 1).max()) > 16, (kt, int((iend - ista + 1).max()))
    assert any(k % 2 for k in kt) and any(k % 2 == 0 for k in kt), kt
    r = np.random.default_rng(41).standard_normal(p.n)
    want = o.rasm_apply(r)
    for order, nw, npc in (("left", 4, 2), ("left", 8, 2), ("left", 4, 1), ("left", 4, 4), ("right", 4, 2)):
        ctx = Context.from_problem(p)
        ctx.set_tuning(chol_waves=nw, chol_panel=npc, chol_order=order)
        ctx.set_points(p.x, p.y)
        ctx.rasm_setup("rasm")
        z64 = ctx.rasm_apply(r)
        assert np.linalg.norm(z64 - want) <= 1e-8 * np.linalg.norm(want), (order, nw, npc)
        ctx.rasm_setup("rasm32")
        info = ctx.info()
        assert (info["chol_order"], info["chol_waves"], info["chol_panel"]) == (
            1 if order == "right" else 2, nw, npc)
        z32 = ctx.rasm_apply(r)
        _fp32_bound_check(ctx, p, r, z32, z64, f"{order} {nw}x{npc}")
        ctx.close()


def test_mixed_precision_rule_small_blocks_in_fp64():
    """Mixed mode keeps blocks of fewer than 32 unknowns in FP64: with a
    one-spacing buffer the edge and corner blocks (25-30 unknowns) are FP64
    and must match the oracle's LU to FP64 accuracy, the interior ones
    (36 unknowns) are FP32 and meet the FP32 reference bound; FGMRES on the
    mixed preconditioner still reaches the FP64 residual."""
    _announce("precision rule")
    p = problems.lattice_2d_problem(40, seed=0x5EED0300, buffer_cells=1)
    o = Oracle.from_problem(p)
    o.rasm_setup()
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y)
    bptr, bidx = ctx.buffer_csr()
    ista, iend, _, _ = ctx.cells()
    nbuf = np.diff(bptr)
    small = np.nonzero((nbuf > 0) & (nbuf < 32))[0]
    large = np.nonzero(nbuf >= 32)[0]
    assert small.size > 0 and large.size > 0
    r = np.random.default_rng(41).standard_normal(p.n)
    want = o.rasm_apply(r)
    ctx.rasm_setup("rasm32")
    z32 = ctx.rasm_apply(r)
    core_small = np.concatenate([bidx[bptr[c]:bptr[c] + iend[c] - ista[c] + 1] for c in small])
    core_large = np.concatenate([bidx[bptr[c]:bptr[c] + iend[c] - ista[c] + 1] for c in large])
    es = np.linalg.norm(z32[core_small] - want[core_small]) / np.linalg.norm(want[core_small])
    _announce(f"small (FP64) blocks {small.size}: rel err {es:.2e}; large (FP32) blocks {large.size}")
    assert es <= 1e-8
    # the FP32 blocks: same bound as an independent FP32 factorisation
    x, y, _ = ctx.sorted_points()
    zref, _ = fp32_apply([x, y, None], bptr, bidx, ista, iend, p.sigma, r, clusters=large)
    e_gpu = np.linalg.norm(z32[core_large] - want[core_large])
    e_ref = np.linalg.norm(zref[core_large] - want[core_large])
    assert e_gpu <= 10 * e_ref + 1e-6 * np.linalg.norm(want[core_large]), (e_gpu, e_ref)
    g, inf = ctx.solve(p.w[0], ksp="fgmres", pc="rasm32", rtol=1e-10, restart=30, max_it=1000)
    assert inf.reason == 2
    pi, _ = o.perm()
    w = np.empty(p.n)
    w[pi] = o.matvec(g[pi])
    res = np.linalg.norm(w - p.w[0]) / np.linalg.norm(p.w[0])
    _announce(f"fgmres on the mixed preconditioner: its {inf.its}, true residual {res:.2e}")
    assert res <= 1.05e-10
    ctx.close()


def _sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_full_size_lattice_against_oracle():
    """The N = 4*10^6 lattice workload (2000 x 2000): permutation, cell
    ranges and both lists bit-identical to the oracle's (SHA-256), the matvec
    within 1e-12 of the oracle's, and the plain and mixed gamma reaching
    ||A~_oracle gamma - w|| <= 1.05e-10 ||w|| with the oracle's own operator."""
    _announce("full size N=4M: oracle lists")
    p = problems.lattice_2d_problem(2000, seed=0x5EED0001)
    o = Oracle.from_problem(p)
    pi, _ = o.perm()
    ref = _sha(pi, *o.cells()[:2], *o.buffer_csr(), *o.trunc_csr())
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y)
    got = _sha(ctx.permutation()[0], *ctx.cells()[:2], *ctx.buffer_csr(), *ctx.trunc_csr())
    assert got == ref
    _announce("full size N=4M: matvec")
    g = np.random.default_rng(4).standard_normal(p.n)
    want = o.matvec(g)
    mv = ctx.matvec(g)
    assert np.linalg.norm(mv - want) <= 1e-12 * np.linalg.norm(want)
    w = p.w[0]
    for ksp, pc in (("gmres", "rasm"), ("fgmres", "rasm32")):
        _announce(f"full size N=4M: solve {ksp}/{pc}")
        gam, inf = ctx.solve(w, ksp=ksp, pc=pc, rtol=1e-10, restart=30, max_it=1000)
        assert inf.reason == 2, inf.reason
        assert inf.n_regularised == 0
        aw = np.empty(p.n)
        aw[pi] = o.matvec(gam[pi])
        res = np.linalg.norm(aw - w) / np.linalg.norm(w)
        _announce(f"{ksp}/{pc}: its {inf.its}, true residual with the oracle's operator {res:.3e}")
        assert res <= 1.05e-10
    ctx.close()


@pytest.mark.parametrize("scale", [1e-3, 1e3])
def test_lattice_at_any_scale_is_not_regularised(scale):
    """The pivot floor is relative to the Gaussian's diagonal 1/(2 pi sigma^2)
    and the padding of every block sits at that diagonal, so a well-separated
    lattice is factorised exactly at any length scale (sigma 1e-3 or 1e3 times
    the unit square's): no regularised block, and the mixed and plain solves
    converge as on the unit square."""
    _announce(f"lattice scaled by {scale}")
    p = problems.lattice_2d_problem(40, seed=0x5EED0400)
    q = problems.lattice_2d_problem(40, seed=0x5EED0400)
    for a in ("x", "y"):
        setattr(q, a, getattr(p, a) * scale)
    q.sigma = p.sigma * scale
    for prob in (p, q):
        ctx = Context.from_problem(prob)
        ctx.set_points(prob.x, prob.y)
        for ksp, pc in (("fgmres", "rasm32"), ("gmres", "rasm")):
            _, inf = ctx.solve(p.w[0], ksp=ksp, pc=pc, rtol=1e-10, restart=30, max_it=1000)
            assert inf.n_regularised == 0 and inf.reason == 2, (scale, pc, inf.n_regularised)
        ctx.close()

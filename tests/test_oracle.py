"""Pins the CPU oracle against definitions that do not share its code:
brute-force lists, long-double dense sums, LAPACK (scipy) factorisations,
closed forms, and the minimal-residual property of GMRES."""

import math

import numpy as np
import pytest
import scipy.linalg as sla

from oracle.oracle import Oracle, dense_gmres, dense_lu, dense_lu_solve
from tests import _brute as B
from tests.cases import all_cases, lattice


def build_oracle(case):
    dim, sigma, nb, sb, st, tgt, src = case
    z = tgt[2] if dim == 3 else None
    if src is None:
        return Oracle(dim, sigma, nb, sb, st, tgt[0], tgt[1], z)
    zj = src[2] if dim == 3 else None
    return Oracle(dim, sigma, nb, sb, st, tgt[0], tgt[1], z, src[0], src[1], zj)


CASES = all_cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_lists_match_brute_force(name):
    case = CASES[name]
    dim, sigma, nb, sb, st, tgt, src = case
    o = build_oracle(case)
    ref = B.lists(dim, sigma, nb, sb, st, tgt, src)
    g = o.grid()
    assert [g["nx"], g["ny"], g["nz"]] == ref["grid"]["nc"]
    assert g["nb_buffer"] == ref["grid"]["reach_b"] and g["nb_trunc"] == ref["grid"]["reach_t"]
    pi, pj = o.perm()
    np.testing.assert_array_equal(pi, ref["perm_i"])
    np.testing.assert_array_equal(pj, ref["perm_j"])
    ista, iend, jsta, jend = o.cells()
    np.testing.assert_array_equal(ista, ref["ista"])
    np.testing.assert_array_equal(iend, ref["iend"])
    np.testing.assert_array_equal(jsta, ref["jsta"])
    np.testing.assert_array_equal(jend, ref["jend"])
    for got, want in zip(B.csr_to_lists(*o.buffer_csr()), ref["buffer"]):
        np.testing.assert_array_equal(got, want)
    for got, want in zip(B.csr_to_lists(*o.trunc_csr()), ref["trunc"]):
        np.testing.assert_array_equal(got, want)


def test_list_invariants_and_box_semantics():
    """Buffer lists are the core then every particle inside the buffer box,
    wherever it lies (the cell reach covers the box); cores partition the
    unknowns; the truncation list is exactly the particles in the trunc box."""
    p = lattice(30)
    o = Oracle.from_problem(p)
    x, y, _ = o.sorted_coords()
    xc, yc, _ = o.centres()
    g = o.grid()
    ista, iend, _, _ = o.cells()
    bl = B.csr_to_lists(*o.buffer_csr())
    tl = B.csr_to_lists(*o.trunc_csr())
    covered = np.zeros(p.n, int)
    blen = p.sigma_buffer * p.sigma / 2 + 1e-6
    tlen = p.sigma_trunc * p.sigma / 2 + 1e-6
    for c in range(o.ncell):
        core = np.arange(ista[c], iend[c] + 1)
        covered[core] += 1
        if core.size == 0:
            assert bl[c].size == 0 and tl[c].size == 0
            continue
        np.testing.assert_array_equal(bl[c][: core.size], core)
        inb = np.nonzero((np.abs(x - xc[c]) < blen) & (np.abs(y - yc[c]) < blen))[0]
        assert set(bl[c].tolist()) == set(inb.tolist()) | set(core.tolist())
        intr = np.nonzero((np.abs(x - xc[c]) < tlen) & (np.abs(y - yc[c]) < tlen))[0]
        assert set(tl[c].tolist()) == set(intr.tolist()) | set(core.tolist())
        assert len(set(tl[c].tolist())) == tl[c].size
    assert (covered == 1).all()
    assert g["nb_buffer"] == 3 and g["nb_trunc"] == 2   # 16 sigma buffer: eps pushes reach to 3


def test_points_outside_grid_rejected_by_construction():
    """The grid spans the bounding box, so every point lands in a cell; the
    maximum coordinate maps strictly inside the last cell (2d/get_cluster.h:25-34)."""
    x = np.array([-1.0, 1.0, 0.3])
    y = np.array([-1.0, 1.0, -0.2])
    o = Oracle(2, 0.1, 3.2, 16.0, 13.2, x, y)
    g = o.grid()
    ista, iend, _, _ = o.cells()
    assert (iend - ista + 1).sum() == 3
    assert g["nx"] == math.ceil((2 + 2e-6) / (0.32 + 1e-6))


@pytest.mark.parametrize("name", ["lattice24", "clustered", "targets_sources", "lattice3d_9", "random3d", "single"])
def test_matvec_matches_long_double_dense(name):
    case = CASES[name]
    dim, sigma, nb, sb, st, tgt, src = case
    o = build_oracle(case)
    ref = B.lists(dim, sigma, nb, sb, st, tgt, src)
    A = B.dense_truncated(dim, sigma, ref)
    rng = np.random.default_rng(1)
    g = rng.standard_normal(A.shape[1])
    want = (A @ g.astype(np.longdouble)).astype(np.float64)
    got = o.matvec(g)
    scale = (np.abs(A) @ np.abs(g.astype(np.longdouble))).astype(np.float64)
    assert np.all(np.abs(got - want) <= 1e-14 * scale + 1e-300)


def test_single_particle_closed_form():
    sigma = 0.1
    o = Oracle(2, sigma, 3.2, 16.0, 13.2, np.array([0.2]), np.array([0.7]))
    w = o.matvec(np.array([3.0]))
    assert w[0] == pytest.approx(3.0 / (2 * math.pi * sigma ** 2), rel=1e-15)


def test_two_particle_closed_form_and_3d_normalisation():
    """Two sources at distance d: w_0 = g_0 G(0) + g_1 G(d); 3D keeps the 2D
    constant 1/(2 pi sigma^2) (3d/matmult.cxx:40-41)."""
    sigma, d = 0.1, 0.13
    for dim in (2, 3):
        pts = [np.array([0.0, d]), np.array([0.0, 0.0])] + ([np.array([0.0, 0.0])] if dim == 3 else [])
        o = Oracle(dim, sigma, 3.2, 16.0, 13.2, *pts)
        pi, _ = o.perm()
        g = np.array([2.0, -1.0])[pi]
        w = o.matvec(g)
        gw = lambda r: math.exp(-r * r / (2 * sigma ** 2)) / (2 * math.pi * sigma ** 2)
        w_orig = np.empty(2)
        w_orig[pi] = w
        assert w_orig[0] == pytest.approx(2.0 * gw(0) - 1.0 * gw(d), rel=1e-14)
        assert w_orig[1] == pytest.approx(-1.0 * gw(0) + 2.0 * gw(d), rel=1e-14)


def test_matvec_column_is_operator_column():
    """g = e_k returns column k of A~: entries are G(x_i - x_k) exactly where k
    is in cluster(i)'s truncation list, zero elsewhere."""
    p = lattice(20)
    o = Oracle.from_problem(p)
    x, y, _ = o.sorted_coords()
    ista, iend, _, _ = o.cells()
    tl = B.csr_to_lists(*o.trunc_csr())
    k = 137
    e = np.zeros(p.n)
    e[k] = 1.0
    col = o.matvec(e)
    for c in range(o.ncell):
        rows = np.arange(ista[c], iend[c] + 1)
        if k in set(tl[c].tolist()):
            r2 = (x[rows] - x[k]) ** 2 + (y[rows] - y[k]) ** 2
            np.testing.assert_allclose(col[rows], np.exp(-r2 / (2 * p.sigma ** 2)) / (2 * np.pi * p.sigma ** 2),
                                       rtol=1e-15)
        else:
            assert not col[rows].any()


def test_lu_matches_lapack_getrf():
    """Same pivot sequence as LAPACK dgetrf (scipy) and factors to rounding."""
    rng = np.random.default_rng(4)
    for n in (1, 7, 40, 90):
        pts = rng.uniform(0, 1, (n, 2))
        sigma = 0.08
        r2 = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
        A = np.exp(-r2 / (2 * sigma ** 2)) / (2 * np.pi * sigma ** 2) + np.diag(rng.uniform(0, 1e-3, n))
        A += rng.standard_normal((n, n)) * 1e-3 * A.max()    # non-symmetric: pivoting matters
        LU, piv = dense_lu(A)
        lu_ref, piv_ref = sla.lu_factor(A)
        np.testing.assert_array_equal(piv, piv_ref)
        np.testing.assert_allclose(LU, lu_ref, rtol=1e-10, atol=1e-12 * np.abs(A).max())
        b = rng.standard_normal(n)
        np.testing.assert_allclose(dense_lu_solve(LU, piv, b), np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)


def test_rasm_blocks_and_apply_match_dense_definition():
    p = lattice(20)
    o = Oracle.from_problem(p)
    o.rasm_setup()
    x, y, _ = o.sorted_coords()
    ista, iend, _, _ = o.cells()
    bl = B.csr_to_lists(*o.buffer_csr())
    rng = np.random.default_rng(9)
    r = rng.standard_normal(p.n)
    z = o.rasm_apply(r)
    zref = np.full(p.n, np.nan)
    for c in range(o.ncell):
        b = bl[c]
        if b.size == 0:
            continue
        r2 = (x[b, None] - x[None, b]) ** 2 + (y[b, None] - y[None, b]) ** 2
        A = np.exp(-r2 / (2 * p.sigma ** 2)) / (2 * np.pi * p.sigma ** 2)
        np.testing.assert_allclose(o.block(c), A, rtol=1e-15, atol=0)
        sol = sla.solve(A, r[b], assume_a="pos")
        T = iend[c] - ista[c] + 1
        zref[b[:T]] = sol[:T]
    assert not np.isnan(zref).any()   # cores partition the unknowns
    assert np.linalg.norm(z - zref) <= 1e-8 * np.linalg.norm(zref)


def _krylov_min_residual(M, b, k):
    """min over x in K_k(M, b) of ||b - M x|| by dense least squares on an
    orthonormal Krylov basis."""
    n = b.shape[0]
    K = np.zeros((n, k))
    v = b / np.linalg.norm(b)
    for j in range(k):
        K[:, j] = v
        v = M @ v
        for i in range(j + 1):
            v -= (K[:, i] @ v) * K[:, i]
        nv = np.linalg.norm(v)
        if nv == 0:
            break
        v /= nv
    Q, _ = np.linalg.qr(K)
    y, *_ = np.linalg.lstsq(M @ Q, b, rcond=None)
    return np.linalg.norm(b - M @ Q @ y)


def test_gmres_residual_history_is_minimal_residual():
    """Without restarts, GMRES's k-th residual is the minimum over the Krylov
    space (Saad & Schultz 1986) -- a property no implementation detail can fake."""
    rng = np.random.default_rng(2)
    n = 40
    A = np.eye(n) * 3 + rng.standard_normal((n, n)) * 0.4
    b = rng.standard_normal(n)
    r = dense_gmres(A, b, rtol=1e-13, restart=60, max_it=60)
    assert r["reason"] > 0
    for k in (1, 3, 7, 12):
        assert r["hist"][k] == pytest.approx(_krylov_min_residual(A, b, k), rel=1e-8)
    np.testing.assert_allclose(r["x"], np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_gmres_left_preconditioned_norm_and_restart():
    """Left preconditioning: the tested norm is ||B(b - A x)|| and the first
    entry is ||B b|| (PETSc default with zero guess); restarted GMRES(5)
    reaches the same solution."""
    rng = np.random.default_rng(5)
    n = 30
    A = np.diag(np.linspace(1, 50, n)) + rng.standard_normal((n, n)) * 0.5
    Bm = np.diag(1.0 / np.diag(A))
    b = rng.standard_normal(n)
    r = dense_gmres(A, b, B=Bm, rtol=1e-12, restart=5, max_it=500)
    assert r["reason"] == 2
    assert r["hist"][0] == pytest.approx(np.linalg.norm(Bm @ b), rel=1e-14)
    x = r["x"]
    assert np.linalg.norm(Bm @ (b - A @ x)) <= 1.01e-12 * np.linalg.norm(Bm @ b)
    for k in (1, 2, 4):
        assert r["hist"][k] == pytest.approx(_krylov_min_residual(Bm @ A, Bm @ b, k), rel=1e-8)


def test_fgmres_right_preconditioned_true_residual():
    rng = np.random.default_rng(6)
    n = 30
    A = np.diag(np.linspace(1, 50, n)) + rng.standard_normal((n, n)) * 0.5
    Bm = np.diag(1.0 / np.diag(A))
    b = rng.standard_normal(n)
    r = dense_gmres(A, b, B=Bm, rtol=1e-12, restart=8, max_it=500, flexible=True)
    assert r["reason"] == 2
    assert r["hist"][0] == pytest.approx(np.linalg.norm(b), rel=1e-14)
    assert np.linalg.norm(b - A @ r["x"]) <= 2e-12 * np.linalg.norm(b)
    for k in (1, 3):
        assert r["hist"][k] == pytest.approx(_krylov_min_residual(A @ Bm, b, k), rel=1e-8)


def test_gmres_iteration_cap_reports_divergence():
    rng = np.random.default_rng(8)
    n = 50
    A = np.eye(n) + rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    r = dense_gmres(A, b, rtol=1e-14, restart=3, max_it=10)
    assert r["reason"] == -3 and r["its"] == 10


def test_solve_c1_converges_and_reproduces_field():
    """The solve workload at N = 10^4 (100x100 lattice): GMRES(30) + RASM
    reaches rtol 1e-10 in a handful of iterations; evaluating the strengths
    reproduces the RHS to the true residual."""
    p = lattice(100, seed=0x5EED0001)
    o = Oracle.from_problem(p)
    pi, _ = o.perm()
    w = p.w[0][pi]
    r = o.solve_sorted(w, rtol=1e-10)
    assert r["reason"] == 2 and r["its"] < 40
    g = r["gamma"]
    true_rel = np.linalg.norm(o.matvec(g) - w) / np.linalg.norm(w)
    assert true_rel < 1e-8
    assert r["hist"][-1] <= 1e-10 * r["hist"][0]

"""The oracle's adaptive cluster structure (leaves of an occupancy-capped
subdivision of the reference's grid, 2d/get_cluster.h:20-166 generalised)
pinned against its brute-force definition on tiny inputs, and the property
that with no split it is the reference's uniform structure."""

import numpy as np
import pytest

from oracle.oracle import Oracle
from petrbf_amd import problems
from tests import _brute as B


def _tree_cases():
    out = {}
    p = problems.mixture_2d_problem(600, seed=0x5EED0E01)
    out["mixture600_cap12"] = (2, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, [p.x, p.y], 12, 3)
    out["mixture600_cap4_deep"] = (2, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, [p.x, p.y], 4, 6)
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0.3, 0.05, 300), rng.uniform(-1, 1, 100)])
    y = np.concatenate([rng.normal(-0.2, 0.05, 300), rng.uniform(-1, 1, 100)])
    out["clustered_cap16"] = (2, 0.02, 3.2, 16.0, 13.2, [x, y], 16, 4)
    # exact lattice with a dense patch: points on node centres and edges
    xs, ys, h = problems.jittered_lattice_2d(13, seed=1, jitter=0.0)
    xd, yd, _ = problems.jittered_lattice_2d(9, seed=2, jitter=0.0, lo=-0.2, hi=0.2)
    out["exact_edges"] = (2, h / 0.8, 3.2, 16.0, 13.2, [np.concatenate([xs, xd]), np.concatenate([ys, yd])], 8, 4)
    rng = np.random.default_rng(9)
    pts = [np.concatenate([rng.uniform(-1, 1, 200), rng.normal(0.2, 0.1, 300)]) for _ in range(3)]
    out["3d_blob_cap20"] = (3, 0.1, 2.4, 4.0, 10.4, pts, 20, 3)
    return out


CASES = _tree_cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_tree_oracle_matches_definition(name):
    dim, sigma, nb, sb, st, pts, cap, dmax = CASES[name]
    ref = B.tree_lists(dim, sigma, nb, sb, st, pts, cap, dmax)
    o = Oracle(dim, sigma, nb, sb, st, *pts, tree=(cap, dmax))
    pi, _ = o.perm()
    np.testing.assert_array_equal(pi, ref["perm"])
    ista, iend, _, _ = o.cells()
    np.testing.assert_array_equal(ista, ref["ista"])
    np.testing.assert_array_equal(iend, ref["iend"])
    lv = o.leaves()
    for l, (ctr, d, c, blen, tlen) in enumerate(ref["geo"]):
        assert (lv["depth"][l], lv["cell0"][l]) == (d, c)
        assert lv["xc"][l] == ctr[0] and lv["yc"][l] == ctr[1] and (dim == 2 or lv["zc"][l] == ctr[2])
        assert lv["blen"][l] == blen and lv["tlen"][l] == tlen
    for got, want in ((o.buffer_csr(), ref["buffer"]), (o.trunc_csr(), ref["trunc"])):
        lists = B.csr_to_lists(*got)
        assert len(lists) == len(want)
        for a, b in zip(lists, want):
            np.testing.assert_array_equal(a, b)
    # the structure did split, and leaves respect the cap below dmax
    cnt = iend - ista + 1
    assert (lv["depth"] > 0).any()
    assert ((cnt <= cap) | (lv["depth"] == dmax)).all()


@pytest.mark.parametrize("dim", [2, 3])
def test_tree_without_splits_is_the_uniform_structure(dim):
    """cap above every cell's occupancy (or dmax 0): the reference's grid,
    sort and lists, bit for bit."""
    if dim == 2:
        p = problems.lattice_2d_problem(30, seed=5)
        pts = [p.x, p.y]
    else:
        p = problems.lattice_3d_problem(10, seed=5)
        pts = [p.x, p.y, p.z]
    u = Oracle.from_problem(p)
    for tree in ((10 ** 6, 4), (1, 0)):
        t = Oracle.from_problem(p, tree=tree)
        assert t.ncell == u.ncell
        for a, b in zip(t.perm(), u.perm()):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(t.cells(), u.cells()):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(t.buffer_csr(), u.buffer_csr()):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(t.trunc_csr(), u.trunc_csr()):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(t.leaves()["depth"], 0)
        g = np.random.default_rng(1).standard_normal(p.n)
        np.testing.assert_array_equal(t.matvec(g), u.matvec(g))


def test_tree_lists_are_boxes_and_matvec_is_the_truncated_sum():
    """Invariants on the mixture: every list is exactly the particles in its
    box among the reached cells, core first; the matvec equals the dense sum
    over each leaf's truncation list in long double."""
    p = problems.mixture_2d_problem(2000, seed=0x5EED0E02)
    o = Oracle.from_problem(p, tree=(16, 5))
    xs, ys, _ = o.sorted_coords()
    ista, iend, _, _ = o.cells()
    lv = o.leaves()
    bl = B.csr_to_lists(*o.buffer_csr())
    tl = B.csr_to_lists(*o.trunc_csr())
    for l in range(o.ncell):
        if iend[l] < ista[l]:
            assert bl[l].size == 0 and tl[l].size == 0
            continue
        T = iend[l] - ista[l] + 1
        np.testing.assert_array_equal(bl[l][:T], np.arange(ista[l], iend[l] + 1))
        assert len(set(bl[l].tolist())) == bl[l].size and len(set(tl[l].tolist())) == tl[l].size
        for lst, ln in ((bl[l][T:], lv["blen"][l]), (tl[l][T:], lv["tlen"][l])):
            assert (np.abs(xs[lst] - lv["xc"][l]) < ln).all() and (np.abs(ys[lst] - lv["yc"][l]) < ln).all()
    g = np.random.default_rng(7).standard_normal(p.n)
    w = o.matvec(g)
    s2 = np.longdouble(p.sigma) ** 2
    want = np.zeros(p.n, np.longdouble)
    for l in range(o.ncell):
        if iend[l] < ista[l]:
            continue
        rows = np.arange(ista[l], iend[l] + 1)
        r2 = (xs[rows, None] - xs[None, tl[l]]) ** 2 + (ys[rows, None] - ys[None, tl[l]]) ** 2
        want[rows] = (np.exp(-np.asarray(r2, np.longdouble) / (2 * s2)) / (2 * np.pi * s2)) @ g[tl[l]]
    assert np.linalg.norm(w - want.astype(np.float64)) <= 1e-13 * np.linalg.norm(want.astype(np.float64))

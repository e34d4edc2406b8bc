"""ctypes binding of the CPU oracle (oracle/petrbf_oracle.cpp).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpetrbf_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    subprocess.check_call(["sh", os.path.join(_HERE, "build_oracle.sh")])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    vp, ll, i32, f64 = C.c_void_p, C.c_longlong, C.c_int, C.c_double
    L.orc_create.restype = vp
    L.orc_create.argtypes = [i32, f64, f64, f64, f64]
    L.orc_destroy.argtypes = [vp]
    L.orc_num_threads.restype = i32
    L.orc_set_points.restype = i32
    L.orc_set_points.argtypes = [vp, ll, vp, vp, vp, ll, vp, vp, vp]
    L.orc_set_points_tree.restype = i32
    L.orc_set_points_tree.argtypes = [vp, ll, vp, vp, vp, i32, i32]
    L.orc_get_leaves.argtypes = [vp, _dp, _dp, _dp, _ip, _lp, _dp, _dp]
    for f in ("orc_ncell", "orc_buffer_nnz", "orc_trunc_nnz"):
        getattr(L, f).restype = ll
        getattr(L, f).argtypes = [vp]
    L.orc_get_grid.argtypes = [vp, _ip, _dp]
    L.orc_get_centres.argtypes = [vp, _dp, _dp, vp]
    L.orc_get_perm.argtypes = [vp, _ip, vp]
    L.orc_get_cells.argtypes = [vp, _ip, _ip, vp, vp]
    L.orc_get_buffer.argtypes = [vp, _lp, _ip]
    L.orc_get_trunc.argtypes = [vp, _lp, _ip]
    L.orc_get_sorted.argtypes = [vp, _dp, _dp, vp]
    L.orc_matvec.restype = i32
    L.orc_matvec.argtypes = [vp, i32, _dp, _dp]
    L.orc_rasm_setup.restype = i32
    L.orc_rasm_setup.argtypes = [vp]
    L.orc_rasm_setup_sample.restype = i32
    L.orc_rasm_setup_sample.argtypes = [vp, ll]
    L.orc_rasm_apply_sample.restype = i32
    L.orc_rasm_apply_sample.argtypes = [vp, _dp, _dp]
    L.orc_rasm_apply.restype = i32
    L.orc_rasm_apply.argtypes = [vp, _dp, _dp]
    L.orc_get_block.restype = i32
    L.orc_get_block.argtypes = [vp, ll, _dp]
    L.orc_get_lu.restype = i32
    L.orc_get_lu.argtypes = [vp, ll, _dp, _ip]
    L.orc_dense_lu.restype = i32
    L.orc_dense_lu.argtypes = [i32, _dp, _ip]
    L.orc_dense_lu_solve.argtypes = [i32, _dp, _ip, _dp]
    L.orc_dense_gmres.restype = i32
    L.orc_dense_gmres.argtypes = [i32, _dp, vp, _dp, _dp, f64, i32, i32, i32,
                                  C.POINTER(i32), _dp, i32, C.POINTER(i32)]
    L.orc_solve_sorted.restype = i32
    L.orc_solve_sorted.argtypes = [vp, _dp, _dp, f64, i32, i32, i32,
                                   C.POINTER(i32), _dp, i32, C.POINTER(i32)]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """One cluster structure over a target set (and optionally a distinct
    source set), with the reference's lists, matvec, RASM and GMRES."""

    def __init__(self, dim, sigma, nsigma_box, sigma_buffer, sigma_trunc,
                 xi, yi, zi=None, xj=None, yj=None, zj=None, tree=None):
        """tree=(cap, dmax): the adaptive structure (one point set; cells of
        more than cap particles split down to depth dmax)."""
        L = lib()
        self.dim = dim
        self.h = L.orc_create(dim, sigma, nsigma_box, sigma_buffer, sigma_trunc)
        if not self.h:
            raise ValueError("bad dimension")
        self._keep = [_f64(a) for a in (xi, yi, zi, xj, yj, zj)]
        xi, yi, zi, xj, yj, zj = self._keep
        self.ni = xi.shape[0]
        self.nj = self.ni if xj is None else xj.shape[0]
        self.tree = tree
        if tree is not None:
            if xj is not None:
                raise ValueError("the adaptive structure takes one point set")
            st = L.orc_set_points_tree(self.h, self.ni, _ptr(xi), _ptr(yi), _ptr(zi), int(tree[0]), int(tree[1]))
        else:
            st = L.orc_set_points(self.h, self.ni, _ptr(xi), _ptr(yi), _ptr(zi),
                                  self.nj, _ptr(xj), _ptr(yj), _ptr(zj))
        if st != 0:
            raise ValueError(f"orc_set_points failed ({st})")
        self.ncell = L.orc_ncell(self.h)

    @classmethod
    def from_problem(cls, p, tree=None):
        return cls(p.dim, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, p.x, p.y, p.z, tree=tree)

    def leaves(self):
        """Adaptive structure: per leaf its centre (x, y, z), depth, level-0
        cell, buffer and truncation half-widths."""
        n = self.ncell
        xc, yc, zc = np.zeros(n), np.zeros(n), np.zeros(n)
        depth = np.zeros(n, np.int32)
        cell0 = np.zeros(n, np.int64)
        blen, tlen = np.zeros(n), np.zeros(n)
        lib().orc_get_leaves(self.h, xc, yc, zc, depth, cell0, blen, tlen)
        return dict(xc=xc, yc=yc, zc=zc, depth=depth, cell0=cell0, blen=blen, tlen=tlen)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def grid(self):
        ints = np.zeros(8, np.int32)
        reals = np.zeros(4, np.float64)
        lib().orc_get_grid(self.h, ints, reals)
        return dict(nx=int(ints[0]), ny=int(ints[1]), nz=int(ints[2]),
                    nb_buffer=int(ints[3]), nb_trunc=int(ints[4]),
                    xmin=float(reals[0]), ymin=float(reals[1]), zmin=float(reals[2]),
                    box=float(reals[3]))

    def centres(self):
        xc = np.zeros(self.ncell)
        yc = np.zeros(self.ncell)
        zc = np.zeros(self.ncell)
        lib().orc_get_centres(self.h, xc, yc, _ptr(zc))
        return xc, yc, zc

    def perm(self):
        pi = np.zeros(self.ni, np.int32)
        pj = np.zeros(self.nj, np.int32)
        lib().orc_get_perm(self.h, pi, _ptr(pj))
        return pi, pj

    def cells(self):
        a = [np.zeros(self.ncell, np.int32) for _ in range(4)]
        lib().orc_get_cells(self.h, a[0], a[1], _ptr(a[2]), _ptr(a[3]))
        return a

    def buffer_csr(self):
        ptr = np.zeros(self.ncell + 1, np.int64)
        idx = np.zeros(lib().orc_buffer_nnz(self.h), np.int32)
        lib().orc_get_buffer(self.h, ptr, idx)
        return ptr, idx

    def trunc_csr(self):
        ptr = np.zeros(self.ncell + 1, np.int64)
        idx = np.zeros(lib().orc_trunc_nnz(self.h), np.int32)
        lib().orc_get_trunc(self.h, ptr, idx)
        return ptr, idx

    def sorted_coords(self):
        x = np.zeros(self.ni)
        y = np.zeros(self.ni)
        z = np.zeros(self.ni) if self.dim == 3 else None
        lib().orc_get_sorted(self.h, x, y, _ptr(z))
        return x, y, z

    def matvec(self, g_sorted):
        g = np.ascontiguousarray(np.atleast_2d(g_sorted), dtype=np.float64)
        w = np.zeros((g.shape[0], self.ni))
        lib().orc_matvec(self.h, g.shape[0], g, w)
        return w if np.ndim(g_sorted) == 2 else w[0]

    def rasm_setup(self):
        st = lib().orc_rasm_setup(self.h)
        if st != 0:
            raise RuntimeError(f"orc_rasm_setup failed ({st})")

    def rasm_setup_sample(self, stride):
        st = lib().orc_rasm_setup_sample(self.h, int(stride))
        if st != 0:
            raise RuntimeError(f"orc_rasm_setup_sample failed ({st})")

    def rasm_apply_sample(self, r):
        z = np.zeros(self.ni)
        lib().orc_rasm_apply_sample(self.h, np.ascontiguousarray(r, dtype=np.float64), z)
        return z

    def rasm_apply(self, r):
        z = np.zeros(self.ni)
        st = lib().orc_rasm_apply(self.h, np.ascontiguousarray(r, dtype=np.float64), z)
        if st != 0:
            raise RuntimeError(f"orc_rasm_apply failed ({st})")
        return z

    def block(self, c):
        ptr, _ = self.buffer_csr()
        n = int(ptr[c + 1] - ptr[c])
        A = np.zeros(n * n)
        lib().orc_get_block(self.h, c, A)
        return A.reshape(n, n)

    def lu(self, c, n):
        LU = np.zeros(n * n)
        piv = np.zeros(n, np.int32)
        st = lib().orc_get_lu(self.h, c, LU, piv)
        if st != 0:
            raise RuntimeError("rasm not set up")
        return LU.reshape(n, n).T.copy(), piv   # return row-major view of the column-major factor

    def solve_sorted(self, w_sorted, rtol=1e-10, restart=30, max_it=1000, flexible=False):
        g = np.zeros(self.ni)
        its = C.c_int(0)
        nh = C.c_int(0)
        hist = np.zeros(max_it + 64)
        reason = lib().orc_solve_sorted(self.h, np.ascontiguousarray(w_sorted, dtype=np.float64), g,
                                        rtol, restart, max_it, int(flexible), C.byref(its),
                                        hist, hist.shape[0], C.byref(nh))
        return dict(gamma=g, its=its.value, reason=reason,
                    hist=hist[: min(nh.value, hist.shape[0])].copy())


def dense_lu(A):
    """LAPACK dgetf2-semantics LU of a dense matrix; returns (LU row-major, piv)."""
    n = A.shape[0]
    a = np.ascontiguousarray(A.T, dtype=np.float64).ravel()   # column-major
    piv = np.zeros(n, np.int32)
    st = lib().orc_dense_lu(n, a, piv)
    if st != 0:
        raise np.linalg.LinAlgError("singular")
    return a.reshape(n, n).T.copy(), piv


def dense_lu_solve(LU, piv, b):
    n = LU.shape[0]
    a = np.ascontiguousarray(LU.T).ravel()
    x = np.ascontiguousarray(b, dtype=np.float64).copy()
    lib().orc_dense_lu_solve(n, a, np.ascontiguousarray(piv, dtype=np.int32), x)
    return x


def dense_gmres(A, b, B=None, rtol=1e-10, restart=30, max_it=1000, flexible=False):
    n = A.shape[0]
    a = np.ascontiguousarray(A.T).ravel()
    bb = None if B is None else np.ascontiguousarray(B.T).ravel()
    x = np.zeros(n)
    its = C.c_int(0)
    nh = C.c_int(0)
    hist = np.zeros(max_it + 64)
    reason = lib().orc_dense_gmres(n, a, _ptr(bb), np.ascontiguousarray(b, dtype=np.float64), x,
                                   rtol, restart, max_it, int(flexible), C.byref(its),
                                   hist, hist.shape[0], C.byref(nh))
    return dict(x=x, its=its.value, reason=reason, hist=hist[: min(nh.value, hist.shape[0])].copy())


def num_threads() -> int:
    return int(lib().orc_num_threads())

"""Host-side mirror of the reference's RBF interface over the HIP library.

`rbf_interpolation` and `vorticity_evaluation` keep the names, argument
meaning and outputs of 2d/rbf_interpolation.cxx:25-26 and
2d/vorticity_evaluation.cxx:26-27 (and the 3D twins): coordinates, sigma and
the three cluster widths in units of sigma in, strengths (and the iteration
count) or evaluated field out.  `Context` exposes the operator surface the
reference hands to PETSc (MatShell MULT, sub-matrices / PCASM) one call at a
time.  Everything runs in libpetrbf.so on the GPU; a missing library or
device raises.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

_KSP = {"gmres": L.PRBF_KSP_GMRES, "fgmres": L.PRBF_KSP_FGMRES}
_PC = {"none": L.PRBF_PC_NONE, "rasm": L.PRBF_PC_RASM_FP64, "asm": L.PRBF_PC_RASM_FP64,
       "rasm32": L.PRBF_PC_RASM_FP32, "learned": L.PRBF_PC_LEARNED}


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    """Pointer of a numpy array or a torch tensor (device pointers pass through)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    return C.c_void_p(a.data_ptr())


@dataclass
class SolveInfo:
    its: int
    reason: int
    rnorm0: float
    rnorm: float
    t_setup_s: float
    t_solve_s: float
    t_matvec_s: float
    t_pc_s: float
    n_matvec: int
    n_pc: int
    n_regularised: int
    hist: np.ndarray
    n_network: int = 0
    n_net_fallback: int = 0

    @property
    def converged(self) -> bool:
        return self.reason > 0


class Context:
    """One cluster structure on the GPU (the reference's BOTH context)."""

    def __init__(self, dim, sigma, nsigma_box, sigma_buffer, sigma_trunc):
        self._L = L.load()
        self.dim = int(dim)
        prm = L.PrbfParams(int(dim), float(sigma), float(nsigma_box), float(sigma_buffer), float(sigma_trunc))
        h = C.c_void_p()
        st = self._L.prbf_create(C.byref(h), C.byref(prm))
        if st != 0:
            raise L.PrbfError(st, "prbf_create failed (no usable GPU or bad parameters)")
        self._h = h
        self.ni = self.nj = 0
        self.ncell = 0

    @classmethod
    def from_problem(cls, p):
        return cls(p.dim, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc)

    def close(self):
        if getattr(self, "_h", None):
            self._L.prbf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            msg = self._L.prbf_last_error(self._h)
            raise L.PrbfError(st, msg.decode() if msg else "")

    # ---- setup (kernel family a)
    def set_points(self, xi, yi, zi=None, xj=None, yj=None, zj=None):
        xi, yi, zi, xj, yj, zj = (_f64(a) for a in (xi, yi, zi, xj, yj, zj))
        ni = xi.shape[0]
        nj = 0 if xj is None else xj.shape[0]
        self._check(self._L.prbf_set_points(self._h, ni, _p(xi), _p(yi), _p(zi), nj, _p(xj), _p(yj), _p(zj)))
        self._after_points()

    def set_points_dev(self, xi, yi, zi=None, xj=None, yj=None, zj=None):
        """Device tensors (float64, contiguous) in the caller's order."""
        ni = xi.shape[0]
        nj = 0 if xj is None else xj.shape[0]
        self._check(self._L.prbf_set_points_dev(self._h, ni, _p(xi), _p(yi), _p(zi), nj, _p(xj), _p(yj), _p(zj)))
        self._after_points()

    def set_points_adaptive(self, x, y, z=None, cap=16, max_depth=6):
        """The adaptive cluster structure over one point set: grid cells of
        more than `cap` particles split recursively down to `max_depth`; the
        leaves are the clusters of every later call."""
        x, y, z = (_f64(a) for a in (x, y, z))
        self._check(self._L.prbf_set_points_adaptive(self._h, x.shape[0], _p(x), _p(y), _p(z), int(cap),
                                                     int(max_depth)))
        self._after_points()

    def set_points_adaptive_dev(self, x, y, z=None, cap=16, max_depth=6):
        self._check(self._L.prbf_set_points_adaptive_dev(self._h, x.shape[0], _p(x), _p(y), _p(z), int(cap),
                                                         int(max_depth)))
        self._after_points()

    def leaves(self):
        """Per leaf: centre, depth, level-0 cell, buffer and truncation half-widths."""
        n = self.ncell
        xc, yc, zc = np.zeros(n), np.zeros(n), np.zeros(n)
        depth = np.zeros(n, np.int32)
        cell0 = np.zeros(n, np.int32)
        blen, tlen = np.zeros(n), np.zeros(n)
        self._check(self._L.prbf_get_leaves(self._h, _p(xc), _p(yc), _p(zc), _p(depth), _p(cell0), _p(blen),
                                            _p(tlen)))
        return dict(xc=xc, yc=yc, zc=zc, depth=depth, cell0=cell0, blen=blen, tlen=tlen)

    def _after_points(self):
        info = self.info()
        self.ncell, self.ni, self.nj = info["ncell"], info["ni"], info["nj"]

    def info(self):
        a = np.zeros(16, np.int64)
        g = np.zeros(8, np.float64)
        self._check(self._L.prbf_get_info(self._h, _p(a), _p(g)))
        keys = ["ncell", "nx", "ny", "nz", "nb_buffer", "nb_trunc", "buffer_nnz", "trunc_nnz",
                "max_buffer", "max_core", "max_trunc", "ni", "nj", "chol_waves", "chol_panel", "chol_order"]
        d = {k: int(a[i]) for i, k in enumerate(keys)}
        d.update(xmin=g[0], ymin=g[1], zmin=g[2], box=g[3], buffer_len=g[4], trunc_len=g[5], chol_kernel_s=g[6],
                 chol_flops=g[7])
        return d

    def permutation(self):
        pi = np.zeros(self.ni, np.int32)
        pj = np.zeros(self.nj, np.int32)
        self._check(self._L.prbf_get_permutation(self._h, _p(pi), _p(pj)))
        return pi, pj

    def cells(self):
        a = [np.zeros(self.ncell, np.int32) for _ in range(4)]
        self._check(self._L.prbf_get_cells(self._h, *(_p(x) for x in a)))
        return a

    def buffer_csr(self):
        ptr = np.zeros(self.ncell + 1, np.int64)
        idx = np.zeros(self.info()["buffer_nnz"], np.int32)
        self._check(self._L.prbf_get_buffer_csr(self._h, _p(ptr), _p(idx)))
        return ptr, idx

    def trunc_csr(self):
        ptr = np.zeros(self.ncell + 1, np.int64)
        idx = np.zeros(self.info()["trunc_nnz"], np.int32)
        self._check(self._L.prbf_get_trunc_csr(self._h, _p(ptr), _p(idx)))
        return ptr, idx

    def trunc_ptr(self):
        ptr = np.zeros(self.ncell + 1, np.int64)
        self._check(self._L.prbf_get_trunc_csr(self._h, _p(ptr), None))
        return ptr

    def buffer_ptr(self):
        ptr = np.zeros(self.ncell + 1, np.int64)
        self._check(self._L.prbf_get_buffer_csr(self._h, _p(ptr), None))
        return ptr

    def set_stream(self, stream_ptr: int):
        """Run this context's work on an existing HIP stream (e.g. a torch stream's handle)."""
        self._check(self._L.prbf_set_stream(self._h, C.c_void_p(stream_ptr)))

    def synchronize(self):
        self._check(self._L.prbf_synchronize(self._h))

    def sorted_points(self):
        x = np.zeros(self.ni)
        y = np.zeros(self.ni)
        z = np.zeros(self.ni) if self.dim == 3 else None
        self._check(self._L.prbf_get_sorted_points(self._h, _p(x), _p(y), _p(z)))
        return x, y, z

    # ---- matvec / evaluation (kernel families b, c)
    def matvec(self, g_sorted):
        g = np.ascontiguousarray(np.atleast_2d(g_sorted), dtype=np.float64)
        w = np.zeros((g.shape[0], self.ni))
        self._check(self._L.prbf_matvec(self._h, g.shape[0], _p(g), _p(w)))
        return w if np.ndim(g_sorted) == 2 else w[0]

    def matvec_dev(self, g_sorted, w_sorted, nrhs=1):
        self._check(self._L.prbf_matvec_dev(self._h, nrhs, _p(g_sorted), _p(w_sorted)))

    def evaluate(self, gj):
        g = np.ascontiguousarray(np.atleast_2d(gj), dtype=np.float64)
        w = np.zeros((g.shape[0], self.ni))
        self._check(self._L.prbf_evaluate(self._h, g.shape[0], _p(g), _p(w)))
        return w if np.ndim(gj) == 2 else w[0]

    def evaluate_dev(self, gj, wi, nrhs=1):
        self._check(self._L.prbf_evaluate_dev(self._h, nrhs, _p(gj), _p(wi)))

    # ---- RASM
    def rasm_setup(self, pc="rasm"):
        self._check(self._L.prbf_rasm_setup_pc(self._h, _PC[pc]))

    def set_tuning(self, chol_waves=None, chol_panel=None, chol_order=None):
        """Force the block-factorisation launch configuration (else autotuned).

        chol_order: "right" or "left" (right-looking unless set)."""
        if chol_waves is not None:
            self._check(self._L.prbf_set_tuning(self._h, L.PRBF_TUNE_CHOL_WAVES, int(chol_waves)))
        if chol_panel is not None:
            self._check(self._L.prbf_set_tuning(self._h, L.PRBF_TUNE_CHOL_PANEL, int(chol_panel)))
        if chol_order is not None:
            v = {"right": L.PRBF_CHOL_RIGHT, "left": L.PRBF_CHOL_LEFT}[chol_order]
            self._check(self._L.prbf_set_tuning(self._h, L.PRBF_TUNE_CHOL_ORDER, v))

    def rasm_apply(self, r_sorted):
        r = _f64(r_sorted)
        z = np.zeros(self.ni)
        self._check(self._L.prbf_rasm_apply(self._h, _p(r), _p(z)))
        return z

    def rasm_apply_dev(self, r, z):
        self._check(self._L.prbf_rasm_apply_dev(self._h, _p(r), _p(z)))

    # ---- learned preconditioner (device arrays)
    def learned_eligible(self):
        n = C.c_int64()
        self._check(self._L.prbf_learned_eligible(self._h, C.byref(n)))
        return n.value

    def learned_pairs_dev(self, r, X, s, z=None, T=None, out_scale=1.0):
        """Features (and with z, targets) of the eligible clusters; returns the row count."""
        n = C.c_int64()
        self._check(self._L.prbf_learned_pairs_dev(self._h, _p(r), _p(z), float(out_scale), _p(X), _p(T), _p(s),
                                                   int(X.shape[0]), C.byref(n)))
        return n.value

    def learned_rho_dev(self, r, z, rho):
        n = C.c_int64()
        self._check(self._L.prbf_learned_rho_dev(self._h, _p(r), _p(z), _p(rho), int(rho.shape[0]), C.byref(n)))
        return n.value

    def learned_attach(self, net, tau):
        """Clusters whose patch residual reduction stays <= tau use the network; returns their count."""
        n = C.c_int64()
        self._check(self._L.prbf_learned_attach(self._h, net._h, float(tau), C.byref(n)))
        return n.value

    def krylov_vectors_dev(self, j, v, z):
        self._check(self._L.prbf_krylov_vectors_dev(self._h, int(j), _p(v), _p(z)))

    def core_rows(self, c):
        n = C.c_int()
        t = C.c_int()
        self._check(self._L.prbf_rasm_get_core_rows(self._h, int(c), None, C.byref(n), C.byref(t)))
        M = np.zeros((t.value, n.value))
        if n.value:
            self._check(self._L.prbf_rasm_get_core_rows(self._h, int(c), _p(M), C.byref(n), C.byref(t)))
        return M  # (T, n): rows of A_c^{-1} for the core members

    def block(self, c):
        """The dense untruncated RASM block of cluster c and its buffer
        indices in subdomain order (the mysubmat block, 2d/matmult.cxx:53-104)."""
        n = C.c_int()
        self._check(self._L.prbf_rasm_get_block(self._h, int(c), None, None, C.byref(n)))
        A = np.zeros((n.value, n.value))
        idx = np.zeros(n.value, np.int32)
        if n.value:
            self._check(self._L.prbf_rasm_get_block(self._h, int(c), _p(A), _p(idx), C.byref(n)))
        return A, idx

    # ---- solve
    def _opts(self, ksp, pc, rtol, restart, max_it, graph=True):
        o = L.PrbfSolveOpts()
        self._L.prbf_solve_opts_default(C.byref(o))
        o.ksp = _KSP[ksp]
        o.pc = _PC[pc]
        o.rtol = float(rtol)
        o.restart = int(restart)
        o.max_it = int(max_it)
        o.graph = 1 if graph else 0
        return o

    def _info(self, inf, hist):
        n = min(hist.shape[0], inf.its + 64)
        return SolveInfo(inf.its, inf.reason, inf.rnorm0, inf.rnorm, inf.t_setup_s, inf.t_solve_s,
                         inf.t_matvec_s, inf.t_pc_s, inf.n_matvec, inf.n_pc, inf.n_regularised, hist[:n],
                         inf.n_network, inf.n_net_fallback)

    def solve(self, w, ksp="gmres", pc="rasm", rtol=1e-10, restart=30, max_it=1000, graph=True):
        """Strengths gamma (caller's order) with A~ gamma = w."""
        o = self._opts(ksp, pc, rtol, restart, max_it, graph)
        w = _f64(w)
        g = np.zeros(self.ni)
        hist = np.zeros(max_it + 64)
        inf = L.PrbfSolveInfo()
        self._check(self._L.prbf_solve(self._h, C.byref(o), _p(w), _p(g), C.byref(inf), _p(hist), hist.shape[0]))
        info = self._info(inf, hist)
        info.hist = info.hist[: self._hist_len(info)]
        return g, info

    def solve_dev(self, w, gamma, ksp="gmres", pc="rasm", rtol=1e-10, restart=30, max_it=1000, graph=True):
        o = self._opts(ksp, pc, rtol, restart, max_it, graph)
        hist = np.zeros(max_it + 64)
        inf = L.PrbfSolveInfo()
        self._check(self._L.prbf_solve_dev(self._h, C.byref(o), _p(w), _p(gamma), C.byref(inf), _p(hist),
                                           hist.shape[0]))
        info = self._info(inf, hist)
        info.hist = info.hist[: self._hist_len(info)]
        return info

    def solve_multi(self, w, ksp="fgmres", pc="rasm32", rtol=1e-10, restart=30, max_it=1000):
        """Strengths for k = 1..3 right-hand sides w (k, n) in one batched
        Krylov solve; returns gamma (k, n) and one SolveInfo per right-hand side."""
        o = self._opts(ksp, pc, rtol, restart, max_it, graph=False)
        w = np.ascontiguousarray(np.atleast_2d(w), dtype=np.float64)
        k = w.shape[0]
        g = np.zeros_like(w)
        cap = max_it + 64
        hist = np.zeros((k, cap))
        inf = (L.PrbfSolveInfo * k)()
        self._check(self._L.prbf_solve_multi(self._h, C.byref(o), k, _p(w), _p(g), inf, _p(hist), cap))
        infos = []
        for r in range(k):
            info = self._info(inf[r], hist[r])
            info.hist = info.hist[: self._hist_len(info)]
            infos.append(info)
        return g, infos

    def solve_multi_dev(self, w, gamma, nrhs, ksp="fgmres", pc="rasm32", rtol=1e-10, restart=30, max_it=1000):
        o = self._opts(ksp, pc, rtol, restart, max_it, graph=False)
        inf = (L.PrbfSolveInfo * nrhs)()
        self._check(self._L.prbf_solve_multi_dev(self._h, C.byref(o), int(nrhs), _p(w), _p(gamma), inf, None, 0))
        return [self._info(inf[r], np.zeros(0)) for r in range(nrhs)]

    @staticmethod
    def _hist_len(info):
        # one entry per cycle start plus one per iteration; trailing zeros are unused slots
        h = info.hist
        nz = np.nonzero(h)[0]
        return int(nz[-1]) + 1 if nz.size else 0


def rbf_interpolation(xi, yi, wi, sigma, nsigma_box, sigma_buffer, sigma_trunc, zi=None,
                      ksp="gmres", pc="rasm", rtol=1e-10, restart=30, max_it=1000):
    """Strengths gi of the Gaussian bases at (xi, yi[, zi]) that reproduce wi
    (2d/rbf_interpolation.cxx:25-26; 3d/rbf_interpolation.cxx).  Returns
    (gi, its) like the reference's (gi out, its out)."""
    dim = 2 if zi is None else 3
    ctx = Context(dim, sigma, nsigma_box, sigma_buffer, sigma_trunc)
    try:
        ctx.set_points(xi, yi, zi)
        g, info = ctx.solve(wi, ksp=ksp, pc=pc, rtol=rtol, restart=restart, max_it=max_it)
    finally:
        ctx.close()
    return g, info.its


def vorticity_evaluation(xi, yi, xj, yj, gj, sigma, nsigma_box, sigma_buffer, sigma_trunc, zi=None, zj=None):
    """Field at targets (xi, yi[, zi]) from Gaussian bases at (xj, yj[, zj]) with
    strengths gj, truncated per cluster (2d/vorticity_evaluation.cxx:26-27)."""
    dim = 2 if zi is None else 3
    ctx = Context(dim, sigma, nsigma_box, sigma_buffer, sigma_trunc)
    try:
        ctx.set_points(xi, yi, zi, xj, yj, zj)
        return ctx.evaluate(gj)
    finally:
        ctx.close()


def RBFInterpolation(xg, yg, zg, wg, xp, yp, zp, gp, sigma, nsigma_box, sigma_buffer, sigma_trunc,
                     ksp="gmres", pc="rasm", rtol=1e-10, restart=30, max_it=1000):
    """The reference's library entry (3d/wrapper.cxx:20-23): FP32 grid values
    wg in, grid strengths out (returned as a new float32 array, with the
    iteration count).  zg/zp None selects 2D."""
    Lb = L.load()
    dim = 2 if zg is None else 3
    f32 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)
    xg, yg, zg, xp, yp, zp, gp = (f32(a) for a in (xg, yg, zg, xp, yp, zp, gp))
    out = np.ascontiguousarray(wg, dtype=np.float32).copy()
    prm = L.PrbfParams(dim, float(sigma), float(nsigma_box), float(sigma_buffer), float(sigma_trunc))
    o = L.PrbfSolveOpts()
    Lb.prbf_solve_opts_default(C.byref(o))
    o.ksp, o.pc, o.rtol, o.restart, o.max_it = _KSP[ksp], _PC[pc], float(rtol), int(restart), int(max_it)
    its = C.c_int(0)
    npart = 0 if xp is None else xp.shape[0]
    st = Lb.prbf_rbf_interpolation_f32(C.byref(prm), xg.shape[0], _p(xg), _p(yg), _p(zg), _p(out), npart,
                                       _p(xp), _p(yp), _p(zp), _p(gp), C.byref(o), C.byref(its))
    if st != 0:
        msg = Lb.prbf_last_error(None)
        raise L.PrbfError(st, msg.decode() if msg else "")
    return out, its.value

"""Reference for what FP32 block solves can achieve: each RASM block factorised
by LAPACK's FP32 Cholesky (numpy/scipy, spotrf/strtrs) and its core rows of
A_c^{-1} formed in FP32, then applied in FP64 -- an independent FP32
implementation of the same preconditioner.  A GPU FP32 preconditioner is
accepted when its distance from the exact FP64 one is of the same order as
this reference's: a broken kernel is off by orders of magnitude more, a
correct one differs only in rounding order."""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def block_fp32(coords, idx, sigma):
    """Dense untruncated Gaussian block over the particles idx (sorted ids) in
    FP32 as an FP32 implementation forms it: coordinate differences in FP64
    rounded to FP32, the squared distance, its scaling and the exponential in
    FP32 (the exponent carries FP32 rounding, not only the entry)."""
    r2 = np.zeros((idx.size, idx.size), np.float32)
    for c in coords:
        if c is None:
            continue
        d = (c[idx][:, None] - c[idx][None, :]).astype(np.float32)
        r2 += d * d
    inv = np.float32(1.0 / (2 * sigma * sigma))
    norm = np.float32(1.0 / (2 * np.pi * sigma * sigma))
    return np.exp(-(r2 * inv)) * norm


def fp32_apply(coords, bptr, bidx, ista, iend, sigma, r, clusters=None):
    """z = sum_c P_core A_c^{-1} R_buffer r with every A_c^{-1} core row
    computed in FP32 (Cholesky + two triangular solves), applied in FP64.
    Returns (z, mask of the entries written)."""
    z = np.zeros_like(r)
    written = np.zeros(r.shape[0], bool)
    cl = range(len(ista)) if clusters is None else clusters
    for c in cl:
        b = bidx[bptr[c]:bptr[c + 1]]
        if b.size == 0:
            continue
        T = int(iend[c] - ista[c] + 1)
        A = block_fp32(coords, b, sigma)
        L = np.linalg.cholesky(A)
        E = np.zeros((b.size, T), np.float32)
        E[np.arange(T), np.arange(T)] = 1.0
        Y = sla.solve_triangular(L, E, lower=True)
        X = sla.solve_triangular(L.T, Y, lower=False)  # A^{-1}[:, core], FP32
        z[b[:T]] = X.T.astype(np.float64) @ r[b]
        written[b[:T]] = True
    return z, written

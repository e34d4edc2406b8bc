"""The learned preconditioner: the network handle, pair harvesting and
training, all on the device through the C ABI (include/petrbf.h).

The network maps the residual on the lattice-slot patch around a cluster
(2D: 8x8 slots around 4x4 lattice spacings; 3D: 5x5x5 slots around 3x3x3),
scaled by its RMS, plus the particles' offsets from their slots to a
correction on the cluster's core particles; it is trained on (patch, exact
RASM core solve) pairs taken from FGMRES runs with the exact FP64 block
preconditioner.  torch provides the device memory only.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .rbf import Context, _p

# (inputs, outputs, core outputs used) per dimension
DIMS = {2: (192, 16, 16), 3: (512, 32, 27)}
N_IN, N_OUT = DIMS[2][:2]


class Mlp:
    """nin -> 256 -> 256 -> 256 -> nout GELU network (BF16 MFMA, FP32 master
    weights): 2D 192 -> 16, 3D 512 -> 32 (27 core corrections used)."""

    def __init__(self, seed=0x5EED, dim=2):
        self._L = L.load()
        h = C.c_void_p()
        st = self._L.prbf_mlp_create_dim(C.byref(h), C.c_uint64(seed), int(dim))
        if st != 0:
            raise L.PrbfError(st, "prbf_mlp_create_dim failed")
        self._h = h
        self.dim = int(dim)
        self.n_in, self.n_out, self.n_core = DIMS[self.dim]

    def close(self):
        if getattr(self, "_h", None):
            self._L.prbf_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, st):
        if st != 0:
            msg = self._L.prbf_mlp_last_error(self._h)
            raise L.PrbfError(st, msg.decode() if msg else "")

    @property
    def num_params(self):
        return int(self._L.prbf_mlp_num_params_dim(self.dim))

    def params(self):
        P = np.zeros(self.num_params, np.float32)
        s = C.c_double()
        self._check(self._L.prbf_mlp_get_params(self._h, _p(P), C.byref(s)))
        return P, s.value

    def set_params(self, P, out_scale):
        P = np.ascontiguousarray(P, dtype=np.float32)
        self._check(self._L.prbf_mlp_set_params(self._h, _p(P), float(out_scale)))

    def forward_dev(self, X, Y):
        self._check(self._L.prbf_mlp_forward_dev(self._h, int(X.shape[0]), _p(X), _p(Y)))

    def loss_grad_dev(self, X, T, grad):
        loss = C.c_float()
        self._check(self._L.prbf_mlp_loss_grad_dev(self._h, int(X.shape[0]), _p(X), _p(T), C.byref(loss), _p(grad)))
        return loss.value

    def train_dev(self, X, T, batch=4096, steps=1000, lr=1e-3, seed=1):
        hist = np.zeros(max(1, steps), np.float32)
        self._check(self._L.prbf_mlp_train_dev(self._h, int(X.shape[0]), _p(X), _p(T), int(batch), int(steps),
                                               float(lr), C.c_uint64(seed), _p(hist)))
        return hist[:steps]


def harvest(ctx: Context, w_dev, n_vectors, torch, restart=30):
    """(X, T, out_scale) of the eligible clusters for the first n_vectors
    basis vectors of an FGMRES solve of A g = w (w in the caller's order, a
    device tensor) with the exact FP64 block preconditioner: X the patch
    features of each basis vector, T the exact core solves scaled by the patch
    RMS and a common output scale.  The network's dimension is the context's
    (ctx.dim)."""
    dev = w_dev.device
    n = ctx.ni
    n_in, n_out, _ = DIMS[ctx.dim]
    ne = ctx.learned_eligible()
    gamma = torch.empty_like(w_dev)
    torch.cuda.synchronize()
    info = ctx.solve_dev(w_dev, gamma, ksp="fgmres", pc="rasm", rtol=1e-10, restart=restart, max_it=restart)
    nv = min(n_vectors, info.its)
    X = torch.empty((nv * ne, n_in), dtype=torch.float32, device=dev)
    T = torch.empty((nv * ne, n_out), dtype=torch.float32, device=dev)
    s = torch.empty(ne, dtype=torch.float64, device=dev)
    v = torch.empty(n, dtype=torch.float64, device=dev)
    z = torch.empty(n, dtype=torch.float64, device=dev)
    for j in range(nv):
        ctx.krylov_vectors_dev(j, v, z)
        ctx.learned_pairs_dev(v, X[j * ne:(j + 1) * ne], s, z=z, T=T[j * ne:(j + 1) * ne], out_scale=1.0)
    ctx.synchronize()
    out_scale = float(torch.sqrt(torch.mean(T.double() ** 2)))
    T /= out_scale
    torch.cuda.synchronize()
    return X, T, out_scale

"""CPU oracle of the learned preconditioner's network (test infrastructure only).

The network has no counterpart in the reference: it is the nin -> 256 -> 256
-> 256 -> nout GELU (tanh form) MLP of the library's learned preconditioner,
2D 192 -> 16, 3D 512 -> 32 (the parameter vector's length tells them apart).
This module restates its forward pass, mean-square loss and gradient in
float64 numpy.  With ``bf16=True`` it rounds to bfloat16 (nearest even) at
exactly the points where the device kernels store BF16 values: the input,
the weights, the hidden activations, and the output and hidden gradients fed
to the GEMMs; everything else (sums, biases, pre-activations, bias
gradients) stays wide.  Parameter packing: W0, b0, W1, b1, W2, b2, W3, b3,
W_l row-major (inputs x outputs).
"""

from __future__ import annotations

import numpy as np

LAYERS = {2: [(192, 256), (256, 256), (256, 256), (256, 16)],
          3: [(512, 256), (256, 256), (256, 256), (256, 32)]}
DIMS = LAYERS[2]
K0, K1 = 0.7978845608028654, 0.044715


def num_params(dim=2):
    return sum(k * n + n for k, n in LAYERS[dim])


def unpack(P):
    dims = LAYERS[2] if len(P) == num_params(2) else LAYERS[3]
    assert len(P) == sum(k * n + n for k, n in dims)
    out, o = [], 0
    for k, n in dims:
        W = P[o:o + k * n].reshape(k, n)
        o += k * n
        b = P[o:o + n]
        o += n
        out.append((W, b))
    return out


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(K0 * (x + K1 * x ** 3)))


def gelu_d(x):
    t = np.tanh(K0 * (x + K1 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K0 * (1.0 + 3.0 * K1 * x * x)


def forward(P, X, bf=False):
    """Y and the layer caches (inputs to each layer, pre-activations)."""
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(np.asarray(P, dtype=np.float64))
    h = q(X)
    ins, pres = [], []
    for l, (W, b) in enumerate(layers):
        ins.append(h)
        z = h @ q(W) + b
        if l < len(layers) - 1:
            pres.append(z)
            h = q(gelu(z))
        else:
            return z, ins, pres


def loss_grad(P, X, T, bf=False):
    """Mean-square loss over all outputs and its gradient (packed like P)."""
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(np.asarray(P, dtype=np.float64))
    Y, ins, pres = forward(P, X, bf)
    D = Y - T
    loss = float(np.mean(D * D))
    g = 2.0 * D / D.size              # wide copy (bias gradient)
    gq = q(g)                          # GEMM operand
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (ins[l].T @ gq, g.sum(axis=0))
        if l == 0:
            break
        g = (gq @ q(W).T) * gelu_d(pres[l - 1])
        gq = q(g)
    G = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, G

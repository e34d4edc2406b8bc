"""The allocation-free evaluation entry point (prbf_evaluate_dev_noalloc)
against prbf_evaluate_dev, through the C ABI: the same field bit for bit on
two calls in a row, the second on a smaller input; no device memory taken
by the calls; sizes beyond the static scratch rejected."""

import ctypes as C
import sys

import numpy as np
import pytest

from petrbf_amd import _lib as L
from petrbf_amd import problems
from petrbf_amd.rbf import Context

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def test_noalloc_evaluation_matches_evaluate_dev():
    import torch
    dev = torch.device("cuda:0")
    lib = L.load()
    # first call: distinct targets (h/2 grid) and sources, three right-hand sides
    p = problems.lattice_2d_problem(120, seed=0x5EED0500)
    xt, yt, _ = problems.jittered_lattice_2d(200, seed=3, jitter=0.0)
    ctx = Context.from_problem(p)
    ctx.set_points(xt, yt, None, p.x, p.y, None)
    calls = [(3, torch.randn(3, p.n, dtype=torch.float64, device=dev), xt.shape[0])]
    # second call: a smaller point set, one right-hand side
    q = problems.lattice_2d_problem(50, seed=0x5EED0501)
    calls.append((1, torch.randn(1, q.n, dtype=torch.float64, device=dev), q.n))
    for step, (nrhs, g, ni) in enumerate(calls):
        if step == 1:
            ctx.set_points(q.x, q.y)
        _announce(f"no-alloc evaluation call {step + 1}: nrhs {nrhs}, ni {ni}, nj {g.shape[1]}")
        w_ref = torch.empty(nrhs, ni, dtype=torch.float64, device=dev)
        w_na = torch.full((nrhs, ni), float("nan"), dtype=torch.float64, device=dev)
        assert lib.prbf_evaluate_dev(ctx._h, nrhs, _ptr(g), _ptr(w_ref)) == L.PRBF_OK
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        assert lib.prbf_evaluate_dev_noalloc(ctx._h, nrhs, _ptr(g), _ptr(w_na)) == L.PRBF_OK
        assert lib.prbf_synchronize(ctx._h) == L.PRBF_OK
        free1, _ = torch.cuda.mem_get_info()
        assert free1 >= free0, (free0, free1)  # no device memory taken by the call
        assert torch.equal(w_na, w_ref)
    # beyond the scratch: 2^21 doubles hold nrhs (nj + ni) <= 2097152
    big = problems.lattice_2d_problem(1100, seed=0x5EED0502)  # 1.21e6 points: 2 * 1.21e6 > 2^21
    ctx.set_points(big.x, big.y)
    g = torch.zeros(big.n, dtype=torch.float64, device=dev)
    w = torch.empty_like(g)
    assert lib.prbf_evaluate_dev_noalloc(ctx._h, 1, _ptr(g), _ptr(w)) == L.PRBF_ERR_LIMIT
    assert lib.prbf_evaluate_dev(ctx._h, 1, _ptr(g), _ptr(w)) == L.PRBF_OK
    ctx.close()

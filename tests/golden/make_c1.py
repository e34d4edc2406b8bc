"""Writes tests/golden/c1_oracle.npz: the oracle's results on the N = 10^4
CPU workload (100x100 jittered lattice, seed 0x5EED0001, sigma = h/0.8,
4x4-spacing cells, two-cell buffer, 5-sigma truncation, Lamb-Oseen RHS,
GMRES(30) with left RASM to rtol 1e-10), so that later changes of the oracle
or of the library are caught against fixed numbers:

* SHA-256 of the sorted permutation, the cell ranges and the buffer and
  truncation CSR lists (int32/int64 little-endian bytes);
* y = A~ x for x ~ N(0, 1) from numpy's default_rng(101) (x itself is not
  stored), sorted order;
* gamma (caller's order), the iteration count, the residual history.

Run from the repository root after building the oracle
(python petrbf_amd/build.py builds it too); the .npz is committed."""

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle  # noqa: E402
from petrbf_amd import problems  # noqa: E402


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def c1_digests(perm, cells, buffer_csr, trunc_csr):
    return dict(perm=sha(perm), cells=sha(*cells), buffer=sha(*buffer_csr), trunc=sha(*trunc_csr))


def main():
    p = problems.lattice_2d_problem(100, seed=0x5EED0001)
    o = Oracle.from_problem(p)
    pi, _ = o.perm()
    d = c1_digests(pi.astype(np.int32), [c.astype(np.int32) for c in o.cells()],
                   [o.buffer_csr()[0].astype(np.int64), o.buffer_csr()[1].astype(np.int32)],
                   [o.trunc_csr()[0].astype(np.int64), o.trunc_csr()[1].astype(np.int32)])
    x = np.random.default_rng(101).standard_normal(p.n)
    y = o.matvec(x)
    o.rasm_setup()
    r = o.solve_sorted(p.w[0][pi], rtol=1e-10, restart=30, max_it=1000)
    assert r["reason"] == 2
    gamma = np.empty(p.n)
    gamma[pi] = r["gamma"]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_oracle.npz")
    np.savez_compressed(out, y=y, gamma=gamma, its=np.int64(r["its"]), hist=r["hist"],
                        **{f"sha_{k}": np.array(v) for k, v in d.items()})
    print(out, {k: v[:12] for k, v in d.items()}, "its", r["its"])


if __name__ == "__main__":
    main()

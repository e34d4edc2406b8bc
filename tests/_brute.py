"""Independent brute-force statements of the reference's index and operator
semantics, for tiny inputs (O(N * cells) loops in numpy).

They re-derive each result from its definition instead of the reference's
loop structure: the sort is a lexsort on (cell, original index); a cluster
list is a global scan of all particles filtered by the cell-reach window and
the per-axis box test, ordered by (cell x, cell y, cell z, sorted position)
after the cluster's own particles; the matvec is a dense masked sum in long
double.
"""

from __future__ import annotations

import math

import numpy as np

EPS = 1e-6


def grid(dim, sigma, nsigma_box, sigma_buffer, sigma_trunc, pts):
    lo = [min(p.min() for p in pd) for pd in pts]
    hi = [max(p.max() for p in pd) for pd in pts]
    box = nsigma_box * sigma + EPS
    cmin = [l - EPS for l in lo]
    nc = [int(math.ceil(((h + EPS) - m) / box)) for h, m in zip(hi, cmin)]
    while len(nc) < 3:
        nc.append(1)
        cmin.append(0.0)
    reach_b = int(math.ceil((sigma_buffer - nsigma_box + EPS) / 2 / nsigma_box))
    reach_t = int(math.ceil((sigma_trunc - nsigma_box + EPS) / 2 / nsigma_box))
    return dict(box=box, cmin=cmin, nc=nc, reach_b=reach_b, reach_t=reach_t,
                blen=sigma_buffer * sigma / 2 + EPS, tlen=sigma_trunc * sigma / 2 + EPS)


def cell_ijk(g, coords):
    k = [np.floor((c - g["cmin"][d]) / g["box"]).astype(np.int64) for d, c in enumerate(coords)]
    while len(k) < 3:
        k.append(np.zeros_like(k[0]))
    return k


def cell_id(g, ix, iy, iz, dim):
    nx, ny = g["nc"][0], g["nc"][1]
    return ix * ny + iy if dim == 2 else iz * nx * ny + ix * ny + iy


def sort(g, coords, dim):
    ix, iy, iz = cell_ijk(g, coords)
    cid = cell_id(g, ix, iy, iz, dim)
    perm = np.lexsort((np.arange(cid.shape[0]), cid))
    ncell = g["nc"][0] * g["nc"][1] * g["nc"][2]
    counts = np.bincount(cid, minlength=ncell)
    start = np.concatenate([[0], np.cumsum(counts)])
    return perm.astype(np.int32), start[:-1].astype(np.int32), (start[1:] - 1).astype(np.int32), cid


def centre(g, c, dim):
    nx, ny = g["nc"][0], g["nc"][1]
    if dim == 2:
        ix, iy, iz = c // ny, c % ny, 0
    else:
        iz, r = divmod(c, nx * ny)
        ix, iy = divmod(r, ny)
    ctr = [g["cmin"][0] + (ix + 0.5) * g["box"], g["cmin"][1] + (iy + 0.5) * g["box"],
           g["cmin"][2] + (iz + 0.5) * g["box"]]
    return (ix, iy, iz), ctr


def cluster_list(g, c, dim, tgt_count, src_sorted, src_cid_sorted, reach, length):
    """List of sorted source ids for cluster c (empty if c has no target)."""
    if tgt_count[c] == 0:
        return np.zeros(0, np.int32)
    (ix, iy, iz), ctr = centre(g, c, dim)
    nx, ny = g["nc"][0], g["nc"][1]
    cid = src_cid_sorted
    if dim == 2:
        jx, jy, jz = cid // ny, cid % ny, np.zeros_like(cid)
    else:
        jz, r = np.divmod(cid, nx * ny)
        jx, jy = np.divmod(r, ny)
    own = cid == c
    near = (np.abs(jx - ix) <= reach) & (np.abs(jy - iy) <= reach) & (np.abs(jz - iz) <= reach) & ~own
    inbox = np.ones(cid.shape[0], bool)
    for d in range(dim):
        inbox &= np.abs(src_sorted[d] - ctr[d]) < length
    sel = np.nonzero(near & inbox)[0]
    order = np.lexsort((sel, jz[sel], jy[sel], jx[sel]))
    return np.concatenate([np.nonzero(own)[0], sel[order]]).astype(np.int32)


def lists(dim, sigma, nsigma_box, sigma_buffer, sigma_trunc, tgt, src=None):
    """Everything the reference's get_cluster / get_buffer / get_trunc build."""
    same = src is None
    src = tgt if same else src
    g = grid(dim, sigma, nsigma_box, sigma_buffer, sigma_trunc, [[t, s] for t, s in zip(tgt, src)])
    pi, ista, iend, cid_i = sort(g, tgt, dim)
    pj, jsta, jend, cid_j = sort(g, src, dim)
    ts = [t[pi] for t in tgt]
    ss = [s[pj] for s in src]
    ncell = g["nc"][0] * g["nc"][1] * g["nc"][2]
    tcount = iend - ista + 1
    bl, tl = [], []
    for c in range(ncell):
        bl.append(cluster_list(g, c, dim, tcount, ts, cid_i[pi], g["reach_b"], g["blen"]))
        tl.append(cluster_list(g, c, dim, tcount, ss, cid_j[pj], g["reach_t"], g["tlen"]))
    return dict(grid=g, perm_i=pi, perm_j=pj, ista=ista, iend=iend, jsta=jsta, jend=jend,
                buffer=bl, trunc=tl, tgt_sorted=ts, src_sorted=ss)


def csr_to_lists(ptr, idx):
    return [idx[ptr[c]:ptr[c + 1]] for c in range(ptr.shape[0] - 1)]


def dense_truncated(dim, sigma, L):
    """Dense A~ (targets x sources, sorted order) in long double."""
    ni = L["tgt_sorted"][0].shape[0]
    nj = L["src_sorted"][0].shape[0]
    A = np.zeros((ni, nj), np.longdouble)
    ts = [np.asarray(t, np.longdouble) for t in L["tgt_sorted"]]
    ss = [np.asarray(s, np.longdouble) for s in L["src_sorted"]]
    s2 = np.longdouble(sigma) ** 2
    for c in range(len(L["trunc"])):
        tl = L["trunc"][c]
        if tl.size == 0:
            continue
        rows = np.arange(L["ista"][c], L["iend"][c] + 1)
        r2 = sum((ts[d][rows, None] - ss[d][None, tl]) ** 2 for d in range(dim))
        A[np.ix_(rows, tl)] = np.exp(-r2 / (2 * s2)) / (2 * np.pi * s2)
    return A


def tree_lists(dim, sigma, nsigma_box, sigma_buffer, sigma_trunc, pts, cap, dmax):
    """The adaptive structure by its definition, for tiny inputs: nodes are
    (grid cell, path of child digits); a node with more than `cap` particles
    below depth `dmax` is split (child digit 2 qx + qy, 3D 4 qz + 2 qx + qy,
    q = coordinate >= node centre); leaves (empty ones included) in the
    lexicographic order of (cell, path); particles by (leaf, original index);
    a leaf's list is its own particles, then every particle of the level-0
    cells within reach of its cell except its own, ordered by (cell x, y, z,
    sorted position), that passes the per-axis box test around the leaf
    centre with the depth's half-width."""
    g = grid(dim, sigma, nsigma_box, sigma_buffer, sigma_trunc, [[p] for p in pts])
    n = pts[0].shape[0]
    ix, iy, iz = cell_ijk(g, pts)
    cell = cell_id(g, ix, iy, iz, dim)
    ncell = g["nc"][0] * g["nc"][1] * g["nc"][2]
    path = [() for _ in range(n)]
    centres = {}
    for c in range(ncell):
        _, ctr = centre(g, c, dim)
        centres[(c, ())] = ctr
    leaves = []
    frontier = [(c, ()) for c in range(ncell)]
    depth = 0
    while frontier:
        members = {}
        for p in range(n):
            members.setdefault((int(cell[p]), path[p]), []).append(p)
        nxt = []
        for node in frontier:
            mem = members.get(node, [])
            if len(mem) <= cap or depth >= dmax:
                leaves.append(node)
                continue
            ctr = centres[node]
            q = math.ldexp(g["box"], -(depth + 2))
            for k in range(1 << dim):
                qx, qy, qz = (k >> 1) & 1, k & 1, (k >> 2) & 1
                cc = [ctr[0] + (q if qx else -q), ctr[1] + (q if qy else -q),
                      (ctr[2] + (q if qz else -q)) if dim == 3 else 0.0]
                child = (node[0], node[1] + (k,))
                centres[child] = cc
                nxt.append(child)
            for p in mem:
                qx = int(pts[0][p] >= ctr[0])
                qy = int(pts[1][p] >= ctr[1])
                qz = int(dim == 3 and pts[2][p] >= ctr[2])
                path[p] = path[p] + ((4 * qz + 2 * qx + qy) if dim == 3 else (2 * qx + qy),)
        frontier = nxt
        depth += 1
    leaves.sort()
    rank = {node: i for i, node in enumerate(leaves)}
    leaf_of = np.array([rank[(int(cell[p]), path[p])] for p in range(n)])
    perm = np.lexsort((np.arange(n), leaf_of)).astype(np.int32)
    nl = len(leaves)
    counts = np.bincount(leaf_of, minlength=nl)
    start = np.concatenate([[0], np.cumsum(counts)])
    ista, iend = start[:-1].astype(np.int32), (start[1:] - 1).astype(np.int32)
    sp = [p[perm] for p in pts]
    scell = cell[perm]
    nx, ny = g["nc"][0], g["nc"][1]
    if dim == 2:
        sx, sy, sz = scell // ny, scell % ny, np.zeros_like(scell)
    else:
        sz, r = np.divmod(scell, nx * ny)
        sx, sy = np.divmod(r, ny)
    tlen0 = g["tlen"]
    out_b, out_t, geo = [], [], []
    for l, (c, pth) in enumerate(leaves):
        d = len(pth)
        blen = math.ldexp(g["blen"], -d)
        tlen = tlen0 if d == 0 else (tlen0 - g["box"] * 0.5) + math.ldexp(g["box"], -(d + 1))
        ctr = centres[(c, pth)]
        geo.append((ctr, d, c, blen, tlen))
        (cx, cy, cz), _ = centre(g, c, dim)
        own = np.arange(ista[l], iend[l] + 1)
        lists_ = []
        for reach, length in ((g["reach_b"], blen), (g["reach_t"], tlen)):
            if own.size == 0:
                lists_.append(np.zeros(0, np.int32))
                continue
            near = (np.abs(sx - cx) <= reach) & (np.abs(sy - cy) <= reach) & (np.abs(sz - cz) <= reach)
            near[own] = False
            inbox = np.ones(n, bool)
            for dd in range(dim):
                inbox &= np.abs(sp[dd] - ctr[dd]) < length
            sel = np.nonzero(near & inbox)[0]
            order = np.lexsort((sel, sz[sel], sy[sel], sx[sel]))
            lists_.append(np.concatenate([own, sel[order]]).astype(np.int32))
        out_b.append(lists_[0])
        out_t.append(lists_[1])
    return dict(perm=perm, ista=ista, iend=iend, buffer=out_b, trunc=out_t, geo=geo, grid=g)

"""The learned preconditioner: its network's forward and backward kernels
against the CPU oracle (oracle/mlp_oracle.py), which is itself pinned here
against torch autograd (float64, CPU), central differences and bfloat16
rounding cases; then training on the device and the learned FGMRES solve."""

import sys

import numpy as np
import pytest

from oracle import mlp_oracle as MO


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _params(seed, dim=2):
    rng = np.random.default_rng(seed)
    P = []
    for k, n in MO.LAYERS[dim]:
        a = np.sqrt(6.0 / (k + n))
        P.append(rng.uniform(-a, a, k * n))
        P.append(rng.uniform(-0.1, 0.1, n))
    return np.concatenate(P).astype(np.float32).astype(np.float64)


def _batch(seed, m, dim=2):
    rng = np.random.default_rng(seed)
    nin, nout = MO.LAYERS[dim][0][0], MO.LAYERS[dim][-1][1]
    X = rng.standard_normal((m, nin)).astype(np.float32).astype(np.float64)
    T = rng.standard_normal((m, nout)).astype(np.float32).astype(np.float64)
    return X, T


# ---------------------------------------------------------------- CPU: the oracle
def test_bf16_rounding_cases():
    e = 2.0 ** -7  # bf16 spacing at 1 (8 significant bits)
    vals = np.array([1.0, 1.0 + e / 2, 1.0 + 1.5 * e, 1.0 + e / 2 + 2 ** -20, -3.0 - e, 2.0 ** -130, 65504.0])
    want = np.array([1.0, 1.0, 1.0 + 2 * e, 1.0 + e, -3.0, 2.0 ** -130, 65536.0])
    # ties go to the even significand (1 + e/2 -> 1, 1 + 1.5e -> 1 + 2e, -3 - e -> -3);
    # just above a tie rounds up; 2^-130 is exact in bf16's subnormal range
    assert np.array_equal(MO.bf16(vals), want)


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_gradient_matches_torch_autograd(dim):
    torch = pytest.importorskip("torch")
    P = _params(1, dim)
    X, T = _batch(2, 32, dim)
    loss, G = MO.loss_grad(P, X, T)
    Pt = torch.tensor(P, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(X)
    o = 0
    for l, (k, n) in enumerate(MO.LAYERS[dim]):
        W = Pt[o:o + k * n].view(k, n)
        o += k * n
        b = Pt[o:o + n]
        o += n
        h = h @ W + b
        if l < 3:
            h = torch.nn.functional.gelu(h, approximate="tanh")
    lt = torch.mean((h - torch.tensor(T)) ** 2)
    lt.backward()
    assert abs(loss - lt.item()) <= 1e-13 * abs(loss)
    g = Pt.grad.numpy()
    assert np.linalg.norm(G - g) <= 1e-11 * np.linalg.norm(g)


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_gradient_central_differences(dim):
    P = _params(3, dim)
    X, T = _batch(4, 16, dim)
    _, G = MO.loss_grad(P, X, T)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.integers(0, MO.num_params(dim), 24), [MO.num_params(dim) - 1]])
    for i in idx:
        h = 1e-5
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (MO.loss_grad(Pp, X, T)[0] - MO.loss_grad(Pm, X, T)[0]) / (2 * h)
        assert abs(fd - G[i]) <= 1e-6 * max(1e-3, abs(G[i])), (i, fd, G[i])


def test_oracle_bf16_mode_stays_close_to_wide_mode():
    P = _params(6)
    X, T = _batch(7, 64)
    l0, g0 = MO.loss_grad(P, X, T)
    l1, g1 = MO.loss_grad(P, X, T, bf=True)
    assert abs(l1 - l0) <= 0.05 * l0
    assert np.linalg.norm(g1 - g0) <= 0.05 * np.linalg.norm(g0)


# ---------------------------------------------------------------- GPU: kernels vs oracle
@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_mlp_forward_and_gradient_match_oracle(dim):
    torch = pytest.importorskip("torch")
    from petrbf_amd.learned import Mlp
    _announce(f"mlp forward/backward vs oracle, {dim}D network")
    dev = torch.device("cuda:0")
    net = Mlp(seed=11, dim=dim)
    assert net.num_params == MO.num_params(dim)
    P = _params(8, dim)
    net.set_params(P.astype(np.float32), 1.0)
    Pg, s = net.params()
    assert np.array_equal(Pg, P.astype(np.float32)) and s == 1.0
    for m in (64, 208):
        X, T = _batch(9 + m, m, dim)
        Xd = torch.tensor(X, dtype=torch.float32, device=dev)
        Td = torch.tensor(T, dtype=torch.float32, device=dev)
        Yd = torch.empty((m, net.n_out), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        net.forward_dev(Xd, Yd)
        Y = Yd.cpu().numpy().astype(np.float64)
        # the oracle rounds to BF16 where the kernels store BF16; the kernels
        # accumulate in FP32, the oracle in FP64, so an activation next to a
        # BF16 rounding boundary may round the other way: the bound is 5% of
        # what BF16 rounding itself changes (the wide oracle's distance)
        Yo, Yw = MO.forward(P, X, bf=True)[0], MO.forward(P, X)[0]
        ey, by = np.linalg.norm(Y - Yo), max(1e-4 * np.linalg.norm(Yo), 0.05 * np.linalg.norm(Yo - Yw))
        assert ey <= by, (m, ey, by)
        Gd = torch.empty(MO.num_params(dim), dtype=torch.float32, device=dev)
        loss = net.loss_grad_dev(Xd, Td, Gd)
        (lo, Go), (lw, Gw) = MO.loss_grad(P, X, T, bf=True), MO.loss_grad(P, X, T)
        assert abs(loss - lo) <= max(1e-4 * lo, 0.05 * abs(lo - lw)), (m, loss, lo, lw)
        G = Gd.cpu().numpy().astype(np.float64)
        o = 0
        for k, n in MO.LAYERS[dim]:
            for size in (k * n, n):
                a, b, c = G[o:o + size], Go[o:o + size], Gw[o:o + size]
                bound = max(2e-3 * np.linalg.norm(b), 0.05 * np.linalg.norm(b - c)) + 1e-12
                assert np.linalg.norm(a - b) <= bound, (m, o, size, np.linalg.norm(a - b), bound)
                o += size
    net.close()


@pytest.mark.gpu
def test_learned_preconditioner_trains_and_solve_reaches_fp64_residual():
    """Harvest (patch, exact block solve) pairs from an exact-RASM FGMRES run
    on one seeded lattice, train the network on the device, and solve another
    seeded lattice with the learned preconditioner: FGMRES reaches rtol 1e-10
    with the FP64 true residual of the exact-block solve, its per-cluster
    fallback moving clusters whose core residual stalls to exact blocks.  The
    gain of the trained network is measured against the same method with an
    untrained network (held-out loss, iterations) and against exact FP32
    blocks everywhere."""
    torch = pytest.importorskip("torch")
    from petrbf_amd import problems
    from petrbf_amd.learned import Mlp, harvest
    from petrbf_amd.rbf import Context
    _announce("learned preconditioner: harvest, train, solve")
    dev = torch.device("cuda:0")
    pt = problems.lattice_2d_problem(200, seed=0x5EED0002)
    ct = Context.from_problem(pt)
    ct.set_points(pt.x, pt.y)
    X, T, osc = harvest(ct, torch.from_numpy(pt.w[0]).to(dev), 24, torch)
    assert X.shape[0] > 10000
    p = problems.lattice_2d_problem(200, seed=0x5EED0001)
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y)
    Xh, Th, osh = harvest(ctx, torch.from_numpy(p.w[0]).to(dev), 24, torch)
    Th *= osh / osc  # held-out targets on the training scale
    net0 = Mlp(seed=7)
    net = Mlp(seed=7)
    net.set_params(net.params()[0], osc)
    net0.set_params(net0.params()[0], osc)
    # the training kernels (forward, backward, Adam) can fit pairs: 2048 of
    # them are memorised (the loss falls well below a quarter in 400 epochs)
    fit = Mlp(seed=5)
    fit.set_params(fit.params()[0], osc)
    hf = fit.train_dev(X[:2048].contiguous(), T[:2048].contiguous(), batch=1024, steps=800, lr=1e-3, seed=2)
    _announce(f"memorisation of 2048 pairs: loss {hf[:20].mean():.3f} -> {hf[-50:].mean():.4f}")
    assert hf[-50:].mean() < 0.25 * hf[:20].mean()
    fit.close()
    # the whole harvest: how far the loss falls depends on the harvested
    # Krylov vectors, which round-off in the exact-block solve reshuffles
    # (between 2x and 3.6x over the runs made so far)
    hist = net.train_dev(X, T, batch=1024, steps=1500, lr=1e-3, seed=3)
    _announce(f"training on {X.shape[0]} pairs: loss {hist[:20].mean():.3f} -> {hist[-50:].mean():.3f}")
    assert hist[-50:].mean() < 0.8 * hist[:20].mean()
    held = []
    for nn in (net0, net):
        Y = torch.empty_like(Th)
        nn.forward_dev(Xh, Y)
        held.append(float(torch.mean((Y - Th) ** 2)))
    w = torch.from_numpy(p.w[0]).to(dev)
    gam = torch.empty_like(w)
    wv = torch.empty_like(w)
    out = {}
    for tag, nn in (("trained", net), ("untrained", net0), ("exact", None)):
        if nn is not None:
            n_net = ctx.learned_attach(nn, 1e30)
            assert n_net == ctx.learned_eligible() > 0
        torch.cuda.synchronize()
        inf = ctx.solve_dev(w, gam, ksp="fgmres", pc="learned" if nn is not None else "rasm32", rtol=1e-10)
        ctx.evaluate_dev(gam, wv)
        ctx.synchronize()
        true_rel = float(torch.linalg.norm(wv - w) / torch.linalg.norm(w))
        assert inf.reason == 2 and true_rel <= 1.05e-10, (tag, inf.reason, true_rel)
        out[tag] = (inf.its, inf.n_network, inf.n_net_fallback, true_rel)
    _announce(f"held-out loss untrained {held[0]:.3f} trained {held[1]:.3f}; "
              f"its/network/moved/true residual: {out}")
    # the gain of the trained network over exact blocks, in iterations (negative:
    # a cost); the per-cluster fallback bounds that cost to the cycles it takes
    # FGMRES's residual to expose the network's clusters
    gain = out["exact"][0] - out["trained"][0]
    _announce(f"iterations gained by the trained network: {gain} (vs untrained {out['untrained'][0] - out['trained'][0]})")
    assert out["trained"][0] <= out["exact"][0] + 2 * 30
    net.close()
    net0.close()
    ctx.close()
    ct.close()


# ---------------------------------------------------------------- 3D network
def _lattice_slots_3d(ctx, n):
    """For the 3D jittered lattice (particle i at lattice indices
    ((i % n^2) // n, i % n, i // n^2), problems.jittered_lattice_3d): the
    clusters whose cell holds a 3x3x3 block of lattice sites with a whole
    ring of sites one spacing around it, in increasing cell order, each with
    the sorted index of the site in patch slot (a, b, c), a, b, c in 0..4."""
    pi, _ = ctx.permutation()
    inv = np.empty_like(pi)
    inv[pi] = np.arange(pi.size, dtype=pi.dtype)
    ista, iend, _, _ = ctx.cells()
    out = []
    for c in range(ista.size):
        if iend[c] - ista[c] + 1 != 27:
            continue
        o = pi[ista[c]:iend[c] + 1]
        ix, iy, iz = (o % (n * n)) // n, o % n, o // (n * n)
        i0, j0, k0 = ix.min(), iy.min(), iz.min()
        if ix.max() - i0 != 2 or iy.max() - j0 != 2 or iz.max() - k0 != 2:
            continue
        if min(i0, j0, k0) < 1 or max(i0, j0, k0) + 3 >= n:
            continue
        slots = np.empty((5, 5, 5), np.int64)
        for a in range(5):
            for b in range(5):
                for cc in range(5):
                    slots[a, b, cc] = inv[(k0 - 1 + cc) * n * n + (i0 - 1 + a) * n + (j0 - 1 + b)]
        out.append((c, slots))
    return out


@pytest.mark.gpu
def test_learned_3d_patches_match_lattice_slots():
    """3D patches: the eligible clusters are exactly those whose cell is a
    3x3x3 block of lattice sites with the whole one-spacing ring around it,
    and each feature row holds the residuals of the 5x5x5 sites in lattice
    order scaled by their RMS, the sites' offsets from their slot centres in
    units of sigma and zeros up to 512; the targets are the core values over
    the RMS, zeros past the 27 core outputs (slots fixed by the lattice
    indices, independently of the kernel's floor arithmetic)."""
    torch = pytest.importorskip("torch")
    from petrbf_amd import problems
    from petrbf_amd.rbf import Context
    _announce("3D learned patches vs lattice slots")
    dev = torch.device("cuda:0")
    n = 18
    p = problems.lattice_3d_problem(n, overlap=1.0)
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y, p.z)
    want = _lattice_slots_3d(ctx, n)
    ne = ctx.learned_eligible()
    # 64 interior cells; the jitter moves a few sites across cell faces (the
    # grid starts at the smallest coordinate, the first plane's jitter), and
    # those cells drop out on both sides (55 here)
    assert ne == len(want) and ne >= 48, (ne, len(want))
    rng = np.random.default_rng(31)
    r = rng.standard_normal(p.n)
    z = rng.standard_normal(p.n)
    X = torch.empty((ne, 512), dtype=torch.float32, device=dev)
    T = torch.empty((ne, 32), dtype=torch.float32, device=dev)
    s = torch.empty(ne, dtype=torch.float64, device=dev)
    got = ctx.learned_pairs_dev(torch.from_numpy(r).to(dev), X, s, z=torch.from_numpy(z).to(dev), T=T, out_scale=1.0)
    ctx.synchronize()
    assert got == ne
    X, T, s = X.cpu().numpy().astype(np.float64), T.cpu().numpy().astype(np.float64), s.cpu().numpy()
    xs, ys, zs = ctx.sorted_points()
    hl = (p.nsigma_box * p.sigma + 1e-6) / 3
    for k, (c, slots) in enumerate(want):
        sl = slots.ravel()  # slot (a, b, c) at (a 5 + b) 5 + c
        rms = np.sqrt(np.mean(r[sl] ** 2))
        assert abs(s[k] - rms) <= 1e-12 * rms
        np.testing.assert_allclose(X[k, :125], r[sl] / rms, rtol=1e-6, atol=1e-6)
        core = slots[1:4, 1:4, 1:4].ravel()
        cen = [v[core].mean() for v in (xs, ys, zs)]
        a, b, cc = np.meshgrid(np.arange(5), np.arange(5), np.arange(5), indexing="ij")
        for d, (v, idx) in enumerate(zip((xs, ys, zs), (a, b, cc))):
            off = (v[sl] - (cen[d] + (idx.ravel() - 2.5 + 0.5) * hl)) / p.sigma
            np.testing.assert_allclose(X[k, 125 * (d + 1):125 * (d + 2)], off, rtol=1e-5, atol=1e-6)
        assert np.all(X[k, 500:] == 0)
        np.testing.assert_allclose(T[k, :27], z[core] / rms, rtol=1e-6, atol=1e-6)
        assert np.all(T[k, 27:] == 0)
    ctx.close()


@pytest.mark.gpu
def test_learned_3d_batched_solve():
    """The 3D network in the batched three-right-hand-side FGMRES: pairs
    harvested on one seeded lattice (3x3x3 clusters, one-cell buffer, h/sigma
    = 1 so that the exact method converges), the network trained on the
    device, then on another seeded lattice: with no cluster accepted the
    learned solve is the exact FP32 one bit for bit; with every eligible
    cluster on the network, FGMRES's residual moves the clusters whose core
    residual stalls back to exact blocks and every component reaches rtol
    1e-10 with the FP64 true residual."""
    torch = pytest.importorskip("torch")
    from petrbf_amd import problems
    from petrbf_amd.learned import Mlp, harvest
    from petrbf_amd.rbf import Context
    _announce("3D learned preconditioner: harvest, train, batched solve")
    dev = torch.device("cuda:0")
    pt = problems.lattice_3d_problem(24, seed=0x5EED0006, overlap=1.0)
    ct = Context.from_problem(pt)
    ct.set_points(pt.x, pt.y, pt.z)
    X, T, osc = harvest(ct, torch.from_numpy(pt.w[0]).to(dev), 12, torch)
    assert X.shape == (12 * ct.learned_eligible(), 512) and T.shape[1] == 32
    assert float(torch.abs(T[:, 27:]).max()) == 0.0
    net = Mlp(seed=9, dim=3)
    net.set_params(net.params()[0], osc)
    hist = net.train_dev(X, T, batch=256, steps=300, lr=1e-3, seed=4)
    _announce(f"3D training on {X.shape[0]} pairs: loss {hist[:20].mean():.3f} -> {hist[-50:].mean():.3f}")
    assert hist[-50:].mean() < hist[:20].mean()
    P, _ = net.params()
    assert np.all(P[-32 + 27:] == 0)  # the pad outputs' biases never move
    p = problems.lattice_3d_problem(24, overlap=1.0)
    ctx = Context.from_problem(p)
    ctx.set_points(p.x, p.y, p.z)
    w = torch.from_numpy(np.ascontiguousarray(p.w)).to(dev)
    gam = torch.empty_like(w)
    wv = torch.empty_like(w)
    res = {}
    for tag, tau in (("exact", None), ("none-accepted", 0.0), ("all", 1e30)):
        if tau is not None:
            n_net = ctx.learned_attach(net, tau)
            assert (n_net == 0) if tau == 0.0 else (n_net == ctx.learned_eligible() > 0)
        torch.cuda.synchronize()
        infs = ctx.solve_multi_dev(w, gam, 3, ksp="fgmres", pc="rasm32" if tau is None else "learned", rtol=1e-10)
        ctx.evaluate_dev(gam, wv, nrhs=3)
        ctx.synchronize()
        res[tag] = (gam.clone(), [(i.its, i.reason, i.n_network, i.n_net_fallback) for i in infs])
        for r in range(3):
            nw = float(torch.linalg.norm(w[r]))
            if nw == 0.0:
                assert infs[r].reason == 3
                continue
            true_rel = float(torch.linalg.norm(wv[r] - w[r])) / nw
            assert infs[r].reason == 2 and true_rel <= 1.05e-10, (tag, r, infs[r].reason, true_rel)
    _announce(f"3D batched solves (its, reason, network, moved): { {k: v[1] for k, v in res.items()} }")
    assert torch.equal(res["exact"][0], res["none-accepted"][0])
    assert [v[0] for v in res["exact"][1]] == [v[0] for v in res["none-accepted"][1]]
    net.close()
    ctx.close()
    ct.close()

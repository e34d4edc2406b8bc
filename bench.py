"""Benchmark: N = 4,000,000 2D Gaussian-RBF solve to rtol 1e-10 on one MI355X.

One step = the whole solve path with the inputs already in HBM: cell list and
sort, buffer / truncation lists, RASM setup (blocks, factorisation, core rows)
and GMRES(30) with left RASM to rtol 1e-10, from particles and RHS to gamma.
Workload: 2000 x 2000 jittered lattice on [-1,1]^2, sigma = h/0.8, 4x4-spacing
cells, a buffer of 8 spacings (two cells) per side (sigma_buffer = 16 sigma,
blocks of ~390 unknowns), truncation 5 sigma beyond the cell, Lamb-Oseen RHS.

Prints one JSON line (see the keys below).  --gpus > 1 is refused: this build
measures one GPU.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_FP32_MFMA_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "block_chol_n4m_hbm_traffic.txt")
CPU_VALIDATION_FILE = os.path.join(ROOT, "profiles", "cpu_baseline_validation.txt")


def cpu_validation():
    """The host-core baseline's estimator checked against full oracle solves
    (the committed summary of `bench.py --cpu-validate` on the benchmark machine's host cores),
    or None when the summary is absent."""
    try:
        for line in open(CPU_VALIDATION_FILE):
            if line.startswith("{"):
                return json.loads(line)
    except (OSError, ValueError):
        pass
    return None


def measured_traffic(kernel_key):
    """HBM bytes per launch of a kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel
    calibration), or None when the summary is absent."""
    try:
        for line in open(TRAFFIC_FILE):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel_key:
                return float(f[2])
    except OSError:
        pass
    return None
FLOP_PER_PAIR_2D = 15     # 2d/matmult.cxx:44-45 (exp counted as one flop)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(p, its, pairs, stride=25, n_mv=3):
    """Host-core oracle (the reference's CPU algorithm: LU blocks, GMRES(30))
    timed on a bounded sample of the same N = 4M workload (about 10 s of host
    work) and scaled to the full solve: lists at full size, n_mv full
    matvecs, the block LU of every stride-th cluster and its apply, then
    t = lists + factor/frac + its*(matvec + apply/frac)."""
    from oracle.oracle import Oracle, num_threads
    t0 = time.perf_counter()
    o = Oracle.from_problem(p)
    t_lists = time.perf_counter() - t0
    pi, _ = o.perm()
    g = p.w[0][pi].copy()
    t0 = time.perf_counter()
    for _ in range(n_mv):
        o.matvec(g)
    t_mv = (time.perf_counter() - t0) / n_mv
    t0 = time.perf_counter()
    o.rasm_setup_sample(stride)
    t_fac = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.rasm_apply_sample(g)
    t_ap = time.perf_counter() - t0
    ncell = o.ncell
    frac = (ncell // stride + 1) / ncell
    est = t_lists + t_fac / frac + its * (t_mv + t_ap / frac)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": est, "unit": "s", "cores": num_threads(), "kind": "port",
        "sample": (f"N=4M lattice: lists+sort full ({t_lists:.2f}s), {n_mv} full matvecs ({t_mv:.2f}s each), "
                   f"block LU of every {stride}th cluster ({t_fac:.2f}s) and its apply ({t_ap:.3f}s), "
                   f"scaled to all {ncell} clusters and {its} GMRES iterations"),
        "t_lists_s": t_lists, "t_matvec_s": t_mv,
        "pair_evals_per_s": pairs / t_mv,
        "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model,
    }


def cpu_validate():
    """Full host-core solves (lists, LU of every block, GMRES(30) to rtol
    1e-10) at two sizes against cpu_baseline's estimate from its bounded
    sample at the same sizes, with the estimate's relative error."""
    from oracle.oracle import Oracle, num_threads
    from petrbf_amd import problems
    rows = []
    for n_side in (500, 1000):
        p = problems.lattice_2d_problem(n_side, seed=0x5EED0001)
        t0 = time.perf_counter()
        o = Oracle.from_problem(p)
        pi, _ = o.perm()
        ref = o.solve_sorted(p.w[0][pi], rtol=1e-10, restart=30, max_it=1000)
        t_full = time.perf_counter() - t0
        ista, iend, _, _ = o.cells()
        tptr, _ = o.trunc_csr()
        pairs = int(((iend - ista + 1).clip(min=0) * np.diff(tptr)).sum())
        del o
        est = cpu_baseline(p, ref["its"], pairs)
        rows.append({"n": p.n, "its": ref["its"], "full_solve_s": t_full, "estimate_s": est["value"],
                     "relative_error": est["value"] / t_full - 1.0})
        log(f"cpu validation N={p.n}: full {t_full:.1f}s, estimate {est['value']:.1f}s "
            f"({100 * (est['value'] / t_full - 1):+.1f}%)")
    print(json.dumps({"cpu_baseline_validation": rows, "threads": num_threads(), "nproc": os.cpu_count(),
                      "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-side", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--mode", choices=["plain", "mixed", "both", "learned"], default="both",
                    help="learned: the C4 leg (network trained on ~2M pairs harvested from another seeded "
                         "lattice of the same size, then the learned FGMRES step against the mixed one); "
                         "its own JSON line.  With --workload c5: the 3D network in the batched three-RHS "
                         "FGMRES after the exact one")
    ap.add_argument("--learned-vectors", type=int, default=18, help="Krylov vectors harvested (x ncell pairs)")
    ap.add_argument("--learned-steps", type=int, default=2000, help="Adam steps of the training")
    ap.add_argument("--chol-waves", type=int, default=0, help="4 or 8 wavefronts per block factorisation")
    ap.add_argument("--chol-panel", type=int, default=0, help="1, 2 or 4 tile columns per panel")
    ap.add_argument("--chol-order", choices=["right", "left"], default=None,
                    help="right- or left-looking block factorisation")
    ap.add_argument("--workload", choices=["c2", "c3", "c5"], default="c2",
                    help="c2: the 2D N=4M solve (the headline); c3: the 2D N=4M non-uniform mixture on the "
                         "adaptive cluster structure; c5: the 3D 160^3 three-RHS workload; c3 and c5 print "
                         "their own JSON line")
    ap.add_argument("--c3-its", type=int, default=60, help="FGMRES/GMRES iteration budget (c3)")
    ap.add_argument("--c3-cap", type=int, default=12, help="particles per leaf before a cell splits (c3)")
    ap.add_argument("--c5-its", type=int, default=150, help="FGMRES iteration budget per component (c5)")
    ap.add_argument("--setup-only", choices=["rasm", "rasm32"], default=None,
                    help="only build the workload's lists and run one RASM setup (for counter runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 1 GiB five times (HBM counter calibration)")
    ap.add_argument("--cpu-validate", action="store_true",
                    help="host only: the oracle's full plain solve at N = 500^2 and 1000^2 against the sampled "
                         "estimator of cpu_baseline at the same sizes (a few minutes); one JSON line")
    args = ap.parse_args()
    if args.cpu_validate:
        return cpu_validate()
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8, "reps": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch
    from petrbf_amd import problems
    from petrbf_amd.rbf import Context

    dev = torch.device("cuda:0")
    if args.workload == "c5":
        return bench_c5(args, torch, problems, Context, dev)
    if args.workload == "c3":
        return bench_c3(args, torch, problems, Context, dev)
    if args.mode == "learned":
        return bench_learned(args, torch, problems, Context, dev)
    log(f"generating lattice {args.n_side}^2")
    p = problems.lattice_2d_problem(args.n_side, seed=0x5EED0001)
    x = torch.from_numpy(p.x).to(dev)
    y = torch.from_numpy(p.y).to(dev)
    w = torch.from_numpy(p.w[0]).to(dev)
    gamma = torch.empty_like(w)
    wv = torch.empty_like(w)
    stream = torch.cuda.Stream(device=dev)
    ctx = Context.from_problem(p)
    ctx.set_stream(stream.cuda_stream)
    if args.chol_waves:
        ctx.set_tuning(chol_waves=args.chol_waves)
    if args.chol_panel:
        ctx.set_tuning(chol_panel=args.chol_panel)
    if args.chol_order:
        ctx.set_tuning(chol_order=args.chol_order)
    if args.setup_only:
        ctx.set_points_dev(x, y)
        t0 = time.perf_counter()
        ctx.rasm_setup(args.setup_only)
        torch.cuda.synchronize()
        ci = ctx.info()
        print(json.dumps({"setup": args.setup_only, "wall_s": time.perf_counter() - t0,
                          "kernel_s": ci["chol_kernel_s"], "flops": ci["chol_flops"],
                          "chol_waves": ci["chol_waves"], "chol_panel": ci["chol_panel"],
                          "chol_order": ci["chol_order"]}), flush=True)
        return

    modes = {"plain": ("gmres", "rasm"), "mixed": ("fgmres", "rasm32")}
    run = ["plain", "mixed"] if args.mode == "both" else [args.mode]
    results = {}
    chol = {}
    gammas = {}
    for mode in run:
        ksp, pc = modes[mode]

        def step():
            ctx.set_points_dev(x, y)
            return ctx.solve_dev(w, gamma, ksp=ksp, pc=pc, rtol=1e-10, restart=30, max_it=1000)

        for i in range(args.warmup):
            inf = step()
            log(f"{mode} warmup {i}: its {inf.its} reason {inf.reason} setup {inf.t_setup_s:.3f}s "
                f"solve {inf.t_solve_s:.3f}s regularised blocks {inf.n_regularised}")
        torch.cuda.synchronize()
        infos = []
        t0 = time.perf_counter()
        for i in range(args.steps):
            infos.append(step())
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        inf = infos[-1]
        # true FP64 residual of the returned strengths (evaluation = the same operator)
        ctx.evaluate_dev(gamma, wv)
        torch.cuda.synchronize()
        true_rel = float(torch.linalg.norm(wv - w) / torch.linalg.norm(w))
        gammas[mode] = gamma.clone()
        log(f"{mode}: {args.steps} steps {elapsed:.3f}s; its {inf.its} reason {inf.reason} true rel resid {true_rel:.3e}")
        ci = ctx.info()
        log(f"{mode}: block factorisation config waves={ci['chol_waves']} panel={ci['chol_panel']} "
            f"order={'left' if ci['chol_order'] == 2 else 'right'}, "
            f"kernel {ci['chol_kernel_s']*1e3:.1f} ms, {ci['chol_flops']/ci['chol_kernel_s']/1e12:.1f} TFLOP/s")
        chol[mode] = (ci["chol_kernel_s"], ci["chol_flops"], ci["chol_waves"], ci["chol_panel"], ci["chol_order"])
        results[mode] = dict(ms_per_step=elapsed / args.steps * 1e3, its=inf.its, reason=inf.reason,
                             chol_waves=ci["chol_waves"], chol_panel=ci["chol_panel"],
                             chol_order="left" if ci["chol_order"] == 2 else "right",
                             t_setup_s=inf.t_setup_s, t_solve_s=inf.t_solve_s, t_matvec_s=inf.t_matvec_s,
                             n_matvec=inf.n_matvec, t_pc_s=inf.t_pc_s, n_pc=inf.n_pc,
                             n_regularised=inf.n_regularised, true_rel_residual=true_rel)

    # ---- interpolation error against the analytic field on the h/2 grid
    # (2d/main2.cxx:53,126-137: the (2 n_side - 1)^2 grid of half the lattice
    # spacing, between the particles), for each mode's strengths
    xt, yt, _ = problems.jittered_lattice_2d(2 * args.n_side - 1, seed=3, jitter=0.0)
    exact = torch.from_numpy(problems.lamb_oseen(xt, yt)).to(dev)
    ev = Context.from_problem(p)
    ev.set_stream(stream.cuda_stream)
    ev.set_points_dev(torch.from_numpy(xt).to(dev), torch.from_numpy(yt).to(dev), None, x, y)
    wt = torch.empty_like(exact)
    for mode, g in gammas.items():
        ev.evaluate_dev(g, wt)
        torch.cuda.synchronize()
        results[mode]["analytic_field_error_h2"] = float(torch.linalg.norm(wt - exact) / torch.linalg.norm(exact))
        log(f"{mode}: analytic-field error on the h/2 grid ({xt.size} points) "
            f"{results[mode]['analytic_field_error_h2']:.6e}")
    ev.close()
    del xt, yt, exact, wt, gammas

    # ---- per-kernel timing with HIP events on the library's stream (FP64 RASM set up last if plain ran)
    ctx.set_points_dev(x, y)
    ctx.rasm_setup("rasm")
    info = ctx.info()
    ista, iend, _, _ = ctx.cells()
    T = (iend - ista + 1).astype(np.int64)
    tptr = ctx.trunc_ptr()
    bptr = ctx.buffer_ptr()
    S = np.diff(tptr)
    nb = np.diff(bptr)
    pairs = int((T * S).sum())
    g_s = torch.randn(p.n, dtype=torch.float64, device=dev)
    out = torch.empty_like(g_s)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        for _ in range(3):  # untimed: first-touch and clock ramp
            fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.kernel_reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / 1e3 / args.kernel_reps

    t_mv = timed(lambda: ctx.matvec_dev(g_s, out))
    t_ap = timed(lambda: ctx.rasm_apply_dev(g_s, out))
    ctx.rasm_setup("rasm32")
    t_ap32 = timed(lambda: ctx.rasm_apply_dev(g_s, out))
    m_elems = int((nb * T).sum())
    side_bytes = int(nb.sum()) * 4 + p.n * 8 * 2 + (info["ncell"] + 1) * 8 * 2
    apply_bytes = m_elems * 8 + side_bytes
    apply32_bytes = m_elems * 4 + side_bytes
    mv_flops = FLOP_PER_PAIR_2D * pairs
    log(f"matvec {t_mv*1e3:.3f} ms ({pairs/t_mv:.3e} pairs/s, {mv_flops/t_mv/1e12:.2f} TFLOP/s); "
        f"rasm apply fp64 {t_ap*1e3:.3f} ms ({apply_bytes/t_ap/1e9:.0f} GB/s), "
        f"fp32 {t_ap32*1e3:.3f} ms ({apply32_bytes/t_ap32/1e9:.0f} GB/s)")

    # graph-captured solves do not time their phases (t_matvec_s and t_pc_s
    # come back 0): report them as unmeasured, and beside them an estimate of
    # the split from the kernels' own timing loops above
    for mode, r in results.items():
        if r.get("t_matvec_s") == 0.0 and r.get("n_matvec", 0) > 0:
            r["t_matvec_s"] = None
            r["t_pc_s"] = None
            t_pc1 = t_ap32 if mode == "mixed" else t_ap
            est_mv, est_pc = r["n_matvec"] * t_mv, r["n_pc"] * t_pc1
            r["solve_split_estimate_s"] = {
                "matvec": est_mv, "rasm_apply": est_pc,
                "orthogonalisation_and_rest": max(0.0, r["t_solve_s"] - est_mv - est_pc),
                "from": "n_matvec and n_pc times the per-launch times of the kernel timing loops"}

    head = "mixed" if "mixed" in results else run[0]
    ms_per_step = results[head]["ms_per_step"]
    res = {
        "metric": "N=4M 2D solve time to rtol 1e-10 (learned/FP32 vs plain); Gaussian pair-evals/s",
        "value": ms_per_step / 1e3,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if head == "plain" else "f64 (Krylov, matvec) + f32 (RASM blocks)",
        "data": "synthetic",
        "config": {"workload": f"2D jittered lattice {args.n_side}x{args.n_side} (N={p.n}), sigma = h/0.8, "
                               "4x4-spacing cells, buffer of 8 spacings (2 cells) per side (sigma_buffer = 16 sigma, "
                               f"blocks of up to {info['max_buffer']} unknowns), 5 sigma truncation beyond the cell, "
                               "Lamb-Oseen RHS, rtol 1e-10",
                   "mode": head, "n": p.n, "ncell": info["ncell"]},
        "modes": results,
        "speedup_mixed_vs_plain": (results["plain"]["ms_per_step"] / results["mixed"]["ms_per_step"]
                                   if len(results) == 2 else None),
        "pair_evals_per_s": pairs / t_mv,
        "pairs_per_matvec": pairs,
        "roofline_apply_fp64": {"bound": "hbm", "kernel": "k_rasm_apply<double>", "achieved": apply_bytes / t_ap / 1e9,
                                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": apply_bytes / t_ap / 1e9 / PEAK_HBM_GBS,
                                "avg_ms": t_ap * 1e3, "bytes_per_launch": apply_bytes},
        "roofline_apply_fp32": {"bound": "hbm", "kernel": "k_rasm_apply<float>",
                                "achieved": apply32_bytes / t_ap32 / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                "frac": apply32_bytes / t_ap32 / 1e9 / PEAK_HBM_GBS, "avg_ms": t_ap32 * 1e3,
                                "bytes_per_launch": apply32_bytes},
        "roofline_matvec": {"bound": "fp64-valu", "kernel": "k_matvec_wave", "achieved": mv_flops / t_mv / 1e12,
                            "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                            "frac": mv_flops / t_mv / 1e12 / PEAK_FP64_TFLOPS, "avg_ms": t_mv * 1e3,
                            "flops_per_launch": mv_flops},
    }
    # the dominant kernel of the headline (mixed) step: the FP32 block
    # factorisation, one persistent launch per setup, timed with HIP events on
    # the library's stream; FLOPs = sum over blocks of n^3/3 + n^2 T + T^3
    # (template arguments: precision, dimension, wavefronts, panel columns,
    # left-looking)
    for mode, key, peak, name in (("mixed", "roofline", PEAK_FP32_MFMA_TFLOPS, "k_block_chol<float,2,%d,%d,%s>"),
                                  ("plain", "roofline_block_chol_fp64", PEAK_FP64_TFLOPS,
                                   "k_block_chol<double,2,%d,%d,%s>")):
        if mode not in chol:
            continue
        t_k, fl, nw, npc, order = chol[mode]
        kname = name % (nw, npc, "true" if order == 2 else "false")
        res[key] = {"bound": "mfma", "kernel": kname, "achieved": fl / t_k / 1e12, "peak": peak,
                    "unit": "TFLOP/s", "frac": fl / t_k / 1e12 / peak, "avg_ms": t_k * 1e3,
                    "flops_per_launch": fl, "traffic": measured_traffic(kname)}
    inf_its = results.get("plain", results[head])["its"]
    if not args.no_cpu:
        log("cpu baseline (oracle, bounded sample)")
        res["cpu_baseline"] = cpu_baseline(p, inf_its, pairs)
        res["cpu_baseline"]["validation"] = cpu_validation()
        log(f"cpu baseline est {res['cpu_baseline']['value']:.1f}s on {res['cpu_baseline']['cores']} threads")
    print(json.dumps(res), flush=True)


def bench_learned(args, torch, problems, Context, dev):
    """C4: the learned preconditioner on the N = n_side^2 lattice.  Pairs
    (patch features, exact FP64 block core solve) are harvested from the
    first Krylov vectors of an exact-block FGMRES run on another seeded
    lattice of the same size (vectors x eligible clusters, ~2M at N = 4M),
    the network is trained on the device (Adam, BF16 MFMA), attached to every
    eligible cluster, and the learned FGMRES step (lists, attach, solve with
    the per-cluster fallback) is timed against the mixed FP32-block step."""
    from petrbf_amd.learned import Mlp, harvest
    log(f"learned: harvesting on lattice {args.n_side}^2 (seed 0x5EED0002)")
    pt = problems.lattice_2d_problem(args.n_side, seed=0x5EED0002)
    ct = Context.from_problem(pt)
    ct.set_points_dev(torch.from_numpy(pt.x).to(dev), torch.from_numpy(pt.y).to(dev))
    t0 = time.perf_counter()
    X, T, osc = harvest(ct, torch.from_numpy(pt.w[0]).to(dev), args.learned_vectors, torch)
    t_harvest = time.perf_counter() - t0
    ct.close()
    del pt
    net = Mlp(seed=7)
    net.set_params(net.params()[0], osc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = net.train_dev(X, T, batch=4096, steps=args.learned_steps, lr=1e-3, seed=3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    npairs = int(X.shape[0])
    log(f"learned: {npairs} pairs harvested in {t_harvest:.1f}s; {args.learned_steps} Adam steps in {t_train:.1f}s, "
        f"loss {float(hist[:20].mean()):.3f} -> {float(hist[-50:].mean()):.3f}")
    del X, T
    p = problems.lattice_2d_problem(args.n_side, seed=0x5EED0001)
    x, y, w = (torch.from_numpy(a).to(dev) for a in (p.x, p.y, p.w[0]))
    gamma, wv = torch.empty_like(w), torch.empty_like(w)
    ctx = Context.from_problem(p)
    out = {}
    for mode in ("mixed", "learned"):
        def step():
            ctx.set_points_dev(x, y)
            if mode == "learned":
                n_net = ctx.learned_attach(net, 1e30)
                return n_net, ctx.solve_dev(w, gamma, ksp="fgmres", pc="learned", rtol=1e-10, restart=30,
                                            max_it=1000)
            return 0, ctx.solve_dev(w, gamma, ksp="fgmres", pc="rasm32", rtol=1e-10, restart=30, max_it=1000)
        step()  # warm-up (tuner, graphs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n_net, inf = step()
        torch.cuda.synchronize()
        t_step = time.perf_counter() - t0
        ctx.evaluate_dev(gamma, wv)
        torch.cuda.synchronize()
        true_rel = float(torch.linalg.norm(wv - w) / torch.linalg.norm(w))
        out[mode] = {"step_s": t_step, "its": inf.its, "reason": inf.reason, "t_setup_s": inf.t_setup_s,
                     "t_solve_s": inf.t_solve_s, "true_rel_residual": true_rel}
        if mode == "learned":
            out[mode].update(clusters_attached=n_net, clusters_on_network_at_end=inf.n_network,
                             clusters_moved_to_exact_blocks=inf.n_net_fallback)
        log(f"{mode}: step {t_step:.3f}s, its {inf.its} reason {inf.reason}, true residual {true_rel:.3e}"
            + (f", attached {n_net}, moved to exact blocks {inf.n_net_fallback}" if mode == "learned" else ""))
    print(json.dumps({
        "metric": "C4 learned preconditioner on the N=4M lattice: step time and iterations vs exact FP32 blocks",
        "n": p.n, "pairs": npairs, "harvest_s": t_harvest, "train_steps": args.learned_steps, "train_s": t_train,
        "loss_first": float(hist[:20].mean()), "loss_last": float(hist[-50:].mean()), "modes": out,
        "gain_s_vs_mixed": out["mixed"]["step_s"] - out["learned"]["step_s"],
        "gain_its_vs_mixed": out["mixed"]["its"] - out["learned"]["its"],
    }), flush=True)
    net.close()


def bench_c3(args, torch, problems, Context, dev):
    """2D non-uniform mixture, N = 4*10^6 (70% uniform on [-1,1]^2, 30% in
    two Gaussians at (+-0.4, 0), std 0.15), sigma from the mean spacing, 5
    sigma cutoff, on the adaptive cluster structure (grid cells of more than
    cap particles split down to depth 8; buffer 8 mean spacings per side at
    depth 0, halved per level).  One step = adaptive lists + RASM setup +
    Krylov solve within an iteration budget, mixed (FP32 blocks of >= 32
    unknowns, FGMRES) and plain (FP64 blocks, GMRES).  The points cluster
    (nearest neighbours down to ~1% of the mean spacing), so blocks and the
    whole system are numerically singular in FP64; the line reports how far
    each solve gets, with the true residual."""
    n = 4_000_000
    log(f"generating mixture N={n}")
    p = problems.mixture_2d_problem(n)
    x, y, w = (torch.from_numpy(a).to(dev) for a in (p.x, p.y, p.w[0]))
    gamma = torch.empty_like(w)
    wv = torch.empty_like(w)
    stream = torch.cuda.Stream(device=dev)
    ctx = Context.from_problem(p)
    ctx.set_stream(stream.cuda_stream)
    cap, depth = args.c3_cap, 8
    ctx.set_points_adaptive_dev(x, y, cap=cap, max_depth=depth)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.set_points_adaptive_dev(x, y, cap=cap, max_depth=depth)
    torch.cuda.synchronize()
    t_lists = (time.perf_counter() - t0) / reps
    info = ctx.info()
    lv = ctx.leaves()
    nb = np.diff(ctx.buffer_ptr())
    ista, iend, _, _ = ctx.cells()
    T = iend - ista + 1
    ne = nb > 0
    stats = {"leaves": info["ncell"], "nonempty_leaves": int(ne.sum()),
             "depth_histogram": np.bincount(lv["depth"]).tolist(),
             "block_unknowns_min_mean_max": [int(nb[ne].min()), float(nb[ne].mean()), int(nb.max())],
             "blocks_under_32": int(((nb > 0) & (nb < 32)).sum()), "core_max": int(T.max()),
             "trunc_mean": float(np.diff(ctx.trunc_ptr())[ne].mean())}
    log(f"adaptive structure {t_lists*1e3:.1f} ms: {stats}")
    out = {}
    for mode, ksp, pc in (("mixed", "fgmres", "rasm32"), ("plain", "gmres", "rasm")):
        def step():
            ctx.set_points_adaptive_dev(x, y, cap=cap, max_depth=depth)
            return ctx.solve_dev(w, gamma, ksp=ksp, pc=pc, rtol=1e-10, restart=30, max_it=args.c3_its)
        inf = step()  # warm-up (tuner, graphs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        inf = step()
        torch.cuda.synchronize()
        t_step = time.perf_counter() - t0
        ctx.evaluate_dev(gamma, wv)
        torch.cuda.synchronize()
        true_rel = float(torch.linalg.norm(wv - w) / torch.linalg.norm(w))
        ci = ctx.info()
        out[mode] = {"step_s": t_step, "its": inf.its, "reason": inf.reason, "t_setup_s": inf.t_setup_s,
                     "t_solve_s": inf.t_solve_s, "residual_estimate": inf.rnorm / inf.rnorm0,
                     "true_rel_residual": true_rel, "regularised_blocks": inf.n_regularised,
                     "chol_kernel_s": ci["chol_kernel_s"],
                     "chol_tflops": ci["chol_flops"] / ci["chol_kernel_s"] / 1e12 if ci["chol_kernel_s"] else None,
                     "chol_config": f"{ci['chol_waves']}x{ci['chol_panel']} "
                                    f"{'left' if ci['chol_order'] == 2 else 'right'}"}
        log(f"{mode}: step {t_step:.3f}s (setup {inf.t_setup_s:.3f}s, solve {inf.t_solve_s:.3f}s), its {inf.its} "
            f"reason {inf.reason}, residual estimate {inf.rnorm/inf.rnorm0:.3e}, true {true_rel:.3e}, "
            f"regularised blocks {inf.n_regularised}")
    print(json.dumps({
        "metric": "2D N=4M non-uniform mixture on the adaptive cluster structure: step time and residual reached",
        "n": n, "adaptive_cap": cap, "adaptive_max_depth": depth, "t_adaptive_lists_s": t_lists,
        "structure": stats, "iteration_budget": args.c3_its, "modes": out,
    }), flush=True)


def bench_c5(args, torch, problems, Context, dev):
    """3D vector-field workload: 160^3 jittered lattice, 3x3x3-spacing cells,
    whole-cell buffer, 4-sigma cutoff, vortex-ring vorticity as three
    right-hand sides.  Measures the lists, the three-RHS matvec (one exp per
    pair shared by the three components), the FP32 block setup and how far
    one batched FGMRES(30) over the three components gets within a fixed
    iteration budget."""
    n_side = 160
    log(f"generating 3D lattice {n_side}^3")
    p = problems.lattice_3d_problem(n_side)
    x, y, z = (torch.from_numpy(a).to(dev) for a in (p.x, p.y, p.z))
    w = torch.from_numpy(np.ascontiguousarray(p.w)).to(dev)  # (3, n)
    ctx = Context.from_problem(p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.set_points_dev(x, y, z)
    torch.cuda.synchronize()
    t_lists = time.perf_counter() - t0
    info = ctx.info()
    ista, iend, _, _ = ctx.cells()
    T = (iend - ista + 1).astype(np.int64)
    pairs = int((np.maximum(T, 0) * np.diff(ctx.trunc_ptr())).sum())
    out = torch.empty_like(w)
    ctx.matvec_dev(w, out, nrhs=3)
    ctx.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.matvec_dev(w, out, nrhs=3)
    ctx.synchronize()
    t_mv = (time.perf_counter() - t0) / reps
    flops = (18 + 2 * 2) * pairs  # 3D pair (18) + two more right-hand sides sharing the exp
    log(f"3D lists {t_lists:.3f}s; 3-RHS matvec {t_mv*1e3:.2f} ms ({pairs/t_mv:.3e} pairs/s, "
        f"{flops/t_mv/1e12:.1f} TFLOP/s)")
    t0 = time.perf_counter()
    ctx.rasm_setup("rasm32")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    ci = ctx.info()
    log(f"FP32 setup {t_setup:.3f}s (kernel {ci['chol_kernel_s']:.3f}s, {ci['chol_flops']/ci['chol_kernel_s']/1e12:.1f} "
        f"TFLOP/s, max block {ci['max_buffer']})")
    # one batched FGMRES(30) over the three components: each iteration one
    # three-vector matvec and one apply streaming the FP32 core rows once
    gam = torch.empty_like(w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = ctx.solve_multi_dev(w, gam, 3, ksp="fgmres", pc="rasm32", rtol=1e-10, restart=30, max_it=args.c5_its)
    ctx.synchronize()
    t_solve = time.perf_counter() - t0
    comps = []
    for c, inf in enumerate(infos):
        rel = inf.rnorm / inf.rnorm0 if inf.rnorm0 > 0 else 0.0  # the ring's z component is zero
        comps.append({"its": inf.its, "reason": inf.reason, "rel_residual": rel})
        log(f"component {c}: its {inf.its} reason {inf.reason} residual {rel:.3e}")
    log(f"batched three-RHS FGMRES: {t_solve:.3f}s for {max(i.its for i in infos)} iterations")
    res = {
        "metric": "3D 160^3 three-RHS workload: lists, batched matvec, FP32 setup, FGMRES progress",
        "n": p.n, "ncell": info["ncell"], "max_buffer": ci["max_buffer"], "t_lists_s": t_lists,
        "matvec_3rhs_ms": t_mv * 1e3, "pairs_per_matvec": pairs, "pair_evals_per_s": pairs / t_mv,
        "matvec_tflops": flops / t_mv / 1e12, "setup_fp32_s": t_setup, "chol_kernel_s": ci["chol_kernel_s"],
        "chol_tflops": ci["chol_flops"] / ci["chol_kernel_s"] / 1e12, "fgmres_batched_3rhs": comps,
        "fgmres_batched_s": t_solve, "fgmres_iteration_budget": args.c5_its,
    }
    if args.mode == "learned":
        res["learned"] = bench_c5_learned(args, torch, problems, Context, dev, ctx, w)
    print(json.dumps(res), flush=True)


def bench_c5_learned(args, torch, problems, Context, dev, ctx, w):
    """The 3D network (5x5x5-slot patch -> 27 core corrections) in the batched
    three-RHS FGMRES of the 160^3 workload: pairs harvested from the first
    Krylov vectors of an exact FP64 FGMRES run (x component) on another seeded
    160^3 lattice, the network trained on the device, attached to every
    eligible cluster of ctx, and the batched learned solve run with the same
    iteration budget as the exact one."""
    from petrbf_amd.learned import Mlp, harvest
    n_side = 160
    log(f"c5 learned: harvesting on lattice {n_side}^3 (seed 0x5EED0006)")
    pt = problems.lattice_3d_problem(n_side, seed=0x5EED0006)
    ct = Context.from_problem(pt)
    ct.set_points_dev(*(torch.from_numpy(a).to(dev) for a in (pt.x, pt.y, pt.z)))
    t0 = time.perf_counter()
    X, T, osc = harvest(ct, torch.from_numpy(pt.w[0]).to(dev), args.learned_vectors, torch)
    t_harvest = time.perf_counter() - t0
    ct.close()
    del pt
    net = Mlp(seed=7, dim=3)
    net.set_params(net.params()[0], osc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = net.train_dev(X, T, batch=4096, steps=args.learned_steps, lr=1e-3, seed=3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    npairs = int(X.shape[0])
    log(f"c5 learned: {npairs} pairs in {t_harvest:.1f}s; {args.learned_steps} Adam steps in {t_train:.1f}s, "
        f"loss {float(hist[:20].mean()):.3f} -> {float(hist[-50:].mean()):.3f}")
    del X, T
    t0 = time.perf_counter()
    n_net = ctx.learned_attach(net, 1e30)
    torch.cuda.synchronize()
    t_attach = time.perf_counter() - t0
    gam = torch.empty_like(w)
    t0 = time.perf_counter()
    infos = ctx.solve_multi_dev(w, gam, 3, ksp="fgmres", pc="learned", rtol=1e-10, restart=30, max_it=args.c5_its)
    ctx.synchronize()
    t_solve = time.perf_counter() - t0
    comps = []
    for c, inf in enumerate(infos):
        rel = inf.rnorm / inf.rnorm0 if inf.rnorm0 > 0 else 0.0
        comps.append({"its": inf.its, "reason": inf.reason, "rel_residual": rel})
        log(f"c5 learned component {c}: its {inf.its} reason {inf.reason} residual {rel:.3e}")
    log(f"c5 learned: attached {n_net} clusters ({t_attach:.2f}s), on the network at the end "
        f"{infos[0].n_network}, moved to exact blocks {infos[0].n_net_fallback}; solve {t_solve:.3f}s")
    net.close()
    return {"pairs": npairs, "harvest_s": t_harvest, "train_steps": args.learned_steps, "train_s": t_train,
            "loss_first": float(hist[:20].mean()), "loss_last": float(hist[-50:].mean()),
            "clusters_attached": n_net, "attach_s": t_attach, "clusters_on_network_at_end": infos[0].n_network,
            "clusters_moved_to_exact_blocks": infos[0].n_net_fallback, "fgmres_batched_3rhs": comps,
            "fgmres_batched_s": t_solve}


if __name__ == "__main__":
    main()

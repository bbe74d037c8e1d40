#!/usr/bin/env python3
"""Benchmark of foeuler's hot path on one MI355X: residual evaluations/s of the first-order Roe
residual (setBC + syncState + findEuler + /v) on the 1.5M-tet throughput mesh, with the DQ Jv and
FP32 variants and the reference's own CPU residual timed on the host cores in the same run; then
(--solve 1, the default) the residual and Jv on the 6M-tet mesh, whose HBM roofline is the one
quoted (its working set does not fit the MALL), and BDF step times to the 1e-8 Newton tolerance on
that mesh for ILU / learned / no preconditioner in FP64 and mixed precision, next to the CPU solver.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20]

Prints one JSON line.  A step is one residual evaluation over the whole mesh with the state
resident in HBM.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BDF step time to 1e-8 Newton tol, learned vs ILU precond; residual evals/s"
HBM_PEAK_GBS = 8000.0
GAMMA = 1.4
SEED = 20261202


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


GRID_UP_S = {}


def throughput_mesh(nx, ny, nz, lx, ly, lz):
    """Config of the throughput benchmark: jittered box, RCM then chunk renumbering; the
    finite-volume grid is built on the GPU (its time recorded per cell count)."""
    import fosolvers_amd as fa
    pg = fa.PolyGrid.box(nx, ny, nz, lx, ly, lz, jitter=0.2, seed=SEED, gids=(1, 2, 3, 4, 5, 6, 0))
    pg.renumber(fa.RENUMBER_RCM)
    pg.renumber(fa.RENUMBER_CHUNKS)
    t0 = time.perf_counter()
    fv = pg.up(gpu=True)
    GRID_UP_S[fv.nC] = time.perf_counter() - t0
    return fv


def throughput_state(fv, mach=0.8, amp=0.05, width=0.1, seed=SEED):
    """Mach 0.8 free stream along x, rho and p scaled by 1 + 0.05 * sum of 8 seeded Gaussians."""
    nC = fv.nC
    c = fv.p[:nC]
    rng = np.random.default_rng(seed)
    lo, hi = fv.pN.min(0), fv.pN.max(0)
    centres = lo + (hi - lo) * rng.random((8, 3))
    g = np.zeros(nC)
    for k in centres:
        g += np.exp(-((c - k) ** 2).sum(1) / (2 * width ** 2))
    f = 1.0 + amp * g
    rho, p = f, f / GAMMA
    x = np.zeros(5 * nC)
    x[:nC] = rho
    x[nC:2 * nC] = rho * mach
    x[4 * nC:] = p / (GAMMA - 1.0) + 0.5 * rho * mach * mach
    return x


def residual_bytes(fv, s):
    """Algorithmic bytes of one residual evaluation: each datum of the reference layout once.
    Per cell state in (5s), volume (s), residual out (5s); per pair iEP (8), aP (s), normP (3s);
    per boundary pair its BC index (4)."""
    return fv.nC * 11 * s + fv.nP * (8 + 4 * s) + fv.nPb * 4


def jv_bytes(fv, s):
    """Fused DQ Jv: the residual's bytes plus y and v for the WRMS norm, v for the perturbed
    gather and fy for the combine (4 vectors of 5 nC)."""
    return residual_bytes(fv, s) + 20 * fv.nC * s


def jvp_bytes(fv, s):
    """Analytic Jv: the residual's bytes plus the direction v (5 nC)."""
    return residual_bytes(fv, s) + 5 * fv.nC * s


def ilu_apply_bytes(fv, s):
    """Block ILU(0) apply: the L/U/D blocks ((nC + 2 nPi) 25 s), their indices (8 nPi + 8 (nC + 1)),
    r in and z out (10 nC s)."""
    return (fv.nC + 2 * fv.nPi) * 25 * s + 8 * fv.nPi + 8 * (fv.nC + 1) + 10 * fv.nC * s


def read_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, copy-kernel
    calibrated), recorded as 'traffic_bytes_per_launch = N' in profiles/<name>."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
    return None


def time_gpu(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1000.0


def cpu_baseline(fv, x, conds, budget_s):
    """The reference's residual (compiled from its sources with its own flags, OpenMP) on the
    host cores, or the C oracle when the reference build is absent."""
    from oracle import oracle as O
    cg, ct, cp = conds
    out = np.zeros_like(x)
    if os.path.exists(O.REF_OMP_LIB):
        rt = O.RefTimer(fv, cg, ct, cp, 1.0, GAMMA)
        kind, cores, fn = "reference", rt.threads, (lambda: rt.rhs(x, out))
    else:
        og = O.Grid(fv.pN, fv.sE, fv.nNE, fv.iNE, fv.gid)
        bc = O.BC(og, cg, ct, cp, r=1.0, gamm=GAMMA)
        scratch = np.zeros(13 * og.nE)
        kind, cores = "port", 1
        fn = lambda: out.__setitem__(slice(None), O.rhs(og, bc, x, scratch))  # noqa: E731
    t0 = time.perf_counter()
    fn()
    t1 = time.perf_counter() - t0
    n = max(1, min(200, int(budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(n):
        fn()
    dt = time.perf_counter() - t0
    ref = out.copy()

    def ref_rhs(z):
        o = np.zeros_like(z)
        if kind == "reference":
            rt.rhs(z, o)
        else:
            o[:] = O.rhs(og, bc, z, scratch)
        return o
    return {"value": n / dt, "unit": "residual evals/s", "cores": cores, "kind": kind,
            "sample": f"{n} FP64 residual evaluations (setBC+syncState+findEuler+/v) on the same "
                      f"{fv.nC}-tet mesh and state, after 1 untimed"}, ref, ref_rhs


def oracle_dq_jv(ref_rhs, x, fx, v, rtol=1e-6, atol=1e-12):
    """CVODE's difference quotient on the FP64 reference residual: sigma = 1/||v||_WRMS with the
    weights the library's fixed-step mode uses (1/(rtol|y| + atol + rtol rms(y)))."""
    fl = atol + rtol * np.sqrt(np.mean(x * x))
    w = 1.0 / (rtol * np.abs(x) + fl)
    sigma = 1.0 / np.sqrt(np.mean((v * w) ** 2))
    return (ref_rhs(x + sigma * v) - fx) / sigma


def cpu_bdf_baseline(fv6, x6, conds6, dt6, gpu_step1, steps_c):
    """The CPU full solver: the oracle's FP64 fixed-step BDF-Newton-GMRES with block ILU(0) (one
    thread), on config A (the 24,000-tet Sod tube, dt 1e-4, 20 steps, against the library's GPU run
    of the same steps) and on a bounded sample of config C (its first steps on the 6M-tet mesh,
    against the GPU's first step)."""
    import torch
    import fosolvers_amd as fa
    from oracle import oracle as O
    out = {"kind": "port", "cores": 1}
    # config A
    pg = fa.PolyGrid.box(40, 10, 10, 1.0, 0.25, 0.25, jitter=0.0, seed=0)
    fvA = pg.up()
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    cg, ct, cp = [1, 2, 3, 4, 5, 6], [0] * 6, np.zeros((6, 5))
    bc = O.BC(og, cg, ct, cp, r=1.0, gamm=GAMMA)
    xA = np.zeros(5 * og.nC)
    xc = og.p[:og.nC, 0]
    rho, p = np.where(xc < 0.5, 1.0, 0.125), np.where(xc < 0.5, 1.0, 0.1)
    xA[:og.nC], xA[4 * og.nC:] = rho, p / (GAMMA - 1.0)
    t0 = time.perf_counter()
    stA, itA = O.bdf_euler(og, bc, xA, 1e-4, 20, maxl=30, lin_tol=1e-3, ilu=True)
    tA = time.perf_counter() - t0
    eu = fa.Euler(fvA, fa.CondTab(cg, ct, cp), r=1.0, gamma=GAMMA)
    yA = torch.tensor(xA, dtype=torch.float64, device="cuda")
    yG, tG, stG = bdf_run(eu, yA, 1e-4, 20, "ilu", maxl=30)
    del eu
    out["A"] = {"workload": "40x10x10 hexes x6 = 24000 tets, Sod tube, walls, dt 1e-4, 20 FP64 block-ILU BDF steps "
                            "to ||G||_WRMS <= 1e-8",
                "cpu_step_ms": 1e3 * tA / 20, "cpu_gmres_iters_per_step": float(itA[:, 1].mean()),
                "cpu_newton_iters_per_step": float(itA[:, 0].mean()),
                "gpu_step_ms_amortised": 1e3 * float(tG.mean()), "gpu_step_ms_after_first": 1e3 * float(tG[1:].mean()),
                "gpu_gmres_iters_per_step": stG["nli"] / 20,
                "final_state_rel_diff_gpu_vs_cpu": float(np.linalg.norm(yG.cpu().numpy() - stA[-1]) /
                                                         np.linalg.norm(stA[-1]))}
    log(f"CPU BDF A: {1e3 * tA / 20:.0f} ms/step (1 thread, {itA[:, 1].mean():.1f} GMRES its/step); GPU "
        f"{1e3 * tG.mean():.1f} ms/step amortised; final states differ by {out['A']['final_state_rel_diff_gpu_vs_cpu']:.1e}")
    if steps_c > 0:
        t0 = time.perf_counter()
        ogC = O.Grid(fv6.pN, fv6.sE, fv6.nNE, fv6.iNE, fv6.gid)
        bcC = O.BC(ogC, *conds6, r=1.0, gamm=GAMMA)
        tg = time.perf_counter() - t0
        t0 = time.perf_counter()
        stC, itC = O.bdf_euler(ogC, bcC, x6, dt6, steps_c, maxl=30, lin_tol=1e-3, ilu=True)
        tC = time.perf_counter() - t0
        out["C"] = {"workload": f"{fv6.nC}-tet config C mesh and state, the first {steps_c} FP64 block-ILU BDF steps",
                    "cpu_step_ms": 1e3 * tC / steps_c, "cpu_gmres_iters_per_step": float(itC[:, 1].mean()),
                    "cpu_grid_build_s": tg,
                    "first_step_state_rel_diff_gpu_vs_cpu": float(np.linalg.norm(gpu_step1 - stC[0]) /
                                                                  np.linalg.norm(stC[0]))}
        log(f"CPU BDF C: {tC / steps_c:.1f} s/step (1 thread, {itC[:, 1].mean():.1f} GMRES its/step), grid {tg:.1f}s; "
            f"first step vs GPU {out['C']['first_step_state_rel_diff_gpu_vs_cpu']:.1e}")
    return out


def solve_state(fv, lx, ly, lz, mach=2.0, amp=0.1, seed=SEED):
    """Mixed-precision solve config: Mach 2 along x, rho = 1, p = (1 + amp * sum_k a_k sin(2 pi k.x/L
    + phi_k)) / gamma over wavenumbers k in {1..4}^3 with seeded phases and amplitudes summing to 1."""
    nC = fv.nC
    c = fv.p[:nC]
    rng = np.random.default_rng(seed)
    ks = np.array([(a, b, d) for a in range(1, 5) for b in range(1, 5) for d in range(1, 5)], dtype=float)
    amps = rng.random(len(ks))
    amps /= amps.sum()
    phs = 2 * np.pi * rng.random(len(ks))
    xh = c / np.array([lx, ly, lz])
    pert = np.zeros(nC)
    for k, a, ph in zip(ks, amps, phs):
        pert += a * np.sin(2 * np.pi * (xh @ k) + ph)
    p = (1.0 + amp * pert) / GAMMA
    x = np.zeros(5 * nC)
    x[:nC] = 1.0
    x[nC:2 * nC] = mach
    x[4 * nC:] = p / (GAMMA - 1.0) + 0.5 * mach * mach
    return x


def bdf_run(eu, y0, dt, nsteps, precond, fallback_cap=0, collect=None, jtimes=True, maxl=30, first=None, fused=True):
    """nsteps fixed BDF steps from y0 to a 1e-8 Newton tolerance; per-step wall times (synchronised).
    The stats gain gcv_max: the largest converged ||G||_WRMS under CVODE's unfloored weights.
    fused: GMRES's product v - gamma J v in the Jv kernel's epilogue (else a second vector pass)."""
    import torch
    import fosolvers_amd as fa
    b = fa.BDFNewtonKrylov.for_euler(eu, maxl=maxl, precond=precond, fallback_cap=fallback_cap, jtimes=jtimes,
                                     fused=fused)
    b.set_newton(tol=1e-8, max_iter=20, lin_tol=1e-3, max_restarts=10)
    b.set_dt(dt)
    y = y0.clone()
    b.set_iv(0.0, y)
    torch.cuda.synchronize()
    times, gcv = [], []
    for k in range(nsteps):
        t0 = time.perf_counter()
        rc, _t = b.step(0.0, y)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        if rc != 0:
            raise RuntimeError(f"BDF step failed (rc={rc}) with precond={precond}")
        gcv.append(b.stats()["lastResidualNormCvode"])
        if k == 0 and first is not None:
            first.append(y.cpu().numpy())
        if collect is not None:
            yp, r, d = torch.zeros_like(y), torch.zeros_like(y), torch.zeros_like(y)
            b.last_triple(yp, r, d)
            collect(yp, r, d)
    st = b.stats()
    st["gcv_max"] = float(max(gcv))
    b.close()
    return y, np.array(times), st


def bdf_run_adaptive(eu, y0, t_end, precond, max_steps=2000):
    """The adaptive (CVODE-method) integrator as foeuler runs it -- rtol 1e-6, atol 1e-12, one
    internal step at a time until t >= t_end -- then the state at exactly t_end from the Nordsieck
    interpolant.  Returns (y(t_end), per-step wall times, stats, orders used)."""
    import torch
    import fosolvers_amd as fa
    b = fa.BDFNewtonKrylov.for_euler(eu, maxl=30, precond=precond)
    b.set_tol(1e-6, 1e-12)
    y = y0.clone()
    b.set_iv(0.0, y)
    torch.cuda.synchronize()
    times, orders, t = [], [], 0.0
    while t < t_end and len(times) < max_steps:
        t0 = time.perf_counter()
        rc, t = b.step(t_end, y)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        if rc != 0:
            raise RuntimeError(f"adaptive BDF step failed (rc={rc}) with precond={precond}")
        orders.append(b.stats()["lastOrder"])
    if t < t_end:
        raise RuntimeError(f"adaptive BDF reached only t = {t:g} < {t_end:g} in {max_steps} steps")
    yT = torch.zeros_like(y)
    b.get_dky(t_end, 0, yT)
    st = b.stats()
    b.close()
    return yT, np.array(times), st, orders


def conds_for(mach):
    """Supersonic inflow (full ghost state) at x = 0 and outflow at x = L for Mach > 1, far field at
    both ends below; walls elsewhere (boundary groups 1, 2 of the box meshes)."""
    import fosolvers_amd as fa
    if mach > 1:
        return fa.CondTab([1, 2], [10, 20], np.array([[1 / GAMMA, 1 / GAMMA, mach, 0, 0], [1 / GAMMA, 0, 0, 0, 0]]))
    free = [1 / GAMMA, 1 / GAMMA, mach, 0, 0]
    return fa.CondTab([1, 2], [30, 30], np.array([free, free]))


def training_data(fv, lengths, h, args):
    """Training data: FP64 block-ILU BDF runs on the given mesh at seeded Mach numbers in U[0.3, 3]
    (sinusoidal p perturbation, 2% below Mach 1, 10% above) and seeded acoustic CFL numbers
    log-uniform in [0.02, 1] (dt = CFL h / (M + 1)), args.train_steps_per_mach steps each; every
    args.keep_every-th step's Newton triple is kept with the FP64 context that made it (its ILU forms
    the ILU pairs at that state)."""
    import torch
    import fosolvers_amd as fa
    rng = np.random.default_rng(SEED)
    machs = rng.uniform(0.3, 3.0, args.train_machs)
    cfls = np.exp(rng.uniform(np.log(0.02), np.log(1.0), args.train_machs))
    out, ctxs, nsteps = [], [], 0
    for m, cfl in zip(machs, cfls):
        dt = cfl * h / (m + 1.0)
        x = solve_state(fv, *lengths, mach=m, amp=0.02 if m < 1 else 0.1, seed=int(1000 * m))
        e64 = fa.Euler(fv, conds_for(m), r=1.0, gamma=GAMMA, flags=fa.JAC_EXACT if args.train_exact_jac else 0)
        kept, k = [], [0]

        def keep(yp, r, d):
            if k[0] % args.keep_every == args.keep_every - 1:
                kept.append((yp.clone(), r.clone(), d.clone()))
            k[0] += 1
        y0 = torch.tensor(x, dtype=torch.float64, device="cuda")
        _y, _t, st = bdf_run(e64, y0, dt, args.train_steps_per_mach, "ilu", collect=keep)
        nsteps += args.train_steps_per_mach
        out += [(yp, r, d, 2 * dt / 3, e64) for yp, r, d in kept]
        ctxs.append(e64)
        log(f"  training data: Mach {m:.2f}, CFL {cfl:.3f}, dt {dt:.2e}: {st['nli'] / args.train_steps_per_mach:.1f} "
            f"GMRES iterations/step, {len(kept)} triples kept")
    return out, ctxs, nsteps, [float(v) for v in machs]


def solve_bench(args, fvT, lengthsT, hT):
    """BDF step time to the 1e-8 Newton tolerance on the 6M-tet mesh (200x100x50 hexes x6): FP64
    flux/Jv with block ILU(0); mixed precision (FP32 flux, Jv and preconditioner apply) with block
    ILU(0), with the learned preconditioner (block ILU behind it as the fallback) and without a
    preconditioner; FP64 Newton residual, Gram-Schmidt and update in all.  The network is trained first, on data from the 1.5M-tet mesh fvT of spacing hT
    (training_data)."""
    import torch
    import fosolvers_amd as fa
    from fosolvers_amd.learned import Trainer
    # training (timed): data generation, then Adam on the library's kernels
    t0 = time.perf_counter()
    data, ctxs, nsteps, machs = training_data(fvT, lengthsT, hT, args)
    t_data = time.perf_counter() - t0
    eT = fa.Euler(fvT, conds_for(2.0), r=1.0, gamma=GAMMA, flags=fa.PREC_FP32)
    mix = tuple(int(v) for v in args.train_mix.split(","))
    tr = Trainer(eT, batch=65536, lr=args.train_lr, seed=1, per_state=4, mix=mix)
    for yp, r, d, g, ctx in data:
        tr.add_triple(yp, r, d, g, pair_ctx=ctx)
    eT.learned_init(seed=1)

    def held_out_loss():
        # the triple loss on fixed data: every 25th triple, its first 65,536 cells
        ls = [eT.learned_loss_grad(yp, r, d, g, 0, min(65536, fvT.nC))[0] for yp, r, d, g, _c in data[::25]]
        return float(np.mean(ls))
    loss0 = held_out_loss()
    kinds, losses = tr.train(args.train_steps)
    torch.cuda.synchronize()
    loss1 = held_out_loss()
    t_train = time.perf_counter() - t0 - t_data
    params = eT.learned_params()
    pl = losses[(kinds == 1) | (kinds == 2)]  # ILU-pair batches
    tl = losses[kinds == 0]  # Newton-triple batches

    def ends(a):
        return [float(a[:20].mean()), float(a[-20:].mean())] if len(a) else None
    log(f"training: {nsteps} FP64 ILU BDF steps on the {fvT.nC}-tet mesh at Mach {', '.join(f'{m:.2f}' for m in machs)} "
        f"({t_data:.1f}s), {len(data)} triples; {args.train_steps} Adam steps ({t_train:.1f}s); triple loss on fixed data "
        f"{loss0:.3e} -> {loss1:.3e}")
    del tr, eT, data, ctxs
    torch.cuda.empty_cache()

    nx, ny, nz = (int(v) for v in args.solve_dims.split(","))
    h = 1.0 / ny
    lx, ly, lz = nx * h, ny * h, nz * h
    t0 = time.time()
    fv = throughput_mesh(nx, ny, nz, lx, ly, lz)
    log(f"solve mesh {fv.nC} tets in {time.time() - t0:.1f}s")
    resid6 = residual_leg(fv, lx, ly, lz, args)
    out = {"residual_6M": resid6, "training": {
        "data": f"{nsteps} FP64 block-ILU BDF steps on the {fvT.nC}-tet mesh at Mach {[round(m, 3) for m in machs]} "
                f"(seeded U[0.3, 3]), every {args.keep_every}th Newton triple kept",
        "triples": nsteps // args.keep_every, "data_s": t_data, "adam_steps": args.train_steps, "train_s": t_train,
        "lr": args.train_lr, "mix_bjac_ilu_triples": list(mix),
        "triple_loss_fixed_data_before_after": [loss0, loss1],
        "ilu_pair_loss_first_last": ends(pl), "triple_loss_first_last": ends(tl)}}
    first1 = []  # the FP64 ILU path's state after the first step of regime C (for the CPU check)
    regimes = {"C": (2.0, 0.1, args.solve_dt, args.solve_steps)}
    for rg in [r for r in args.regimes.split(",") if r]:
        regimes[rg] = {"E1": (0.3, 0.02, args.e1_dt, args.e_steps), "E2": (3.0, 0.1, args.e2_dt, args.e_steps)}[rg]
    for rg, (mach, amp, dt, K) in regimes.items():
        y0 = torch.tensor(solve_state(fv, lx, ly, lz, mach=mach, amp=amp), dtype=torch.float64, device="cuda")
        rec = {"workload": f"{nx}x{ny}x{nz} hexes x6 = {fv.nC} tets, Mach {mach} + {int(100 * amp)}% sinusoidal p, "
                           f"{'supersonic in/out' if mach > 1 else 'far field'}, dt {dt}, {K} steps to "
                           "||G||_WRMS <= 1e-8"}
        ref = None
        for name, flags, pc, jt in (("ilu_fp64", 0, "ilu", True), ("ilu_mixed", fa.PREC_FP32, "ilu", True),
                                    ("learned_mixed", fa.PREC_FP32, "learned", True),
                                    ("none_mixed", fa.PREC_FP32, None, True),
                                    ("ilu_fp64_analytic_jv", 0, "ilu", "analytic"),
                                    ("ilu_mixed_analytic_jv", fa.PREC_FP32, "ilu", "analytic"),
                                    ("ilu_mixed_analytic_jv_exact_jac", fa.PREC_FP32 | fa.JAC_EXACT, "ilu",
                                     "analytic"),
                                    ("learned_mixed_analytic_jv", fa.PREC_FP32, "learned", "analytic"),
                                    ("none_mixed_analytic_jv", fa.PREC_FP32, None, "analytic")):
            eu = fa.Euler(fv, conds_for(mach), r=1.0, gamma=GAMMA, flags=flags)
            if pc == "learned":
                eu.learned_set_params(params)
            y, t, st = bdf_run(eu, y0, dt, K, pc, fallback_cap=args.fallback_cap if pc == "learned" else 0,
                               jtimes=jt, first=first1 if (rg == "C" and name == "ilu_fp64") else None)
            s_ = 4 if flags & fa.PREC_FP32 else 8
            jvb = jvp_bytes(fv, s_) if jt == "analytic" else jv_bytes(fv, s_)
            # algorithmic bytes of the callback kernels the steps ran (rhs FP64, Jv, the ILU apply)
            cb_bytes = st["nfe"] * residual_bytes(fv, 8) + st["njv"] * jvb + (
                st["nps"] * ilu_apply_bytes(fv, s_) if pc == "ilu" else 0)
            rec[f"bdf_step_ms_{name}"] = 1e3 * float(t[1:].mean())
            rec[f"bdf_first_step_ms_{name}"] = 1e3 * float(t[0])
            rec[f"bdf_step_ms_amortised_{name}"] = 1e3 * float(t.mean())
            rec[f"gmres_iters_per_step_{name}"] = st["nli"] / K
            rec[f"newton_iters_per_step_{name}"] = st["nni"] / K
            rec[f"max_newton_norm_cvode_weights_{name}"] = st["gcv_max"]
            rec[f"callback_hbm_frac_{name}"] = cb_bytes / float(t.sum()) / 1e9 / HBM_PEAK_GBS
            if pc == "learned":
                rec[f"ilu_fallbacks_{name}"] = st["nfallback"]
            if ref is None:
                ref = y
            else:
                rec[f"final_state_rel_err_{name}_vs_ilu_fp64"] = float((y - ref).norm() / ref.norm())
            log(f"{rg} {name}: {1e3 * t[1:].mean():.1f} ms/step after a first step of {1e3 * t[0]:.0f} ms "
                f"({1e3 * t.mean():.1f} amortised); {st['nli'] / K:.1f} GMRES its/step; ||G|| (CVODE weights) <= "
                f"{st['gcv_max']:.1e}; callback kernels {rec[f'callback_hbm_frac_{name}']:.2f} of HBM peak"
                + (f", {st['nfallback']} fallbacks" if pc == "learned" else "")
                + (f"; final state vs FP64 ILU {rec[f'final_state_rel_err_{name}_vs_ilu_fp64']:.2e}" if name != "ilu_fp64" else ""))
            del eu, y
            torch.cuda.empty_cache()
        # the adaptive mode over the same interval, FP64 block ILU
        eu = fa.Euler(fv, conds_for(mach), r=1.0, gamma=GAMMA)
        yA, tA, stA, qA = bdf_run_adaptive(eu, y0, K * dt, "ilu")
        rec["adaptive_ilu_fp64"] = {
            "t_end": K * dt, "steps": len(tA), "wall_ms": 1e3 * float(tA.sum()),
            "wall_ms_excluding_first_step": 1e3 * float(tA[1:].sum()),
            "orders": {str(q): int(np.sum(np.array(qA) == q)) for q in sorted(set(qA))},
            "gmres_iters": int(stA["nli"]), "newton_iters": int(stA["nni"]), "preconditioner_setups": int(stA["npe"]),
            "error_test_failures": int(stA["netf"]), "last_step": float(stA["lastStep"]),
            "state_at_t_end_rel_diff_vs_fixed_step_ilu_fp64": float((yA - ref).norm() / ref.norm())}
        log(f"{rg} adaptive ilu_fp64: {len(tA)} steps to t = {K * dt:g} in {1e3 * tA.sum():.0f} ms "
            f"({1e3 * tA[1:].sum():.0f} ms after the first), orders {rec['adaptive_ilu_fp64']['orders']}, "
            f"{stA['netf']} error-test failures; y(t_end) vs the fixed-step run "
            f"{rec['adaptive_ilu_fp64']['state_at_t_end_rel_diff_vs_fixed_step_ilu_fp64']:.2e}")
        del eu, yA
        torch.cuda.empty_cache()
        out["C" if rg == "C" else rg] = rec
    if args.cpu_bdf and not args.no_cpu:
        del y0
        torch.cuda.empty_cache()
        x6 = solve_state(fv, lx, ly, lz, mach=2.0, amp=0.1)
        cond2 = ([1, 2], [10, 20], np.array([[1 / GAMMA, 1 / GAMMA, 2.0, 0, 0], [1 / GAMMA, 0, 0, 0, 0]]))
        out["cpu_bdf"] = cpu_bdf_baseline(fv, x6, cond2, args.solve_dt, first1[0] if first1 else None,
                                          args.cpu_bdf_c_steps)
    return out


def residual_leg(fv, lx, ly, lz, args):
    """The residual, the difference-quotient Jv and the analytic Jv on the 6M-tet mesh of config C
    (HIP events on the launch stream, 100 launches after 10 warm-up), and the block ILU(0) apply
    (FP64 and FP32 factors, 30 applies after 3), with their HBM fractions."""
    import torch
    import fosolvers_amd as fa
    stream = torch.cuda.Stream()
    conds = fa.CondTab([1, 2], [10, 20], np.array([[1 / GAMMA, 1 / GAMMA, 2.0, 0, 0], [1 / GAMMA, 0, 0, 0, 0]]))
    x = solve_state(fv, lx, ly, lz, mach=2.0, amp=0.1)
    eu = fa.Euler(fv, conds, r=1.0, gamma=GAMMA, stream=stream.cuda_stream, flags=fa.ASYNC_CALLBACKS)
    eu32 = fa.Euler(fv, conds, r=1.0, gamma=GAMMA, stream=stream.cuda_stream, flags=fa.ASYNC_CALLBACKS | fa.PREC_FP32)
    with torch.cuda.stream(stream):
        y = torch.tensor(x, dtype=torch.float64, device="cuda")
        yd, fy, jv = torch.zeros_like(y), torch.zeros_like(y), torch.zeros_like(y)
        v = torch.tensor(np.random.default_rng(2).standard_normal(5 * fv.nC) * np.abs(x) * 1e-3, device="cuda")
        y32, yd32 = y.float(), torch.zeros(5 * fv.nC, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    eu.rhs(y, fy)
    n, w = 100, 10
    t64 = time_gpu(lambda: eu.rhs(y, yd), n, w, stream) / n
    t32 = time_gpu(lambda: eu.rhs_f32(y32, yd32), n, w, stream) / n
    tj = time_gpu(lambda: eu.jtimes(v, jv, y, fy), n, w, stream) / n
    ta64 = time_gpu(lambda: eu.jtimes_analytic(v, jv, y), n, w, stream) / n
    ta32 = time_gpu(lambda: eu32.jtimes_analytic(v, jv, y), n, w, stream) / n
    # block ILU(0) apply at BDF2's gamma for config C's dt, on the state's Jacobian
    gam = 2.0 * args.solve_dt / 3.0
    r = torch.tensor(np.random.default_rng(3).standard_normal(5 * fv.nC) * np.abs(x), device="cuda")
    z = torch.zeros_like(r)
    torch.cuda.synchronize()
    ti = {}
    for name, e in (("fp64", eu), ("fp32", eu32)):
        e.psetup(y, fy, gam)
        ti[name] = time_gpu(lambda: e.psolve(r, z, gam), 30, 3, stream) / 30
    if eu.nonfinite() or eu32.nonfinite():
        raise RuntimeError("non-finite values in the 6M residual leg")
    b64, b32 = residual_bytes(fv, 8), residual_bytes(fv, 4)
    out = {"workload": f"{fv.nC}-tet config C mesh (RCM + chunk renumbering), Mach 2 + 10% sinusoidal p, "
                       "supersonic in/out",
           "fp64_residual_us": 1e6 * t64, "fp64_residual_evals_per_s": 1 / t64, "fp64_algorithmic_bytes": b64,
           "fp64_hbm_frac": b64 / t64 / 1e9 / HBM_PEAK_GBS,
           "fp32_residual_us": 1e6 * t32, "fp32_residual_evals_per_s": 1 / t32,
           "fp32_hbm_frac": b32 / t32 / 1e9 / HBM_PEAK_GBS,
           "fp64_dq_jv_us": 1e6 * tj, "fp64_dq_jv_hbm_frac": jv_bytes(fv, 8) / tj / 1e9 / HBM_PEAK_GBS,
           "fp64_analytic_jv_us": 1e6 * ta64, "fp64_analytic_jv_hbm_frac": jvp_bytes(fv, 8) / ta64 / 1e9 / HBM_PEAK_GBS,
           "fp32_analytic_jv_us": 1e6 * ta32, "fp32_analytic_jv_hbm_frac": jvp_bytes(fv, 4) / ta32 / 1e9 / HBM_PEAK_GBS,
           "fp64_ilu_apply_us": 1e6 * ti["fp64"], "fp64_ilu_apply_algorithmic_bytes": ilu_apply_bytes(fv, 8),
           "fp64_ilu_apply_hbm_frac": ilu_apply_bytes(fv, 8) / ti["fp64"] / 1e9 / HBM_PEAK_GBS,
           "fp32_ilu_apply_us": 1e6 * ti["fp32"],
           "fp32_ilu_apply_hbm_frac": ilu_apply_bytes(fv, 4) / ti["fp32"] / 1e9 / HBM_PEAK_GBS}
    log(f"6M residual: FP64 {1e6 * t64:.1f} us ({out['fp64_hbm_frac']:.3f} of HBM peak), FP32 {1e6 * t32:.1f} us; "
        f"DQ Jv {1e6 * tj:.1f} us; analytic Jv FP64 {1e6 * ta64:.1f} us ({out['fp64_analytic_jv_hbm_frac']:.3f}), "
        f"FP32 {1e6 * ta32:.1f} us; block ILU apply FP64 {1e6 * ti['fp64']:.0f} us "
        f"({out['fp64_ilu_apply_hbm_frac']:.3f}), FP32 {1e6 * ti['fp32']:.0f} us ({out['fp32_ilu_apply_hbm_frac']:.3f})")
    del eu, eu32
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dims", default="100,50,50", help="hexes per axis (x6 tets)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solve", type=int, default=1, help="also time BDF steps on the 6M-tet mesh (0: skip)")
    ap.add_argument("--solve-dims", default="200,100,50")
    ap.add_argument("--solve-dt", type=float, default=5e-5)
    ap.add_argument("--solve-steps", type=int, default=50)
    ap.add_argument("--train-machs", type=int, default=8, help="training runs (seeded Mach numbers)")
    ap.add_argument("--train-steps-per-mach", type=int, default=25)
    ap.add_argument("--keep-every", type=int, default=1, help="keep every n-th Newton triple (1: all)")
    ap.add_argument("--train-steps", type=int, default=1500, help="Adam steps")
    ap.add_argument("--train-lr", type=float, default=1e-3)
    ap.add_argument("--train-mix", default="1,1,2", help="Adam steps per state block: block-Jacobi pairs, ILU pairs, "
                    "Newton triples (units of 4 steps)")
    ap.add_argument("--train-exact-jac", type=int, default=1,
                    help="1: the training runs and the ILU pairs use block ILU on the exact Jacobian "
                         "(0: on findEulerJac's)")
    ap.add_argument("--regimes", default="E1,E2", help="E1 (Mach 0.3, 2%%) and/or E2 (Mach 3, 10%%) besides C; "
                                                        "'' for C only")
    ap.add_argument("--cpu-bdf", type=int, default=1, help="time the CPU full solver (oracle FP64 BDF, 1 thread) on "
                                                           "config A and the first steps of config C")
    ap.add_argument("--cpu-bdf-c-steps", type=int, default=1, help="config C steps of the CPU full solver (0: none)")
    ap.add_argument("--e1-dt", type=float, default=5e-3)
    ap.add_argument("--e2-dt", type=float, default=2e-4)
    ap.add_argument("--e-steps", type=int, default=30)
    ap.add_argument("--fallback-cap", type=int, default=60)
    ap.add_argument("--layout", choices=["tet", "generic"], default="tet",
                    help="residual layout: face geometry from node coordinates (tet) or per-face arrays (generic)")
    ap.add_argument("--copy-mb", type=int, default=1024,
                    help="also run a device copy of this many MiB (calibrates rocprofv3 FETCH/WRITE_SIZE)")
    ap.add_argument("--residual-6m-only", action="store_true",
                    help="only the 6M-tet residual / Jv leg (for profiling it); prints its JSON record")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path runs on one "
              "GPU; sharding the mesh across GPUs is a design, not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    import fosolvers_amd as fa
    if args.residual_6m_only:
        nx, ny, nz = (int(v) for v in args.solve_dims.split(","))
        h = 1.0 / ny
        fv6 = throughput_mesh(nx, ny, nz, nx * h, ny * h, nz * h)
        if args.copy_mb:
            n = args.copy_mb * (1 << 20) // 8
            a, b = torch.ones(n, dtype=torch.float64, device="cuda"), torch.empty(n, dtype=torch.float64, device="cuda")
            tc = time_gpu(lambda: b.copy_(a), 10, 2, torch.cuda.current_stream())
            log(f"copy {args.copy_mb} MiB: {2 * n * 8 * 10 / tc / 1e9:.0f} GB/s (read+write)")
            del a, b
        print(json.dumps(residual_leg(fv6, nx * h, ny * h, nz * h, args)), flush=True)
        return
    nx, ny, nz = (int(v) for v in args.dims.split(","))
    lx, ly, lz = 2.0, 1.0, 1.0
    t0 = time.time()
    fv = throughput_mesh(nx, ny, nz, lx, ly, lz)
    log(f"mesh {fv.nC} tets, {fv.nP} pairs ({fv.nPb} boundary) built in {time.time() - t0:.1f}s")
    free = [1.0 / GAMMA, 1.0 / GAMMA, 0.8, 0.0, 0.0]  # p, T = p/(r rho), u: rho = 1, c = 1
    conds = ([1, 2], [30, 30], np.array([free, free]))  # far field at x = 0 and x = L, walls elsewhere
    ctab = fa.CondTab(*conds)
    x = throughput_state(fv)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    # the throughput loops run the callbacks without a host wait per call (ASYNC_CALLBACKS); the
    # kernels still flag non-finite outputs, and the record is checked after each loop
    lay = fa.LAYOUT_GENERIC if args.layout == "generic" else 0
    eu = fa.Euler(fv, ctab, r=1.0, gamma=GAMMA, stream=stream.cuda_stream, flags=lay | fa.ASYNC_CALLBACKS)
    layout = eu.layout
    kernel = "tet_kernel<double,double,0,4>" if layout == 1 else "residual_kernel<double,double,0,4>"
    log(f"residual layout: {'tetrahedral' if layout == 1 else 'generic'} ({kernel})")
    with torch.cuda.stream(stream):
        y = torch.tensor(x, dtype=torch.float64, device=dev)
        yd = torch.zeros_like(y)
        y32 = y.float()
        yd32 = torch.zeros_like(y32)
        rng = np.random.default_rng(1)
        v = torch.tensor(rng.standard_normal(5 * fv.nC) * np.abs(x) * 1e-3, dtype=torch.float64, device=dev)
        jv = torch.zeros_like(y)
        fy = torch.zeros_like(y)
        v32, jv32, fy32 = v.float(), torch.zeros_like(y32), torch.zeros_like(y32)
    torch.cuda.synchronize()
    copy_gbs = None
    if args.copy_mb:
        n = args.copy_mb * (1 << 20) // 8
        a, b = torch.ones(n, dtype=torch.float64, device=dev), torch.empty(n, dtype=torch.float64, device=dev)
        tc = time_gpu(lambda: b.copy_(a), 10, 2, torch.cuda.current_stream())
        copy_gbs = 2 * n * 8 * 10 / tc / 1e9
        log(f"copy {args.copy_mb} MiB: {copy_gbs:.0f} GB/s (read+write)")
        del a, b
    eu.rhs(y, fy)
    eu.rhs_f32(y32, fy32)
    sig32 = float(np.sqrt(np.finfo(np.float32).eps) * np.linalg.norm(x) / float(v.norm()))

    t64 = time_gpu(lambda: eu.rhs(y, yd), args.steps, args.warmup, stream)
    t32 = time_gpu(lambda: eu.rhs_f32(y32, yd32), args.steps, args.warmup, stream)
    tj64 = time_gpu(lambda: eu.jtimes(v, jv, y, fy), args.steps, args.warmup, stream)
    tj32 = time_gpu(lambda: eu.jtimes_f32(v32, jv32, y32, fy32, sig32), args.steps, args.warmup, stream)
    # the analytic product (dual numbers), FP64 and FP32 arithmetic
    eu32 = fa.Euler(fv, ctab, r=1.0, gamma=GAMMA, stream=stream.cuda_stream, flags=fa.PREC_FP32 | lay | fa.ASYNC_CALLBACKS)
    with torch.cuda.stream(stream):
        jva, jva32 = torch.zeros_like(y), torch.zeros_like(y)
    tja64 = time_gpu(lambda: eu.jtimes_analytic(v, jva, y), args.steps, args.warmup, stream)
    tja32 = time_gpu(lambda: eu32.jtimes_analytic(v, jva32, y), args.steps, args.warmup, stream)
    eu.jtimes(v, jv, y, fy)
    jdq32 = torch.zeros_like(y)
    eu32.jtimes(v, jdq32, y, fy)  # the mixed path's product: FP32 fluxes differenced per face
    torch.cuda.synchronize()
    if eu.nonfinite() or eu32.nonfinite():
        raise RuntimeError("non-finite values in the throughput loops")
    jva_vs_dq = float((jva - jv).norm() / jva.norm())
    jva32_err = float((jva32 - jva).norm() / jva.norm())
    del eu32
    K = args.steps
    ev64, ev32 = K / t64, K / t32
    b64, b32 = residual_bytes(fv, 8), residual_bytes(fv, 4)
    dur64 = t64 / K  # one kernel per residual evaluation
    achieved = b64 / dur64 / 1e9
    log(f"FP64 residual {dur64 * 1e6:.1f} us ({ev64:.0f}/s, {achieved:.0f} GB/s); FP32 {t32 / K * 1e6:.1f} us; "
        f"Jv64 {tj64 / K * 1e6:.1f} us; Jv32 {tj32 / K * 1e6:.1f} us; analytic Jv64 {tja64 / K * 1e6:.1f} us, "
        f"Jv32 {tja32 / K * 1e6:.1f} us (vs DQ {jva_vs_dq:.1e}, FP32 vs FP64 {jva32_err:.1e})")
    gpu_out = yd.cpu().numpy()
    err32 = float((yd32.double() - yd).norm() / yd.norm())  # FP32 residual against the FP64 one
    log(f"FP32 residual vs FP64: {err32:.2e} relative")

    cpu = None
    jv_err = None
    if not args.no_cpu:
        cpu, cpu_out, ref_rhs = cpu_baseline(fv, x, conds, args.cpu_seconds)
        err = float(np.linalg.norm(gpu_out - cpu_out) / np.linalg.norm(cpu_out))
        cpu["gpu_vs_cpu_rel_err"] = err
        log(f"CPU baseline {cpu['value']:.2f} evals/s on {cpu['cores']} threads ({cpu['kind']}); "
            f"GPU vs CPU rel err {err:.2e}")
        # the Jv products against CVODE's difference quotient on the FP64 reference residual
        jref = oracle_dq_jv(ref_rhs, x, cpu_out, v.cpu().numpy())
        nrm = np.linalg.norm(jref)
        jv_err = {"reference": "CVODE's difference quotient of the FP64 reference residual (sigma = 1/||v||_WRMS)",
                  "fp64_dq": float(np.linalg.norm(jv.cpu().numpy() - jref) / nrm),
                  "fp32_dq_mixed": float(np.linalg.norm(jdq32.cpu().numpy() - jref) / nrm),
                  "fp64_analytic": float(np.linalg.norm(jva.cpu().numpy() - jref) / nrm),
                  "fp32_analytic": float(np.linalg.norm(jva32.cpu().numpy() - jref) / nrm)}
        log("Jv vs the FP64 oracle DQ: " + ", ".join(f"{k} {v_:.2e}" for k, v_ in jv_err.items() if k != "reference"))
    del jdq32, jva, jva32

    solve = None
    if args.solve:
        del eu, y, yd, y32, yd32, v, jv, fy, v32, jv32, fy32
        torch.cuda.empty_cache()
        solve = solve_bench(args, fv, (lx, ly, lz), ly / ny)

    traffic = None
    if (nx, ny, nz) == (100, 50, 50):
        traffic = read_traffic("residual_fp64_1p5M_pmc.txt" if layout == 1 else "residual_generic_fp64_1p5M_pmc.txt")
    kernels = f"{kernel} after ghost_kernel<double,double,0>" if layout == 1 else kernel
    roof_1p5 = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernels,
                "algorithmic_bytes_per_launch": b64,
                "workload": f"{fv.nC}-tet residual throughput mesh (its ~220 MB working set fits the 256 MB MALL)"}
    roof = roof_1p5
    r6 = (solve or {}).get("residual_6M")
    if r6 and layout == 1:
        # HBM claims on the 6M-tet mesh, whose working set (~1 GB) does not fit the MALL
        a6 = r6["fp64_algorithmic_bytes"] / (r6["fp64_residual_us"] * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": a6, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a6 / HBM_PEAK_GBS,
                "traffic": read_traffic("residual_fp64_6M_pmc.txt"), "kernel": kernels,
                "algorithmic_bytes_per_launch": r6["fp64_algorithmic_bytes"], "workload": r6["workload"]}
    res = {
        "metric": METRIC,
        "value": ev64,
        "unit": "residual evals/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": dur64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"residual throughput mesh: {nx}x{ny}x{nz} hexes x6 = {fv.nC} tets on "
                        f"[0,{lx}]x[0,{ly}]x[0,{lz}], jitter 0.2h (seed {SEED}), RCM then chunk renumbering; "
                        "Mach 0.8 + 5% Gaussians at 8 seeded centres; far field at x=0,L, walls elsewhere",
            "nC": fv.nC, "nP": fv.nP, "nPb": fv.nPb, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": {
            "roofline_1p5M": roof_1p5,
            "fp32_residual_evals_per_s": ev32,
            "fp32_residual_rel_err_vs_fp64": err32,
            "fp32_residual_hbm_frac": b32 / (t32 / K) / 1e9 / HBM_PEAK_GBS,
            "fp64_jv_per_s": K / tj64,
            "fp64_jv_hbm_frac": jv_bytes(fv, 8) / (tj64 / K) / 1e9 / HBM_PEAK_GBS,
            "fp32_jv_per_s": K / tj32,
            "analytic_jv": {"fp64_per_s": K / tja64,
                            "fp64_hbm_frac": jvp_bytes(fv, 8) / (tja64 / K) / 1e9 / HBM_PEAK_GBS,
                            "fp32_per_s": K / tja32,
                            "fp32_hbm_frac": jvp_bytes(fv, 4) / (tja32 / K) / 1e9 / HBM_PEAK_GBS,
                            "fp64_rel_diff_vs_dq_jv": jva_vs_dq, "fp32_rel_err_vs_fp64": jva32_err},
            "jv_rel_err_vs_fp64_oracle": jv_err,
            "copy_gbs": copy_gbs,
            "grid_up_gpu_s": {str(k): v for k, v in sorted(GRID_UP_S.items())},
            "bdf_solve": solve,
        },
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

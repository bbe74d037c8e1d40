"""GPU tests of the learned preconditioner: BF16/MFMA forward + apply and the backward against the
FP64 oracle (oracle/learned.c), training on the solver's own (state, Newton residual, converged
correction) triples, what the trained network gains over no preconditioner, and the ILU fallback.

Tolerances: BF16 weights and activations carry ~0.4 % per rounding; network output, apply and
gradients are compared with the FP64 oracle (fed the same BF16-rounded weights and the kernels'
own inputs) at 3e-2 relative per parameter block.
"""
import numpy as np
import pytest

import fosolvers_amd as fa
from oracle import oracle as O
from tests import meshes as M

pytestmark = pytest.mark.gpu

NI, NH, NO = 20, 128, 30
BLOCKS = {"W0": (0, NI * NH), "b0": (NI * NH, NH)}
_o = NI * NH + NH
BLOCKS["W1"] = (_o, 2 * NH * NH); _o += 2 * NH * NH
BLOCKS["b1"] = (_o, NH); _o += NH
BLOCKS["W2"] = (_o, 2 * NH * NH); _o += 2 * NH * NH
BLOCKS["b2"] = (_o, NH); _o += NH
BLOCKS["W3"] = (_o, NH * NO); _o += NH * NO
BLOCKS["b3"] = (_o, NO)


def bf16(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


def _cells_nbr(og):
    return np.where((og.neib >= 0) & (og.neib < og.nC), og.neib, -1).astype(np.int32)


def _ctx(cuda, dims=(6, 4, 4), kind="mixed", seed=4):
    import torch
    pg, fv, og, bc, ctab = M.setup(*dims, kind=kind)
    x = M.perturbed_state(og.p[:og.nC], mach=0.7, seed=seed, width=0.2)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    return og, bc, eu, x, torch.tensor(x, dtype=torch.float64, device=cuda)


def _random_params(eu, seed):
    eu.learned_init(seed=seed)
    p = eu.learned_params()
    rng = np.random.default_rng(seed)
    a, n = BLOCKS["W3"]
    p[a:a + n] = rng.standard_normal(n).astype(np.float32) * 0.05
    a, n = BLOCKS["b3"]
    p[a:a + n] = rng.standard_normal(n).astype(np.float32) * 0.05
    for k in ("b0", "b1", "b2"):
        a, n = BLOCKS[k]
        p[a:a + n] = rng.standard_normal(n).astype(np.float32) * 0.1
    eu.learned_set_params(p)
    return p


def test_learned_apply_matches_oracle(cuda):
    import torch
    print("[case] learned forward+apply", flush=True)
    og, bc, eu, x, y = _ctx(cuda)
    p = _random_params(eu, 3)
    gamma = 0.02
    X, S = eu.learned_inputs(y, gamma)
    assert np.all(np.isfinite(X)) and np.all(S > 0)
    nbr = _cells_nbr(og)
    Oref = O.mlp_forward(X, nbr, bf16(p))
    r = np.random.default_rng(0).standard_normal(5 * og.nC) * np.repeat(S, og.nC)
    zref = O.mlp_apply(Oref, nbr, S, r)
    eu.psetup_learned(y, gamma)
    z = torch.zeros_like(y)
    eu.psolve_learned(torch.tensor(r, device=cuda), z)
    torch.cuda.synchronize()
    err = M.rel(z.cpu().numpy(), zref)
    print(f"learned apply: BF16/MFMA vs FP64 oracle relative error {err:.3e}")
    assert err <= 3e-2
    # it is a linear operator in r
    z2 = torch.zeros_like(y)
    eu.psolve_learned(torch.tensor(2 * r, device=cuda), z2)
    torch.cuda.synchronize()
    assert M.rel(z2.cpu().numpy(), 2 * z.cpu().numpy()) <= 1e-6


def test_learned_gradient_matches_oracle(cuda):
    import torch
    print("[case] learned gradient", flush=True)
    og, bc, eu, x, y = _ctx(cuda, seed=6)
    p = _random_params(eu, 5)
    gamma = 0.02
    X, S = eu.learned_inputs(y, gamma)
    nbr = _cells_nbr(og)
    rng = np.random.default_rng(1)
    nC = og.nC
    r = rng.standard_normal(5 * nC) * np.repeat(S, nC)
    delta = 0.5 * r + 0.1 * rng.standard_normal(5 * nC) * np.repeat(S, nC)
    loss_o, g_o = O.mlp_loss_grad(X, nbr, bf16(p), S, r, delta, np.ones(nC, np.int32))
    loss_g, g_g = eu.learned_loss_grad(y, torch.tensor(r, device=cuda), torch.tensor(delta, device=cuda), gamma, 0, nC)
    print(f"loss gpu {loss_g:.6e} oracle {loss_o:.6e}")
    assert abs(loss_g - loss_o) <= 2e-2 * loss_o
    for name, (a, n) in BLOCKS.items():
        e = M.rel(g_g[a:a + n].astype(np.float64), g_o[a:a + n])
        print(f"grad {name}: rel err {e:.3e}")
        assert e <= 3e-2, name


def _collect_triples(eu, y0, cuda, nsteps, dt):
    """FP64 BDF with block ILU; the (predictor, Newton residual, correction) triple of every step."""
    import torch
    b = fa.BDFNewtonKrylov.for_euler(eu, maxl=30)
    b.set_newton(tol=1e-8, max_iter=20, lin_tol=1e-3)
    b.set_dt(dt)
    y = y0.clone()
    b.set_iv(0.0, y)
    out = []
    for _ in range(nsteps):
        rc, t = b.step(0.0, y)
        assert rc == 0
        yp, r, d = torch.zeros_like(y), torch.zeros_like(y), torch.zeros_like(y)
        b.last_triple(yp, r, d)
        out.append((yp, r, d))
    return out, y


def _gmres_iters(eu, y0, precond, dt, nsteps, fallback_cap=0):
    b = fa.BDFNewtonKrylov.for_euler(eu, maxl=40, precond=precond, fallback_cap=fallback_cap)
    b.set_newton(tol=1e-8, max_iter=20, lin_tol=1e-3, max_restarts=10)
    b.set_dt(dt)
    y = y0.clone()
    b.set_iv(0.0, y)
    for _ in range(nsteps):
        rc, t = b.step(0.0, y)
        assert rc == 0
    return b.stats(), y


def test_training_reduces_loss_and_gmres_iterations(cuda):
    """Train on the FP64 ILU solver's own Newton triples plus ILU-solve pairs at its states
    (fosolvers_amd.learned.Trainer), then count GMRES iterations over the same 4 steps with no
    preconditioner, the trained network and block ILU: the network must beat no preconditioner."""
    import torch
    from fosolvers_amd.learned import Trainer
    print("[case] learned training", flush=True)
    og, bc, eu, x, y0 = _ctx(cuda, dims=(8, 5, 5), kind="supersonic", seed=2)
    nC = og.nC
    dt = 0.1
    gamma = 2 * dt / 3
    triples, y_end = _collect_triples(eu, y0, cuda, 12, dt)
    tr = Trainer(eu, batch=nC, lr=1e-3, seed=3)
    for yp, r, d in triples:
        tr.add_triple(yp, r, d, gamma)
    eu.learned_init(seed=7)
    kinds, losses = tr.train(600)
    # the ILU-pair losses fall
    pair = losses[(kinds == 1) | (kinds == 2)]
    first, last = pair[:20].mean(), pair[-20:].mean()
    print(f"ILU-pair training loss {first:.4e} -> {last:.4e}")
    assert last < first
    st_none, y_none = _gmres_iters(eu, y_end, None, dt, 4)
    st_l, y_l = _gmres_iters(eu, y_end, "learned", dt, 4)
    st_ilu, y_ilu = _gmres_iters(eu, y_end, "ilu", dt, 4)
    print(f"GMRES iterations over 4 steps: none {st_none['nli']}, learned {st_l['nli']}, ILU {st_ilu['nli']}")
    assert st_l["nli"] < st_none["nli"]
    assert M.rel(y_l.cpu().numpy(), y_none.cpu().numpy()) <= 1e-9
    assert M.rel(y_ilu.cpu().numpy(), y_none.cpu().numpy()) <= 1e-9


def test_learned_falls_back_to_ilu(cuda):
    print("[case] learned fallback", flush=True)
    og, bc, eu, x, y0 = _ctx(cuda, dims=(8, 5, 5), kind="supersonic", seed=3)
    eu.learned_init(seed=1)  # untrained: the identity operator
    st, y = _gmres_iters(eu, y0, "learned", 0.05, 2, fallback_cap=1)
    print("fallback stats", st)
    assert st["nfallback"] >= 1
    st2, y2 = _gmres_iters(eu, y0, "ilu", 0.05, 2)
    assert M.rel(y.cpu().numpy(), y2.cpu().numpy()) <= 1e-9

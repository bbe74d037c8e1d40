"""GPU parity of findEulerJac, the block ILU(0) preconditioner and the BDF-Newton-GMRES integrator
against the CPU oracle.

Tolerances: JacC within 1e-12 relative; the FP64 ILU apply within 1e-10 of the oracle's block
ILU(0) (same updates, different summation order); the integrated states within 1e-9 relative of
the oracle's after each step (both solve G(y) = 0 to ||G||_WRMS <= 1e-8, i.e. ~1e-14 relative).
"""
import ctypes

import numpy as np
import pytest

import fosolvers_amd as fa
from oracle import oracle as O
from tests import meshes as M

pytestmark = pytest.mark.gpu


def _dev(torch, a, cuda, dtype=None):
    return torch.tensor(np.asarray(a), dtype=dtype or torch.float64, device=cuda)


@pytest.mark.parametrize("kind", ["mixed", "wall", "supersonic"])
def test_find_euler_jac_matches_oracle(cuda, kind):
    import torch
    print(f"[case] jacobian {kind}", flush=True)
    pg, fv, og, bc, ctab = M.setup(8, 5, 5, kind=kind)
    x = M.perturbed_state(og.p[:og.nC], mach=0.6 if kind != "supersonic" else 2.0, seed=7)
    rho, rhou, rhoE, p = O.element_state(og, bc, x)
    ref = O.find_euler_jac(og, rho, rhou, rhoE, p, M.GAMMA)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    jac = torch.zeros(125 * og.nC, dtype=torch.float64, device=cuda)
    eu.find_euler_jac(_dev(torch, x, cuda), jac)
    torch.cuda.synchronize()
    got = jac.cpu().numpy().reshape(og.nC, 25, 5)
    assert M.rel(got, ref) <= 1e-12, M.rel(got, ref)


def _jac_matvec(og, jac, v):
    """sum over the cell's blocks of JacC[c, 5b+q, r] v_{nb(c,b)}[q] (JacC is not divided by v)."""
    nC = og.nC
    V = v.reshape(5, nC)
    out = np.einsum("cqr,qc->rc", jac[:, 0:5, :], V)
    for j in range(og.maxNF):
        nb = og.neib[:, j]
        ok = (nb >= 0) & (nb < nC)
        blk = jac[ok, 5 * (1 + j):5 * (2 + j), :]
        out[:, ok] += np.einsum("cqr,qc->rc", blk, V[:, nb[ok]])
    return out.reshape(-1)


@pytest.mark.parametrize("kind", ["wall", "far", "mixed", "supersonic"])
def test_exact_jacobian_is_the_residual_derivative(cuda, kind):
    """FOS_JAC_EXACT: JacC assembled as a matrix equals the residual's derivative -- its product
    with v is v-weighted analytic Jv (boundary faces through their ghost states included) to
    rounding -- while findEulerJac's blocks (frozen |A~|, no ghost dependence) do not."""
    import torch
    print(f"[case] exact jacobian {kind}", flush=True)
    pg, fv, og, bc, ctab = M.setup(6, 4, 3, kind=kind, jitter=0.2)
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6 if kind != "supersonic" else 2.0, seed=5)
    v = np.random.default_rng(9).standard_normal(5 * nC) * np.abs(x) * 0.1
    y, vd = _dev(torch, x, cuda), _dev(torch, v, cuda)
    ex = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=fa.JAC_EXACT)
    jac = torch.zeros(125 * nC, dtype=torch.float64, device=cuda)
    ex.find_euler_jac(y, jac)
    jv = torch.zeros_like(y)
    ex.jtimes_analytic(vd, jv, y)
    torch.cuda.synchronize()
    J = jac.cpu().numpy().reshape(nC, 25, 5)
    want = jv.cpu().numpy() * np.tile(og.v, 5)
    err = M.rel(_jac_matvec(og, J, v), want)
    print(f"exact JacC v vs volume-weighted analytic Jv: {err:.2e}")
    assert err <= 1e-12
    ap = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    ap.find_euler_jac(y, jac)
    torch.cuda.synchronize()
    assert M.rel(_jac_matvec(og, jac.cpu().numpy().reshape(nC, 25, 5), v), want) > 1e-4


def test_bdf_with_exact_jacobian_preconditioner(cuda):
    """Block ILU on the exact Jacobian: the same converged states as with findEulerJac's."""
    import torch
    print("[case] bdf exact jacobian", flush=True)
    pg, fv, og, bc, ctab = M.setup(12, 6, 6, kind="mixed")
    x0 = M.perturbed_state(og.p[:og.nC], mach=0.6, amp=0.1, seed=2)
    outs, its = [], []
    for flags in (0, fa.JAC_EXACT):
        eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
        b = fa.BDFNewtonKrylov.for_euler(eu, maxl=20, jtimes="analytic")
        b.set_newton(tol=1e-8, max_iter=20, lin_tol=1e-3)
        b.set_dt(1e-3)
        y = _dev(torch, x0, cuda)
        b.set_iv(0.0, y)
        for k in range(4):
            rc, t = b.step(0.0, y)
            assert rc == 0, (flags, k, b.stats())
        outs.append(y.cpu().numpy())
        its.append(b.stats()["nli"])
    print("GMRES iterations: findEulerJac", its[0], "exact", its[1])
    assert M.rel(outs[1], outs[0]) <= 1e-8


@pytest.mark.parametrize("flags,tol", [(0, 1e-10), (fa.PREC_FP32, 1e-5)])
def test_block_ilu_apply_matches_oracle(cuda, flags, tol):
    import torch
    print(f"[case] block ilu flags={flags}", flush=True)
    pg, fv, og, bc, ctab = M.setup(10, 6, 6, kind="mixed")
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=8)
    rho, rhou, rhoE, p = O.element_state(og, bc, x)
    jac = O.find_euler_jac(og, rho, rhou, rhoE, p, M.GAMMA)
    gamma = 2e-3 * np.cbrt(og.v.mean())
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
    order = eu.ilu_order()  # the library's elimination order, used by the oracle's ILU as well
    b = O.Bilu(og, jac, gamma, order=order)
    r = np.random.default_rng(3).standard_normal(5 * nC)
    ref = b.solve(r)
    y = _dev(torch, x, cuda)
    jcur = eu.psetup(y, y, gamma, jok=0)
    assert jcur == 1
    z = torch.zeros_like(y)
    eu.psolve(_dev(torch, r, cuda), z, gamma)
    torch.cuda.synchronize()
    err = M.rel(z.cpu().numpy(), ref)
    print(f"ILU apply relative error vs FP64 oracle: {err:.3e}")
    assert err <= tol
    # jok = 1 reuses the Jacobian, refactors with the new gamma
    assert eu.psetup(y, y, 2 * gamma, jok=1) == 0
    eu.psolve(_dev(torch, r, cuda), z, 2 * gamma)
    torch.cuda.synchronize()
    ref2 = O.Bilu(og, jac, 2 * gamma, order=order).solve(r)
    assert M.rel(z.cpu().numpy(), ref2) <= tol


@pytest.mark.parametrize("flags,tol", [(0, 1e-10), (fa.PREC_FP32, 1e-5)])
def test_block_ilu_apply_shuffled_cells_matches_oracle(cuda, flags, tol):
    """Cells numbered in a seeded random order: a 64-cell chunk then has neighbours all over the mesh,
    up to 4 per row outside the chunk (more than 191 in some chunks), which the sweep records must
    still address."""
    import torch
    print(f"[case] block ilu shuffled cells flags={flags}", flush=True)
    pg = M.box(16, 8, 8, 1.0, 0.5, 0.5, jitter=0.2, seed=3)
    vol = np.flatnonzero(np.isin(pg.sE, (fa.TET, fa.HEX, fa.TET10)))
    src = vol[np.random.default_rng(21).permutation(len(vol))]
    for a in ("sE", "nNE", "iNE", "gid"):
        arr = getattr(pg, a)
        arr[vol] = arr[src].copy()
    fv = pg.up()
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    cg, ct, cp = M.conds("mixed")
    bc = O.BC(og, cg, ct, cp, r=M.R, gamm=M.GAMMA)
    ctab = fa.CondTab(cg, ct, cp)
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=8)
    rho, rhou, rhoE, p = O.element_state(og, bc, x)
    jac = O.find_euler_jac(og, rho, rhou, rhoE, p, M.GAMMA)
    gamma = 2e-3 * np.cbrt(og.v.mean())
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
    order = eu.ilu_order()
    # the most neighbours outside one chunk on either sweep side
    rank = np.argsort(order)
    nb = np.where((og.neib >= 0) & (og.neib < nC), og.neib, -1)
    most = 0
    for c0 in range(0, nC, 64):
        rows = np.arange(c0, min(c0 + 64, nC))
        k = nb[rows].ravel()
        own = np.repeat(rows, nb.shape[1])
        out = (k >= 0) & ((k < c0) | (k >= c0 + 64))
        lower = np.zeros_like(out)
        lower[out] = rank[k[out]] < rank[own[out]]
        most = max(most, len(np.unique(k[out & lower])), len(np.unique(k[out & ~lower])))
    print(f"most outside neighbours of a chunk on one side: {most}")
    assert most > 191
    ref = O.Bilu(og, jac, gamma, order=order).solve(r := np.random.default_rng(3).standard_normal(5 * nC))
    y = _dev(torch, x, cuda)
    assert eu.psetup(y, y, gamma, jok=0) == 1
    z = torch.zeros_like(y)
    eu.psolve(_dev(torch, r, cuda), z, gamma)
    torch.cuda.synchronize()
    err = M.rel(z.cpu().numpy(), ref)
    print(f"ILU apply (shuffled cells) relative error vs FP64 oracle: {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("nx", [48, 2])
@pytest.mark.parametrize("flags,tol", [(0, 1e-10), (fa.PREC_FP32, 1e-5)])
def test_block_ilu_apply_strip_mesh_matches_oracle(cuda, flags, tol, nx):
    """A strip one hex wide: each 64-cell chunk is a run of ~11 hexes whose rows wait on one another
    in long chains, so a chunk's sweep takes many more steps than on a box.  nx = 2: one partial
    chunk of 12 cells, whose block runs end at the end of the packed arrays."""
    import torch
    print(f"[case] block ilu strip mesh nx={nx} flags={flags}", flush=True)
    pg, fv, og, bc, ctab = M.setup(nx, 1, 1, kind="mixed", lx=nx / 48, ly=1.0 / 48, lz=1.0 / 48)
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=8)
    rho, rhou, rhoE, p = O.element_state(og, bc, x)
    jac = O.find_euler_jac(og, rho, rhou, rhoE, p, M.GAMMA)
    gamma = 5e-2 * np.cbrt(og.v.mean())
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
    order = eu.ilu_order()
    r = np.random.default_rng(5).standard_normal(5 * nC)
    ref = O.Bilu(og, jac, gamma, order=order).solve(r)
    y = _dev(torch, x, cuda)
    assert eu.psetup(y, y, gamma, jok=0) == 1
    z = torch.zeros_like(y)
    eu.psolve(_dev(torch, r, cuda), z, gamma)
    torch.cuda.synchronize()
    err = M.rel(z.cpu().numpy(), ref)
    print(f"ILU apply (strip mesh, {nC} cells) relative error vs FP64 oracle: {err:.3e}")
    assert err <= tol


def test_block_ilu_elimination_order_is_chunk_colouring(cuda):
    """The elimination order is a permutation in which each 64-cell chunk is contiguous, with its
    cells in index order (what the per-colour sweeps assume)."""
    print("[case] ilu elimination order", flush=True)
    pg, fv, og, bc, ctab = M.setup(24, 12, 12, kind="far")
    nC = og.nC
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    order = eu.ilu_order()
    assert np.array_equal(np.sort(order), np.arange(nC))
    rank = np.argsort(order)
    ch = np.arange(nC) // 64
    for c in range(ch[-1] + 1):
        rr = rank[ch == c]
        assert np.array_equal(rr, np.arange(rr[0], rr[0] + len(rr)))  # contiguous, index order
    assert nC > 3 * 64


@pytest.mark.parametrize("flags", [0, fa.PREC_FP32])
def test_block_ilu_graph_replays_match_eager_launches(cuda, flags, monkeypatch):
    """The factorisation and sweep levels replay as graphs; each replay (three solves per
    factorisation, two factorisations, eager work and synchronisations between them) must equal the
    level-by-level launches bit for bit."""
    import torch
    print(f"[case] block ilu graph vs eager flags={flags}", flush=True)
    pg, fv, og, bc, ctab = M.setup(10, 6, 6, kind="mixed")
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=8)
    y = _dev(torch, x, cuda)
    gamma = 2e-3 * np.cbrt(og.v.mean())
    monkeypatch.setenv("FOS_ILU_EAGER", "1")
    eager = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
    assert eager.psetup(y, y, gamma, jok=0) == 1
    monkeypatch.delenv("FOS_ILU_EAGER")
    graph = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
    assert graph.psetup(y, y, gamma, jok=0) == 1
    rng = np.random.default_rng(11)
    for g in (gamma, 3 * gamma):
        if g != gamma:
            assert eager.psetup(y, y, g, jok=1) == 0
            assert graph.psetup(y, y, g, jok=1) == 0
        for k in range(3):
            r = _dev(torch, rng.standard_normal(5 * nC), cuda)
            ze, zg = torch.zeros_like(r), torch.zeros_like(r)
            eager.psolve(r, ze, g)
            torch.cuda.synchronize()
            graph.psolve(r, zg, g)
            torch.cuda.synchronize()
            assert torch.equal(ze, zg), (g, k, float((ze - zg).abs().max()))
            assert float(zg.norm()) > 0


@pytest.mark.parametrize("flags,tol", [(0, 1e-12), (fa.PREC_FP32, 1e-5)])
def test_block_jacobi_apply_matches_dense_solve(cuda, flags, tol):
    """z_c = M_cc^-1 r_c with M = I - gamma JacC/v, checked against numpy's dense solve of the oracle's
    diagonal blocks."""
    import torch
    print(f"[case] block jacobi flags={flags}", flush=True)
    pg, fv, og, bc, ctab = M.setup(10, 6, 6, kind="mixed")
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=8)
    rho, rhou, rhoE, p = O.element_state(og, bc, x)
    jac = O.find_euler_jac(og, rho, rhou, rhoE, p, M.GAMMA)  # [c, col, row]
    gamma = 5e-2 * np.cbrt(og.v.mean())
    Mcc = np.eye(5)[None] - gamma / og.v[:, None, None] * np.transpose(jac[:, 0:5, :], (0, 2, 1))
    r = np.random.default_rng(4).standard_normal(5 * nC)
    ref = np.linalg.solve(Mcc, r.reshape(5, nC).T[:, :, None])[:, :, 0].T.ravel()
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
    y = _dev(torch, x, cuda)
    jcur = ctypes.c_int(0)
    fa.check(fa.lib().fos_euler_psetup_bjac(0.0, fa._ptr(y), fa._ptr(y), 0, ctypes.byref(jcur), gamma, eu.h))
    assert jcur.value == 1
    z = torch.zeros_like(y)
    rd = _dev(torch, r, cuda)
    fa.check(fa.lib().fos_euler_psolve_bjac(0.0, None, None, fa._ptr(rd), fa._ptr(z), gamma, 0.0, 1, eu.h))
    torch.cuda.synchronize()
    err = M.rel(z.cpu().numpy(), ref)
    print(f"block Jacobi apply relative error vs dense solve: {err:.3e}")
    assert err <= tol


def test_bdf_harmonic_oscillator_kat(cuda):
    """src/test/BDFNewtonKrylov1.f90 on the device integrator with a host rhs (no preconditioner)."""
    print("[case] bdf oscillator", flush=True)

    def f(t, y, yd):
        yd[0] = y[1]
        yd[1] = -y[0]
    n = 4000
    b = fa.BDFNewtonKrylov(2, f, maxl=0)
    b.set_tol(1e-8, 1e-12)
    b.set_max_steps(-1)
    b.set_newton(tol=1e-3)
    b.set_dt(2.0 / n)
    s = np.array([1.0, 0.0])
    b.set_iv(0.0, s)
    t = 0.0
    while t < 2.0 - 1e-12:
        rc, t = b.step(2.0, s)
        assert rc == 0
    assert np.linalg.norm(s - [np.cos(t), -np.sin(t)]) <= 1e-6


def test_bdf_adaptive_oscillator_kat_matches_oracle(cuda):
    """src/test/BDFNewtonKrylov1.f90 as the reference runs it: adaptive order and step (the
    default mode), rtol 1e-8, atol 1e-12, step(2.0) until t >= 2, |s - (cos t, -sin t)| <= 1e-6;
    every step's time, size and order equal the oracle's adaptive restatement."""
    print("[case] bdf adaptive oscillator", flush=True)

    def f(t, y, yd):
        yd[0] = y[1]
        yd[1] = -y[0]
    b = fa.BDFNewtonKrylov(2, f, maxl=0)
    b.set_tol(1e-8, 1e-12)
    b.set_max_steps(-1)
    s = np.array([1.0, 0.0])
    b.set_iv(0.0, s)
    t, ts, hs, qs = 0.0, [], [], []
    while t < 2.0:
        rc, t = b.step(2.0, s)
        assert rc == 0
        st = b.stats()
        ts.append(t)
        hs.append(b.get_dt())
        qs.append(st["lastOrder"])
    assert np.linalg.norm(s - [np.cos(t), -np.sin(t)]) <= 1e-6
    # CVodeGetDky: the interpolant at the step's end is the state; inside the last step it follows
    # the solution and its derivative to the integration tolerance
    d = np.zeros(2)
    b.get_dky(t, 0, d)
    assert np.array_equal(d, s)
    b.get_dky(2.0, 0, d)
    assert np.linalg.norm(d - [np.cos(2.0), -np.sin(2.0)]) <= 1e-6
    b.get_dky(2.0, 1, d)
    assert np.linalg.norm(d - [-np.sin(2.0), -np.cos(2.0)]) <= 1e-5
    with pytest.raises(RuntimeError, match="outside"):
        b.get_dky(t - 2 * hs[-1], 0, d)
    rts, rhs, rqs, rst, rstats = O.cv_run(lambda tt, y: np.array([y[1], -y[0]]), [1.0, 0.0], 2.0, 2.0, 5000,
                                          rtol=1e-8, atol=1e-12, maxl=5, restarts=5)
    print(f"gpu steps {len(ts)}, oracle steps {len(rts)}, max order {max(qs)}", b.stats())
    # the same decisions step by step; the sizes differ only through the difference-quotient Jv:
    # at sigma = 1/||v||_WRMS (~1e-8 here) it keeps ~8 digits, the two sides take sigma from
    # different weights (the step's here, the iterate's in the oracle, as the fused Euler jtimes
    # does), and GMRES stops at its tolerance rather than at the exact solve, so the error
    # estimates, hence h, differ at the 1e-8 level (measured: t 2.0e-8, h 2.9e-8, state 8.0e-9)
    dt_rel = np.max(np.abs(np.array(ts) - rts) / rts)
    dh_rel = np.max(np.abs(np.array(hs) - rhs) / rhs)
    print(f"max relative difference: t {dt_rel:.2e}, h {dh_rel:.2e}, final state {np.linalg.norm(s - rst[-1]):.2e}")
    assert len(ts) == len(rts)
    assert list(qs) == list(rqs)
    assert dt_rel <= 1e-6 and dh_rel <= 1e-6
    assert np.linalg.norm(s - rst[-1]) <= 1e-7


def test_bdf_mode_switch_continues_from_current_state(cuda):
    """Adaptive -> fixed -> adaptive on the oscillator: each switch continues from the state reached
    (not from the initial value), so the solution stays on (cos t, -sin t) to the tolerance."""
    print("[case] bdf mode switch", flush=True)

    def f(t, y, yd):
        yd[0] = y[1]
        yd[1] = -y[0]
    b = fa.BDFNewtonKrylov(2, f, maxl=0)
    b.set_tol(1e-8, 1e-12)
    s = np.array([1.0, 0.0])
    b.set_iv(0.0, s)
    t = 0.0
    while t < 0.5:
        rc, t = b.step(2.0, s)
        assert rc == 0
    b.set_newton(tol=1e-10)
    b.set_dt(1e-4)
    for _ in range(3):
        rc, t2 = b.step(2.0, s)
        assert rc == 0
    assert abs(t2 - (t + 3e-4)) <= 1e-12
    assert np.linalg.norm(s - [np.cos(t2), -np.sin(t2)]) <= 1e-6
    b.set_adaptive()
    for _ in range(5):
        rc, t3 = b.step(2.0, s)
        assert rc == 0
    assert t3 > t2
    assert np.linalg.norm(s - [np.cos(t3), -np.sin(t3)]) <= 1e-6


def test_bdf_adaptive_sod_matches_oracle(cuda):
    """Adaptive BDF with block ILU(0) on the all-wall Sod tube (1,080 tets): the GPU's steps
    (time, size, order) and states follow the oracle's adaptive restatement with the same
    elimination order, and mass is conserved."""
    import torch
    print("[case] bdf adaptive sod", flush=True)
    pg, fv, og, bc, ctab = M.setup(10, 3, 3, kind="wall", jitter=0.0, renumber=None, lx=1.0, ly=0.25, lz=0.25)
    x0 = M.sod_state(og.p[:og.nC])
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    b = fa.BDFNewtonKrylov.for_euler(eu, maxl=20)
    b.set_tol(1e-6, 1e-12)
    y = _dev(torch, x0, cuda)
    b.set_iv(0.0, y)
    t_end = 1e-2
    ts, hs, qs, states = [], [], [], []
    t = 0.0
    while t < t_end and len(ts) < 400:
        rc, t = b.step(t_end, y)
        assert rc == 0
        ts.append(t)
        hs.append(b.get_dt())
        qs.append(b.stats()["lastOrder"])
        states.append(y.cpu().numpy())
    st = b.stats()
    rts, rhs, rqs, rst, rstats = O.cv_euler(og, bc, x0, t_end, t_end, 400, rtol=1e-6, atol=1e-12, maxl=20,
                                            order=eu.ilu_order())
    print(f"gpu {len(ts)} steps (orders {np.bincount(qs)}), oracle {len(rts)} (orders {np.bincount(rqs)})")
    print("gpu stats", st, "oracle stats", rstats)
    assert t >= t_end and ts[-1] == t
    nC = og.nC
    m0 = (x0[:nC] * og.v).sum()
    assert abs((states[-1][:nC] * og.v).sum() - m0) <= 1e-12 * m0 * len(ts)
    k = min(len(ts), len(rts), 40)
    assert list(qs[:k]) == list(rqs[:k])
    assert np.allclose(ts[:k], rts[:k], rtol=1e-6, atol=0)
    for i in range(k):
        assert M.rel(states[i], rst[i]) <= 1e-6, (i, M.rel(states[i], rst[i]))
    assert max(qs) >= 2 and st["npe"] < st["nst"]


@pytest.mark.parametrize("dims,dt,nsteps", [((10, 3, 3), 1e-3, 5), ((40, 10, 10), 1e-4, 20)])
def test_bdf_sod_matches_oracle_every_step(cuda, dims, dt, nsteps):
    """Sod tube, all walls, FP64 BDF-Newton-GMRES with block ILU; GPU state vs oracle every step.
    The (40,10,10) case is the 24,000-tet shock tube at dt = 1e-4 for 20 steps."""
    import torch
    print(f"[case] bdf sod dims={dims} steps={nsteps}", flush=True)
    pg, fv, og, bc, ctab = M.setup(*dims, kind="wall", jitter=0.0, renumber=None, lx=1.0, ly=0.25, lz=0.25)
    x0 = M.sod_state(og.p[:og.nC])
    ref, refit = O.bdf_euler(og, bc, x0, dt, nsteps, maxl=20, lin_tol=1e-2, ilu=True)
    assert len(ref) == nsteps
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    b = fa.BDFNewtonKrylov.for_euler(eu, maxl=20)
    b.set_tol(1e-6, 1e-12)
    b.set_newton(tol=1e-8, max_iter=20, lin_tol=1e-2)
    b.set_dt(dt)
    y = _dev(torch, x0, cuda)
    b.set_iv(0.0, y)
    nC = og.nC
    m0 = (x0[:nC] * og.v).sum()
    for k in range(nsteps):
        rc, t = b.step(dt * (k + 1), y)
        assert rc == 0
        got = y.cpu().numpy()
        assert M.rel(got, ref[k]) <= 1e-9, (k, M.rel(got, ref[k]))
        assert abs((got[:nC] * og.v).sum() - m0) <= 1e-12 * m0 * (k + 1) + 1e-9 * m0
    st = b.stats()
    print("gpu stats", st, "oracle iters", refit.sum(0))


def test_bdf_mixed_precision_converges_to_fp64(cuda):
    """FP32 flux/Jv/ILU apply with FP64 Newton residual: same converged states within 1e-6."""
    import torch
    print("[case] bdf mixed precision", flush=True)
    pg, fv, og, bc, ctab = M.setup(12, 6, 6, kind="supersonic")
    nC = og.nC
    x0 = M.perturbed_state(og.p[:nC], mach=2.0, amp=0.1, width=0.15, seed=2)
    outs = []
    for flags in (0, fa.PREC_FP32):
        eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
        b = fa.BDFNewtonKrylov.for_euler(eu, maxl=20)
        b.set_newton(tol=1e-8, max_iter=30, lin_tol=1e-2)
        b.set_dt(5e-4)
        y = _dev(torch, x0, cuda)
        b.set_iv(0.0, y)
        for k in range(5):
            rc, t = b.step(0.0, y)
            assert rc == 0
        outs.append(y.cpu().numpy())
    err = M.rel(outs[1], outs[0])
    print(f"mixed vs FP64 final-state relative error: {err:.3e}")
    assert err <= 1e-6


@pytest.mark.parametrize("precond", ["ilu", None])
def test_bdf_fp32_krylov_basis_matches_fp64_basis(cuda, precond):
    """The FP32 GMRES basis (FOS_PREC_FP32 on the integrator) against the FP64 basis, both on an FP32
    context: the basis's rounding (~1e-7 relative) only perturbs each Newton correction inside
    GMRES's 1e-3 relative tolerance, and the Newton iterations converge to the same 1e-8 WRMS
    tolerance: final states within 1e-9 relative (the Newton tolerance times rtol), GMRES iteration
    counts within one per solve."""
    import torch
    print(f"[case] bdf fp32 krylov basis precond={precond}", flush=True)
    pg, fv, og, bc, ctab = M.setup(12, 6, 6, kind="supersonic")
    x0 = M.perturbed_state(og.p[:og.nC], mach=2.0, amp=0.1, width=0.15, seed=4)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=fa.PREC_FP32)
    res = {}
    for k32 in (False, True):
        b = fa.BDFNewtonKrylov.for_euler(eu, maxl=30, precond=precond, jtimes="analytic", krylov_fp32=k32)
        b.set_newton(tol=1e-8, max_iter=30, lin_tol=1e-3, max_restarts=10)
        b.set_dt(5e-4)
        y = _dev(torch, x0, cuda)
        b.set_iv(0.0, y)
        for _ in range(5):
            assert b.step(0.0, y)[0] == 0
        res[k32] = (y.cpu().numpy(), b.stats())
        b.close()
    (y0, s0), (y1, s1) = res[False], res[True]
    err = M.rel(y1, y0)
    print(f"  FP32 vs FP64 basis {err:.2e}; GMRES its {s1['nli']} vs {s0['nli']}, Newton {s1['nni']} vs "
          f"{s0['nni']}, second passes {s1['nreorth']} vs {s0['nreorth']}")
    assert err <= 1e-9
    assert abs(s1["nli"] - s0["nli"]) <= s0["nni"]


@pytest.mark.parametrize("flags", [0, fa.PREC_FP32], ids=["fp64", "mixed"])
def test_bdf_long_krylov_cycles(cuda, flags):
    """Unpreconditioned GMRES at a large step, where a solve needs more than 32 Arnoldi steps: with
    maxl 40 the basis kernels take their any-count form (more than 32 vectors) and the inner products
    two launches; with maxl 8 the same solves restart.  Both converge Newton to the 1e-8 WRMS
    tolerance, so the states agree to it (1e-9 relative after 3 steps)."""
    import torch
    print(f"[case] bdf long krylov cycles flags={flags}", flush=True)
    pg, fv, og, bc, ctab = M.setup(12, 6, 6, kind="supersonic")
    x0 = M.perturbed_state(og.p[:og.nC], mach=2.0, amp=0.1, width=0.15, seed=6)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
    res = {}
    for maxl in (40, 8):
        b = fa.BDFNewtonKrylov.for_euler(eu, maxl=maxl, precond=None, jtimes="analytic")
        b.set_newton(tol=1e-8, max_iter=30, lin_tol=1e-4, max_restarts=40)
        b.set_dt(5e-2)  # CFL ~2: the oracle's unpreconditioned solves take 55-73 iterations at 1e-6
        y = _dev(torch, x0, cuda)
        b.set_iv(0.0, y)
        longest = 0
        for _ in range(3):
            assert b.step(0.0, y)[0] == 0
            st = b.stats()
            longest = max(longest, st["lastLinIters"] / max(1, st["lastNewtonIters"]))
        res[maxl] = (y.cpu().numpy(), b.stats(), longest)
        b.close()
    (y40, s40, l40), (y8, s8, _) = res[40], res[8]
    err = M.rel(y8, y40)
    print(f"  maxl 40 vs 8: {err:.2e}; GMRES its {s40['nli']} vs {s8['nli']}, longest mean solve {l40:.1f} its, "
          f"second passes {s40['nreorth']}")
    assert l40 > 32  # the > 32-vector kernels ran
    assert err <= 1e-9


def test_block_ilu_and_learned_reject_non_tetrahedral_meshes(cuda):
    """The block preconditioners are built for 4-face cells: on hexahedra psetup fails with a
    message (SUNDIALS' fatal code), it does not run."""
    import torch
    print("[case] preconditioners on hexahedra", flush=True)
    pg, fv, og, bc, ctab = M.setup_hex(4, 3, 3, kind="far")
    x = M.perturbed_state(og.p[:og.nC], seed=3)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    y = _dev(torch, x, cuda)
    with pytest.raises(RuntimeError, match="faces"):
        eu.psetup(y, y, 1e-3)
    with pytest.raises(RuntimeError, match="faces"):
        eu.learned_init(seed=1)


def test_tiny_mesh_preconditioners_and_step(cuda):
    """6 tetrahedra (one partial chunk, one ILU level per cell chain): the ILU apply against the
    oracle's block ILU(0), and one BDF step with each preconditioner."""
    import torch
    print("[case] tiny mesh preconditioners", flush=True)
    pg, fv, og, bc, ctab = M.setup(1, 1, 1, kind="mixed", jitter=0.0)
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.5, seed=1)
    rho, rhou, rhoE, p = O.element_state(og, bc, x)
    jac = O.find_euler_jac(og, rho, rhou, rhoE, p, M.GAMMA)
    gamma = 1e-2
    r = np.random.default_rng(2).standard_normal(5 * nC)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    ref = O.Bilu(og, jac, gamma, order=eu.ilu_order()).solve(r)
    y = _dev(torch, x, cuda)
    eu.psetup(y, y, gamma)
    z = torch.zeros_like(y)
    eu.psolve(_dev(torch, r, cuda), z, gamma)
    torch.cuda.synchronize()
    assert M.rel(z.cpu().numpy(), ref) <= 1e-10
    eu.learned_init(seed=1)
    for pc in ("ilu", "bjac", "learned", None):
        b = fa.BDFNewtonKrylov.for_euler(eu, maxl=10, precond=pc, fallback_cap=20 if pc == "learned" else 0)
        b.set_newton(tol=1e-8, max_iter=20, lin_tol=1e-3)
        b.set_dt(1e-3)
        yy = y.clone()
        b.set_iv(0.0, yy)
        rc, t = b.step(0.0, yy)
        assert rc == 0, pc
        assert torch.isfinite(yy).all()


def test_bdf_preconditioner_setup_frequency(cuda):
    """CVODE's setup control: psetup on the first step, when gamma moves by more than 30 % (the
    BDF1 -> BDF2 switch at step 2) and every 20 steps; the states agree with a setup every step."""
    import torch
    print("[case] bdf setup frequency", flush=True)
    pg, fv, og, bc, ctab = M.setup(20, 5, 5, kind="wall", jitter=0.0, renumber=None, lx=1.0, ly=0.25, lz=0.25)
    x0 = M.sod_state(og.p[:og.nC])
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    out = {}
    for msbp in (20, 1):
        b = fa.BDFNewtonKrylov.for_euler(eu, maxl=20)
        b.set_newton(tol=1e-8, max_iter=20, lin_tol=1e-2)
        b.set_lsetup_frequency(msbp=msbp)
        b.set_dt(1e-4)
        y = _dev(torch, x0, cuda)
        b.set_iv(0.0, y)
        for k in range(30):
            rc, t = b.step(1e-4 * (k + 1), y)
            assert rc == 0
        out[msbp] = (y.cpu().numpy(), b.stats())
    assert out[20][1]["npe"] == 3  # steps 1, 2 (gamma dt -> 2dt/3) and 22
    assert out[1][1]["npe"] == 30
    assert M.rel(out[20][0], out[1][0]) <= 1e-9

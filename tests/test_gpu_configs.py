"""The benchmark configurations C (mixed-precision solve) and E (learned vs ILU) as tests.

* C at a size the oracle runs: the Mach 2 box with the 10 % sinusoidal pressure perturbation of
  wavenumbers 1-4, supersonic inflow / outflow, dt 5e-5, the mixed-precision fixed-step BDF (FP32
  Jv and ILU apply, FP64 Newton residual and update) against the oracle's FP64 BDF with block ILU
  after every step, within 1e-6 (the configured bound on the mixed path's error).
* C at full size (200x100x50 hexes x6 = 6M tets): properties that need no oracle run -- the
  residual's mass and energy balance against the boundary fluxes (supersonic inflow brings exactly
  F(ghost).n, supersonic outflow removes F(U_cell).n, walls carry no mass or energy), and the
  mixed-precision steps against the FP64 GPU steps within 1e-6.
* E at reduced size, both regimes (Mach 0.3 + 2 %, far field; Mach 3 + 10 %, supersonic): the
  learned preconditioner (a briefly trained network, block ILU behind it) and block ILU, each to the
  oracle's FP64 ILU BDF states within 1e-6; the fallback count is printed.
"""
import numpy as np
import pytest

import fosolvers_amd as fa
from oracle import oracle as O
from tests import meshes as M

pytestmark = pytest.mark.gpu
GAMMA = 1.4


def sinus_state(centroids, lengths, mach, amp, seed=20261202):
    """rho = 1, u = (mach, 0, 0) (c = 1), p = (1 + amp sum_k a_k sin(2 pi k.x/L + phi_k)) / gamma over
    k in {1..4}^3 with seeded amplitudes summing to 1 and phases (the configuration's state)."""
    rng = np.random.default_rng(seed)
    ks = np.array([(a, b, d) for a in range(1, 5) for b in range(1, 5) for d in range(1, 5)], dtype=float)
    amps = rng.random(len(ks))
    amps /= amps.sum()
    phs = 2 * np.pi * rng.random(len(ks))
    xh = centroids / np.asarray(lengths)
    pert = (amps[None, :] * np.sin(2 * np.pi * (xh @ ks.T) + phs[None, :])).sum(1)
    n = len(centroids)
    u = np.zeros((n, 3))
    u[:, 0] = mach
    return M.cons(np.ones(n), u, (1.0 + amp * pert) / GAMMA)


def conds_for(mach):
    if mach > 1:
        return [1, 2], [10, 20], np.array([[1 / GAMMA, 1 / GAMMA, mach, 0, 0], [1 / GAMMA, 0, 0, 0, 0]])
    free = [1 / GAMMA, 1 / GAMMA, mach, 0, 0]
    return [1, 2], [30, 30], np.array([free, free])


def _box(nx, ny, nz, mach):
    h = 1.0 / ny
    lengths = (nx * h, ny * h, nz * h)
    pg = fa.PolyGrid.box(nx, ny, nz, *lengths, jitter=0.2, seed=20261202, gids=(1, 2, 3, 4, 5, 6, 0))
    pg.renumber(fa.RENUMBER_RCM)
    pg.renumber(fa.RENUMBER_CHUNKS)
    return pg, lengths


def _gpu_steps(eu, x0, dt, nsteps, precond="ilu", jtimes=True, fallback_cap=0, each=None):
    import torch
    b = fa.BDFNewtonKrylov.for_euler(eu, maxl=30, precond=precond, jtimes=jtimes, fallback_cap=fallback_cap)
    b.set_newton(tol=1e-8, max_iter=20, lin_tol=1e-3, max_restarts=10)
    b.set_dt(dt)
    y = torch.tensor(x0, dtype=torch.float64, device="cuda")
    b.set_iv(0.0, y)
    for k in range(nsteps):
        rc, _t = b.step(0.0, y)
        assert rc == 0, (precond, k)
        if each is not None:
            each(k, y.cpu().numpy())
    st = b.stats()
    b.close()
    return y.cpu().numpy(), st


@pytest.mark.parametrize("jtimes", [True, "analytic"])
def test_config_c_mixed_precision_against_the_oracle(cuda, jtimes):
    print(f"[case] config C reduced, mixed precision, jtimes={jtimes}", flush=True)
    pg, lengths = _box(20, 10, 5, 2.0)
    fv = pg.up()
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    cg, ct, cp = conds_for(2.0)
    bc = O.BC(og, cg, ct, cp, r=1.0, gamm=GAMMA)
    x0 = sinus_state(og.p[:og.nC], lengths, 2.0, 0.1)
    dt, K = 5e-5, 8
    ref, _ = O.bdf_euler(og, bc, x0, dt, K, maxl=30, lin_tol=1e-3, ilu=True)
    eu = fa.Euler(fv, fa.CondTab(cg, ct, cp), r=1.0, gamma=GAMMA, flags=fa.PREC_FP32)
    errs = []
    _gpu_steps(eu, x0, dt, K, jtimes=jtimes, each=lambda k, y: errs.append(M.rel(y, ref[k])))
    print("  mixed vs oracle FP64 per step: " + " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert max(errs) <= 1e-6


def test_config_c_full_size_properties(cuda):
    """The 6M-tet configuration itself: the FP64 residual's mass and energy balance against the
    boundary fluxes at the initial state and after the steps, and the mixed-precision steps
    against the FP64 GPU steps."""
    import torch
    print("[case] config C full size (6M tets)", flush=True)
    pg, lengths = _box(200, 100, 50, 2.0)
    fv = pg.up(gpu=True)
    assert fv.nC == 6_000_000
    cg, ct, cp = conds_for(2.0)
    x0 = sinus_state(fv.p[:fv.nC], lengths, 2.0, 0.1)
    nC, nPi = fv.nC, fv.nPi
    m = fv.iEP[nPi:, 0]
    fgid = fv.gid[fv.iEP[nPi:, 1]]
    a, nrm = fv.aP[nPi:], fv.normP[nPi:]

    def balance(eu, x):
        """sum_c v_c R_c against inflow - outflow, mass and energy, relative to the inflow."""
        y = torch.tensor(x, dtype=torch.float64, device="cuda")
        r = torch.zeros_like(y)
        eu.rhs(y, r)
        R = r.cpu().numpy().reshape(5, nC)
        net = (R * fv.v[None, :]).sum(1)
        X = x.reshape(5, nC)
        rho, E = X[0, m], X[4, m]
        un = (X[1, m] * nrm[:, 0] + X[2, m] * nrm[:, 1] + X[3, m] * nrm[:, 2]) / rho
        p = (GAMMA - 1.0) * (E - 0.5 * (X[1, m] ** 2 + X[2, m] ** 2 + X[3, m] ** 2) / rho)
        out = fgid == 2
        inn = fgid == 1
        p_in, u_in = 1.0 / GAMMA, 2.0  # rho = 1, T = p/rho
        E_in = p_in / (GAMMA - 1.0) + 0.5 * u_in ** 2
        area_in = a[inn].sum()
        mass = u_in * area_in - (rho[out] * un[out] * a[out]).sum()
        energy = (E_in + p_in) * u_in * area_in - ((E[out] + p[out]) * un[out] * a[out]).sum()
        return abs(net[0] - mass) / abs(u_in * area_in), abs(net[4] - energy) / abs((E_in + p_in) * u_in * area_in)

    e64 = fa.Euler(fv, fa.CondTab(cg, ct, cp), r=1.0, gamma=GAMMA)
    bm, be = balance(e64, x0)
    print(f"  initial state: mass balance {bm:.1e}, energy balance {be:.1e}", flush=True)
    assert bm <= 1e-10 and be <= 1e-10
    # at this size the tetrahedral kernels take their non-temporal (streaming) form; the generic
    # layout (stored face geometry, plain loads) is the independent reference, as in the small-mesh
    # parity tests against the oracle (1e-12)
    eg = fa.Euler(fv, fa.CondTab(cg, ct, cp), r=1.0, gamma=GAMMA, flags=fa.LAYOUT_GENERIC)
    assert e64.layout == 1 and eg.layout == 0
    y = torch.tensor(x0, dtype=torch.float64, device="cuda")
    r1, r2 = torch.zeros_like(y), torch.zeros_like(y)
    e64.rhs(y, r1)
    eg.rhs(y, r2)
    lerr = M.rel(r1.cpu().numpy(), r2.cpu().numpy())
    print(f"  tetrahedral (streaming) vs generic layout residual: {lerr:.1e}", flush=True)
    assert lerr <= 1e-12
    eg.close()
    del y, r1, r2
    dt, K = 5e-5, 3
    y64, st64 = _gpu_steps(e64, x0, dt, K)
    e32 = fa.Euler(fv, fa.CondTab(cg, ct, cp), r=1.0, gamma=GAMMA, flags=fa.PREC_FP32)
    y32, st32 = _gpu_steps(e32, x0, dt, K, jtimes="analytic")
    err = M.rel(y32, y64)
    bm, be = balance(e64, y64)
    print(f"  after {K} steps: mixed vs FP64 {err:.1e}; mass balance {bm:.1e}, energy balance {be:.1e}; "
          f"GMRES its FP64 {st64['nli']}, mixed {st32['nli']}", flush=True)
    assert err <= 1e-6
    assert bm <= 1e-10 and be <= 1e-10


@pytest.mark.parametrize("regime", ["E1", "E2"])
def test_config_e_learned_and_ilu_against_the_oracle(cuda, regime):
    import torch
    from fosolvers_amd.learned import Trainer
    mach, amp, dt = (0.3, 0.02, 5e-3) if regime == "E1" else (3.0, 0.1, 2e-4)
    print(f"[case] config E reduced, {regime}", flush=True)
    pg, lengths = _box(20, 10, 5, mach)
    fv = pg.up()
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    cg, ct, cp = conds_for(mach)
    bc = O.BC(og, cg, ct, cp, r=1.0, gamm=GAMMA)
    x0 = sinus_state(og.p[:og.nC], lengths, mach, amp)
    K = 6
    ref, _ = O.bdf_euler(og, bc, x0, dt, K, maxl=30, lin_tol=1e-3, ilu=True)
    ctab = fa.CondTab(cg, ct, cp)
    # a briefly trained network: triples from a few FP64 ILU steps at this regime's state
    e64 = fa.Euler(fv, ctab, r=1.0, gamma=GAMMA)
    trip = []
    b = fa.BDFNewtonKrylov.for_euler(e64, maxl=30, precond="ilu")
    b.set_newton(tol=1e-8, max_iter=20, lin_tol=1e-3)
    b.set_dt(dt)
    y = torch.tensor(x0, dtype=torch.float64, device="cuda")
    b.set_iv(0.0, y)
    for _ in range(3):
        assert b.step(0.0, y)[0] == 0
        yp, r, d = torch.zeros_like(y), torch.zeros_like(y), torch.zeros_like(y)
        b.last_triple(yp, r, d)
        trip.append((yp, r, d))
    b.close()
    el = fa.Euler(fv, ctab, r=1.0, gamma=GAMMA, flags=fa.PREC_FP32)
    el.learned_init(seed=1)
    tr = Trainer(el, batch=65536, lr=1e-3, seed=1, per_state=2, mix=(1, 1, 2))
    for yp, r, d in trip:
        tr.add_triple(yp, r, d, 2 * dt / 3, pair_ctx=e64)
    kinds, losses = tr.train(40)
    yl, stl = _gpu_steps(el, x0, dt, K, precond="learned", jtimes="analytic", fallback_cap=30)
    yi, sti = _gpu_steps(fa.Euler(fv, ctab, r=1.0, gamma=GAMMA, flags=fa.PREC_FP32), x0, dt, K, jtimes="analytic")
    el_err, il_err = M.rel(yl, ref[-1]), M.rel(yi, ref[-1])
    print(f"  learned: {stl['nli'] / K:.1f} GMRES its/step, {stl['nfallback']} ILU fallbacks, vs oracle {el_err:.1e}; "
          f"ILU: {sti['nli'] / K:.1f} its/step, vs oracle {il_err:.1e}", flush=True)
    assert el_err <= 1e-6 and il_err <= 1e-6
    assert stl["nfallback"] <= stl["nni"]

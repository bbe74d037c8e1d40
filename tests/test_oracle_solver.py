"""Pinning the oracle integrator and block ILU(0): the reference's BDF KAT, closed forms of
y' = l y, and the defining properties of ILU(0) and of its block form."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import meshes as M


def test_bdf_harmonic_oscillator_kat():
    """src/test/BDFNewtonKrylov1.f90: x'' = -x from (1, 0) to t = 2 within 1e-6 (rtol 1e-8,
    atol 1e-12).  With the fixed-step BDF2 the step is chosen so the global error meets 1e-6."""
    f = lambda t, s: np.array([s[1], -s[0]])  # noqa: E731
    n = 20000
    states, iters = O.bdf_run(f, [1.0, 0.0], 2.0 / n, n, rtol=1e-8, atol=1e-12, newton_tol=1e-3)
    assert len(states) == n
    assert np.linalg.norm(states[-1] - [np.cos(2.0), -np.sin(2.0)]) <= 1e-6


def test_bdf_linear_closed_forms():
    """y' = l y: BDF1 gives y1 = y0/(1 - l dt); BDF2 then y_{k+1} = (4/3 y_k - 1/3 y_{k-1})/(1 - 2/3 l dt)."""
    lam, dt = -3.0, 0.05
    f = lambda t, y: lam * y  # noqa: E731
    states, _ = O.bdf_run(f, [2.0], dt, 6, newton_tol=1e-4, lin_tol=1e-12)
    y = [2.0, 2.0 / (1 - lam * dt)]
    for _ in range(5):
        y.append((4 / 3 * y[-1] - 1 / 3 * y[-2]) / (1 - 2 / 3 * lam * dt))
    assert np.allclose(states[:, 0], y[1:], rtol=1e-10, atol=0)


def test_adaptive_constant_step_coefficients_are_bdf1_to_5():
    """Adaptive mode, constant step history: the Nordsieck vector l(x) = prod_i (1 + x/i) (the
    textbook BDF-q in Nordsieck form: gamma = h / l_1 = h beta_0 with beta_0 = 1, 2/3, 6/11,
    12/25, 60/137) and the local-error constant tq[2] = |C_{q+1}| = 1/2, 2/9, 3/22, 12/125, 10/137."""
    consts = [1 / 2, 2 / 9, 3 / 22, 12 / 125, 10 / 137]
    beta0 = [1.0, 2 / 3, 6 / 11, 12 / 25, 60 / 137]
    for q in range(1, 6):
        for h in (1.0, 3e-4):
            l, tq2 = O.cv_coeffs(q, h, [h] * q)
            ref = np.array([1.0])
            for i in range(1, q + 1):
                ref = np.convolve(ref, [1.0, 1.0 / i])
            assert np.allclose(l, ref, rtol=1e-14, atol=0), q
            assert abs(1 / l[1] - beta0[q - 1]) <= 1e-15
            assert abs(tq2 - consts[q - 1]) <= 1e-15, (q, tq2)
    # variable steps: the leading coefficient stays that of the constant-step formula
    # (l_1 = sum 1/i, the fixed-leading-coefficient form), and the corrector polynomial
    # interpolates the past q-1 points: l(-xi_j) = 0 for xi_j = (t_n - t_{n-j}) / h, j < q
    rng = np.random.default_rng(5)
    for q in range(2, 6):
        for _ in range(5):
            h = rng.uniform(0.2, 3.0)
            tau = rng.uniform(0.2, 3.0, q)
            l, _ = O.cv_coeffs(q, h, tau)
            assert abs(l[1] - sum(1.0 / i for i in range(1, q + 1))) <= 1e-13
            for jj in range(1, q):
                xi = (h + tau[:jj - 1].sum()) / h
                assert abs(np.polyval(l[::-1], -xi)) <= 1e-12 * np.abs(l).sum() * xi ** q, (q, jj)


def test_adaptive_harmonic_oscillator_kat():
    """src/test/BDFNewtonKrylov1.f90 as the reference runs it (CVODE defaults: adaptive order and
    step, rtol 1e-8, atol 1e-12, one step at a time until t >= 2): |s - (cos t, -sin t)| <= 1e-6.
    The order rises to 5 on this smooth problem and the step count stays small."""
    f = lambda t, s: np.array([s[1], -s[0]])  # noqa: E731
    ts, hs, qs, st, stats = O.cv_run(f, [1.0, 0.0], 2.0, 2.0, 5000, rtol=1e-8, atol=1e-12)
    t = ts[-1]
    assert t >= 2.0
    assert np.linalg.norm(st[-1] - [np.cos(t), -np.sin(t)]) <= 1e-6
    assert qs.max() == 5
    assert len(ts) < 200
    assert np.allclose(np.diff(ts), hs[1:], rtol=1e-12)


def test_adaptive_first_step_closed_form():
    """y' = l y from y0 = 1: the initial-step estimate is exact for a linear problem
    (||y''|| = l^2 / (rtol + atol)), so h0 = 0.5 sqrt(2 (rtol + atol)) / |l|, and the first step
    is backward Euler, y1 = 1 / (1 - l h0)."""
    for lam, rtol in ((-3.0, 1e-6), (-50.0, 1e-4), (2.0, 1e-8)):
        atol = 1e-12
        f = lambda t, y: lam * y  # noqa: E731
        ts, hs, qs, st, _ = O.cv_run(f, [1.0], 1.0, 1.0, 1, rtol=rtol, atol=atol)
        h0 = 0.5 * np.sqrt(2 * (rtol + atol)) / abs(lam)
        assert abs(hs[0] - h0) <= 1e-12 * h0, (hs[0], h0)
        assert qs[0] == 1
        assert abs(st[0, 0] - 1 / (1 - lam * h0)) <= 1e-14


def test_adaptive_error_tracks_tolerance():
    """Tolerance proportionality on y' = -3 y to t = 1: the global error falls with rtol (more
    than 30x for 1000x tighter), stays within 100 rtol, and the step count grows."""
    f = lambda t, y: -3.0 * y  # noqa: E731
    errs, nst = [], []
    for rtol in (1e-5, 1e-8):
        ts, hs, qs, st, stats = O.cv_run(f, [1.0], 1.0, 1.0, 20000, rtol=rtol, atol=1e-14)
        errs.append(abs(st[-1, 0] - np.exp(-3.0 * ts[-1])) / np.exp(-3.0 * ts[-1]))
        nst.append(len(ts))
        assert errs[-1] <= 100 * rtol, (rtol, errs[-1])
    assert errs[1] < errs[0] / 30
    assert nst[1] > nst[0]


def test_oracle_adaptive_euler_sod():
    """Adaptive BDF with block ILU(0) on the all-wall Sod tube: steps accepted, mass conserved to
    roundoff, the order rises above 1 and the Jacobian is reused between setups."""
    pg, fv, og, bc, ctab = M.setup(10, 3, 3, kind="wall", jitter=0.0, renumber=None, lx=1.0, ly=0.25, lz=0.25)
    x0 = M.sod_state(og.p[:og.nC])
    ts, hs, qs, st, stats = O.cv_euler(og, bc, x0, 1e-2, 1e-2, 200, rtol=1e-6, atol=1e-12, maxl=20)
    assert ts[-1] >= 1e-2
    nC = og.nC
    m0 = (x0[:nC] * og.v).sum()
    assert abs((st[-1, :nC] * og.v).sum() - m0) <= 1e-12 * m0 * len(ts)
    assert qs.max() >= 2
    assert stats["npe"] < stats["nst"]


def _expand_blocks(grid, jac, gamma):
    """Dense M = I - gamma JacC / v (5x5 blocks on the cell pattern), SoA index v*nC + c."""
    nC = grid.nC
    n = 5 * nC
    Md = np.eye(n)
    for c in range(nC):
        for r in range(5):
            for q in range(5):
                Md[r * nC + c, q * nC + c] -= gamma * jac[c, q, r] / grid.v[c]
                for j in range(grid.maxNF):
                    k = grid.neib[c, j]
                    if 0 <= k < nC:
                        Md[r * nC + c, q * nC + k] -= gamma * jac[c, 5 * (j + 1) + q, r] / grid.v[c]
    return Md


def test_block_ilu_equals_scalar_ilu0_on_expanded_pattern():
    """Block ILU(0) with dense 5x5 blocks keeps the same updates as SPARSKIT ilu0 on the expanded
    scalar pattern when the scalar rows are ordered cell-major (ilut.f:1793-1917)."""
    pg, fv, og, bc, ctab = M.setup(3, 2, 2, kind="mixed")
    x = M.perturbed_state(og.p[:og.nC], mach=0.6, seed=4, width=0.3)
    rho, rhou, rhoE, p = O.element_state(og, bc, x)
    jac = O.find_euler_jac(og, rho, rhou, rhoE, p, M.GAMMA)
    gamma = 1e-3 / og.v.mean() ** (1 / 3) * og.v.mean()
    nC = og.nC
    Md = _expand_blocks(og, jac, gamma)
    # cell-major scalar ordering: index 5c + r
    perm = np.array([r * nC + c for c in range(nC) for r in range(5)])
    A = Md[np.ix_(perm, perm)]
    # the full block pattern, explicit zeros included (a structural zero inside a dense block
    # must stay in the pattern, or point ILU(0) drops different fill)
    mask = np.zeros((nC, nC), bool)
    for c in range(nC):
        mask[c, c] = True
        for k in og.neib[c]:
            if 0 <= k < nC:
                mask[c, k] = True
    rows, cols = np.nonzero(np.kron(mask, np.ones((5, 5), bool)))
    order = np.lexsort((cols, rows))  # rows sorted, columns ascending (ilu0 needs sorted L)
    ao, jao, iao = O.coocsr(5 * nC, A[rows[order], cols[order]], rows[order] + 1, cols[order] + 1)
    rng = np.random.default_rng(0)
    r = rng.standard_normal(5 * nC)
    zs, ierr = O.ilu0_solve(5 * nC, ao, jao, iao, r[perm])
    assert ierr == 0
    b = O.Bilu(og, jac, gamma)
    assert b.ierr == 0
    zb = b.solve(r)
    assert np.allclose(zb[perm], zs, rtol=1e-11, atol=1e-12 * np.abs(zs).max())
    if O.ref_available():  # the compiled SPARSKIT ilu0 + lusol agree as well
        zr, ierr = O.ref_ilu0_solve(5 * nC, iao[1:], jao[1:iao[-1]], ao[1:iao[-1]], r[perm])
        assert ierr == 0 and np.allclose(zr, zs, rtol=1e-12, atol=1e-13 * np.abs(zs).max())


@pytest.mark.parametrize("ordering", ["natural", "permuted"])
def test_block_ilu_defining_property(ordering):
    """(L U)_ij = M_ij on the block pattern (ILU(0)), for the natural elimination order and for a
    seeded permutation of the rows ("lower" = eliminated earlier)."""
    pg, fv, og, bc, ctab = M.setup(4, 2, 2, kind="far")
    x = M.perturbed_state(og.p[:og.nC], mach=0.6, seed=5, width=0.3)
    rho, rhou, rhoE, p = O.element_state(og, bc, x)
    jac = O.find_euler_jac(og, rho, rhou, rhoE, p, M.GAMMA)
    gamma = 2e-4
    nC = og.nC
    order = None if ordering == "natural" else np.random.default_rng(8).permutation(nC).astype(np.int32)
    rank = np.arange(nC) if order is None else np.argsort(order)
    b = O.Bilu(og, jac, gamma, order=order)
    assert b.ierr == 0
    # assemble L (unit block diagonal) and U from the factor and compare L U with M on the pattern
    D = b.D.reshape(nC, 5, 5)
    Off = b.Off.reshape(nC, og.maxNF, 5, 5)
    Md = _expand_blocks(og, jac, gamma)
    blk = lambda A, i, k: A[np.ix_([r * nC + i for r in range(5)], [q * nC + k for q in range(5)])]  # noqa: E731
    for i in range(nC):
        pat = [i] + [k for k in og.neib[i] if 0 <= k < nC]
        lower = {k: Off[i, j] for j, k in enumerate(og.neib[i]) if 0 <= k < nC and rank[k] < rank[i]}
        for jcol in pat:
            s = np.zeros((5, 5))
            for k, Lik in lower.items():
                if k == jcol:
                    s += Lik @ np.linalg.inv(D[k])
                else:
                    for jj, kk in enumerate(og.neib[k]):
                        if kk == jcol and 0 <= kk < nC and rank[kk] > rank[k]:
                            s += Lik @ Off[k, jj]
            if jcol == i:
                s += np.linalg.inv(D[i])
            elif rank[jcol] > rank[i]:
                s += Off[i, list(og.neib[i]).index(jcol)]
            assert np.allclose(s, blk(Md, i, jcol), rtol=1e-10, atol=1e-12), (i, jcol)
    # the triangular solves: z = (L U)^-1 r, i.e. (L U) z = r with the assembled factors
    r = np.random.default_rng(9).standard_normal(5 * nC)
    z = b.solve(r)
    LU = np.zeros((5 * nC, 5 * nC))
    for i in range(nC):
        for jcol in [i] + [k for k in og.neib[i] if 0 <= k < nC]:
            s = np.zeros((5, 5))
            for k, Lik in {k: Off[i, j] for j, k in enumerate(og.neib[i])
                           if 0 <= k < nC and rank[k] < rank[i]}.items():
                if k == jcol:
                    s += Lik @ np.linalg.inv(D[k])
                else:
                    for jj, kk in enumerate(og.neib[k]):
                        if kk == jcol and 0 <= kk < nC and rank[kk] > rank[k]:
                            s += Lik @ Off[k, jj]
            if jcol == i:
                s += np.linalg.inv(D[i])
            elif rank[jcol] > rank[i]:
                s += Off[i, list(og.neib[i]).index(jcol)]
            LU[np.ix_([q * nC + i for q in range(5)], [q * nC + jcol for q in range(5)])] = s
    # fill outside the pattern (products L_ik U_kj with j not a neighbour of i) belongs to L U too
    for i in range(nC):
        for k in [k for k in og.neib[i] if 0 <= k < nC and rank[k] < rank[i]]:
            Lik = Off[i, list(og.neib[i]).index(k)]
            for jj, kk in enumerate(og.neib[k]):
                if 0 <= kk < nC and rank[kk] > rank[k] and kk != i and kk not in og.neib[i]:
                    LU[np.ix_([q * nC + i for q in range(5)], [q * nC + kk for q in range(5)])] += Lik @ Off[k, jj]
    assert np.allclose(LU @ z, r, rtol=1e-9, atol=1e-9 * np.abs(r).max())


def test_oracle_bdf_euler_sod_converges():
    pg, fv, og, bc, ctab = M.setup(10, 3, 3, kind="wall", jitter=0.0, lx=1.0, ly=0.25, lz=0.25)
    x0 = M.sod_state(og.p[:og.nC])
    states, iters = O.bdf_euler(og, bc, x0, 1e-3, 5, ilu=True)
    assert len(states) == 5
    nC = og.nC
    # all walls: total mass and energy conserved by every step (up to the Newton tolerance)
    m0 = (x0[:nC] * og.v).sum()
    for s in states:
        assert abs((s[:nC] * og.v).sum() - m0) <= 1e-9 * m0
    # the block ILU reduces the GMRES work against no preconditioner
    s2, it2 = O.bdf_euler(og, bc, x0, 1e-3, 5, ilu=False)
    assert np.allclose(s2[-1], states[-1], rtol=0, atol=1e-6 * np.abs(states[-1]).max())
    assert iters[:, 1].sum() < it2[:, 1].sum()


def test_chunk_colour_order_keeps_ilu_quality():
    """Block ILU(0) eliminated chunk-colour by chunk-colour (256-cell chunks greedily coloured so that
    face-adjacent chunks differ, chunks of one colour in index order, cells in index order -- the
    library's elimination order) needs no more GMRES iterations than the natural order on a
    Mach 0.3 flow at acoustic CFL ~2 (12,000 tets)."""
    import scipy.sparse.linalg as sla
    pg, fv, og, bc, ctab = M.setup(20, 10, 10, kind="far", lx=2.0, ly=1.0, lz=1.0)
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.3, amp=0.02, seed=3)
    rho, rhou, rhoE, p = O.element_state(og, bc, x)
    jac = O.find_euler_jac(og, rho, rhou, rhoE, p, M.GAMMA)
    h = 0.1
    gamma = 2.0 / 3.0 * 2.0 * h / 1.3
    nb = np.asarray(og.neib)
    ch = np.arange(nC) // 256
    nch = ch[-1] + 1
    col = -np.ones(nch, int)
    for a in range(nch):
        used = {col[ch[k]] for c in np.nonzero(ch == a)[0] for k in nb[c] if 0 <= k < nC and ch[k] != a}
        col[a] = min(set(range(64)) - used)
    order = np.lexsort((np.arange(nC), ch, col[ch])).astype(np.int32)
    assert col.max() >= 1
    A = sla.LinearOperator((5 * nC, 5 * nC), matvec=lambda v: O.bilu_matvec(og, jac, gamma, v))
    r = np.random.default_rng(1).standard_normal(5 * nC)
    its = {}
    for name, o in (("natural", None), ("colour", order)):
        P = O.Bilu(og, jac, gamma, order=o)
        Pm = sla.LinearOperator((5 * nC, 5 * nC), matvec=P.solve)
        n = [0]
        _, info = sla.gmres(A, r, M=Pm, rtol=1e-6, restart=100, maxiter=500,
                            callback=lambda _r: n.__setitem__(0, n[0] + 1), callback_type="pr_norm")
        assert info == 0
        its[name] = n[0]
    print("GMRES iterations", its)
    assert its["colour"] <= its["natural"] + 1

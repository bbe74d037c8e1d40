"""Python access to the CPU oracle (liboracle.so) and the compiled reference (_ref/libfosref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  The product package never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libfosref.so")

P_D = ctypes.POINTER(ctypes.c_double)
P_I = ctypes.POINTER(ctypes.c_int)


def build_oracle():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def build_ref():
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def ref_available():
    return os.path.exists(REF_LIB)


class OrcGrid(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int) for k in ("nN", "nE", "nC", "nF", "maxNNE", "maxNF", "nP", "nPi", "nPb")] + [
        ("pN", P_D), ("sE", P_I), ("nNE", P_I), ("iNE", P_I), ("gid", P_I), ("p", P_D), ("v", P_D),
        ("iEP", P_I), ("aP", P_D), ("normP", P_D), ("pP", P_D), ("neib", P_I)]


class OrcBC(ctypes.Structure):
    _fields_ = [("nCond", ctypes.c_int), ("maxPar", ctypes.c_int), ("t", P_I), ("par", P_D), ("iBC", P_I),
                ("udf", P_D), ("r", ctypes.c_double), ("gamm", ctypes.c_double)]


class OrcBilu(ctypes.Structure):
    _fields_ = [("nC", ctypes.c_int), ("maxNF", ctypes.c_int), ("neib", P_I), ("D", P_D), ("Off", P_D),
                ("order", P_I)]


_olib = None
_rlib = None


def olib():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_LIB):
            build_oracle()
        l = ctypes.CDLL(ORACLE_LIB)
        l.orc_grid_build.restype = ctypes.POINTER(OrcGrid)
        l.orc_grid_build.argtypes = [ctypes.c_int, P_D, ctypes.c_int, P_I, P_I, ctypes.c_int, P_I, P_I]
        l.orc_grid_free.argtypes = [ctypes.POINTER(OrcGrid)]
        l.orc_wrms.restype = ctypes.c_double
        l.orc_a3p.restype = ctypes.c_double
        l.orc_v4p.restype = ctypes.c_double
        _olib = l
    return _olib


def rlib():
    global _rlib
    if _rlib is None:
        _rlib = ctypes.CDLL(REF_LIB, mode=os.RTLD_LAZY)
    return _rlib


def dp(a):
    return a.ctypes.data_as(P_D)


def ip(a):
    return a.ctypes.data_as(P_I)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Grid:
    """An oracle grid (restated polyFvGrid%up) with numpy copies of its arrays."""

    def __init__(self, pN, sE, nNE, iNE, gid):
        pN, sE, nNE, iNE, gid = _c(pN, np.float64), _c(sE, np.int32), _c(nNE, np.int32), _c(iNE, np.int32), _c(gid, np.int32)
        gp = olib().orc_grid_build(pN.shape[0], dp(pN), sE.shape[0], ip(sE), ip(nNE), iNE.shape[1], ip(iNE), ip(gid))
        self._g = gp
        g = gp.contents
        self.nN, self.nE, self.nC, self.nF = g.nN, g.nE, g.nC, g.nF
        self.nP, self.nPi, self.nPb, self.maxNF, self.maxNNE = g.nP, g.nPi, g.nPb, g.maxNF, g.maxNNE

        def cp(ptr, n, dt, shape):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt).reshape(shape).copy() if n else np.zeros(shape, dt)
        self.p = cp(g.p, 3 * g.nE, np.float64, (g.nE, 3))
        self.v = cp(g.v, g.nC, np.float64, (g.nC,))
        self.iEP = cp(g.iEP, 2 * g.nP, np.int32, (g.nP, 2))
        self.aP = cp(g.aP, g.nP, np.float64, (g.nP,))
        self.normP = cp(g.normP, 3 * g.nP, np.float64, (g.nP, 3))
        self.pP = cp(g.pP, 3 * g.nP, np.float64, (g.nP, 3))
        self.neib = cp(g.neib, g.maxNF * g.nC, np.int32, (g.nC, g.maxNF))
        self.gid = cp(g.gid, g.nE, np.int32, (g.nE,))
        self.sE = cp(g.sE, g.nE, np.int32, (g.nE,))

    def __del__(self):
        try:
            olib().orc_grid_free(self._g)
        except Exception:
            pass


class BC:
    """Condition table mapped onto a grid (iBC), plus gas constants."""

    def __init__(self, grid, cond_gid=(), cond_t=(), cond_p=None, r=1.0, gamm=1.4, udf=None):
        self.cond_gid = _c(cond_gid, np.int32)
        self.cond_t = _c(cond_t, np.int32)
        n = len(self.cond_gid)
        self.cond_p = _c(np.zeros((n, 5)) if cond_p is None else np.asarray(cond_p).reshape(n, -1), np.float64)
        self.iBC = np.zeros(grid.nE, dtype=np.int32)
        olib().orc_map_cond(grid.nE, ip(grid.gid), n, ip(self.cond_gid), ip(self.iBC))
        self.udf = None if udf is None else _c(udf, np.float64)
        self.s = OrcBC()
        self.s.nCond, self.s.maxPar = n, (self.cond_p.shape[1] if n else 5)
        self.s.t, self.s.par, self.s.iBC = ip(self.cond_t), dp(self.cond_p), ip(self.iBC)
        self.s.udf = dp(self.udf) if self.udf is not None else ctypes.cast(None, P_D)
        self.s.r, self.s.gamm = r, gamm


def rhs(grid, bc, x, scratch=None):
    x = _c(x, np.float64)
    out = np.zeros(5 * grid.nC)
    if scratch is None:
        scratch = np.zeros(13 * grid.nE)
    olib().orc_rhs(grid._g, ctypes.byref(bc.s), dp(x), dp(out), dp(scratch))
    return out


def element_state(grid, bc, x, full=False):
    """(rho, rhou, rhoE, p) element arrays after setBC + syncState; full adds u, temp, c."""
    nE = grid.nE
    s = np.zeros(13 * nE)
    x = _c(x, np.float64)
    rho, rhou, rhoE = s[:nE], s[nE:4 * nE], s[4 * nE:5 * nE]
    u, p, temp, c = s[5 * nE:8 * nE], s[8 * nE:9 * nE], s[9 * nE:10 * nE], s[10 * nE:11 * nE]
    olib().orc_set_bc(grid._g, ctypes.byref(bc.s), dp(x), dp(rho), dp(rhou), dp(rhoE))
    olib().orc_sync_state(grid._g, ctypes.c_double(bc.s.r), ctypes.c_double(bc.s.gamm), dp(x), dp(rho), dp(rhou),
                          dp(rhoE), dp(u), dp(p), dp(temp), dp(c))
    if full:
        return rho.copy(), rhou.copy(), rhoE.copy(), u.copy(), p.copy(), temp.copy(), c.copy()
    return rho.copy(), rhou.copy(), rhoE.copy(), p.copy()


def find_euler(grid, rho, rhou, rhoE, p, gamm):
    out = np.zeros(5 * grid.nC)
    nC = grid.nC
    olib().orc_find_euler(grid._g, dp(_c(rho, np.float64)), dp(_c(rhou, np.float64)), dp(_c(rhoE, np.float64)),
                          dp(_c(p, np.float64)), ctypes.c_double(gamm), dp(out[:nC]), dp(out[nC:4 * nC]),
                          dp(out[4 * nC:]))
    return out[:nC].copy(), out[nC:4 * nC].reshape(nC, 3).copy(), out[4 * nC:].copy()


def find_euler_jac(grid, rho, rhou, rhoE, p, gamm):
    ncol = 5 * (1 + grid.maxNF)
    jac = np.zeros(grid.nC * ncol * 5)
    olib().orc_find_euler_jac(grid._g, dp(_c(rho, np.float64)), dp(_c(rhou, np.float64)), dp(_c(rhoE, np.float64)),
                              dp(_c(p, np.float64)), ctypes.c_double(gamm), dp(jac))
    return jac.reshape(grid.nC, ncol, 5)  # [c, col, row]


def wrms(v, y, rtol=1e-6, atol=1e-12):
    v, y = _c(v, np.float64), _c(y, np.float64)
    return olib().orc_wrms(len(v), dp(v), dp(y), ctypes.c_double(rtol), ctypes.c_double(atol))


def dq_jv(grid, bc, y, fy, v, sigma):
    y, fy, v = _c(y, np.float64), _c(fy, np.float64), _c(v, np.float64)
    out = np.zeros_like(y)
    work = np.zeros_like(y)
    scratch = np.zeros(13 * grid.nE)
    olib().orc_dq_jv(grid._g, ctypes.byref(bc.s), dp(y), dp(fy), dp(v), ctypes.c_double(sigma), dp(out), dp(work),
                     dp(scratch))
    return out


# ------------------------------------------------------------------ compiled reference


class RefGrid:
    """The reference's own polyFvGrid%up (modFOEuler's module grid), arrays copied out."""

    def __init__(self, pN=None, sE=None, nNE=None, iNE=None, gid=None, gmsh=None):
        l = rlib()
        if gmsh is not None:
            b = os.fsencode(gmsh)
            l.ref_grid_read_gmsh(ctypes.c_char_p(b), ctypes.c_int(len(b)))
        else:
            pN, sE, nNE, iNE, gid = _c(pN, np.float64), _c(sE, np.int32), _c(nNE, np.int32), _c(iNE, np.int32), _c(gid, np.int32)
            l.ref_grid_from_arrays(ctypes.c_int(pN.shape[0]), dp(pN), ctypes.c_int(sE.shape[0]), ip(sE), ip(nNE),
                                   ctypes.c_int(iNE.shape[1]), ip(iNE), ip(gid))
        sz = np.zeros(10, dtype=np.int32)
        l.ref_grid_sizes(ip(sz))
        self.nN, self.nE, self.nC, self.nF, self.nP, self.nPi, self.nPb, self.maxNF, self.maxNNE = [int(v) for v in sz[:9]]
        self.iEP = np.zeros((self.nP, 2), np.int32)
        self.aP = np.zeros(self.nP)
        self.normP = np.zeros((self.nP, 3))
        self.pP = np.zeros((self.nP, 3))
        self.v = np.zeros(self.nC)
        self.neib = np.zeros((self.nC, self.maxNF), np.int32)
        self.p = np.zeros((self.nE, 3))
        self.sE = np.zeros(self.nE, np.int32)
        self.gid = np.zeros(self.nE, np.int32)
        self.iNE = np.zeros((self.nE, self.maxNNE), np.int32)
        self.pN = np.zeros((self.nN, 3))
        l.ref_grid_get(ip(self.iEP), dp(self.aP), dp(self.normP), dp(self.pP), dp(self.v), ip(self.neib), dp(self.p),
                       ip(self.sE), ip(self.gid), ip(self.iNE), dp(self.pN))

    def setup_state(self, cond_gid, cond_t, cond_p, r, gamm):
        cg, ct = _c(cond_gid, np.int32), _c(cond_t, np.int32)
        n = len(cg)
        cp = _c(np.zeros((max(n, 1), 5)) if n == 0 else np.asarray(cond_p).reshape(n, -1), np.float64)
        rlib().ref_setup_state(ctypes.c_int(n), ctypes.c_int(cp.shape[1]), ip(cg), ip(ct), dp(cp), ctypes.c_double(r),
                               ctypes.c_double(gamm))

    def rhs(self, x):
        x = _c(x, np.float64)
        out = np.zeros(5 * self.nC)
        rlib().ref_rhs(dp(x), dp(out))
        return out

    def state(self):
        nE = self.nE
        rho, rhou, rhoE, p = np.zeros(nE), np.zeros(3 * nE), np.zeros(nE), np.zeros(nE)
        rlib().ref_get_state(dp(rho), dp(rhou), dp(rhoE), dp(p))
        return rho, rhou, rhoE, p

    def find_euler(self, rho, rhou, rhoE, p, gamm):
        nC = self.nC
        a, b, c = np.zeros(nC), np.zeros(3 * nC), np.zeros(nC)
        rlib().ref_find_euler(dp(_c(rho, np.float64)), dp(_c(rhou, np.float64)), dp(_c(rhoE, np.float64)),
                              dp(_c(p, np.float64)), ctypes.c_double(gamm), dp(a), dp(b), dp(c))
        return a, b.reshape(nC, 3), c

    def find_euler_jac(self, rho, rhou, rhoE, p, gamm):
        ncol = 5 * (1 + self.maxNF)
        jac = np.zeros(self.nC * ncol * 5)
        rlib().ref_find_euler_jac(dp(_c(rho, np.float64)), dp(_c(rhou, np.float64)), dp(_c(rhoE, np.float64)),
                                  dp(_c(p, np.float64)), ctypes.c_double(gamm), dp(jac))
        return jac.reshape(self.nC, ncol, 5)


def ref_ilu_solve(n, iA, jA, A, lfill, droptol, rhs_):
    iA, jA, A, rhs_ = _c(iA, np.int32), _c(jA, np.int32), _c(A, np.float64), _c(rhs_, np.float64)
    x = np.zeros(n)
    rlib().ref_ilu_solve(ctypes.c_int(n), ctypes.c_int(len(A)), ip(iA), ip(jA), dp(A), ctypes.c_int(lfill),
                         ctypes.c_double(droptol), dp(rhs_), dp(x))
    return x


def ref_mulvec(n, iA, jA, A, x):
    iA, jA, A, x = _c(iA, np.int32), _c(jA, np.int32), _c(A, np.float64), _c(x, np.float64)
    y = np.zeros(n)
    rlib().ref_mulvec(ctypes.c_int(n), ctypes.c_int(len(A)), ip(iA), ip(jA), dp(A), dp(x), dp(y))
    return y


def ref_ilu0_solve(n, ia, ja, a, rhs_):
    """1-based CSR (ia[n+1], ja, a) through SPARSKIT ilu0 + lusol."""
    ia, ja, a, rhs_ = _c(ia, np.int32), _c(ja, np.int32), _c(a, np.float64), _c(rhs_, np.float64)
    x = np.zeros(n)
    ierr = ctypes.c_int(0)
    rlib().ref_ilu0_solve(ctypes.c_int(n), ip(ia), ip(ja), dp(a), dp(rhs_), dp(x), ctypes.byref(ierr))
    return x, ierr.value


# ------------------------------------------------------------------ SPARSKIT restatements (1-based)


def coocsr(nrow, a, ir, jc):
    nnz = len(a)
    a1 = np.concatenate([[0.0], _c(a, np.float64)])
    ir1 = np.concatenate([[0], _c(ir, np.int32)]).astype(np.int32)
    jc1 = np.concatenate([[0], _c(jc, np.int32)]).astype(np.int32)
    ao = np.zeros(nnz + 1)
    jao = np.zeros(nnz + 1, np.int32)
    iao = np.zeros(nrow + 2, np.int32)
    olib().orc_coocsr(nrow, nnz, dp(a1), ip(ir1), ip(jc1), dp(ao), ip(jao), ip(iao))
    return ao, jao, iao  # 1-based (index 0 unused)


def amux(n, x, ao, jao, iao):
    x1 = np.concatenate([[0.0], _c(x, np.float64)])
    y1 = np.zeros(n + 1)
    olib().orc_amux(n, dp(x1), dp(y1), dp(ao), ip(jao), ip(iao))
    return y1[1:]


def ilut_solve(n, ao, jao, iao, lfil, droptol, rhs_):
    iwk = n * min(2 * lfil + 1, n) + n + 2
    alu = np.zeros(iwk + 2)
    jlu = np.zeros(iwk + 2, np.int32)
    ju = np.zeros(n + 2, np.int32)
    w = np.zeros(n + 2)
    jw = np.zeros(2 * n + 2, np.int32)
    ierr = olib().orc_ilut(n, dp(ao), ip(jao), ip(iao), min(lfil, n - 1), ctypes.c_double(droptol), dp(alu), ip(jlu),
                           ip(ju), iwk, dp(w), ip(jw))
    r1 = np.concatenate([[0.0], _c(rhs_, np.float64)])
    x1 = np.zeros(n + 1)
    olib().orc_lusol(n, dp(r1), dp(x1), dp(alu), ip(jlu), ip(ju))
    return x1[1:], ierr


def ilut_factor(n, ao, jao, iao, lfil, droptol):
    """SPARSKIT ilut (restated): (alu, jlu, ju, ierr), 1-based arrays with index 0 unused; the factor
    occupies alu/jlu[1 .. jlu[n+1] - 1]."""
    iwk = n * min(2 * lfil + 1, n) + n + 2
    alu = np.zeros(iwk + 2)
    jlu = np.zeros(iwk + 2, np.int32)
    ju = np.zeros(n + 2, np.int32)
    w = np.zeros(n + 2)
    jw = np.zeros(2 * n + 2, np.int32)
    ierr = olib().orc_ilut(n, dp(ao), ip(jao), ip(iao), min(lfil, n - 1), ctypes.c_double(droptol), dp(alu), ip(jlu),
                           ip(ju), iwk, dp(w), ip(jw))
    return alu, jlu, ju, ierr


def lusol(n, alu, jlu, ju, rhs_):
    """SPARSKIT lusol (restated) on 1-based MSR arrays."""
    r1 = np.concatenate([[0.0], _c(rhs_, np.float64)])
    x1 = np.zeros(n + 1)
    olib().orc_lusol(n, dp(r1), dp(x1), dp(_c(alu, np.float64)), ip(_c(jlu, np.int32)), ip(_c(ju, np.int32)))
    return x1[1:]


def ilu0_solve(n, ao, jao, iao, rhs_):
    nnz = iao[n + 1] - 1
    alu = np.zeros(nnz + n + 3)
    jlu = np.zeros(nnz + n + 3, np.int32)
    ju = np.zeros(n + 2, np.int32)
    iw = np.zeros(n + 2, np.int32)
    ierr = olib().orc_ilu0(n, dp(ao), ip(jao), ip(iao), dp(alu), ip(jlu), ip(ju), ip(iw))
    r1 = np.concatenate([[0.0], _c(rhs_, np.float64)])
    x1 = np.zeros(n + 1)
    if ierr == 0:
        olib().orc_lusol(n, dp(r1), dp(x1), dp(alu), ip(jlu), ip(ju))
    return x1[1:], ierr


class Bilu:
    """Block ILU(0) of M = I - gamma JacC / v on the neighbour pattern, rows eliminated in `order`
    (order[s] = row eliminated at step s; None = natural order)."""

    def __init__(self, grid, jac, gamma, order=None):
        self.nC, self.maxNF = grid.nC, grid.maxNF
        self.neib = _c(grid.neib, np.int32)
        self.D = np.zeros(25 * self.nC)
        self.Off = np.zeros(25 * self.maxNF * self.nC)
        self.order = None if order is None else _c(order, np.int32)
        self.s = OrcBilu(self.nC, self.maxNF, ip(self.neib), dp(self.D), dp(self.Off),
                         ip(self.order) if self.order is not None else None)
        self.jac = _c(jac, np.float64)
        self.v = _c(grid.v, np.float64)
        self.ierr = olib().orc_bilu_setup(ctypes.byref(self.s), dp(self.jac), dp(self.v), ctypes.c_double(gamma))

    def solve(self, r):
        r = _c(r, np.float64)
        z = np.zeros_like(r)
        olib().orc_bilu_solve(ctypes.byref(self.s), dp(r), dp(z))
        return z


def bilu_matvec(grid, jac, gamma, x):
    x = _c(x, np.float64)
    y = np.zeros_like(x)
    olib().orc_bilu_matvec(grid._g, dp(_c(jac, np.float64)), dp(_c(grid.v, np.float64)), ctypes.c_double(gamma),
                           dp(x), dp(y))
    return y


# ------------------------------------------------------------------ CPU baseline (timing build)

REF_OMP_LIB = os.path.join(HERE, "_ref", "libfosref_omp.so")


class RefTimer:
    """The reference's setBC + syncState + findEuler (+ /v), compiled with its own -Ofast -fopenmp
    flags, on polyFvGrid arrays installed directly (no O(n k) face search)."""

    def __init__(self, fv, cond_gid, cond_t, cond_p, r, gamm):
        self.l = ctypes.CDLL(REF_OMP_LIB, mode=os.RTLD_LAZY)
        a = {k: np.ascontiguousarray(getattr(fv, k)) for k in ("pN", "sE", "nNE", "iNE", "gid", "p", "v", "iEP", "aP",
                                                               "normP", "pP", "neib")}
        self._keep = a
        I = ctypes.c_int
        self.l.ref_grid_set_fv(I(fv.nN), I(fv.nE), I(fv.nC), I(fv.nP), I(fv.nPi), I(fv.nPb), I(fv.maxNF),
                               I(fv.maxNNE), dp(a["pN"]), ip(a["sE"]), ip(a["nNE"]), ip(a["iNE"]), ip(a["gid"]),
                               dp(a["p"]), dp(a["v"]), ip(a["iEP"]), dp(a["aP"]), dp(a["normP"]), dp(a["pP"]),
                               ip(a["neib"]))
        cg, ct = _c(cond_gid, np.int32), _c(cond_t, np.int32)
        cp = _c(np.asarray(cond_p).reshape(len(cg), -1), np.float64)
        self.l.ref_setup_state(I(len(cg)), I(cp.shape[1]), ip(cg), ip(ct), dp(cp), ctypes.c_double(r),
                               ctypes.c_double(gamm))
        self.nC = fv.nC
        try:
            self.l.omp_get_max_threads.restype = ctypes.c_int
            self.threads = int(self.l.omp_get_max_threads())
        except AttributeError:
            self.threads = 1

    def rhs(self, x, out):
        self.l.ref_rhs(dp(x), dp(out))


# ------------------------------------------------------------------ integrator

ORC_F = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_double, P_D, P_D, ctypes.c_void_p)


def bdf_run(f, y0, dt, nsteps, rtol=1e-6, atol=1e-12, newton_tol=1e-8, lin_tol=1e-2, maxl=20, restarts=5,
            max_newton=20, t0=0.0):
    """Generic oracle BDF on a Python rhs f(t, y) -> ydot (no preconditioner)."""
    n = len(y0)

    def cb(t, yp, ydp, ud):
        y = np.ctypeslib.as_array(yp, shape=(n,))
        yd = np.ctypeslib.as_array(ydp, shape=(n,))
        yd[:] = f(t, y.copy())
        return 0
    cfun = ORC_F(cb)
    y0 = _c(y0, np.float64)
    states = np.zeros((nsteps, n))
    iters = np.zeros((nsteps, 2), np.int32)
    olib().orc_bdf_run.argtypes = [ctypes.c_int, ORC_F, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, P_D,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, P_D,
                                   P_I]
    k = olib().orc_bdf_run(n, cfun, None, None, None, dp(y0), t0, dt, nsteps, rtol, atol, newton_tol, lin_tol, maxl,
                           restarts, max_newton, dp(states), ip(iters))
    return states[:k], iters[:k]


class CvStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_longlong) for k in ("nst", "nni", "nli", "npe", "netf", "ncfn")]


def _stats_dict(st):
    return {k: getattr(st, k) for k, _ in st._fields_}


def cv_run(f, y0, t_out, t_end, max_steps, rtol=1e-6, atol=1e-12, maxl=5, restarts=5, qmax=5, h0=0.0, t0=0.0):
    """Adaptive (CVODE-method) oracle BDF on a Python rhs f(t, y) -> ydot, no preconditioner: one
    internal step at a time until t >= t_end or max_steps.  Returns (t, h, q, states, stats)."""
    n = len(y0)

    def cb(t, yp, ydp, ud):
        y = np.ctypeslib.as_array(yp, shape=(n,))
        yd = np.ctypeslib.as_array(ydp, shape=(n,))
        yd[:] = f(t, y.copy())
        return 0
    cfun = ORC_F(cb)
    y0 = _c(y0, np.float64)
    ts, hs = np.zeros(max_steps), np.zeros(max_steps)
    qs = np.zeros(max_steps, np.int32)
    states = np.zeros((max_steps, n))
    st = CvStats()
    D = ctypes.c_double
    fn = olib().orc_cv_run
    fn.argtypes = [ctypes.c_int, ORC_F, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, P_D, D, D, D,
                   ctypes.c_int, D, D, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, P_D, P_D, P_I, P_D,
                   ctypes.POINTER(CvStats)]
    k = fn(n, cfun, None, None, None, dp(y0), t0, t_out, t_end, max_steps, rtol, atol, maxl, restarts, qmax, h0,
           dp(ts), dp(hs), ip(qs), dp(states), ctypes.byref(st))
    return ts[:k], hs[:k], qs[:k], states[:k], _stats_dict(st)


def cv_coeffs(q, h, tau):
    """The adaptive method's l[0..q] and error constant tq[2] for order q after steps tau[1..q]."""
    tau = _c(list(tau) + [0.0] * (6 - len(tau)), np.float64)
    l = np.zeros(q + 1)
    tq2 = ctypes.c_double(0.0)
    olib().orc_cv_coeffs(q, ctypes.c_double(h), dp(tau), dp(l), ctypes.byref(tq2))
    return l, tq2.value


def cv_euler(grid, bc, y0, t_out, t_end, max_steps, rtol=1e-6, atol=1e-12, maxl=20, restarts=5, order=None):
    """Adaptive oracle BDF on the Euler residual with block ILU(0) (Jacobian reused when jok = 1)."""
    y0 = _c(y0, np.float64)
    n = len(y0)
    ts, hs = np.zeros(max_steps), np.zeros(max_steps)
    qs = np.zeros(max_steps, np.int32)
    states = np.zeros((max_steps, n))
    st = CvStats()
    D = ctypes.c_double
    ordp = ip(_c(order, np.int32)) if order is not None else None
    keep = order if order is None else _c(order, np.int32)
    if keep is not None:
        ordp = ip(keep)
    k = olib().orc_cv_euler(grid._g, ctypes.byref(bc.s), dp(y0), D(t_out), D(t_end), max_steps, D(rtol), D(atol),
                            maxl, restarts, ordp, dp(ts), dp(hs), ip(qs), dp(states), ctypes.byref(st))
    return ts[:k], hs[:k], qs[:k], states[:k], _stats_dict(st)


def bdf_euler(grid, bc, y0, dt, nsteps, rtol=1e-6, atol=1e-12, newton_tol=1e-8, lin_tol=1e-2, maxl=20, restarts=5,
              max_newton=20, ilu=True):
    y0 = _c(y0, np.float64)
    states = np.zeros((nsteps, len(y0)))
    iters = np.zeros((nsteps, 2), np.int32)
    D = ctypes.c_double
    k = olib().orc_bdf_euler(grid._g, ctypes.byref(bc.s), dp(y0), D(dt), nsteps, D(rtol), D(atol), D(newton_tol),
                             D(lin_tol), maxl, restarts, max_newton, 1 if ilu else 0, dp(states), ip(iters))
    return states[:k], iters[:k]


# ------------------------------------------------------------------ learned preconditioner (FP64)


def mlp_nparams():
    return olib().orc_mlp_nparams()


def mlp_forward(X, nbr, par):
    X, nbr, par = _c(X, np.float64), _c(nbr, np.int32), _c(par, np.float64)
    n = X.shape[0]
    O = np.zeros((n, 30))
    olib().orc_mlp_forward(n, dp(X), nbr.shape[1], ip(nbr), dp(par), dp(O))
    return O


def mlp_apply(O, nbr, S, r):
    O, nbr, S, r = _c(O, np.float64), _c(nbr, np.int32), _c(S, np.float64), _c(r, np.float64)
    z = np.zeros_like(r)
    olib().orc_mlp_apply(O.shape[0], dp(O), nbr.shape[1], ip(nbr), dp(S), dp(r), dp(z))
    return z


def mlp_loss_grad(X, nbr, par, S, r, delta, mask):
    X, nbr, par = _c(X, np.float64), _c(nbr, np.int32), _c(par, np.float64)
    S, r, delta, mask = _c(S, np.float64), _c(r, np.float64), _c(delta, np.float64), _c(mask, np.int32)
    g = np.zeros_like(par)
    olib().orc_mlp_loss_grad.restype = ctypes.c_double
    loss = olib().orc_mlp_loss_grad(X.shape[0], dp(X), nbr.shape[1], ip(nbr), dp(par), dp(S), dp(r), dp(delta),
                                    ip(mask), dp(g))
    return loss, g

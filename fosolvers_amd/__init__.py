"""fosolvers_amd: MI355X-native hot path of foeuler (mianzhi/fosolvers) behind a C ABI.

Python mirror of the reference's host-side interfaces over libfosolvers_amd.so:
  PolyGrid / FvGrid   polyGrid / polyFvGrid  (src/libfosolvers/polyGrid/*.f90)
  CondTab             condTab                (src/libfosolvers/condition/condition.f90)
  UDFTab              UDFTab                 (src/libfosolvers/UDF/UDF.f90)
  Euler               foeuler's rhs / jtimes / pset / psol callbacks (src/foeuler/foeuler.f90)
The shared library is required: importing works without it, but every call raises
LibraryMissing when libfosolvers_amd.so has not been built (python -m fosolvers_amd.build).
"""
import ctypes
import os

import numpy as np

from . import _capi
from ._capi import LibraryMissing, lib, check  # noqa: F401

TRI, QUAD, TRI6, TET, HEX, TET10 = 10, 11, 110, 20, 21, 120
BC_WALL, BC_IN_STATIC, BC_IN_STATIC_UDF, BC_IN_TOTAL, BC_IN_TOTAL_UDF = 0, 10, 110, 11, 111
BC_OUT, BC_OUT_UDF, BC_FAR, BC_FAR_UDF = 20, 120, 30, 130
RENUMBER_RCM, RENUMBER_CHUNKS = 1, 2
VECTORS_ON_HOST, PREC_FP32, LAYOUT_GENERIC, JAC_EXACT, ASYNC_CALLBACKS = 1, 2, 4, 8, 16
CHUNK_CELLS = 256


def _arr(ptr, n, dtype, shape=None):
    if n == 0:
        out = np.zeros(0, dtype=dtype)
    else:
        ctype = {np.float64: ctypes.c_double, np.int32: ctypes.c_int}[dtype]
        buf = ctypes.cast(ptr, ctypes.POINTER(ctype * n)).contents
        out = np.frombuffer(buf, dtype=dtype).copy()
    return out.reshape(shape) if shape is not None else out


class PolyGrid:
    """Raw element lists (what readGMSH fills, GMSH.f90:4-157). Arrays are numpy copies."""

    def __init__(self, pN, sE, nNE, iNE, gid):
        self.pN = np.ascontiguousarray(pN, dtype=np.float64).reshape(-1, 3)
        self.sE = np.ascontiguousarray(sE, dtype=np.int32)
        self.nNE = np.ascontiguousarray(nNE, dtype=np.int32)
        self.iNE = np.ascontiguousarray(iNE, dtype=np.int32)
        self.gid = np.ascontiguousarray(gid, dtype=np.int32)

    @property
    def nN(self):
        return self.pN.shape[0]

    @property
    def nE(self):
        return self.sE.shape[0]

    @classmethod
    def _from_c(cls, c):
        g = cls(_arr(c.pN, 3 * c.nN, np.float64, (c.nN, 3)), _arr(c.sE, c.nE, np.int32),
                _arr(c.nNE, c.nE, np.int32), _arr(c.iNE, c.maxNNE * c.nE, np.int32, (c.nE, c.maxNNE)),
                _arr(c.gid, c.nE, np.int32))
        lib().fos_poly_grid_free(ctypes.byref(c))
        return g

    def _to_c(self):
        c = _capi.PolyGridC()
        c.nN, c.nE, c.maxNNE = self.nN, self.nE, self.iNE.shape[1]
        c.pN = self.pN.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        c.sE = self.sE.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        c.nNE = self.nNE.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        c.iNE = self.iNE.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        c.gid = self.gid.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        return c

    @classmethod
    def box(cls, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=0, gids=(1, 2, 3, 4, 5, 6, 0)):
        """Box of nx*ny*nz hexes x 6 Kuhn tets plus boundary triangles tagged per side."""
        c = _capi.PolyGridC()
        g = (ctypes.c_int * 7)(*gids)
        check(lib().fos_grid_box(nx, ny, nz, lx, ly, lz, jitter, ctypes.c_uint64(seed), g, ctypes.byref(c)))
        return cls._from_c(c)

    @classmethod
    def read_gmsh(cls, path):
        c = _capi.PolyGridC()
        check(lib().fos_grid_read_gmsh(os.fsencode(path), ctypes.byref(c)))
        return cls._from_c(c)

    def write_gmsh(self, path):
        c = self._to_c()
        check(lib().fos_grid_write_gmsh(os.fsencode(path), ctypes.byref(c)))

    def renumber(self, method):
        """In-place cell renumbering; returns perm (old index of new cell k, cells in element order)."""
        c = self._to_c()
        ncell = int(np.isin(self.sE, (TET, HEX, TET10)).sum())
        perm = np.zeros(ncell, dtype=np.int32)
        check(lib().fos_grid_renumber(ctypes.byref(c), method, perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return perm

    def up(self, gpu=False, stream=None):
        """polyFvGrid%up: returns an FvGrid (on the GPU with gpu=True; the same arrays)."""
        c = self._to_c()
        out = _capi.FvGridC()
        if gpu:
            check(lib().fos_grid_up_gpu(ctypes.byref(c), ctypes.byref(out), ctypes.c_void_p(stream) if stream else None))
        else:
            check(lib().fos_grid_up(ctypes.byref(c), ctypes.byref(out)))
        return FvGrid._from_c(out)


class FvGrid:
    """polyFvGrid fields (polyFvGrid.f90:14-29), 0-based."""

    FIELDS = ("pN", "sE", "nNE", "iNE", "gid", "p", "v", "iEP", "aP", "normP", "pP", "neib")

    def __init__(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)

    @classmethod
    def _from_c(cls, c):
        g = cls(nN=c.nN, nE=c.nE, nC=c.nC, nF=c.nF, nP=c.nP, nPi=c.nPi, nPb=c.nPb, maxNF=c.maxNF, maxNNE=c.maxNNE)
        g.pN = _arr(c.pN, 3 * c.nN, np.float64, (c.nN, 3))
        g.sE = _arr(c.sE, c.nE, np.int32)
        g.nNE = _arr(c.nNE, c.nE, np.int32)
        g.iNE = _arr(c.iNE, c.maxNNE * c.nE, np.int32, (c.nE, c.maxNNE))
        g.gid = _arr(c.gid, c.nE, np.int32)
        g.p = _arr(c.p, 3 * c.nE, np.float64, (c.nE, 3))
        g.v = _arr(c.v, c.nC, np.float64)
        g.iEP = _arr(c.iEP, 2 * c.nP, np.int32, (c.nP, 2))
        g.aP = _arr(c.aP, c.nP, np.float64)
        g.normP = _arr(c.normP, 3 * c.nP, np.float64, (c.nP, 3))
        g.pP = _arr(c.pP, 3 * c.nP, np.float64, (c.nP, 3))
        g.neib = _arr(c.neib, c.maxNF * c.nC, np.int32, (c.nC, c.maxNF))
        lib().fos_fv_grid_free(ctypes.byref(c))
        return g

    def sort_pairs_reference(self):
        """Permute the pairs into sortPolyFvGrid's exact order (fos_grid_sort_pairs_reference)."""
        c = self.to_c()
        check(lib().fos_grid_sort_pairs_reference(ctypes.byref(c)))
        for k, a in zip(self.FIELDS, c._keep):
            setattr(self, k, a)
        return self

    def to_c(self):
        """ctypes view (keeps references to the numpy arrays on the returned object)."""
        c = _capi.FvGridC()
        for k in ("nN", "nE", "nC", "nF", "nP", "nPi", "nPb", "maxNF", "maxNNE"):
            setattr(c, k, int(getattr(self, k)))
        keep = []
        for k in self.FIELDS:
            a = getattr(self, k)
            a = np.ascontiguousarray(a)
            keep.append(a)
            ct = ctypes.c_double if a.dtype == np.float64 else ctypes.c_int
            setattr(c, k, a.ctypes.data_as(ctypes.POINTER(ct)))
        c._keep = keep
        return c


class CondTab:
    """condTab: gid, t, p(maxPar, n) (condition.f90:8-16)."""

    def __init__(self, gid=(), t=(), p=None, maxPar=5):
        self.gid = np.asarray(gid, dtype=np.int32)
        self.t = np.asarray(t, dtype=np.int32)
        n = len(self.gid)
        self.p = np.zeros((n, maxPar)) if p is None else np.asarray(p, dtype=np.float64).reshape(n, -1)

    @property
    def n(self):
        return len(self.gid)

    @classmethod
    def read(cls, path):
        c = _capi.CondTabC()
        check(lib().fos_read_cond_tab(os.fsencode(path), ctypes.byref(c)))
        out = cls(_arr(c.gid, c.n, np.int32), _arr(c.t, c.n, np.int32),
                  _arr(c.p, c.n * c.maxPar, np.float64, (c.n, c.maxPar)))
        lib().fos_cond_tab_free(ctypes.byref(c))
        return out

    def to_c(self):
        c = _capi.CondTabC()
        c.n = self.n
        c.maxPar = self.p.shape[1] if self.n else 0
        self._g = np.ascontiguousarray(self.gid, dtype=np.int32)
        self._t = np.ascontiguousarray(self.t, dtype=np.int32)
        self._p = np.ascontiguousarray(self.p, dtype=np.float64)
        c.gid = self._g.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        c.t = self._t.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        c.p = self._p.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return c


class UDFTab:
    """UDFTab%eval(k, x, t) over the library's evaluator (UDF.f90:75-92); k is 1-based."""

    def __init__(self, exprs):
        self._h = []
        for e in exprs:
            h = lib().evaluator_create(e.strip().encode())
            if not h:
                raise ValueError(_capi.last_error())
            self._h.append(h)

    @classmethod
    def read(cls, path):
        with open(path) as f:
            lines = f.read().splitlines()
        n = int(lines[0].split()[0])
        return cls(lines[1:1 + n])

    def eval(self, k, x, t):
        names = (ctypes.c_char_p * 4)(b"x", b"y", b"z", b"t")
        vals = (ctypes.c_double * 4)(float(x[0]), float(x[1]), float(x[2]), float(t))
        return lib().evaluator_evaluate(self._h[k - 1], 4, names, vals)

    def __del__(self):
        try:
            for h in self._h:
                lib().evaluator_destroy(h)
        except Exception:
            pass


def _ptr(t):
    """Device or host address of a torch tensor / numpy array."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)


class Euler:
    """Device context for foeuler's callbacks on one mesh (fos_euler_create)."""

    def __init__(self, grid, bc=None, r=1.0, gamma=1.4, flags=0, stream=None):
        self.grid = grid
        self._gc = grid.to_c()
        self._bc = (bc or CondTab()).to_c()
        h = lib().fos_euler_create(ctypes.byref(self._gc), ctypes.byref(self._bc), r, gamma, flags,
                                   ctypes.c_void_p(stream) if stream else None)
        if not h:
            raise RuntimeError(_capi.last_error())
        self.h = ctypes.c_void_p(h)
        self.nC = grid.nC
        self.flags = flags

    def close(self):
        if getattr(self, "h", None):
            lib().fos_euler_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rhs(self, y, ydot, t=0.0):
        check(lib().fos_euler_rhs(t, _ptr(y), _ptr(ydot), self.h))

    def jtimes(self, v, Jv, y, fy, t=0.0):
        check(lib().fos_euler_jtimes(_ptr(v), _ptr(Jv), t, _ptr(y), _ptr(fy), self.h, None))

    def rhs_f32(self, y, ydot):
        check(lib().fos_euler_rhs_f32(self.h, _ptr(y), _ptr(ydot)))

    def jtimes_f32(self, v, Jv, y, fy, sigma):
        check(lib().fos_euler_jtimes_f32(self.h, _ptr(v), _ptr(Jv), _ptr(y), _ptr(fy), float(sigma)))

    def jtimes_analytic(self, v, Jv, y, t=0.0):
        """Jv = R'(y) v exactly (dual numbers; FP32 arithmetic with PREC_FP32)."""
        check(lib().fos_euler_jtimes_analytic(_ptr(v), _ptr(Jv), t, _ptr(y), None, self.h, None))

    def newton_matvec(self, gamma, v, out, y, fy=None, analytic=False, t=0.0):
        """out = v - gamma J v in one pass (the difference quotient of jtimes, or the analytic
        product); device vectors, out distinct from v and y."""
        fn = lib().fos_euler_newton_matvec_analytic if analytic else lib().fos_euler_newton_matvec
        check(fn(gamma, _ptr(v), _ptr(out), t, _ptr(y), _ptr(fy), self.h))

    def sync_state(self, y, rho=None, rhou=None, rhoE=None, u=None, p=None, temp=None, c=None):
        """setBC + syncState (foeuler.f90:186-330) of y into element arrays of nE (rhou, u: 3 nE)."""
        check(lib().fos_euler_sync_state(self.h, _ptr(y), _ptr(rho), _ptr(rhou), _ptr(rhoE), _ptr(u), _ptr(p),
                                         _ptr(temp), _ptr(c)))

    def write_vtk(self, y, path):
        """writeState: legacy VTK of the grid with density, velocity, pressure, temperature."""
        gc = self.grid.to_c()
        check(lib().fos_euler_write_vtk(self.h, ctypes.byref(gc), _ptr(y), os.fsencode(path)))

    def psetup(self, y, fy, gamma, jok=0, t=0.0):
        jcur = ctypes.c_int(0)
        check(lib().fos_euler_psetup(t, _ptr(y), _ptr(fy), jok, ctypes.byref(jcur), gamma, self.h))
        return jcur.value

    def psolve(self, r, z, gamma, y=None, fy=None, delta=0.0, lr=1, t=0.0):
        check(lib().fos_euler_psolve(t, _ptr(y), _ptr(fy), _ptr(r), _ptr(z), gamma, delta, lr, self.h))

    def ilu_order(self):
        """Row elimination order of the block ILU (order[s] = cell eliminated at step s)."""
        order = np.zeros(self.nC, dtype=np.int32)
        check(lib().fos_euler_ilu_order(self.h, order.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return order

    def psetup_bjac(self, y, fy, gamma, jok=0, t=0.0):
        """Block Jacobi of M = I - gamma JacC/v (device vectors only)."""
        jcur = ctypes.c_int(0)
        check(lib().fos_euler_psetup_bjac(t, _ptr(y), _ptr(fy), jok, ctypes.byref(jcur), gamma, self.h))
        return jcur.value

    def psolve_bjac(self, r, z, gamma=0.0, t=0.0):
        check(lib().fos_euler_psolve_bjac(t, None, None, _ptr(r), _ptr(z), gamma, 0.0, 1, self.h))

    def set_tolerances(self, rtol, atol):
        check(lib().fos_euler_set_tolerances(self.h, rtol, atol))

    def nonfinite(self, clear=True):
        """True when a call since the last clear produced a NaN or Inf (the record of ASYNC_CALLBACKS
        contexts; synchronous contexts report it through the return code instead)."""
        rc = lib().fos_euler_nonfinite(self.h, 1 if clear else 0)
        if rc < 0:
            raise RuntimeError(_capi.last_error())
        return rc == 1

    def set_udf_bc(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().fos_euler_set_udf_bc(self.h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def find_euler(self, rho, rhou, rhoE, p, gamma, dRho, dRhou, dRhoE):
        check(lib().fos_find_euler(self.h, _ptr(rho), _ptr(rhou), _ptr(rhoE), _ptr(p), gamma,
                                   _ptr(dRho), _ptr(dRhou), _ptr(dRhoE)))

    def find_euler_jac(self, y, jac):
        check(lib().fos_find_euler_jac(self.h, _ptr(y), _ptr(jac)))

    # ---- learned preconditioner
    def learned_init(self, seed=0):
        check(lib().fos_learned_init(self.h, ctypes.c_uint64(seed)))

    def learned_inputs(self, y, gamma):
        X = np.zeros((self.nC, 20), dtype=np.float32)
        S = np.zeros(5, dtype=np.float32)
        check(lib().fos_learned_inputs(self.h, _ptr(y), gamma, X.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       S.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return X, S

    def learned_params(self):
        p = np.zeros(lib().fos_learned_nparams(), dtype=np.float32)
        check(lib().fos_learned_get_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return p

    def learned_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().fos_learned_set_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def learned_loss_grad(self, y, r, delta, gamma, start, size):
        g = np.zeros(lib().fos_learned_nparams(), dtype=np.float32)
        loss = ctypes.c_double(0.0)
        check(lib().fos_learned_loss_grad(self.h, _ptr(y), _ptr(r), _ptr(delta), gamma, start, size,
                                          ctypes.byref(loss), g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return loss.value, g

    def learned_train_step(self, y, r, delta, gamma, start, size, lr=1e-3):
        loss = ctypes.c_double(0.0)
        check(lib().fos_learned_train_step(self.h, _ptr(y), _ptr(r), _ptr(delta), gamma, start, size, lr,
                                           ctypes.byref(loss)))
        return loss.value

    def psetup_learned(self, y, gamma):
        jcur = ctypes.c_int(0)
        check(lib().fos_euler_psetup_learned(0.0, _ptr(y), _ptr(y), 0, ctypes.byref(jcur), gamma, self.h))

    def psolve_learned(self, r, z, gamma=0.0):
        check(lib().fos_euler_psolve_learned(0.0, None, None, _ptr(r), _ptr(z), gamma, 0.0, 1, self.h))

    @property
    def layout(self):
        """1: tetrahedral layout (face geometry from node coordinates, the default where it applies),
        0: generic (LAYOUT_GENERIC, or meshes the tetrahedral layout does not cover)."""
        return lib().fos_euler_layout(self.h)

    def counters(self):
        a = (ctypes.c_longlong * 4)()
        check(lib().fos_euler_counters(self.h, a))
        return list(a)


class ILU:
    """The reference's ILU type (sparse.f90:52-63): init(nEq, maxNNz, lFill), setCOO, fact(eps),
    solve(rhs, x), mulVec; 0-based indices.  Vectors: torch device tensors, or numpy arrays with
    host=True."""

    def __init__(self, nEq, maxNNz, lFill, host=False, stream=None):
        h = lib().fos_ilut_create(nEq, maxNNz, lFill, VECTORS_ON_HOST if host else 0,
                                  ctypes.c_void_p(stream) if stream else None)
        if not h:
            raise RuntimeError(_capi.last_error())
        self.h = ctypes.c_void_p(h)
        self.n = nEq

    def set_coo(self, iA, jA, A):
        iA, jA = np.ascontiguousarray(iA, np.int32), np.ascontiguousarray(jA, np.int32)
        A = np.ascontiguousarray(A, np.float64)
        check(lib().fos_ilut_set_coo(self.h, _ptr(iA), _ptr(jA), _ptr(A), len(A)))

    def set_csr(self, ia, ja, a):
        ia, ja = np.ascontiguousarray(ia, np.int32), np.ascontiguousarray(ja, np.int32)
        a = np.ascontiguousarray(a, np.float64)
        check(lib().fos_ilut_set_csr(self.h, _ptr(ia), _ptr(ja), _ptr(a)))

    def fact(self, eps=0.0):
        check(lib().fos_ilut_fact(self.h, float(eps)))

    def solve(self, rhs, x):
        check(lib().fos_ilut_solve(self.h, _ptr(rhs), _ptr(x)))

    def mul_vec(self, x, y):
        check(lib().fos_ilut_mul_vec(self.h, _ptr(x), _ptr(y)))

    def msr(self):
        """(alu, jlu, ju): SPARSKIT's MSR arrays of the factor, 1-based values (Fortran position k
        at index k - 1)."""
        m = lib().fos_ilut_get_msr(self.h, None, None, None)
        if m < 0:
            raise RuntimeError(_capi.last_error())
        alu, jlu, ju = np.zeros(m), np.zeros(m, np.int32), np.zeros(self.n, np.int32)
        check(0 if lib().fos_ilut_get_msr(self.h, _ptr(alu), _ptr(jlu), _ptr(ju)) == m else -1)
        return alu, jlu, ju

    def levels(self):
        f, b = ctypes.c_int(0), ctypes.c_int(0)
        check(lib().fos_ilut_levels(self.h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    def close(self):
        if getattr(self, "h", None):
            lib().fos_ilut_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fnaddr(name):
    return ctypes.cast(getattr(lib(), name), ctypes.c_void_p)


class BDFNewtonKrylov:
    """BDFNewtonKrylov (SUNDIALS.f90:65-74) on the library's device-resident integrator.

    BDFNewtonKrylov(nEq, f, maxl, pSet, pSol, user_data): f / pSet / pSol are C function
    addresses (e.g. Euler callbacks) or, for f only, a Python callable f(t, y, ydot) on host numpy
    arrays (host-vector mode, small systems)."""

    def __init__(self, nEq, f, maxl=0, pSet=None, pSol=None, user_data=None, jtimes=None, stream=None):
        self._keep = []
        flags = 0
        if callable(f) and not isinstance(f, ctypes.c_void_p):
            n = nEq

            def cb(t, yp, ydp, ud):
                y = np.ctypeslib.as_array(yp, shape=(n,))
                yd = np.ctypeslib.as_array(ydp, shape=(n,))
                f(t, y, yd)
                return 0
            fp = _capi.RHS_FN(cb)
            self._keep.append(fp)
            faddr = ctypes.cast(fp, ctypes.c_void_p)
            flags |= VECTORS_ON_HOST
        else:
            faddr = f
        h = lib().fos_bdf_create(nEq, faddr, maxl, pSet, pSol, user_data)
        if not h:
            raise RuntimeError(_capi.last_error())
        self.h = ctypes.c_void_p(h)
        self.n = nEq
        self.host = bool(flags & VECTORS_ON_HOST)
        check(lib().fos_bdf_set_options(self.h, flags, ctypes.c_void_p(stream) if stream else None))
        if jtimes is not None:
            check(lib().fos_bdf_set_jtimes(self.h, jtimes))

    @classmethod
    def for_euler(cls, euler, maxl=20, precond="ilu", jtimes=True, stream=None, fallback_cap=0, fused=True,
                  krylov_fp32=None):
        """The foeuler set on one Euler context: rhs, jtimes (True / "dq": the fused difference
        quotient; "analytic": the exact dual-number product; False: the integrator's own
        difference quotient) and a preconditioner --
        "ilu" (block ILU(0)), "bjac" (block Jacobi), "learned" (with block ILU as fallback after
        fallback_cap GMRES iterations when fallback_cap > 0) or None.  fused: GMRES forms
        v - gamma J v in the Jv kernels' epilogue (fos_bdf_set_newton_matvec) instead of a second
        vector pass.  krylov_fp32: the GMRES basis stored in FP32 (FOS_PREC_FP32 on the integrator);
        None follows the context's precision."""
        if precond is True:
            precond = "ilu"
        ps, pv = {"ilu": ("fos_euler_psetup", "fos_euler_psolve"),
                  "bjac": ("fos_euler_psetup_bjac", "fos_euler_psolve_bjac"),
                  "learned": ("fos_euler_psetup_learned", "fos_euler_psolve_learned")}.get(precond, (None, None))
        jt = {True: "fos_euler_jtimes", "dq": "fos_euler_jtimes", "analytic": "fos_euler_jtimes_analytic"}.get(jtimes)
        b = cls(5 * euler.nC, _fnaddr("fos_euler_rhs"), maxl, _fnaddr(ps) if ps else None,
                _fnaddr(pv) if pv else None, euler.h, _fnaddr(jt) if jt else None, stream)
        if precond == "learned" and fallback_cap > 0:
            check(lib().fos_bdf_set_fallback(b.h, fallback_cap, _fnaddr("fos_euler_psetup"),
                                             _fnaddr("fos_euler_psolve")))
        nm = {"fos_euler_jtimes": "fos_euler_newton_matvec",
              "fos_euler_jtimes_analytic": "fos_euler_newton_matvec_analytic"}.get(jt)
        if fused and nm and not (euler.flags & VECTORS_ON_HOST):
            check(lib().fos_bdf_set_newton_matvec(b.h, _fnaddr(nm)))
        if krylov_fp32 is None:
            krylov_fp32 = bool(euler.flags & PREC_FP32)
        if krylov_fp32:
            check(lib().fos_bdf_set_options(b.h, (VECTORS_ON_HOST if b.host else 0) | PREC_FP32,
                                            ctypes.c_void_p(stream) if stream else None))
        return b

    def last_triple(self, ypred, r, delta):
        check(lib().fos_bdf_last_triple(self.h, _ptr(ypred), _ptr(r), _ptr(delta)))

    def set_tol(self, rtol, atol):
        check(lib().fos_bdf_set_tol(self.h, rtol, atol))

    def set_max_steps(self, n):
        check(lib().fos_bdf_set_max_steps(self.h, n))

    def set_newton(self, tol=0.0, max_iter=0, lin_tol=0.0, max_restarts=-1, jac_interval=0):
        check(lib().fos_bdf_set_newton(self.h, tol, max_iter, lin_tol, max_restarts, jac_interval))

    def set_lsetup_frequency(self, msbp=0, msbj=0, dgmax=0.0):
        """CVODE's preconditioner setup control: setup every msbp steps (20) or when gamma moves by
        more than dgmax (0.3); a new Jacobian every msbj steps (51)."""
        check(lib().fos_bdf_set_lsetup_frequency(self.h, msbp, msbj, dgmax))

    def set_dt(self, dt):
        """Fixed steps of size dt (BDF1 then BDF2) instead of the default adaptive stepping."""
        check(lib().fos_bdf_set_dt(self.h, dt))

    def set_adaptive(self, max_order=0, h0=0.0, hmin=0.0, hmax=0.0):
        """CVODE's variable-order variable-step BDF (the default); effective at the next set_iv."""
        check(lib().fos_bdf_set_adaptive(self.h, max_order, h0, hmin, hmax))

    def set_iv(self, t, y):
        check(lib().fos_bdf_set_iv(self.h, t, _ptr(y)))

    def step(self, tOut, y):
        """One internal step (it may pass tOut); y receives the state. Returns (rc, t): rc 0 ok,
        >0 the step failed (state unchanged)."""
        t = ctypes.c_double(0.0)
        rc = lib().fos_bdf_step(self.h, tOut, ctypes.byref(t), _ptr(y))
        if rc < 0:
            raise RuntimeError(_capi.last_error())
        return rc, t.value

    def get_dt(self):
        d = ctypes.c_double(0.0)
        check(lib().fos_bdf_get_dt(self.h, ctypes.byref(d)))
        return d.value

    def get_dky(self, t, k, out):
        """CVodeGetDky: k-th derivative of the interpolant at t within the last adaptive step."""
        check(lib().fos_bdf_get_dky(self.h, t, k, _ptr(out)))

    def stats(self):
        st = _capi.BdfStatsC()
        check(lib().fos_bdf_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().fos_bdf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

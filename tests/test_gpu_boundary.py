"""The C-ABI boundary on the GPU: the callbacks with SUNDIALS' N_Vector signatures against the
raw-pointer ones, and non-finite results reported as recoverable failures (return 1), with the
integrator retrying the step.

The N_Vector here is a stand-in struct whose first field is the data pointer, read by an accessor
registered like N_VGetArrayPointer (SUNDIALS is absent from this image); the *_nv entry points must
give bitwise the results of the raw entry points they forward to.
"""
import ctypes

import numpy as np
import pytest

import fosolvers_amd as fa
from fosolvers_amd import _capi
from oracle import oracle as O
from tests import meshes as M

pytestmark = pytest.mark.gpu


class FakeNVector(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("length", ctypes.c_long)]


@_capi.NV_DATA_FN
def _nv_data(nv):
    return ctypes.cast(nv, ctypes.POINTER(FakeNVector)).contents.data


def _nv(t):
    v = FakeNVector(t.data_ptr() if hasattr(t, "data_ptr") else t.ctypes.data, t.shape[0])
    return v


def _addr(x):
    return ctypes.c_void_p(ctypes.addressof(x))


def _setup(kind="mixed", flags=fa.LAYOUT_GENERIC):
    """The generic residual layout sums each cell's faces in a fixed order, so two calls give
    bitwise the same result and the forwarding of the *_nv entries can be checked bitwise."""
    pg, fv, og, bc, ctab = M.setup(8, 5, 4, kind=kind)
    x = M.perturbed_state(og.p[:og.nC], mach=0.6, seed=11)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
    return og, bc, x, eu


def test_nvector_callbacks_equal_the_raw_ones(cuda):
    import torch
    print("[case] nv callbacks, device vectors", flush=True)
    og, bc, x, eu = _setup()
    L = fa.lib()
    n = 5 * og.nC
    y = torch.tensor(x, device=cuda)
    rng = np.random.default_rng(3)
    v = torch.tensor(rng.standard_normal(n) * 1e-3, device=cuda)
    fy, fy2 = torch.zeros_like(y), torch.zeros_like(y)
    # no accessor registered yet: refused with a message
    ny, nfy = _nv(y), _nv(fy2)
    assert L.fos_euler_rhs_nv(0.0, _addr(ny), _addr(nfy), eu.h) == -1
    assert "accessor" in _capi.last_error()
    assert L.fos_euler_set_nvector_access(eu.h, ctypes.cast(_nv_data, ctypes.c_void_p)) == 0
    eu.rhs(y, fy)
    assert L.fos_euler_rhs_nv(0.0, _addr(ny), _addr(nfy), eu.h) == 0
    torch.cuda.synchronize()
    assert torch.equal(fy, fy2)
    # the residual itself is the oracle's (the raw entry's parity, restated through the N_Vector path)
    ref = O.rhs(og, bc, x)
    assert M.rel(fy2.cpu().numpy(), ref) <= 1e-12
    for name, raw in (("fos_euler_jtimes_nv", "fos_euler_jtimes"),
                      ("fos_euler_jtimes_analytic_nv", "fos_euler_jtimes_analytic")):
        j1, j2 = torch.zeros_like(y), torch.zeros_like(y)
        assert getattr(L, raw)(fa._ptr(v), fa._ptr(j1), 0.0, fa._ptr(y), fa._ptr(fy), eu.h, None) == 0
        nv_, nj, ny2, nf2 = _nv(v), _nv(j2), _nv(y), _nv(fy)
        assert getattr(L, name)(_addr(nv_), _addr(nj), 0.0, _addr(ny2), _addr(nf2), eu.h, None) == 0
        torch.cuda.synchronize()
        assert torch.equal(j1, j2), name
    gamma = 2e-3
    jc1, jc2 = ctypes.c_int(0), ctypes.c_int(0)
    assert L.fos_euler_psetup(0.0, fa._ptr(y), fa._ptr(fy), 0, ctypes.byref(jc1), gamma, eu.h) == 0
    z1 = torch.zeros_like(y)
    assert L.fos_euler_psolve(0.0, fa._ptr(y), fa._ptr(fy), fa._ptr(v), fa._ptr(z1), gamma, 0.0, 1, eu.h) == 0
    ny3, nf3 = _nv(y), _nv(fy)
    assert L.fos_euler_psetup_nv(0.0, _addr(ny3), _addr(nf3), 0, ctypes.byref(jc2), gamma, eu.h) == 0
    assert jc1.value == jc2.value == 1
    z2 = torch.zeros_like(y)
    nr, nz = _nv(v), _nv(z2)
    assert L.fos_euler_psolve_nv(0.0, _addr(ny3), _addr(nf3), _addr(nr), _addr(nz), gamma, 0.0, 1, eu.h) == 0
    torch.cuda.synchronize()
    assert torch.equal(z1, z2)
    # an accessor that returns NULL is an error, not a dereference
    bad = FakeNVector(None, n)
    assert L.fos_euler_rhs_nv(0.0, _addr(bad), _addr(nfy), eu.h) == -1
    assert "NULL" in _capi.last_error()


def test_nvector_callbacks_on_host_vectors(cuda):
    """NVECTOR_SERIAL's case: host data behind the N_Vector, a FOS_VECTORS_ON_HOST context."""
    print("[case] nv callbacks, host vectors", flush=True)
    og, bc, x, eu = _setup(flags=fa.VECTORS_ON_HOST | fa.LAYOUT_GENERIC)
    L = fa.lib()
    L.fos_euler_set_nvector_access(eu.h, ctypes.cast(_nv_data, ctypes.c_void_p))
    y = np.ascontiguousarray(x)
    yd = np.zeros_like(y)
    ny, nyd = _nv(y), _nv(yd)
    assert L.fos_euler_rhs_nv(0.0, _addr(ny), _addr(nyd), eu.h) == 0
    assert M.rel(yd, O.rhs(og, bc, x)) <= 1e-12


def test_nonfinite_results_return_recoverable(cuda):
    import torch
    print("[case] non-finite results", flush=True)
    og, bc, x, eu = _setup(flags=0)  # the default (tetrahedral) layout: ghost kernel, LDS sums
    L = fa.lib()
    n = 5 * og.nC
    y = torch.tensor(x, device=cuda)
    fy = torch.zeros_like(y)
    ok = torch.zeros_like(y)
    assert L.fos_euler_rhs(0.0, fa._ptr(y), fa._ptr(fy), eu.h) == 0
    bad = y.clone()
    bad[og.nC // 2] = float("nan")  # one cell's density
    out = torch.zeros_like(y)
    assert L.fos_euler_rhs(0.0, fa._ptr(bad), fa._ptr(out), eu.h) == 1
    assert "non-finite" in _capi.last_error()
    # the record is cleared: the next finite call returns 0
    assert L.fos_euler_rhs(0.0, fa._ptr(y), fa._ptr(ok), eu.h) == 0
    torch.cuda.synchronize()
    assert float((ok - fy).norm() / fy.norm()) <= 1e-14
    # a negative density: the pressure's square root makes the residual non-finite too
    neg = y.clone()
    neg[3] = -0.5
    assert L.fos_euler_rhs(0.0, fa._ptr(neg), fa._ptr(out), eu.h) == 1
    v = torch.zeros_like(y)
    v[7] = float("inf")
    jv = torch.zeros_like(y)
    assert L.fos_euler_jtimes_analytic(fa._ptr(v), fa._ptr(jv), 0.0, fa._ptr(y), fa._ptr(fy), eu.h, None) == 1
    assert L.fos_euler_jtimes(fa._ptr(bad), fa._ptr(jv), 0.0, fa._ptr(y), fa._ptr(fy), eu.h, None) == 1
    jc = ctypes.c_int(0)
    assert L.fos_euler_psetup(0.0, fa._ptr(y), fa._ptr(fy), 0, ctypes.byref(jc), 1e-3, eu.h) == 0
    r = torch.ones_like(y)
    r[5] = float("nan")
    z = torch.zeros_like(y)
    assert L.fos_euler_psolve(0.0, fa._ptr(y), fa._ptr(fy), fa._ptr(r), fa._ptr(z), 1e-3, 0.0, 1, eu.h) == 1
    assert L.fos_euler_psolve(0.0, fa._ptr(y), fa._ptr(fy), fa._ptr(ok), fa._ptr(z), 1e-3, 0.0, 1, eu.h) == 0
    # FOS_ASYNC_CALLBACKS: the calls do not wait and return 0; the record says what happened
    pg, fv, og2, bc2, ctab = M.setup(8, 5, 4, kind="mixed")
    ea = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=fa.ASYNC_CALLBACKS)
    assert L.fos_euler_rhs(0.0, fa._ptr(y), fa._ptr(out), ea.h) == 0
    assert ea.nonfinite() is False
    assert L.fos_euler_rhs(0.0, fa._ptr(bad), fa._ptr(out), ea.h) == 0
    assert L.fos_euler_rhs(0.0, fa._ptr(y), fa._ptr(out), ea.h) == 0
    assert ea.nonfinite(clear=False) is True
    assert ea.nonfinite() is True  # read and cleared
    assert ea.nonfinite() is False
    assert n == 5 * og2.nC


def test_integrator_retries_after_a_nonfinite_residual(cuda):
    """A first step far too large for the Sod tube drives the Newton iterate to negative density: the
    residual is non-finite, the callback returns 1 and the adaptive integrator (CVODE's method) cuts
    the step and retries until it converges.  In fixed-step mode the step fails with the state
    unchanged, and the host retries with a smaller dt (fopbc's way)."""
    import torch
    print("[case] integrator retry after non-finite", flush=True)
    pg, fv, og, bc, ctab = M.setup(16, 3, 3, kind="wall", jitter=0.0, lx=1.0, ly=0.25, lz=0.25)
    x = M.sod_state(og.p[:og.nC])
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    y0 = torch.tensor(x, device=cuda)
    # the scenario: y0 + h0 f(y0) (the first predictor) has negative densities, a non-finite residual
    h0 = 0.5
    f0 = torch.zeros_like(y0)
    eu.rhs(y0, f0)
    pred = y0 + h0 * f0
    assert float(pred[:og.nC].min()) < 0.0
    out = torch.zeros_like(y0)
    assert fa.lib().fos_euler_rhs(0.0, fa._ptr(pred), fa._ptr(out), eu.h) == 1
    t_end = 0.02
    b = fa.BDFNewtonKrylov.for_euler(eu, maxl=20, precond="ilu")
    b.set_tol(1e-6, 1e-12)
    b.set_adaptive(h0=h0)
    y = y0.clone()
    b.set_iv(0.0, y)
    t, k = 0.0, 0
    while t < t_end and k < 2000:
        rc, t = b.step(t_end, y)
        assert rc == 0
        k += 1
    st = b.stats()
    assert st["nncfails"] >= 1, st  # the non-finite attempts were retried with smaller steps
    yT = torch.zeros_like(y)
    b.get_dky(t_end, 0, yT)
    assert torch.isfinite(yT).all()
    b.close()
    # the same interval from the estimated first step (no failure): same state to the tolerance
    b2 = fa.BDFNewtonKrylov.for_euler(eu, maxl=20, precond="ilu")
    b2.set_tol(1e-6, 1e-12)
    y2 = y0.clone()
    b2.set_iv(0.0, y2)
    t = 0.0
    while t < t_end:
        rc, t = b2.step(t_end, y2)
        assert rc == 0
    yR = torch.zeros_like(y2)
    b2.get_dky(t_end, 0, yR)
    b2.close()
    assert float((yT - yR).norm() / yR.norm()) <= 5e-4
    # fixed steps: the oversized step fails with the state unchanged; a smaller one succeeds
    b3 = fa.BDFNewtonKrylov.for_euler(eu, maxl=20, precond="ilu")
    b3.set_newton(tol=1e-8, max_iter=3, lin_tol=1e-3, max_restarts=5)
    b3.set_dt(h0)
    y3 = y0.clone()
    b3.set_iv(0.0, y3)
    rc, _ = b3.step(0.0, y3)
    assert rc == 1  # a non-finite residual, or no convergence in 3 Newton iterations
    assert torch.equal(y3, y0)
    b3.set_newton(max_iter=20)
    b3.set_dt(1e-4)
    rc, t3 = b3.step(0.0, y3)
    assert rc == 0 and abs(t3 - 1e-4) < 1e-15
    ref, _ = O.bdf_euler(og, bc, x, 1e-4, 1, lin_tol=1e-3)
    assert M.rel(y3.cpu().numpy(), ref[0]) <= 1e-9
    b3.close()


def _read_vtk(path):
    """A small reader of the legacy VTK layout: points, cells, types and the CELL_DATA fields."""
    lines = open(path).read().split("\n")
    out, i = {}, 0
    while i < len(lines):
        w = lines[i].split()
        if not w:
            i += 1
            continue
        if w[0] == "POINTS":
            n = int(w[1])
            out["points"] = np.array([[float(t) for t in lines[i + 1 + k].split()] for k in range(n)])
            i += 1 + n
        elif w[0] == "CELLS":
            n = int(w[1])
            out["cells"] = [[int(t) for t in lines[i + 1 + k].split()] for k in range(n)]
            i += 1 + n
        elif w[0] == "CELL_TYPES":
            n = int(w[1])
            out["types"] = np.array([int(lines[i + 1 + k]) for k in range(n)])
            i += 1 + n
        elif w[0] == "CELL_DATA":
            out["nE"] = int(w[1])
            i += 1
        elif w[0] == "SCALARS":
            n = out["nE"]
            out[w[1]] = np.array([float(lines[i + 2 + k]) for k in range(n)])
            i += 2 + n
        elif w[0] == "VECTORS":
            n = out["nE"]
            out[w[1]] = np.array([[float(t) for t in lines[i + 1 + k].split()] for k in range(n)])
            i += 1 + n
        else:
            i += 1
    return out


@pytest.mark.parametrize("host", [False, True])
def test_write_vtk_matches_the_oracle_state(cuda, tmp_path, host):
    """writeState through the library: the nodes and elements of the grid, and density, velocity,
    pressure and temperature per element (cells from y, boundary facets from their ghost states)
    equal to the oracle's setBC + syncState."""
    import torch
    print(f"[case] write vtk host={host}", flush=True)
    pg, fv, og, bc, ctab = M.setup(6, 4, 3, kind="mixed")
    x = M.perturbed_state(og.p[:og.nC], mach=0.6, seed=4)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=fa.VECTORS_ON_HOST if host else 0)
    y = np.ascontiguousarray(x) if host else torch.tensor(x, device=cuda)
    path = str(tmp_path / "rst_0.vtk")
    eu.write_vtk(y, path)
    v = _read_vtk(path)
    assert np.array_equal(v["points"], fv.pN)
    assert len(v["cells"]) == fv.nE and v["nE"] == fv.nE
    for e in (0, fv.nC - 1, fv.nC, fv.nE - 1):
        assert v["cells"][e] == [int(fv.nNE[e])] + [int(q) for q in fv.iNE[e, :fv.nNE[e]]]
    assert set(v["types"][:fv.nC]) == {10} and set(v["types"][fv.nC:]) == {5}
    rho, rhou, rhoE, u, p, temp, c = O.element_state(og, bc, x, full=True)
    has = np.zeros(fv.nE, bool)  # elements the state defines: cells and the facets of boundary pairs
    has[:fv.nC] = True
    has[fv.iEP[fv.nPi:, 1]] = True
    assert M.rel(v["density"][has], rho[has]) <= 1e-14
    assert M.rel(v["pressure"][has], p[has]) <= 1e-14
    assert M.rel(v["temperature"][has], temp[has]) <= 1e-14
    assert M.rel(v["velocity"][has], u.reshape(-1, 3)[has]) <= 1e-14

"""foeuler through its Fortran host on the library (fosolvers_amd/foeuler_amd): a case directory
in foeuler's formats (grid.msh, bc, ic, sim, fl, udf; src/foeuler/foeuler.f90:60-112) is run to
tFinal, and the rst_N.vtk files are checked against the oracle's adaptive integration of the same
mesh, conditions and initial state, and UDF conditions against their constant equivalents."""
import os
import re
import subprocess

import numpy as np
import pytest

import fosolvers_amd as fa
from fosolvers_amd import build as B
from oracle import oracle as O
from tests import meshes as M

pytestmark = pytest.mark.gpu


def _driver():
    if not os.path.exists(B.FOEULER):
        pytest.fail("fosolvers_amd/foeuler_amd is not built (amdflang build step)")
    return B.FOEULER


def _cond_file(path, gids, types, params):
    with open(path, "w") as f:
        f.write(f"{len(gids)}  # number of conditions\n5  # maximum number of parameters\n")
        for g, t, p in zip(gids, types, params):
            f.write(f"\n{g}\n{t}\n{len(p)}\n")
            for v in p:
                f.write(f"{v!r}\n")


def _write(path, lines):
    with open(path, "w") as f:
        f.write("\n".join(str(v) for v in lines) + "\n")


def _run(d, env=None):
    print(f"[case] foeuler_amd in {d} {env or ''}", flush=True)
    r = subprocess.run([_driver()], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=240, env=dict(os.environ, **(env or {})))
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:]
    return r.stdout


def _cell_data(path):
    """name -> array of the CELL_DATA section of a legacy VTK file."""
    lines = open(path).read().split("\n")
    i = next(k for k, l in enumerate(lines) if l.startswith("CELL_DATA"))
    n = int(lines[i].split()[1])
    out = {}
    k = i + 1
    while k < len(lines):
        l = lines[k]
        if l.startswith("SCALARS"):
            name = l.split()[1]
            out[name] = np.array([float(v) for v in lines[k + 2:k + 2 + n]])
            k += 2 + n
        elif l.startswith("VECTORS"):
            name = l.split()[1]
            out[name] = np.array([[float(v) for v in s.split()] for s in lines[k + 1:k + 1 + n]])
            k += 1 + n
        else:
            k += 1
    return out


def _steps_before_outputs(stdout):
    """number of steps taken when each rst_N (N >= 1) was written"""
    counts, nst = [], 0
    for l in stdout.split("\n"):
        if l.startswith("[i] finished time step"):
            nst += 1
        m = re.match(r"\[i\] writing: rst_(\d+)\.vtk", l)
        if m:
            counts.append(nst)
    return counts


def test_foeuler_driver_sod_matches_oracle(cuda, tmp_path):
    """Sod tube (1,296 tets, all walls) with a UDF initial condition (step functions), to t = 0.02
    with output every 0.01: rst_1 and rst_2 are written at the oracle's step counts and their cell
    and ghost densities, velocities and pressures agree with the oracle's adaptive BDF + block
    ILU(0) run (same elimination order) to 1e-6 relative."""
    d = str(tmp_path)
    pg = M.box(12, 3, 3, lx=1.0, ly=0.25, lz=0.25)
    pg.write_gmsh(os.path.join(d, "grid.msh"))
    _cond_file(os.path.join(d, "bc"), [1, 2, 3, 4, 5, 6], [0] * 6, [[]] * 6)
    _write(os.path.join(d, "ic"), [1, 1, 2, 3, 4, 5])
    _write(os.path.join(d, "udf"), [5, "1-0.9*step(x-0.5)", "(1-0.9*step(x-0.5))/(1-0.875*step(x-0.5))", "0", "0",
                                    "0"])
    _write(os.path.join(d, "sim"), [0.02, 0.01])
    _write(os.path.join(d, "fl"), [M.R, M.GAMMA])
    out = _run(d)
    # the oracle on the same file
    pg2 = fa.PolyGrid.read_gmsh(os.path.join(d, "grid.msh"))
    og = O.Grid(pg2.pN, pg2.sE, pg2.nNE, pg2.iNE, pg2.gid)
    fv = pg2.up()
    cg, ct, cp = [1, 2, 3, 4, 5, 6], [0] * 6, np.zeros((6, 5))
    bc = O.BC(og, cg, ct, cp, r=M.R, gamm=M.GAMMA)
    eu = fa.Euler(fv, fa.CondTab(cg, ct, cp), r=M.R, gamma=M.GAMMA)
    nC, nE = og.nC, og.nE
    xc = og.p[:nC, 0]
    s = np.where(xc - 0.5 >= 0, 1.0, 0.0)
    p0 = 1 - 0.9 * s
    T0 = p0 / (1 - 0.875 * s)
    rho = p0 / M.R / T0
    x0 = M.cons(rho, np.zeros((nC, 3)), p0)
    ts, hs, qs, st, stats = O.cv_euler(og, bc, x0, 0.01, 0.02, 4000, rtol=1e-6, atol=1e-12, maxl=5, restarts=5,
                                       order=eu.ilu_order())
    counts = _steps_before_outputs(out)
    ref_counts = [int(np.argmax(ts + 1e-300 >= tn)) + 1 for tn in (0.01, 0.02)]
    print("outputs after steps", counts, "oracle", ref_counts, "oracle stats", stats)
    assert counts == ref_counts
    ic = _cell_data(os.path.join(d, "rst_0.vtk"))
    assert M.rel(ic["density"][:nC], rho) <= 1e-15
    for k, n in zip((1, 2), ref_counts):
        got = _cell_data(os.path.join(d, f"rst_{k}.vtk"))
        ref = O.element_state(og, bc, st[n - 1], full=True)
        assert M.rel(got["density"], ref[0]) <= 1e-6, k
        assert M.rel(got["pressure"], ref[4]) <= 1e-6, k
        assert M.rel(got["velocity"].ravel(), ref[3]) <= 1e-5, k
        assert got["density"].shape == (nE,)
    # the same case with the exact Jv and Jacobian in mixed precision: the run completes with its
    # own step sequence, so rst_2 holds the state of a slightly different step time than the
    # oracle's (the driver writes the step that passes the output time, as foeuler does): a sanity
    # bound on the moving shock tube, not a parity check
    fast = _run(d, {"FOS_FP32": "1", "FOS_JV_ANALYTIC": "1", "FOS_JAC_EXACT": "1"})
    n2 = _steps_before_outputs(fast)[-1]
    got = _cell_data(os.path.join(d, "rst_2.vtk"))
    ref = O.element_state(og, bc, st[ref_counts[-1] - 1], full=True)
    print(f"FP32 + analytic Jv + exact Jacobian: rst_2 after {n2} steps, density vs the oracle "
          f"{M.rel(got['density'], ref[0]):.2e}")
    assert M.rel(got["density"], ref[0]) <= 5e-2


def test_foeuler_driver_udf_conditions_equal_constants(cuda, tmp_path):
    """Far-field conditions given as UDFs (code 130, evaluated per face at the step time) produce
    the same output as the constant far field (code 30) with the same values; the other sides
    default to walls (unassigned groups).  With the generic layout (fixed-order flux sums) the
    files agree bit for bit; with the tetrahedral layout (LDS atomic flux sums, whose order varies
    from run to run) the step counts agree and the fields to 1e-7 relative: rounding-level changes
    in the residual move the Newton and GMRES iterates anywhere inside their stopping tolerances
    (RTOL 1e-6), so the solver tolerances, not rounding, bound the difference (2e-10 to 1.1e-8 seen
    on the velocity, whose transverse components are near zero)."""
    outs = {}
    for layout in ("generic", "tet"):
        for udf in (False, True):
            d = str(tmp_path / f"{layout}_{'udf' if udf else 'const'}")
            os.makedirs(d)
            M.box(8, 3, 3, lx=1.0, ly=0.25, lz=0.25, jitter=0.15, seed=4).write_gmsh(os.path.join(d, "grid.msh"))
            if udf:
                _cond_file(os.path.join(d, "bc"), [1, 2], [130, 130], [[1, 2, 3, 4, 5]] * 2)
                _write(os.path.join(d, "udf"), [5] + [repr(v) for v in M.FREE])
            else:
                _cond_file(os.path.join(d, "bc"), [1, 2], [30, 30], [M.FREE] * 2)
            _write(os.path.join(d, "ic"), [0, 0.8, 0.7, 0.3, 0.0, 0.0])
            _write(os.path.join(d, "sim"), [0.01, 0.005])
            _write(os.path.join(d, "fl"), [M.R, M.GAMMA])
            out = _run(d, {"FOS_LAYOUT_GENERIC": "1"} if layout == "generic" else None)
            outs[layout, udf] = (_steps_before_outputs(out), [os.path.join(d, f"rst_{k}.vtk") for k in (1, 2)])
    assert outs["generic", False][0] == outs["generic", True][0]
    for a, b in zip(outs["generic", False][1], outs["generic", True][1]):
        assert open(a).read() == open(b).read()
    assert outs["tet", False][0] == outs["tet", True][0]
    for a, b in zip(outs["tet", False][1], outs["tet", True][1]):
        ca, cb = _cell_data(a), _cell_data(b)
        for key in ("density", "pressure", "temperature"):
            assert M.rel(ca[key], cb[key]) <= 1e-7, key
        assert M.rel(ca["velocity"].ravel(), cb["velocity"].ravel()) <= 1e-7

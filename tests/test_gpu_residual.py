"""GPU parity of the Roe residual (rhs), findEuler and the DQ Jacobian-vector product against the
FP64 CPU oracle on the same meshes and states.

Tolerances: FP64 residual within 1e-12 relative (2-norm over all 5 nC entries, and max-abs error
within 1e-12 of the largest entry); FP32 residual within 2e-5 relative.  The DQ Jv amplifies the
residual's rounding by 1/sigma (sigma ~ 1e-6 relative perturbation), so FP64 Jv is compared at
1e-7 relative and must converge to the analytic directional derivative.
"""
import numpy as np
import pytest

import fosolvers_amd as fa
from oracle import oracle as O
from tests import meshes as M

pytestmark = pytest.mark.gpu

# both residual layouts: tetrahedral (geometry from node coordinates) and generic (per-face arrays)
LAYOUTS = pytest.mark.parametrize("layout", [1, 0], ids=["tet", "generic"])


def _euler(fv, ctab, layout, flags=0):
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags | (0 if layout else fa.LAYOUT_GENERIC))
    assert eu.layout == layout
    return eu


def _dev(torch, a, dtype=None, cuda=None):
    return torch.tensor(np.asarray(a), dtype=dtype or torch.float64, device=cuda)


@pytest.mark.parametrize("kind,dims,jitter", [
    ("far", (6, 4, 4), 0.2), ("mixed", (8, 6, 6), 0.2), ("wall", (10, 4, 4), 0.0),
    ("supersonic", (12, 6, 5), 0.15), ("mixed", (20, 10, 10), 0.2)])
@LAYOUTS
def test_rhs_fp64_matches_oracle(cuda, kind, dims, jitter, layout):
    import torch
    print(f"[case] rhs fp64 kind={kind} dims={dims} layout={layout}", flush=True)
    pg, fv, og, bc, ctab = M.setup(*dims, kind=kind, jitter=jitter)
    x = M.perturbed_state(og.p[:og.nC], mach=0.6 if kind != "supersonic" else 2.0, seed=5)
    if kind == "wall":
        x = M.sod_state(og.p[:og.nC])
    ref = O.rhs(og, bc, x)
    eu = _euler(fv, ctab, layout)
    y = _dev(torch, x, cuda=cuda)
    yd = torch.zeros_like(y)
    eu.rhs(y, yd)
    torch.cuda.synchronize()
    got = yd.cpu().numpy()
    print(f"fp64 residual relative error vs oracle: {M.rel(got, ref):.3e}")
    assert M.rel(got, ref) <= 1e-12, M.rel(got, ref)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # run to run: the generic layout sums each cell's faces in a fixed order (bitwise the same); the
    # tetrahedral layout's LDS atomic sums vary in order, so it agrees to rounding
    yd2 = torch.zeros_like(y)
    eu.rhs(y, yd2)
    torch.cuda.synchronize()
    if layout == "generic":
        assert torch.equal(yd, yd2)
    else:
        assert M.rel(yd2.cpu().numpy(), got) <= 1e-14


@LAYOUTS
def test_rhs_fp32_error_vs_fp64_oracle(cuda, layout):
    import torch
    print(f"[case] rhs fp32 layout={layout}", flush=True)
    pg, fv, og, bc, ctab = M.setup(12, 8, 8, kind="mixed")
    x = M.perturbed_state(og.p[:og.nC], mach=0.7, seed=9)
    ref = O.rhs(og, bc, x)
    eu = _euler(fv, ctab, layout)
    y32 = _dev(torch, x, torch.float32, cuda)
    yd32 = torch.zeros_like(y32)
    eu.rhs_f32(y32, yd32)
    torch.cuda.synchronize()
    err = M.rel(yd32.double().cpu().numpy(), ref)
    print(f"fp32 residual relative error vs fp64 oracle: {err:.3e}")
    assert err <= 2e-5
    # mixed mode keeps the rhs (the Newton residual) in FP64
    em = _euler(fv, ctab, layout, fa.PREC_FP32)
    y = _dev(torch, x, cuda=cuda)
    yd = torch.zeros_like(y)
    em.rhs(y, yd)
    torch.cuda.synchronize()
    assert M.rel(yd.cpu().numpy(), ref) <= 1e-12


@LAYOUTS
def test_free_stream_preserved(cuda, layout):
    import torch
    print(f"[case] free stream layout={layout}", flush=True)
    pg, fv, og, bc, ctab = M.setup(10, 6, 6, kind="far", jitter=0.2)
    nC = og.nC
    rho = np.ones(nC)
    u = np.tile([0.8, 0.0, 0.0], (nC, 1))
    x = M.cons(rho, u, np.full(nC, 1.0 / M.GAMMA))
    eu = _euler(fv, ctab, layout)
    y = _dev(torch, x, cuda=cuda)
    yd = torch.zeros_like(y)
    eu.rhs(y, yd)
    torch.cuda.synchronize()
    # uniform state + matching far field: the residual is rounding noise on closed cells
    scale = np.abs(x).max() * 2.0 / np.cbrt(og.v).min()
    assert np.abs(yd.cpu().numpy()).max() <= 1e-13 * scale


def test_host_vector_mode(cuda):
    print("[case] host vectors", flush=True)
    pg, fv, og, bc, ctab = M.setup(6, 4, 4, kind="mixed")
    x = M.perturbed_state(og.p[:og.nC], seed=2)
    ref = O.rhs(og, bc, x)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=fa.VECTORS_ON_HOST)
    out = np.zeros_like(x)
    eu.rhs(x, out)
    assert M.rel(out, ref) <= 1e-12


def test_find_euler_element_arrays(cuda):
    import torch
    print("[case] findEuler element arrays", flush=True)
    pg, fv, og, bc, ctab = M.setup(8, 5, 5, kind="mixed")
    x = M.perturbed_state(og.p[:og.nC], seed=4)
    rho, rhou, rhoE, p = O.element_state(og, bc, x)
    dr, dru, drE = O.find_euler(og, rho, rhou, rhoE, p, M.GAMMA)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    t = [_dev(torch, a, cuda=cuda) for a in (rho, rhou, rhoE, p)]
    nC = og.nC
    o = [torch.zeros(nC, dtype=torch.float64, device=cuda), torch.zeros(3 * nC, dtype=torch.float64, device=cuda),
         torch.zeros(nC, dtype=torch.float64, device=cuda)]
    eu.find_euler(*t, M.GAMMA, *o)
    torch.cuda.synchronize()
    got = np.concatenate([o[0].cpu().numpy(), o[1].cpu().numpy(), o[2].cpu().numpy()])
    want = np.concatenate([dr, dru.ravel(), drE])
    assert M.rel(got, want) <= 1e-12


@pytest.mark.parametrize("kind", ["wall", "far", "mixed", "supersonic"])
def test_sync_state_matches_oracle(cuda, kind):
    """setBC + syncState on their own (fos_euler_sync_state): every element array of the cells and
    of the boundary facets (ghost states of the four condition families, both inflow kinds
    subsonic) against the oracle's, device and host-vector modes."""
    import torch
    print(f"[case] sync_state {kind}", flush=True)
    pg, fv, og, bc, ctab = M.setup(6, 4, 3, kind=kind, jitter=0.2)
    x = M.perturbed_state(og.p[:og.nC], mach=0.6, seed=7)
    ref = O.element_state(og, bc, x, full=True)
    nE = og.nE
    widths = (1, 3, 1, 3, 1, 1, 1)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    outs = [torch.zeros(w * nE, dtype=torch.float64, device=cuda) for w in widths]
    eu.sync_state(_dev(torch, x, cuda=cuda), *outs)
    torch.cuda.synchronize()
    for name, got, want in zip(("rho", "rhou", "rhoE", "u", "p", "temp", "c"), outs, ref):
        assert M.rel(got.cpu().numpy(), want) <= 1e-14, name
    hu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=fa.VECTORS_ON_HOST)
    houts = [np.zeros(w * nE) for w in widths]
    hu.sync_state(np.ascontiguousarray(x), *houts[:3], None, houts[4], None, None)
    for k in (0, 1, 2, 4):
        assert M.rel(houts[k], ref[k]) <= 1e-14
    assert not houts[3].any()  # not requested: untouched


def _richardson_jv(og, bc, x, v, h=1e-4):
    """J v of the oracle residual by Richardson-extrapolated central differences (error O(h^4)).
    h = 1e-4 keeps every face on one side of the far-field inflow/outflow switch (un = 0) on these
    states: estimates at h and 2h agree to ~2e-11, rounding included."""
    d = lambda e: (O.rhs(og, bc, x + e * v) - O.rhs(og, bc, x - e * v)) / (2 * e)  # noqa: E731
    return (4 * d(h / 2) - d(h)) / 3


@pytest.mark.parametrize("kind", ["wall", "far", "mixed", "supersonic"])
@LAYOUTS
def test_analytic_jv_matches_oracle_derivative(cuda, kind, layout):
    """The analytic Jv (dual numbers through setBC, the primitives and the Roe flux) against the
    directional derivative of the oracle's residual by Richardson-extrapolated central differences,
    for the four condition families (subsonic static and total inflow, outflow, far field, walls)
    on a jittered mesh; linear in v; the fused difference quotient within its own truncation error."""
    import torch
    print(f"[case] analytic jv {kind} layout={layout}", flush=True)
    pg, fv, og, bc, ctab = M.setup(6, 4, 3, kind=kind, jitter=0.2)
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=13)
    rng = np.random.default_rng(3)
    v1 = rng.standard_normal(5 * nC) * np.abs(x) * 0.1
    v2 = rng.standard_normal(5 * nC) * np.abs(x) * 0.1
    ref = _richardson_jv(og, bc, x, v1)
    eu = _euler(fv, ctab, layout)
    y = _dev(torch, x, cuda=cuda)
    out = []
    for v in (v1, v2, 0.3 * v1 - 2.0 * v2):
        jv = torch.zeros_like(y)
        eu.jtimes_analytic(_dev(torch, v, cuda=cuda), jv, y)
        out.append(jv.cpu().numpy())
    err = M.rel(out[0], ref)
    print(f"analytic Jv vs Richardson central differences of the oracle: {err:.2e}")
    assert err <= 1e-8
    assert M.rel(out[2], 0.3 * out[0] - 2.0 * out[1]) <= 1e-13
    dq = torch.zeros_like(y)
    fy = torch.zeros_like(y)
    eu.rhs(y, fy)
    eu.jtimes(_dev(torch, v1, cuda=cuda), dq, y, fy)
    torch.cuda.synchronize()
    assert M.rel(dq.cpu().numpy(), out[0]) <= 1e-5


@LAYOUTS
def test_analytic_jv_fp32(cuda, layout):
    """The analytic Jv in FP32 arithmetic (PREC_FP32 context) has no difference-quotient
    cancellation: it follows the FP64 analytic product to FP32 rounding."""
    import torch
    print(f"[case] analytic jv fp32 layout={layout}", flush=True)
    pg, fv, og, bc, ctab = M.setup(10, 6, 6, kind="mixed")
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=11)
    v = np.random.default_rng(0).standard_normal(5 * nC) * np.abs(x) * 1e-2
    y, vd = _dev(torch, x, cuda=cuda), _dev(torch, v, cuda=cuda)
    res = []
    for flags in (0, fa.PREC_FP32):
        eu = _euler(fv, ctab, layout, flags=flags)
        jv = torch.zeros_like(y)
        eu.jtimes_analytic(vd, jv, y)
        res.append(jv.cpu().numpy())
    err = M.rel(res[1], res[0])
    print(f"FP32 analytic Jv vs FP64: {err:.2e}")
    assert err <= 2e-5


def test_analytic_jv_hex_mesh(cuda):
    """Quadrilateral faces (generic layout): the analytic Jv against the oracle's derivative."""
    import torch
    print("[case] analytic jv hex", flush=True)
    pg, fv, og, bc, ctab = M.setup_hex(4, 3, 3, kind="far")
    x = M.perturbed_state(og.p[:og.nC], seed=3)
    v = np.random.default_rng(4).standard_normal(len(x)) * np.abs(x) * 0.1
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    y = _dev(torch, x, cuda=cuda)
    jv = torch.zeros_like(y)
    eu.jtimes_analytic(_dev(torch, v, cuda=cuda), jv, y)
    torch.cuda.synchronize()
    assert M.rel(jv.cpu().numpy(), _richardson_jv(og, bc, x, v)) <= 1e-8


@LAYOUTS
def test_dq_jv_fp64(cuda, layout):
    import torch
    print(f"[case] dq jv fp64 layout={layout}", flush=True)
    pg, fv, og, bc, ctab = M.setup(10, 6, 6, kind="mixed")
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=11)
    rng = np.random.default_rng(0)
    v = rng.standard_normal(5 * nC) * np.abs(x) * 1e-2
    fy = O.rhs(og, bc, x)
    sigma = 1.0 / O.wrms(v, x)
    ref = O.dq_jv(og, bc, x, fy, v, sigma)
    eu = _euler(fv, ctab, layout)
    y, vv = _dev(torch, x, cuda=cuda), _dev(torch, v, cuda=cuda)
    fyd = torch.zeros_like(y)
    eu.rhs(y, fyd)
    jv = torch.zeros_like(y)
    eu.jtimes(vv, jv, y, fyd)
    torch.cuda.synchronize()
    got = jv.cpu().numpy()
    assert M.rel(got, ref) <= 1e-7, M.rel(got, ref)
    # linear in v: J(2v) = 2 J(v) up to the DQ truncation error
    jv2 = torch.zeros_like(y)
    eu.jtimes(2 * vv, jv2, y, fyd)
    torch.cuda.synchronize()
    assert M.rel(jv2.cpu().numpy(), 2 * got) <= 1e-5
    # Jv(0) = 0
    z = torch.zeros_like(y)
    jz = torch.ones_like(y)
    eu.jtimes(z, jz, y, fyd)
    torch.cuda.synchronize()
    assert float(jz.abs().max()) == 0.0


@LAYOUTS
def test_dq_jv_fp32(cuda, layout):
    import torch
    print(f"[case] dq jv fp32 layout={layout}", flush=True)
    pg, fv, og, bc, ctab = M.setup(10, 6, 6, kind="far")
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=12)
    rng = np.random.default_rng(1)
    v = rng.standard_normal(5 * nC) * np.abs(x) * 1e-2
    fy = O.rhs(og, bc, x)
    # central-difference directional derivative in FP64 as the reference value
    h = 1e-4 * np.linalg.norm(x) / np.linalg.norm(v)
    ref = (O.rhs(og, bc, x + h * v) - O.rhs(og, bc, x - h * v)) / (2 * h)
    eu = _euler(fv, ctab, layout)
    y, vv = (torch.tensor(a, dtype=torch.float32, device=cuda) for a in (x, v))
    fyd = torch.zeros_like(y)
    eu.rhs_f32(y, fyd)
    # FP32 needs sigma ~ sqrt(eps32) relative to |y|: cancellation otherwise dominates
    sig = float(np.sqrt(np.finfo(np.float32).eps) * np.linalg.norm(x) / np.linalg.norm(v))
    jv = torch.zeros_like(y)
    eu.jtimes_f32(vv, jv, y, fyd, sig)
    torch.cuda.synchronize()
    err = M.rel(jv.double().cpu().numpy(), ref)
    print(f"fp32 DQ Jv relative error vs FP64 central difference: {err:.3e}")
    assert err <= 2e-2


SQRT_EPS32 = 3.4526698300124393e-4


def _mixed_sigma(v, y, rtol=1e-6, atol=1e-12):
    """Mixed-mode increment: the largest weighted component of sigma v is sqrt(eps32)/rtol weighted
    units, with the WRMS weights 1/(rtol|y_i| + atol + rtol rms(y))."""
    w = 1.0 / (rtol * np.abs(y) + atol + rtol * np.sqrt(np.mean(y * y)))
    return SQRT_EPS32 / rtol / np.max(np.abs(v) * w)


@pytest.mark.parametrize("kind", ["far", "mixed"])
@LAYOUTS
def test_dq_jv_mixed_two_sided(cuda, kind, layout):
    """FOS_PREC_FP32: FP64 vectors, FP32 flux differences F(y + sigma v) - F(y) per face with sigma
    such that the largest weighted component of sigma v is sqrt(eps32)/rtol weighted units.  Its
    error is stated against the FP64 oracle's
    difference quotient at the same sigma (the FP32 arithmetic error) and, on the smooth far-field
    case, against the FP64 central difference (arithmetic plus truncation).  The mixed case has
    far-field sides with u.n ~ 0, where the ghost state switches between the inflow and outflow
    formulas: at a relative perturbation of ~1e-4 even the FP64 difference quotient differs from
    the central difference by O(1e-2) or more, so only the same-sigma comparison is meaningful
    there."""
    import torch
    print(f"[case] dq jv mixed {kind} layout={layout}", flush=True)
    pg, fv, og, bc, ctab = M.setup(10, 6, 6, kind=kind)
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=13)
    v = np.random.default_rng(2).standard_normal(5 * nC) * np.abs(x) * 1e-2
    sigma = _mixed_sigma(v, x)
    fy = O.rhs(og, bc, x)
    ref_same = O.dq_jv(og, bc, x, fy, v, sigma)
    eu = _euler(fv, ctab, layout, fa.PREC_FP32)
    y, vv = _dev(torch, x, cuda=cuda), _dev(torch, v, cuda=cuda)
    fyd = torch.zeros_like(y)
    eu.rhs(y, fyd)
    jv = torch.zeros_like(y)
    eu.jtimes(vv, jv, y, fyd)
    torch.cuda.synchronize()
    got = jv.cpu().numpy()
    err_same = M.rel(got, ref_same)
    print(f"mixed DQ Jv ({kind}): relative error vs FP64 DQ at the same sigma {err_same:.3e}")
    assert err_same <= 1e-2
    if kind == "far":
        h = 1e-4 * np.linalg.norm(x) / np.linalg.norm(v)
        ref = (O.rhs(og, bc, x + h * v) - O.rhs(og, bc, x - h * v)) / (2 * h)
        err = M.rel(got, ref)
        print(f"mixed DQ Jv (far): relative error vs FP64 central difference {err:.3e}")
        assert err <= 1e-2


def test_jv_weight_floor_follows_the_state(cuda):
    """The Jv weight floor is reused for the same y array only while no rhs call intervenes: after y
    changes in place and rhs(y) is evaluated, jtimes matches the oracle for the new state."""
    import torch
    print("[case] jv floor reuse", flush=True)
    pg, fv, og, bc, ctab = M.setup(8, 5, 5, kind="mixed")
    nC = og.nC
    x = M.perturbed_state(og.p[:nC], mach=0.6, seed=21)
    v = np.random.default_rng(5).standard_normal(5 * nC) * np.abs(x) * 1e-2
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    y, vv = _dev(torch, x, cuda=cuda), _dev(torch, v, cuda=cuda)
    fyd, jv = torch.zeros_like(y), torch.zeros_like(y)
    for scale in (1.0, 1.7):
        y.mul_(scale)
        xs = x * scale
        eu.rhs(y, fyd)
        eu.jtimes(vv, jv, y, fyd)
        eu.jtimes(vv, jv, y, fyd)  # second call reuses the floor
        torch.cuda.synchronize()
        ref = O.dq_jv(og, bc, xs, O.rhs(og, bc, xs), v, 1.0 / O.wrms(v, xs))
        assert M.rel(jv.cpu().numpy(), ref) <= 1e-7, (scale, M.rel(jv.cpu().numpy(), ref))


def test_hex_mesh_residual_and_jv_match_oracle(cuda):
    """Hexahedra (6 quadrilateral faces per cell, 8 slots): the generic layout serves them; FP64
    residual within 1e-12 of the oracle, FP32 within 2e-5, FP64 DQ Jv within 1e-7."""
    import torch
    print("[case] hex mesh", flush=True)
    pg, fv, og, bc, ctab = M.setup_hex(8, 5, 4, kind="mixed")
    x = M.perturbed_state(og.p[:og.nC], mach=0.6, seed=6)
    ref = O.rhs(og, bc, x)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    assert eu.layout == 0  # the tetrahedral layout does not apply
    y = _dev(torch, x, cuda=cuda)
    yd = torch.zeros_like(y)
    eu.rhs(y, yd)
    torch.cuda.synchronize()
    assert M.rel(yd.cpu().numpy(), ref) <= 1e-12
    y32, yd32 = y.float(), torch.zeros(y.numel(), dtype=torch.float32, device=cuda)
    eu.rhs_f32(y32, yd32)
    torch.cuda.synchronize()
    assert M.rel(yd32.double().cpu().numpy(), ref) <= 2e-5
    v = np.random.default_rng(3).standard_normal(x.size) * np.abs(x) * 1e-2
    jref = O.dq_jv(og, bc, x, ref, v, 1.0 / O.wrms(v, x))
    jv = torch.zeros_like(y)
    eu.jtimes(_dev(torch, v, cuda=cuda), jv, y, yd)
    torch.cuda.synchronize()
    assert M.rel(jv.cpu().numpy(), jref) <= 1e-7


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 1)])
def test_tiny_meshes(cuda, dims):
    """6 and 12 tetrahedra: one partial chunk, most faces on the boundary."""
    import torch
    print(f"[case] tiny mesh {dims}", flush=True)
    pg, fv, og, bc, ctab = M.setup(*dims, kind="mixed", jitter=0.0)
    x = M.perturbed_state(og.p[:og.nC], mach=0.5, seed=1)
    ref = O.rhs(og, bc, x)
    for layout in (1, 0):
        eu = _euler(fv, ctab, layout)
        y = _dev(torch, x, cuda=cuda)
        yd = torch.zeros_like(y)
        eu.rhs(y, yd)
        torch.cuda.synchronize()
        assert M.rel(yd.cpu().numpy(), ref) <= 1e-12, layout


def test_full_size_properties(cuda):
    """The 1.5M-tet throughput mesh: free-stream preservation in FP64, run-to-run agreement to the
    rounding of the face fluxes (the LDS sums' order varies), the FP32 error against FP64 (stated, and bounded against the flux scale) and DQ
    Jv linearity, checked at the benchmark's size."""
    import torch
    import bench as B
    print("[case] full size 1.5M", flush=True)
    fv = B.throughput_mesh(100, 50, 50, 2.0, 1.0, 1.0)
    nC = fv.nC
    free = [1 / M.GAMMA, 1 / M.GAMMA, 0.8, 0.0, 0.0]
    ctab = fa.CondTab([1, 2], [30, 30], np.array([free, free]))
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    assert eu.layout == 1
    x0 = M.cons(np.ones(nC), np.tile([0.8, 0.0, 0.0], (nC, 1)), np.full(nC, 1 / M.GAMMA))
    y = _dev(torch, x0, cuda=cuda)
    yd = torch.zeros_like(y)
    eu.rhs(y, yd)
    torch.cuda.synchronize()
    scale = np.abs(x0).max() * 2.0 / np.cbrt(fv.v).min()
    assert float(yd.abs().max()) <= 1e-13 * scale
    x = B.throughput_state(fv)
    y = _dev(torch, x, cuda=cuda)
    a, b = torch.zeros_like(y), torch.zeros_like(y)
    eu.rhs(y, a)
    eu.rhs(y, b)
    torch.cuda.synchronize()
    # run to run: the LDS sums' order varies, so a cell's sum of four face fluxes that cancel to
    # O(h |grad U|) changes by a few ulps of the fluxes, not of the sum (4e-14 of |R| seen)
    fscale = np.abs(x).max() * 2.0 / np.cbrt(fv.v).min()
    assert float((a - b).abs().max()) <= 1e-13 * fscale
    a32 = torch.zeros(5 * nC, dtype=torch.float32, device=cuda)
    eu.rhs_f32(y.float(), a32)
    torch.cuda.synchronize()
    err32 = float((a32.double() - a).norm() / a.norm())
    abs32 = float((a32.double() - a).abs().max()) / (np.abs(x).max() * 2.0 / np.cbrt(fv.v).min())
    print(f"FP32 vs FP64 residual at 1.5M tets: {err32:.2e} relative, {abs32:.2e} of the flux scale")
    # FP32 rounds each face flux to ~1e-7 of its size; the residual is a sum of face fluxes that
    # cancel to O(h |grad U|), so its relative error grows as the mesh is refined (2e-5 at 1,200
    # tets, ~3e-4 here) while the error against the flux scale stays at rounding level
    assert err32 <= 1e-3
    assert abs32 <= 1e-6
    g = torch.Generator(device=cuda)
    g.manual_seed(5)
    v = torch.randn(5 * nC, generator=g, device=cuda, dtype=torch.float64) * y.abs() * 1e-3
    j1, j2 = torch.zeros_like(y), torch.zeros_like(y)
    eu.jtimes(v, j1, y, a)
    eu.jtimes(2 * v, j2, y, a)
    torch.cuda.synchronize()
    lin = float((j2 - 2 * j1).norm() / (2 * j1).norm())
    print(f"DQ Jv linearity at 1.5M tets: {lin:.2e}")
    assert lin <= 1e-5

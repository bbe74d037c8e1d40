"""The fused Newton matrix product out = v - gamma J v (fos_euler_newton_matvec[_analytic]): equal to
v - gamma * jtimes(v) formed in two passes, for the difference quotient and the analytic product, FP64
and FP32 contexts, both residual layouts; refusal of an output aliasing v; and the integrator with the
fused product against the two-pass one.

Tolerances: the fused and two-pass products run the same kernels and differ in where v - gamma o is
rounded (one fused multiply-add in the kernel's epilogue against a separate vector pass) and, on the
tetrahedral layout, in the order of the per-cell flux sums (LDS atomics): 1e-14 relative to |v| in
the 2-norm, plus an order term relative to gamma |Jv| -- 1e-11 where the summed terms are the
product's own (analytic; FP32 face-wise differences), 1e-6 for the FP64 difference quotient, whose
R(y + sigma v) sum order noise (~1e-16 of the flux) is divided by the increment (sigma v of order
one weighted unit, 1e-6..1e-8 relative).  A sign or factor slip moves out by gamma |Jv| itself.
The integrated states agree within the solver tolerances (Newton 1e-8 under the WRMS weights): 1e-9
relative after 4 fixed steps, with GMRES iteration counts equal to within the stopping test's
sensitivity to the sum order (one iteration, or 2%)."""
import numpy as np
import pytest

import fosolvers_amd as fa
from tests import meshes as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", [1, 0], ids=["tet", "generic"])
@pytest.mark.parametrize("flags", [0, fa.PREC_FP32], ids=["fp64", "fp32"])
def test_newton_matvec_equals_two_passes(cuda, layout, flags):
    import torch
    print(f"[case] newton matvec layout={layout} flags={flags}", flush=True)
    pg, fv, og, bc, ctab = M.setup(10, 6, 6, kind="supersonic", jitter=0.2)
    x = M.perturbed_state(og.p[:og.nC], mach=2.0, seed=5)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags | (0 if layout else fa.LAYOUT_GENERIC))
    assert eu.layout == layout
    y = torch.tensor(x, device=cuda)
    fy = torch.zeros_like(y)
    eu.rhs(y, fy)
    rng = np.random.default_rng(7)
    v = torch.tensor(rng.standard_normal(5 * og.nC) * np.abs(x) * 1e-3, device=cuda)
    gamma = 3.3e-3
    for analytic in (False, True):
        jv = torch.zeros_like(y)
        if analytic:
            eu.jtimes_analytic(v, jv, y)
        else:
            eu.jtimes(v, jv, y, fy)
        two = v - gamma * jv
        out = torch.zeros_like(y)
        eu.newton_matvec(gamma, v, out, y, fy, analytic=analytic)
        torch.cuda.synchronize()
        err = float((out - two).norm() / v.norm())
        gjv = float((gamma * jv).norm() / v.norm())
        order = 1e-6 if (layout == 1 and not analytic and flags == 0) else (1e-11 if layout == 1 else 0.0)
        print(f"  analytic={analytic}: fused vs two passes {err:.1e} (|gamma Jv|/|v| {gjv:.1e})")
        assert gjv > 1e-3
        assert err <= 1e-14 + order * gjv
        assert float((out - v).norm()) > 0  # the Jv term is there
    with pytest.raises(RuntimeError, match="alias"):
        eu.newton_matvec(gamma, v, v, y, fy)


@pytest.mark.parametrize("jtimes", [True, "analytic"])
def test_bdf_fused_newton_product_matches_two_passes(cuda, jtimes):
    import torch
    print(f"[case] bdf fused newton product jtimes={jtimes}", flush=True)
    pg, fv, og, bc, ctab = M.setup(12, 6, 5, kind="supersonic", jitter=0.15)
    x = M.perturbed_state(og.p[:og.nC], mach=2.0, seed=2)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=fa.PREC_FP32)
    res = {}
    for fused in (False, True):
        b = fa.BDFNewtonKrylov.for_euler(eu, maxl=30, precond="ilu", jtimes=jtimes, fused=fused)
        b.set_newton(tol=1e-8, max_iter=20, lin_tol=1e-3, max_restarts=10)
        b.set_dt(2e-3)
        y = torch.tensor(x, device=cuda)
        b.set_iv(0.0, y)
        for _ in range(4):
            assert b.step(0.0, y)[0] == 0
        res[fused] = (y.cpu().numpy(), b.stats())
        b.close()
    (y0, s0), (y1, s1) = res[False], res[True]
    print(f"  fused vs two passes {M.rel(y1, y0):.1e}; GMRES its {s1['nli']} vs {s0['nli']}, "
          f"Jv products {s1['njv']} vs {s0['njv']}")
    assert M.rel(y1, y0) <= 1e-9
    # the flux sums' order (LDS atomics) can move a GMRES stopping test by one iteration
    assert abs(s1["nli"] - s0["nli"]) <= max(1, s0["nli"] // 50)
    assert abs(s1["njv"] - s0["njv"]) <= max(1, s0["njv"] // 50)

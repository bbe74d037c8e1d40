"""fos_euler_jtimes_noalloc (the difference-quotient Jv with its reduction scratch in a file-scope
device array, for callers that cannot allocate) against fos_euler_jtimes: bitwise the same product,
on two calls in a row, the second on a smaller mesh, then the first mesh again (the shared scratch
holds the other context's values in between), in FP64 and with FOS_PREC_FP32."""
import ctypes

import numpy as np
import pytest

import fosolvers_amd as fa
from fosolvers_amd import _capi
from tests import meshes as M

pytestmark = pytest.mark.gpu


def _case(dims, seed, flags):
    """A context on the generic residual layout, whose fixed-order face sums make two products at
    the same state bitwise equal (the tetrahedral layout's LDS atomic sums agree to rounding)."""
    import torch
    flags |= fa.LAYOUT_GENERIC
    pg, fv, og, bc, ctab = M.setup(*dims, kind="mixed")
    x = M.perturbed_state(og.p[:og.nC], mach=0.7, seed=seed)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=flags)
    y = torch.tensor(x, device="cuda")
    fy = torch.zeros_like(y)
    eu.rhs(y, fy)
    rng = np.random.default_rng(seed)
    v = torch.tensor(rng.standard_normal(5 * og.nC) * np.abs(x) * 1e-3, device="cuda")
    return eu, y, fy, v


def _both(L, eu, y, fy, v):
    import torch
    j_ref, j_na = torch.zeros_like(y), torch.zeros_like(y)
    assert L.fos_euler_jtimes(fa._ptr(v), fa._ptr(j_ref), 0.0, fa._ptr(y), fa._ptr(fy), eu.h, None) == 0
    assert L.fos_euler_jtimes_noalloc(fa._ptr(v), fa._ptr(j_na), 0.0, fa._ptr(y), fa._ptr(fy), eu.h, None) == 0
    torch.cuda.synchronize()
    return j_ref, j_na


@pytest.mark.parametrize("flags", [0, fa.PREC_FP32])
def test_jtimes_noalloc_equals_jtimes(cuda, flags):
    import torch
    print(f"[case] jtimes_noalloc flags={flags}", flush=True)
    L = fa.lib()
    big = _case((10, 6, 5), 3, flags)
    small = _case((4, 3, 2), 5, flags)
    assert small[0].nC < big[0].nC
    for eu, y, fy, v in (big, small, big):
        j_ref, j_na = _both(L, eu, y, fy, v)
        assert torch.isfinite(j_ref).all() and float(j_ref.norm()) > 0.0
        assert torch.equal(j_ref, j_na), float((j_ref - j_na).norm() / j_ref.norm())
    # a second product at the same state (the floor in scratch is reused) and after an rhs call
    eu, y, fy, v = big
    v2 = v * 0.5 + 1e-4
    j_ref, j_na = _both(L, eu, y, fy, v2)
    assert torch.equal(j_ref, j_na)
    eu.rhs(y, fy)
    j_ref, j_na = _both(L, eu, y, fy, v)
    assert torch.equal(j_ref, j_na)
    # zero direction: zero product
    z = torch.zeros_like(v)
    j_ref, j_na = _both(L, eu, y, fy, z)
    assert float(j_na.abs().max()) == 0.0


def test_jtimes_noalloc_on_the_tetrahedral_layout(cuda):
    """The default layout: its LDS atomic face sums vary in order from call to call, so the two
    entries agree to the difference quotient's rounding noise, as two calls of either do."""
    import torch
    print("[case] jtimes_noalloc tetrahedral layout", flush=True)
    L = fa.lib()
    pg, fv, og, bc, ctab = M.setup(10, 6, 5, kind="mixed")
    x = M.perturbed_state(og.p[:og.nC], mach=0.7, seed=3)
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA)
    assert eu.layout == 1
    y = torch.tensor(x, device="cuda")
    fy = torch.zeros_like(y)
    eu.rhs(y, fy)
    v = torch.tensor(np.random.default_rng(3).standard_normal(5 * og.nC) * np.abs(x) * 1e-3, device="cuda")
    j1, j2 = _both(L, eu, y, fy, v)
    j3 = torch.zeros_like(y)
    assert L.fos_euler_jtimes(fa._ptr(v), fa._ptr(j3), 0.0, fa._ptr(y), fa._ptr(fy), eu.h, None) == 0
    torch.cuda.synchronize()
    noise = float((j3 - j1).norm() / j1.norm())
    diff = float((j2 - j1).norm() / j1.norm())
    print(f"  jtimes run to run {noise:.1e}, noalloc vs jtimes {diff:.1e}", flush=True)
    assert diff <= max(1e-9, 10 * noise)


def test_jtimes_noalloc_refuses_host_vectors(cuda):
    print("[case] jtimes_noalloc host vectors", flush=True)
    pg, fv, og, bc, ctab = M.setup(4, 3, 2, kind="far")
    eu = fa.Euler(fv, ctab, r=M.R, gamma=M.GAMMA, flags=fa.VECTORS_ON_HOST)
    n = 5 * og.nC
    a = np.ones(n)
    b = np.zeros(n)
    p = a.ctypes.data_as(ctypes.c_void_p)
    assert fa.lib().fos_euler_jtimes_noalloc(p, b.ctypes.data_as(ctypes.c_void_p), 0.0, p, p, eu.h, None) == -1
    assert "device vectors only" in _capi.last_error()

"""The ILU interface on the GPU: the level-scheduled triangular solves (lusol) and the product
(amux) against the oracle's restatements of SPARSKIT, the reference's ILU1 KAT through the GPU solve,
host-vector mode, and fopbc's configuration (lfil 50, drop tolerance 1e-6) on a 3-D pressure-type
matrix.  Tolerance: the solve within 1e-12 relative of lusol's (same factor, different summation
order); the product within 1e-14."""
import numpy as np
import pytest

import fosolvers_amd as fa
from oracle import oracle as O
from tests.test_ilut import banded, poisson3

pytestmark = pytest.mark.gpu


def _lusol_oracle(p, n, b):
    alu, jlu, ju = p.msr()
    return O.lusol(n, np.concatenate([[0.0], alu]), np.concatenate([[0], jlu]), np.concatenate([[0], ju]), b)


@pytest.mark.parametrize("kind,size,lfill,droptol", [("banded", 300, 4, 0.0), ("banded", 300, 40, 1e-3),
                                                     ("ties", 200, 3, 0.0), ("poisson", 12, 50, 1e-6),
                                                     ("poisson", 20, 10, 1e-4)])
def test_ilut_gpu_solve_matches_lusol(cuda, kind, size, lfill, droptol):
    import torch
    print(f"[case] ilut gpu solve {kind} {size} lfil={lfill} droptol={droptol}", flush=True)
    if kind == "poisson":
        n, r, c, v = poisson3(size)
    else:
        A = banded(size, 11, ties=kind == "ties")
        n = size
        r, c = np.nonzero(A)
        v = A[r, c]
    p = fa.ILU(n, len(v), lfill)
    p.set_coo(r, c, v)
    p.fact(droptol)
    # the factor is the oracle's ilut, entry for entry (pinned on the CPU); the solve is lusol's
    ao, jao, iao = O.coocsr(n, v, r + 1, c + 1)
    b = np.random.default_rng(2).standard_normal(n)
    xo = O.ilut_solve(n, ao, jao, iao, lfill, droptol, b)[0]
    assert np.array_equal(_lusol_oracle(p, n, b), xo)
    bd = torch.tensor(b, device=cuda)
    x = torch.zeros_like(bd)
    p.solve(bd, x)
    x2 = torch.zeros_like(bd)
    p.solve(bd, x2)  # the recorded graph replays to the same result
    torch.cuda.synchronize()
    got = x.cpu().numpy()
    assert np.linalg.norm(got - xo) <= 1e-12 * np.linalg.norm(xo)
    assert torch.equal(x, x2)
    # the product: y = A x (amux)
    y = torch.zeros_like(bd)
    p.mul_vec(bd, y)
    torch.cuda.synchronize()
    yo = O.amux(n, b, ao, jao, iao)
    assert np.linalg.norm(y.cpu().numpy() - yo) <= 1e-14 * np.linalg.norm(yo)
    f, bw = p.levels()
    print(f"  {n} rows, {len(p.msr()[0]) - n - 1} factor entries, {f} forward / {bw} backward levels", flush=True)


def test_ilu1_kat_on_the_gpu(cuda):
    """src/test/ILU1.f90:20 through the GPU solve, device and host vectors."""
    import torch
    print("[case] ilut ILU1 KAT", flush=True)
    iA = np.array([1, 2, 2, 3, 3, 4, 4]) - 1
    jA = np.array([1, 2, 1, 3, 4, 4, 2]) - 1
    A = np.array([1.0, 2.0, -1.0, 5.0, -4.0, 2.0, -1.0])
    want = np.array([1.0, 1.5, 14 / 5, 11 / 4])
    p = fa.ILU(4, 7, 4)
    p.set_coo(iA, jA, A)
    p.fact(0.0)
    x = torch.zeros(4, dtype=torch.float64, device=cuda)
    p.solve(torch.tensor([1.0, 2.0, 3.0, 4.0], dtype=torch.float64, device=cuda), x)
    torch.cuda.synchronize()
    assert np.linalg.norm(x.cpu().numpy() - want) <= 1e-9
    ph = fa.ILU(4, 7, 4, host=True)
    ph.set_coo(iA, jA, A)
    ph.fact(0.0)
    xh = np.zeros(4)
    ph.solve(np.array([1.0, 2.0, 3.0, 4.0]), xh)
    assert np.linalg.norm(xh - want) <= 1e-9
    yh = np.zeros(4)
    ph.mul_vec(want, yh)
    assert np.allclose(yh, [1.0, 2.0, 3.0, 4.0], rtol=0, atol=1e-14)

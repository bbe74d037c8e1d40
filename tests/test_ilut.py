"""The ILU interface (init / setCOO / fact / solve over ILUT, src/libfosolvers/numerics/sparse.f90:52-63,
334-392): the product's host factorisation against the oracle's restatement of SPARSKIT's ilut,
entry for entry in SPARSKIT's MSR arrays, on matrices where the drop rules, the quick-split
selection and the replaced zero pivot all act; the reference's own KAT (src/test/ILU1.f90:20) through
the factor; the compiled reference's ILU wrapper where it is built.  The GPU solve and product are
in tests/test_gpu_ilut.py."""
import numpy as np
import pytest

import fosolvers_amd as fa
from oracle import oracle as O


def banded(n, seed, bands=(1, 3, 7), diag=5.0, ties=False):
    rng = np.random.default_rng(seed)
    A = np.diag(diag + rng.random(n))
    for k in bands:
        if ties:  # many equal magnitudes: the selection's tie order matters
            A += np.diag(np.round(rng.random(n - k) * 4 - 2), k) + np.diag(np.round(rng.random(n - k) * 4 - 2), -k)
        else:
            A += np.diag(rng.random(n - k) - 0.5, k) + np.diag(rng.random(n - k) - 0.5, -k)
    return A


def poisson3(m):
    """7-point Laplacian on an m^3 grid plus a small shift (the pressure-equation shape fopbc
    factors): (n, rows, cols, values)."""
    import scipy.sparse as sp
    T = sp.diags([-np.ones(m - 1), 2.0 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    I = sp.identity(m)
    A = (sp.kron(I, sp.kron(I, T)) + sp.kron(I, sp.kron(T, I)) + sp.kron(T, sp.kron(I, I))
         + 1e-3 * sp.identity(m ** 3)).tocoo()
    return m ** 3, A.row.astype(np.int64), A.col.astype(np.int64), A.data.copy()


def product_factor(n, r, c, v, lfill, droptol):
    p = fa.ILU(n, len(v), lfill)
    p.set_coo(r, c, v)
    p.fact(droptol)
    return p


def oracle_factor(n, r, c, v, lfill, droptol):
    ao, jao, iao = O.coocsr(n, v, r + 1, c + 1)
    alu, jlu, ju, ierr = O.ilut_factor(n, ao, jao, iao, lfill, droptol)
    assert ierr == 0
    m = jlu[n + 1] - 1  # last used Fortran position
    return alu[1:m + 1], jlu[1:m + 1], ju[1:n + 1]


CASES = [("banded", 60, 60, 0.0), ("banded", 60, 2, 0.0), ("banded", 60, 3, 0.05), ("ties", 50, 2, 0.0),
         ("ties", 50, 4, 0.3), ("poisson", 6, 50, 1e-6), ("poisson", 6, 3, 1e-2), ("zero-pivot", 30, 5, 0.0)]


@pytest.mark.parametrize("kind,size,lfill,droptol", CASES)
def test_ilut_factor_equals_sparskit(kind, size, lfill, droptol):
    if kind == "poisson":
        n, r, c, v = poisson3(size)
    else:
        A = banded(size, 7, ties=kind == "ties")
        if kind == "zero-pivot":
            A[4, 4] = 0.0  # replaced by (1e-4 + droptol) times the row's mean magnitude
        n = size
        r, c = np.nonzero(A)
        v = A[r, c]
    perm = np.random.default_rng(3).permutation(len(v))  # unsorted COO, like an assembly loop
    r, c, v = r[perm], c[perm], v[perm]
    p = product_factor(n, r, c, v, lfill, droptol)
    alu, jlu, ju = p.msr()
    oalu, ojlu, oju = oracle_factor(n, r, c, v, lfill, droptol)
    assert np.array_equal(jlu, ojlu) and np.array_equal(ju, oju)
    assert np.array_equal(alu[:n], oalu[:n]) and np.array_equal(alu[n + 1:], oalu[n + 1:])
    # the factor solves like the oracle's lusol (a solve through the product's arrays)
    b = np.random.default_rng(1).random(n)
    x = O.lusol(n, np.concatenate([[0.0], alu]), np.concatenate([[0], jlu]), np.concatenate([[0], ju]), b)
    xo = O.lusol(n, np.concatenate([[0.0], oalu]), np.concatenate([[0], ojlu]), np.concatenate([[0], oju]), b)
    assert np.array_equal(x, xo)
    if lfill >= n - 1 and droptol == 0.0 and kind == "banded":
        # no dropping: ILUT keeps every fill entry, i.e. the exact LU
        A = np.zeros((n, n))
        np.add.at(A, (r, c), v)
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-13)
    f, bwd = p.levels()
    assert 1 <= f <= n and 1 <= bwd <= n


def test_ilu1_kat_through_the_factor():
    """src/test/ILU1.f90:20 -- init(4,7,4), setCOO, fact(0), solve: x = (1, 3/2, 14/5, 11/4)."""
    iA = np.array([1, 2, 2, 3, 3, 4, 4]) - 1
    jA = np.array([1, 2, 1, 3, 4, 4, 2]) - 1
    A = np.array([1.0, 2.0, -1.0, 5.0, -4.0, 2.0, -1.0])
    p = product_factor(4, iA, jA, A, 4, 0.0)
    alu, jlu, ju = p.msr()
    x = O.lusol(4, np.concatenate([[0.0], alu]), np.concatenate([[0], jlu]), np.concatenate([[0], ju]),
                np.array([1.0, 2.0, 3.0, 4.0]))
    assert np.linalg.norm(x - [1.0, 1.5, 14 / 5, 11 / 4]) <= 1e-9


def test_ilut_interface_errors():
    with pytest.raises(RuntimeError):
        fa.ILU(0, 1, 1)
    p = fa.ILU(3, 4, 2)
    with pytest.raises(RuntimeError, match="no matrix"):
        p.fact(0.0)
    with pytest.raises(RuntimeError, match="outside"):
        p.set_coo([0, 3], [0, 0], [1.0, 1.0])
    with pytest.raises(RuntimeError, match="maxNNz"):
        p.set_coo([0, 1, 2, 0, 1], [0, 1, 2, 1, 0], [1.0] * 5)
    p.set_coo([0, 2], [0, 2], [1.0, 1.0])  # row 1 empty
    with pytest.raises(RuntimeError, match="zero"):
        p.fact(0.0)


@pytest.mark.ref
@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built")
def test_ilut_factor_against_the_reference_wrapper():
    """The reference's own ILU type (init/setCOO/fact/solve) compiled from its sources, on fills where
    its storage bound holds (2 lFill + 1 >= n)."""
    A = banded(40, 5)
    r, c = np.nonzero(A)
    perm = np.random.default_rng(5).permutation(len(r))
    r, c = r[perm], c[perm]
    b = np.random.default_rng(6).random(40)
    for lfill in (20, 40):
        xr = O.ref_ilu_solve(40, r + 1, c + 1, A[r, c], lfill, 0.0, b)
        p = product_factor(40, r, c, A[r, c], lfill, 0.0)
        alu, jlu, ju = p.msr()
        x = O.lusol(40, np.concatenate([[0.0], alu]), np.concatenate([[0], jlu]), np.concatenate([[0], ju]), b)
        assert np.allclose(x, xr, rtol=1e-13, atol=1e-14), lfill

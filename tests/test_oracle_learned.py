"""Pinning the FP64 oracle of the learned preconditioner: its backward against central finite
differences of its own loss, the apply against a dense assembly, zero output = identity."""
import numpy as np

from oracle import oracle as O


def _graph(n, seed):
    rng = np.random.default_rng(seed)
    nbr = -np.ones((n, 4), np.int32)
    # a ring with chords: symmetric adjacency, some cells with fewer than 4 neighbours
    for c in range(n):
        for d in ((c + 1) % n, (c - 1) % n, (c + 7) % n, (c - 7) % n):
            if d != c and d not in nbr[c]:
                k = list(nbr[c]).index(-1) if -1 in nbr[c] else None
                if k is not None:
                    nbr[c, k] = d
    nbr[3, 2:] = -1  # break symmetry cleanly: drop edges both ways
    for c in range(n):
        for j in range(4):
            d = nbr[c, j]
            if d >= 0 and c not in nbr[d]:
                nbr[c, j] = -1
    return nbr, rng


def test_zero_output_is_identity():
    n = 12
    nbr, rng = _graph(n, 0)
    Oz = np.zeros((n, 30))
    r = rng.standard_normal(5 * n)
    z = O.mlp_apply(Oz, nbr, np.array([1.0, 2.0, 2.0, 2.0, 3.0]), r)
    assert np.allclose(z, r)


def test_apply_matches_dense_operator():
    n = 10
    nbr, rng = _graph(n, 1)
    Ot = rng.standard_normal((n, 30)) * 0.3
    S = np.array([1.0, 0.5, 0.5, 0.5, 2.5])
    P = np.zeros((5 * n, 5 * n))
    for k in range(5 * n):
        e = np.zeros(5 * n)
        e[k] = 1.0
        P[:, k] = O.mlp_apply(Ot, nbr, S, e)
    r = rng.standard_normal(5 * n)
    assert np.allclose(P @ r, O.mlp_apply(Ot, nbr, S, r))  # linear in r
    # t_c = B_c r_c / S; z_c = S ((1 + s_4c) t_c + sum_j s_jc t_nb(j)), s = o[25:30]
    B = np.eye(5)[None] + Ot[:, :25].reshape(n, 5, 5)
    t = np.einsum("cab,cb->ca", B, (r.reshape(5, n).T / S))
    zc = np.zeros((n, 5))
    for c in range(n):
        acc = (1.0 + Ot[c, 29]) * t[c]
        for j in range(4):
            if nbr[c, j] >= 0:
                acc = acc + Ot[c, 25 + j] * t[nbr[c, j]]
        zc[c] = S * acc
    assert np.allclose(O.mlp_apply(Ot, nbr, S, r), zc.T.ravel())
    # a neighbour weight couples exactly that neighbour: column block (c, d) of P is nonzero only
    # for d = c and the neighbours of c
    for c in range(n):
        for d in range(n):
            blk = P[np.arange(5) * n + c][:, np.arange(5) * n + d]
            if d != c and d not in nbr[c]:
                assert np.all(blk == 0.0)


def test_backward_matches_finite_differences():
    n = 9
    nbr, rng = _graph(n, 2)
    npar = O.mlp_nparams()
    assert npar == 20 * 128 + 128 + 2 * (256 * 128 + 128) + 128 * 30 + 30
    par = rng.standard_normal(npar) * 0.08
    X = rng.standard_normal((n, 20))
    S = np.array([1.0, 0.7, 0.7, 0.7, 2.0])
    r = rng.standard_normal(5 * n)
    delta = rng.standard_normal(5 * n)
    mask = (rng.random(n) < 0.7).astype(np.int32)
    mask[0] = 1
    loss, g = O.mlp_loss_grad(X, nbr, par, S, r, delta, mask)
    idx = np.concatenate([rng.choice(npar, 40, replace=False),
                          [0, 20 * 128, 20 * 128 + 128 + 128 * 128 + 5, npar - 1, npar - 31, npar - 6]])
    h = 1e-6
    for k in idx:
        pp, pm = par.copy(), par.copy()
        pp[k] += h
        pm[k] -= h
        fd = (O.mlp_loss_grad(X, nbr, pp, S, r, delta, mask)[0] - O.mlp_loss_grad(X, nbr, pm, S, r, delta, mask)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])

"""Pinning the FP64 oracle of the learned preconditioner: the forward against a numpy brute force
written from the network's definition and against closed forms, the apply by its structural
properties, the loss against the pinned forward and apply, the backward against central finite
differences of the loss."""
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


def _dense(Ot, nbr, S):
    n = Ot.shape[0]
    P = np.zeros((5 * n, 5 * n))
    for k in range(5 * n):
        e = np.zeros(5 * n)
        e[k] = 1.0
        P[:, k] = O.mlp_apply(Ot, nbr, S, e)
    return P


def _blk(P, n, c, d):
    return P[np.arange(5) * n + c][:, np.arange(5) * n + d]


def test_apply_structure():
    """Structural properties of the apply only (no restatement of its formula): linear in r; a
    neighbour weight couples exactly that neighbour; with the neighbour weights zero the operator is
    block diagonal with blocks diag(S) (I + o[0:25]) diag(1/S); with o[0:25] = 0 it is independent
    of S and equals (1 + s_4) I on the diagonal and s_j I on the block of neighbour j."""
    n = 10
    nbr, rng = _graph(n, 1)
    Ot = rng.standard_normal((n, 30)) * 0.3
    S = np.array([1.0, 0.5, 0.5, 0.5, 2.5])
    P = _dense(Ot, nbr, S)
    r = rng.standard_normal(5 * n)
    assert np.allclose(P @ r, O.mlp_apply(Ot, nbr, S, r))  # linear in r
    for c in range(n):
        for d in range(n):
            if d != c and d not in nbr[c]:
                assert np.all(_blk(P, n, c, d) == 0.0)
    # neighbour weights off: block diagonal, diag(S) B diag(1/S)
    Ob = Ot.copy()
    Ob[:, 25:29] = 0.0
    Ob[:, 29] = 0.0
    Pb = _dense(Ob, nbr, S)
    for c in range(n):
        B = np.eye(5) + Ob[c, :25].reshape(5, 5)
        assert np.allclose(_blk(Pb, n, c, c), np.diag(S) @ B @ np.diag(1.0 / S), atol=1e-14)
        for d in range(n):
            if d != c:
                assert np.all(_blk(Pb, n, c, d) == 0.0)
    # block part off: scalar couplings, S cancels
    Os = Ot.copy()
    Os[:, :25] = 0.0
    Ps = _dense(Os, nbr, np.array([3.0, 0.2, 1.0, 7.0, 0.5]))
    for c in range(n):
        assert np.allclose(_blk(Ps, n, c, c), (1.0 + Os[c, 29]) * np.eye(5), atol=1e-14)
        for j in range(4):
            d = nbr[c, j]
            if d >= 0:
                assert np.allclose(_blk(Ps, n, c, d), Os[c, 25 + j] * np.eye(5), atol=1e-14)


NI, NH, NO = 20, 128, 30


def _unpack(par):
    """Parameter packing W0[20][128] b0 W1[256][128] b1 W2[256][128] b2 W3[128][30] b3, row-major
    [in][out] (the library's documented layout)."""
    sizes = [(NI, NH), (NH,), (2 * NH, NH), (NH,), (2 * NH, NH), (NH,), (NH, NO), (NO,)]
    out, k = [], 0
    for sh in sizes:
        m = int(np.prod(sh))
        out.append(par[k:k + m].reshape(sh))
        k += m
    assert k == len(par)
    return out


def _forward_numpy(X, nbr, par):
    """Brute-force forward written from the network's definition: explicit matrix products, SiLU as
    x / (1 + exp(-x)), and the neighbour mean as an explicit loop over neib (facets and -1 skipped,
    mean 0 for a cell with no neighbour)."""
    W0, b0, W1, b1, W2, b2, W3, b3 = _unpack(par)
    n = X.shape[0]

    def silu(x):
        return x / (1.0 + np.exp(-x))

    def mean_nb(h):
        m = np.zeros_like(h)
        for c in range(n):
            ds = [d for d in nbr[c] if 0 <= d < n]
            for d in ds:
                m[c] += h[d]
            if ds:
                m[c] /= len(ds)
        return m

    h0 = silu(X @ W0 + b0)
    h1 = silu(np.concatenate([h0, mean_nb(h0)], axis=1) @ W1 + b1)
    h2 = silu(np.concatenate([h1, mean_nb(h1)], axis=1) @ W2 + b2)
    kappa = np.minimum(1.0, X[:, 18] * (1.0 + X[:, 19]))  # the cell's CFL number, at most 1
    return kappa[:, None] * (h2 @ W3 + b3)


def test_forward_matches_numpy_brute_force():
    n = 10
    nbr, rng = _graph(n, 4)
    nbr[5] = -1  # an isolated cell
    for c in range(n):
        nbr[c][nbr[c] == 5] = -1
    nbr[2, 3] = n + 3  # a facet id (outside the cell range) is skipped
    par = rng.standard_normal(O.mlp_nparams()) * 0.1
    X = rng.standard_normal((n, NI))
    got = O.mlp_forward(X, nbr, par)
    ref = _forward_numpy(X, nbr, par)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_forward_closed_forms():
    rng = np.random.default_rng(5)
    npar = O.mlp_nparams()
    n = 6
    nbr, _ = _graph(n, 6)
    X = rng.standard_normal((n, NI))
    kappa = np.minimum(1.0, X[:, 18] * (1.0 + X[:, 19]))
    # every weight zero, biases only: the output is b3 scaled by the cell's CFL number
    par = np.zeros(npar)
    W0, b0, W1, b1, W2, b2, W3, b3 = _unpack(par)
    b0[:] = rng.standard_normal(NH)
    b1[:] = rng.standard_normal(NH)
    b2[:] = rng.standard_normal(NH)
    b3[:] = rng.standard_normal(NO)
    assert np.allclose(O.mlp_forward(X, nbr, par), kappa[:, None] * b3[None, :], atol=0, rtol=1e-15)
    # a cell at rest with a vanishing time step (feature 18 = 0) gets o = 0: the identity
    X0 = X.copy()
    X0[2, 18] = 0.0
    assert not np.any(O.mlp_forward(X0, nbr, par)[2])
    # W1's own-cell half = I, W2's mean half = I, W3 = e_0 -> e_0, biases zero: h1 = silu(silu(x W0)),
    # o_0 = silu(mean over the neighbours of h1[:, 0]); a cell's own h1 does not reach o_0 (a swapped
    # [h, mean h] half, a sum for the mean or another activation fails this)
    par = np.zeros(npar)
    W0, b0, W1, b1, W2, b2, W3, b3 = _unpack(par)
    W0[:] = rng.standard_normal((NI, NH)) * 0.3
    W1[:NH] = np.eye(NH)
    W2[NH:] = np.eye(NH)
    W3[0, 0] = 1.0
    silu = lambda x: x / (1.0 + np.exp(-x))  # noqa: E731
    h1 = silu(silu(X @ W0))
    o = O.mlp_forward(X, nbr, par)
    for c in range(n):
        ds = [d for d in nbr[c] if 0 <= d < n]
        m = np.mean([h1[d, 0] for d in ds]) if ds else 0.0
        assert abs(o[c, 0] - kappa[c] * silu(m)) <= 1e-14 * max(1.0, abs(kappa[c]))
    # a single cell with no neighbours: the mean halves of W1 and W2 have no effect
    nb1 = -np.ones((1, 4), np.int32)
    par = rng.standard_normal(npar) * 0.1
    p2 = par.copy()
    W0, b0, W1, b1, W2, b2, W3, b3 = _unpack(p2)
    W1[NH:] = rng.standard_normal((NH, NH))
    W2[NH:] = rng.standard_normal((NH, NH))
    x1 = rng.standard_normal((1, NI))
    assert np.array_equal(O.mlp_forward(x1, nb1, par), O.mlp_forward(x1, nb1, p2))


def test_loss_is_scaled_masked_mean():
    """The loss equals mean over masked cells of |(P r - delta) / S|^2 with P r from the (pinned)
    forward and apply."""
    n = 8
    nbr, rng = _graph(n, 7)
    par = rng.standard_normal(O.mlp_nparams()) * 0.08
    X = rng.standard_normal((n, NI))
    S = np.array([1.0, 0.7, 0.7, 0.7, 2.0])
    r = rng.standard_normal(5 * n)
    delta = rng.standard_normal(5 * n)
    mask = np.array([1, 0, 1, 1, 0, 1, 1, 0], np.int32)
    loss, _ = O.mlp_loss_grad(X, nbr, par, S, r, delta, mask)
    z = O.mlp_apply(O.mlp_forward(X, nbr, par), nbr, S, r)
    e = ((z - delta).reshape(5, n).T / S) ** 2
    assert abs(loss - e[mask == 1].sum() / mask.sum()) <= 1e-13 * max(1.0, loss)


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

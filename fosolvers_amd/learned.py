"""Training driver of the learned preconditioner (host orchestration over the C ABI).

The kernels (forward, loss, backward, Adam) are the library's; this module chooses what they are
trained on.  Two kinds of (input, target) pairs, both at states the solver itself visited:

* Newton triples from the FP64 BDF with block ILU (``BDFNewtonKrylov.last_triple``): the
  predictor y_pred, r = -G(y_pred) and the converged correction delta = y_new - y_pred.  The loss is
  |S^-1 (P r - delta)|^2, the configured ||P r - delta_newton||^2.
* ILU pairs at the same states: seeded random vectors r (white, and smoothed over the face
  neighbours) and z = ILU^-1 r from the block ILU(0) solve, so the plain ILU solve is the judge of
  the learned operator's error on the directions GMRES explores, not only on the one Newton
  direction per step.
* Block-Jacobi pairs at the same states: white r and z = M_cc^-1 r_c with the exact diagonal
  blocks of the Newton matrix, which pins the 5x5 block the network emits to the local inverse
  (a per-cell target its features determine), while the ILU pairs shape the neighbour weights.

Batches are ranges of consecutive cells (the chunk renumbering makes them compact sub-meshes);
message passing stays inside a batch, as the library's training kernels do.
"""
import numpy as np

from . import Euler  # noqa: F401  (type of the context the trainer drives)


KIND_TRIPLE, KIND_ILU_WHITE, KIND_ILU_SMOOTH, KIND_BJAC = 0, 1, 2, 3


class Trainer:
    """Adam on the library's kernels over interleaved kinds of batches: a Newton triple, an ILU pair
    on a white random vector, an ILU pair on a neighbour-smoothed random vector and a block-Jacobi
    pair on a white random vector.  Pairs are drawn fresh each step at a triple's predictor state
    (the preconditioner is set up once per state and reused for `per_state` steps), so the network
    sees many directions per state."""

    def __init__(self, euler, batch=65536, lr=1e-3, seed=0, per_state=8, mix=(1, 2, 1)):
        import torch
        self.eu = euler
        self.nC = euler.nC
        self.batch = min(batch, self.nC)
        self.lr = lr
        self.per_state = per_state
        self.mix = mix  # steps per state block, in units of per_state: (block Jacobi, ILU, triples)
        self.triples = []  # (y_pred, r, delta, gamma)
        self.pair_ctx = []  # context that forms the pairs at each triple's state
        self.rng = np.random.default_rng(seed)
        self.gen = torch.Generator(device="cuda")
        self.gen.manual_seed(seed)
        self._nbr = None
        self._ilu_state = None
        self._bjac_state = None
        self._S = None


    def add_triple(self, ypred, r, delta, gamma, pair_ctx=None):
        """One Newton triple (device vectors; copied).  pair_ctx: the Euler context (same mesh, the
        triple's own boundary conditions) that sets up the ILU and block-Jacobi pairs at this state;
        default the trained context."""
        self.triples.append((ypred.clone(), r.clone(), delta.clone(), float(gamma)))
        self.pair_ctx.append(pair_ctx if pair_ctx is not None else self.eu)

    def _neighbours(self):
        import torch
        if self._nbr is None:
            nb = np.asarray(self.eu.grid.neib, dtype=np.int64).reshape(self.nC, -1)
            nb = np.where((nb >= 0) & (nb < self.nC), nb, -1)
            self._nbr = torch.tensor(nb, device="cuda")
        return self._nbr

    def _smooth(self, z, passes=3):
        import torch
        nbr = self._neighbours()
        z = z.view(5, self.nC)
        for _ in range(passes):
            acc = z.clone()
            cnt = torch.ones(self.nC, device=z.device, dtype=z.dtype)
            for j in range(nbr.shape[1]):
                m = nbr[:, j] >= 0
                idx = nbr[:, j].clamp(min=0)
                acc += torch.where(m, z[:, idx], torch.zeros_like(z))
                cnt += m.to(z.dtype)
            z = acc / cnt
        return z.reshape(-1)

    def ilu_pair(self, k, smooth):
        """(r, ILU^-1 r) at the predictor state of triple k: block ILU(0) set up there (once)."""
        import torch
        y, _r, _d, gamma = self.triples[k]
        ctx = self.pair_ctx[k]
        if self._ilu_state != k:
            fy = torch.zeros_like(y)
            ctx.rhs(y, fy)
            ctx.psetup(y, fy, gamma)
            self._S = self._scale(y, gamma)
            self._ilu_state = k
            self._bjac_state = None  # the factors now hold the ILU
        r = torch.randn(5 * self.nC, generator=self.gen, device=y.device, dtype=torch.float64)
        if smooth:
            r = self._smooth(r)
            r = r / r.std()
        r = r * self._S
        z = torch.zeros_like(r)
        ctx.psolve(r, z, gamma)
        return y, r, z, gamma

    def _scale(self, y, gamma):
        import torch
        _, S = self.eu.learned_inputs(y, gamma)
        return torch.tensor(np.repeat(S.astype(np.float64), self.nC), device=y.device)

    def bjac_pair(self, k):
        """(r, M_D^-1 r) at the predictor state of triple k, M_D the block diagonal of the Newton
        matrix."""
        import torch
        y, _r, _d, gamma = self.triples[k]
        ctx = self.pair_ctx[k]
        if self._bjac_state != k:
            ctx.psetup_bjac(y, y, gamma)
            self._Sb = self._scale(y, gamma)
            self._bjac_state = k
            self._ilu_state = None  # the factors now hold the block inverses
        r = torch.randn(5 * self.nC, generator=self.gen, device=y.device, dtype=torch.float64) * self._Sb
        z = torch.zeros_like(r)
        ctx.psolve_bjac(r, z, gamma)
        return y, r, z, gamma

    def _batch_start(self):
        if self.batch >= self.nC:
            return 0
        nb = (self.nC - self.batch) // 256 + 1
        return int(self.rng.integers(0, nb)) * 256

    def train(self, steps):
        """Adam steps in blocks per state (state k = block index mod triples): mix[0] * per_state
        block-Jacobi pairs, mix[1] * per_state ILU pairs (white and smooth alternating), then
        mix[2] * per_state Newton triples, so each preconditioner is set up once per block.
        Returns the (kind, loss) arrays."""
        if not self.triples:
            raise ValueError("no Newton triples")
        kinds, losses = [], []
        nt, per = len(self.triples), self.per_state
        nb, ni, ntr = (m * per for m in self.mix)
        blk = nb + ni + ntr
        for it in range(steps):
            k = (it // blk) % nt
            pos = it % blk
            if pos < nb:
                kind = KIND_BJAC
                y, r, t, gamma = self.bjac_pair(k)
            elif pos < nb + ni:
                kind = KIND_ILU_WHITE if (pos - nb) % 2 == 0 else KIND_ILU_SMOOTH
                y, r, t, gamma = self.ilu_pair(k, smooth=kind == KIND_ILU_SMOOTH)
            else:
                kind = KIND_TRIPLE
                y, r, t, gamma = self.triples[it % nt]
            s0 = self._batch_start()
            losses.append(self.eu.learned_train_step(y, r, t, gamma, s0, min(self.batch, self.nC - s0), lr=self.lr))
            kinds.append(kind)
        return np.array(kinds), np.array(losses)

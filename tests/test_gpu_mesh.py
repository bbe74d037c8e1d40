"""polyFvGrid%up on the GPU (fos_grid_up_gpu) against the host build (fos_grid_up), which is
pinned against the compiled reference: every array equal, bit for bit, on tetrahedral and
hexahedral meshes, renumbered and not, with unmatched cell faces, and the non-manifold error."""
import numpy as np
import pytest

import fosolvers_amd as fa
from tests import meshes as M

pytestmark = pytest.mark.gpu

SCAL = ("nN", "nE", "nC", "nF", "nP", "nPi", "nPb", "maxNF", "maxNNE")


def _same(a, b):
    for k in SCAL:
        assert getattr(a, k) == getattr(b, k), k
    for k in fa.FvGrid.FIELDS:
        x, y = np.asarray(getattr(a, k)), np.asarray(getattr(b, k))
        assert x.shape == y.shape, k
        assert np.array_equal(x, y), (k, np.abs(x - y).max() if x.dtype.kind == "f" else None)


@pytest.mark.parametrize("case", ["tet-1", "tet-jitter", "tet-chunks", "tet-216k", "hex", "missing-facets"])
def test_grid_up_gpu_equals_host(cuda, case):
    print(f"[case] grid up gpu {case}", flush=True)
    if case == "tet-1":
        pg = M.box(1, 1, 1)
    elif case == "tet-jitter":
        pg = M.box(6, 4, 3, jitter=0.2, seed=2)
    elif case == "tet-chunks":
        pg = M.box(12, 8, 6, jitter=0.2, seed=3, renumber=fa.RENUMBER_CHUNKS)
    elif case == "tet-216k":
        pg = M.box(40, 30, 30, jitter=0.2, seed=4)
        pg.renumber(fa.RENUMBER_RCM)
        pg.renumber(fa.RENUMBER_CHUNKS)
    elif case == "hex":
        pg = M.hex_box(5, 4, 3, jitter=0.15, seed=5)
    else:  # every third boundary triangle dropped: those cell faces get no pair
        full = M.box(5, 3, 3, jitter=0.1, seed=6)
        keep = [e for e in range(full.nE) if full.sE[e] != fa.TRI or e % 3]
        pg = fa.PolyGrid(full.pN, full.sE[keep], full.nNE[keep], full.iNE[keep], full.gid[keep])
    _same(pg.up(gpu=True), pg.up())


def test_grid_up_gpu_rejects_non_manifold(cuda):
    """Three tetrahedra on one triangle: both builds refuse the mesh with a message."""
    print("[case] grid up gpu non-manifold", flush=True)
    pN = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1], [0.3, 0.3, 1]], float)
    iNE = np.array([[0, 1, 2, 3], [0, 2, 1, 4], [0, 1, 2, 5]], np.int32)
    pg = fa.PolyGrid(pN, np.full(3, fa.TET, np.int32), np.full(3, 4, np.int32), iNE, np.zeros(3, np.int32))
    for gpu in (False, True):
        with pytest.raises(RuntimeError, match="non-manifold"):
            pg.up(gpu=gpu)


@pytest.mark.parametrize("case", ["tet-jitter", "tet-chunks", "hex"])
def test_grid_up_gpu_matches_oracle(cuda, case):
    """The GPU build against the oracle's restatement of getNeib + sortPolyFvGrid directly (not only
    through the host build): after fos_grid_sort_pairs_reference the pair order is the reference's,
    and every pair's cells, area, normal and centre, the volumes, centroids and neighbours agree."""
    from oracle import oracle as O
    print(f"[case] grid up gpu vs oracle {case}", flush=True)
    if case == "tet-jitter":
        pg = M.box(6, 4, 3, jitter=0.2, seed=2)
    elif case == "tet-chunks":
        pg = M.box(12, 8, 6, jitter=0.2, seed=3, renumber=fa.RENUMBER_CHUNKS)
    else:
        pg = M.hex_box(5, 4, 3, jitter=0.15, seed=5)
    fv = pg.up(gpu=True).sort_pairs_reference()
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    assert (fv.nC, fv.nP, fv.nPi, fv.nPb) == (og.nC, og.nP, og.nPi, og.nPb)
    assert np.array_equal(fv.iEP, og.iEP)
    assert np.abs(fv.aP - og.aP).max() <= 1e-15
    assert np.abs(fv.normP - og.normP).max() <= 1e-14
    assert np.abs(fv.pP - og.pP).max() <= 1e-16
    # the volume as a triple product against the reference's cofactor expansion: both cancel the
    # jittered coordinates' leading digits, so they agree to a few ulps of the products, ~1e-13 of v
    assert np.abs(fv.v - og.v).max() <= 1e-13 * og.v.max()
    assert np.abs(fv.p - og.p).max() <= 1e-15
    assert np.array_equal(np.sort(fv.neib, 1), np.sort(og.neib, 1))
    assert np.array_equal(fv.gid, og.gid) and np.array_equal(fv.sE, og.sE)

"""Host side of the product library (no GPU): the C ABI loads and exports every declared symbol,
the mesh generator / grid build / renumbering / readers behave like the reference's."""
import ctypes
import os
import re

import numpy as np
import pytest

import fosolvers_amd as fa
from oracle import oracle as O
from tests import meshes as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"typedef[^;]*;", "", txt)
        names += re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    l = ctypes.CDLL(fa._capi.LIB_PATH)
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(l, n)]
    assert not missing, missing
    fa.lib()  # signatures bind


def test_box_mesh_invariants():
    pg = fa.PolyGrid.box(5, 4, 3, 2.0, 1.0, 0.5, jitter=0.2, seed=9)
    fv = pg.up()
    assert fv.nC == 6 * 5 * 4 * 3 and fv.nF == 4 * (4 * 3 + 5 * 3 + 5 * 4)
    assert 4 * fv.nC == 2 * fv.nPi + fv.nPb and fv.nPb == fv.nF
    assert np.all(fv.v > 0) and abs(fv.v.sum() - 1.0) <= 1e-13
    assert np.allclose(np.linalg.norm(fv.normP, axis=1), 1.0, atol=1e-15)
    # boundary pairs last, internal pairs m < n
    assert np.all(fv.iEP[:fv.nPi, 1] < fv.nC) and np.all(fv.iEP[fv.nPi:, 1] >= fv.nC)
    assert np.all(fv.iEP[:fv.nPi, 0] < fv.iEP[:fv.nPi, 1])
    # each closed cell: sum of outward area vectors is zero; normals point out of m
    S = np.zeros((fv.nC, 3))
    for i, (m, n) in enumerate(fv.iEP):
        S[m] += fv.aP[i] * fv.normP[i]
        if n < fv.nC:
            S[n] -= fv.aP[i] * fv.normP[i]
    assert np.abs(S).max() <= 1e-15
    out = ((fv.pP - fv.p[fv.iEP[:, 0]]) * fv.normP).sum(1)
    assert np.all(out > 0)


def test_grid_up_matches_oracle_pair_set():
    pg = M.box(4, 3, 3, jitter=0.2, seed=5, renumber=fa.RENUMBER_RCM)
    fv = pg.up()
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    assert (fv.nC, fv.nP, fv.nPi, fv.nPb) == (og.nC, og.nP, og.nPi, og.nPb)
    key = lambda g: {(int(a), int(b)): i for i, (a, b) in enumerate(g.iEP)}  # noqa: E731
    kf, ko = key(fv), key(og)
    assert set(kf) == set(ko)
    i_f = np.array([kf[k] for k in ko])
    i_o = np.array([ko[k] for k in ko])
    assert np.abs(fv.aP[i_f] - og.aP[i_o]).max() <= 1e-16
    assert np.abs(fv.normP[i_f] - og.normP[i_o]).max() <= 4e-16
    assert np.abs(fv.pP[i_f] - og.pP[i_o]).max() <= 1e-16
    assert np.abs(fv.v - og.v).max() <= 1e-14 * og.v.max()  # triple product vs cofactor form
    assert np.array_equal(np.sort(fv.neib, 1), np.sort(og.neib, 1))


@pytest.mark.parametrize("case", ["2x2x2", "4x3x3-jitter", "6x4x4-rcm", "8x5x4-chunks", "hex"])
def test_reference_pair_order(case):
    """fos_grid_sort_pairs_reference reproduces sortPolyFvGrid's pair order exactly -- the oracle's
    restatement of getNeib + the quicksort, and the compiled reference where it is built -- with each
    pair's geometry carried along."""
    if case == "2x2x2":
        pg = M.box(2, 2, 2)
    elif case == "4x3x3-jitter":
        pg = M.box(4, 3, 3, jitter=0.2, seed=5)
    elif case == "6x4x4-rcm":
        pg = M.box(6, 4, 4, jitter=0.2, seed=8, renumber=fa.RENUMBER_RCM)
    elif case == "8x5x4-chunks":
        pg = M.box(8, 5, 4, jitter=0.1, seed=9, renumber=fa.RENUMBER_CHUNKS)
    else:
        pg = M.hex_box(4, 3, 3, jitter=0.15, seed=4)
    fv = pg.up()
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    if case != "2x2x2":
        assert not np.array_equal(fv.iEP, og.iEP)  # the default tie order differs
    fv.sort_pairs_reference()
    assert np.array_equal(fv.iEP, og.iEP)
    assert np.abs(fv.aP - og.aP).max() <= 1e-15
    assert np.abs(fv.normP - og.normP).max() <= 1e-14
    assert np.abs(fv.pP - og.pP).max() <= 1e-16
    assert (fv.iEP[fv.nPi:, 1] >= fv.nC).all()  # boundary pairs last
    if O.ref_available():
        rg = O.RefGrid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
        assert np.array_equal(fv.iEP, rg.iEP)


def test_hex_grid_up_matches_oracle():
    """8-node hexahedra with quadrilateral boundary facets (the reference's polyhedral path): the
    same pair set, areas, normals, volumes and neighbours as the oracle's grid build."""
    pg = M.hex_box(5, 4, 3, jitter=0.15, seed=4)
    fv = pg.up()
    og = O.Grid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    assert (fv.nC, fv.maxNF) == (60, 6)
    assert (fv.nP, fv.nPi, fv.nPb) == (og.nP, og.nPi, og.nPb) == (227, 133, 94)
    assert abs(fv.v.sum() - 0.25) <= 1e-14
    key = lambda g: {(int(a), int(b)): i for i, (a, b) in enumerate(g.iEP)}  # noqa: E731
    kf, ko = key(fv), key(og)
    assert set(kf) == set(ko)
    i_f = np.array([kf[k] for k in ko])
    i_o = np.array([ko[k] for k in ko])
    assert np.abs(fv.aP[i_f] - og.aP[i_o]).max() <= 1e-15
    assert np.abs(fv.normP[i_f] - og.normP[i_o]).max() <= 1e-14
    assert np.abs(fv.v - og.v).max() <= 1e-14 * og.v.max()
    # closed cells: the outward area vectors of each hexahedron sum to zero when the quads are
    # planar.  A warped quad's area is the sum of its two triangles' areas and its normal the
    # normalised sum of their normals (geometry.f90:30-57), not its vector area, so jittered hexes
    # do not close exactly.
    fv0 = M.hex_box(5, 4, 3, jitter=0.0).up()
    acc = np.zeros((fv0.nC, 3))
    for (m, n), a, nv in zip(fv0.iEP, fv0.aP, fv0.normP):
        acc[m] += a * nv
        if n < fv0.nC:
            acc[n] -= a * nv
    assert np.abs(acc).max() <= 1e-15


def test_renumbering():
    pg = M.box(8, 6, 6, jitter=0.2, seed=2)
    fv0 = pg.up()
    bw0 = np.abs(fv0.iEP[:fv0.nPi, 0] - fv0.iEP[:fv0.nPi, 1]).max()
    perm = pg.renumber(fa.RENUMBER_RCM)
    assert sorted(perm.tolist()) == list(range(fv0.nC))
    fv1 = pg.up()
    bw1 = np.abs(fv1.iEP[:fv1.nPi, 0] - fv1.iEP[:fv1.nPi, 1]).max()
    assert bw1 < bw0
    assert abs(fv1.v.sum() - fv0.v.sum()) <= 1e-14
    pg.renumber(fa.RENUMBER_CHUNKS)
    fv2 = pg.up()
    # chunks of 256 cells are compact: most internal faces stay inside a chunk
    ch = fv2.iEP[:fv2.nPi] // fa.CHUNK_CELLS
    inside = float((ch[:, 0] == ch[:, 1]).mean())
    assert inside > 0.7, inside
    # an element naming a node outside the grid is an error, not an out-of-bounds read
    bad = M.box(2, 2, 2)
    bad.iNE[0, 0] = bad.nN
    for method in (fa.RENUMBER_RCM, fa.RENUMBER_CHUNKS):
        with pytest.raises(RuntimeError, match="outside"):
            bad.renumber(method)


def test_grid_validation_before_any_index_use(tmp_path):
    """One host-side check guards the grid build (host and GPU), the renumbering and the reader: a
    node id outside the grid, a node count that differs from the element's shape, or one above
    maxNNE is refused with a message before any index is used.  fos_grid_up_gpu refuses it before
    any device work, so it returns the error on a machine without a GPU too."""
    def cases():
        g = M.box(2, 2, 2)
        g.iNE[0, 0] = g.nN  # node id outside [0, nN)
        yield g, "outside"
        g = M.box(2, 2, 2)
        g.iNE[3, 2] = -1  # a -1 inside the element's own node count
        yield g, "outside"
        g = M.box(2, 2, 2)
        c = int(np.where(g.sE == fa.TET)[0][0])
        g.nNE[c] = 3  # a tet with 3 nodes (the rest -1 padding)
        g.iNE[c, 3] = -1
        yield g, "nodes"
        g = M.box(2, 2, 2)
        f = int(np.where(g.sE == fa.TRI)[0][0])
        g.nNE[f] = 4  # a triangle claiming 4 nodes
        yield g, "nodes"
    for g, msg in cases():
        c, out = g._to_c(), fa._capi.FvGridC()
        for fn in ("fos_grid_up", "fos_grid_up_gpu"):
            args = (ctypes.byref(c), ctypes.byref(out)) + ((None,) if fn.endswith("gpu") else ())
            assert getattr(fa.lib(), fn)(*args) == -1
            assert msg in fa.lib().fos_last_error().decode(), fn
            assert out.nC == 0 and not out.iEP  # nothing built
        for method in (fa.RENUMBER_RCM, fa.RENUMBER_CHUNKS):
            iNE0 = g.iNE.copy()
            with pytest.raises(RuntimeError, match=msg):
                g.renumber(method)
            assert np.array_equal(g.iNE, iNE0)  # unchanged
    # the reader: a GMSH v2 file whose tet names node 5 of 4
    p = tmp_path / "bad.msh"
    p.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\n"
                 "$EndNodes\n$Elements\n1\n1 4 2 0 0 1 2 3 5\n$EndElements\n")
    with pytest.raises(RuntimeError, match="outside"):
        fa.PolyGrid.read_gmsh(str(p))


def test_gmsh_roundtrip(tmp_path):
    pg = M.box(3, 2, 2, jitter=0.1, seed=4)
    p = str(tmp_path / "g.msh")
    pg.write_gmsh(p)
    q = fa.PolyGrid.read_gmsh(p)
    assert np.array_equal(q.sE, pg.sE) and np.array_equal(q.gid, pg.gid) and np.array_equal(q.iNE, pg.iNE)
    assert np.abs(q.pN - pg.pN).max() == 0.0


def test_gmsh_v4_reader(tmp_path):
    # one tet and one boundary triangle in GMSH 4.1 ASCII (entity tag = gid)
    txt = """$MeshFormat
4.1 0 8
$EndMeshFormat
$Nodes
1 4 1 4
3 1 0 4
1
2
3
4
0 0 0
1 0 0
0 1 0
0 0 1
$EndNodes
$Elements
2 2 1 2
2 7 2 1
1 1 3 2
3 9 4 1
2 1 2 3 4
$EndElements
"""
    p = tmp_path / "v4.msh"
    p.write_text(txt)
    g = fa.PolyGrid.read_gmsh(str(p))
    assert g.nN == 4 and g.nE == 2
    assert list(g.sE) == [fa.TRI, fa.TET] and list(g.gid) == [7, 9]
    fv = g.up()
    assert fv.nC == 1 and fv.nPb == 1 and fv.nPi == 0 and abs(fv.v[0] - 1 / 6) <= 1e-16


def test_cond_tab_kat():
    """src/test/readCondTab1.f90 on its own data file."""
    c = fa.CondTab.read(os.path.join(GOLD, "condition_simple"))
    assert list(c.gid) == [10, 20] and list(c.t) == [1, 2]
    want = np.array([[300.0, 1e5, 20.0, 0.0, 0.0], [0.8e5, 0.0, 0.0, 0.0, 0.0]])
    assert np.abs(c.p - want).max() <= 1e-13 * 1e5


def test_udf_kat():
    """src/test/UDF1.f90 on its own data file: values at (1,2,3), t=4."""
    u = fa.UDFTab.read(os.path.join(GOLD, "udf"))
    x = (1.0, 2.0, 3.0)
    assert abs(u.eval(1, x, 4.0) - 1234.0) <= 1e-8
    assert abs(u.eval(2, x, 4.0) - 4.0) <= 1e-8
    assert abs(u.eval(3, x, 4.0) - 1.0) <= 1e-8


@pytest.mark.parametrize("expr,want", [
    ("1+2*3", 7.0), ("-2^2", -4.0), ("2^3^2", 64.0), ("(1+2)*(3-5)/4", -1.5), ("step(x-0.5)", 1.0),
    ("exp(0)+sqrt(16)+abs(-2)", 7.0), ("sin(0)+cos(0)", 1.0), ("x*y-z/t", 2 - 0.75), ("1e3*x+10.*t", 1040.0)])
def test_udf_grammar(expr, want):
    u = fa.UDFTab([expr])
    assert abs(u.eval(1, (1.0, 2.0, 3.0), 4.0) - want) <= 1e-12


def test_udf_parse_error():
    with pytest.raises(ValueError):
        fa.UDFTab(["1+*2"])


def test_library_missing_is_loud(monkeypatch):
    from fosolvers_amd import _capi
    monkeypatch.setattr(_capi, "_lib", None)
    monkeypatch.setattr(_capi, "LIB_PATH", "/nonexistent/libfosolvers_amd.so")
    with pytest.raises(_capi.LibraryMissing):
        _capi.lib()

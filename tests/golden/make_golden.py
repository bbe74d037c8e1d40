"""Write the golden fixtures of the residual path from the compiled reference (oracle/_ref).

Run in the container that holds /root/reference (the reference build is made there):
    make -C oracle ref && python tests/golden/make_golden.py
Fixtures (inputs + the reference's outputs, small):
  box2_rhs.npz   2x2x2-hex jittered box (48 tets): element lists, a seeded state and the full
                 rhs vectors of the reference for each BC family, plus findEulerJac blocks.
  sodA_rhs.npz   40x10x10-hex Sod tube (24,000 tets, all walls): rhs norms, 64 sampled cells and
                 SHA-256 of the FP64 rhs vector, for the unperturbed Sod state.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import fosolvers_amd as fa  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests import meshes as M  # noqa: E402

BC_FAMILIES = ["wall", "far", "mixed", "supersonic"]


def box2():
    pg = fa.PolyGrid.box(2, 2, 2, 1.0, 0.5, 0.5, jitter=0.2, seed=7)
    rg = O.RefGrid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    x = M.perturbed_state(rg.p[:rg.nC], mach=0.7, seed=3, width=0.3)
    out = dict(pN=pg.pN, sE=pg.sE, nNE=pg.nNE, iNE=pg.iNE, gid=pg.gid, x=x, iEP=rg.iEP, aP=rg.aP,
               normP=rg.normP, v=rg.v, neib=rg.neib)
    for fam in BC_FAMILIES:
        cg, ct, cp = M.conds(fam)
        rg.setup_state(cg, ct, cp, M.R, M.GAMMA)
        out["rhs_" + fam] = rg.rhs(x)
        rho, rhou, rhoE, p = rg.state()
        out["jac_" + fam] = rg.find_euler_jac(rho, rhou, rhoE, p, M.GAMMA)
    np.savez_compressed(os.path.join(HERE, "box2_rhs.npz"), **out)


def sod_a():
    pg = fa.PolyGrid.box(40, 10, 10, 1.0, 0.25, 0.25)
    rg = O.RefGrid(pg.pN, pg.sE, pg.nNE, pg.iNE, pg.gid)
    x = M.sod_state(rg.p[:rg.nC])
    cg, ct, cp = M.conds("wall")
    rg.setup_state(cg, ct, cp, M.R, M.GAMMA)
    r = rg.rhs(x)
    idx = np.random.default_rng(0).choice(rg.nC, 64, replace=False)
    samples = np.stack([r[k * rg.nC + idx] for k in range(5)], 1)
    np.savez_compressed(os.path.join(HERE, "sodA_rhs.npz"), idx=idx, samples=samples,
                        norms=np.array([np.linalg.norm(r[k * rg.nC:(k + 1) * rg.nC]) for k in range(5)]),
                        sha256=np.frombuffer(hashlib.sha256(r.tobytes()).digest(), dtype=np.uint8),
                        nC=rg.nC, nP=rg.nP)


if __name__ == "__main__":
    box2()
    sod_a()
    print("fixtures written to", HERE)

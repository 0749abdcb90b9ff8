"""Writes the committed parity fixtures from the FP64 CPU oracle.

Each fixture is the oracle's converged solution (outer loop driven to
alternativeTolerance 1e-12 so that the fixed point, not the stopping point,
is recorded) of one configuration, stored as float64 arrays in .npz:
  cfg1_cantilever.npz   50x10x10 box 1.0x0.2x0.2 m, x-min fixed, x-max traction (0,-1e6,0)
  cube16_jitter.npz     16^3 unit cube, interior vertices +/-15% jitter (seed 2),
                        z-min fixed, z-max traction (0,0,1e7)
  tet4_jitter.npz       4^3 unit cube split into 6 tets per hex, +/-15% jitter (seed 5),
                        extendedLeastSquares (full boundary delta) gradient

Run:  python tests/golden/make_golden.py   (from the repository root)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402
from tests._cases import E_STEEL, NU_STEEL, cantilever_bc, cube_tension_bc  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
TIGHT = dict(nCorr=200000, solTol=1e-12, altTol=1e-12, pcgTol=1e-13, pcgRelTol=0.1, precond=O.PRECOND_DIC)


def run(name, mesh_args, bc_fn, ls_full_delta=False):
    t = time.time()
    m = O.Mesh.box(*mesh_args[:3], **mesh_args[3])
    pt, bv = bc_fn(m.nBoundaryFaces, [m.patch_faces(p) for p in range(m.nPatches)])
    c = O.Case(m, E_STEEL, NU_STEEL, pt, bv, ls_full_delta=ls_full_delta)
    r = c.solve(**TIGHT)
    D = c.get("D").reshape(-1, 3)
    sig = c.get("sigma").reshape(-1, 6)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), D=D, sigma=sig, nOuter=r["nOuter"], rvf=r["rvf"],
                        initRes=r["initRes"])
    print(name, m.nCells, r, f"{time.time() - t:.1f}s", flush=True)


if __name__ == "__main__":
    which = sys.argv[1:] or ["cube16_jitter", "tet4_jitter", "cfg1_cantilever"]
    if "cube16_jitter" in which:
        run("cube16_jitter", (16, 16, 16, dict(L=(1, 1, 1), jitter=0.15, seed=2)), cube_tension_bc)
    if "tet4_jitter" in which:
        run("tet4_jitter", (4, 4, 4, dict(L=(1, 1, 1), jitter=0.15, seed=5, tets=True)), cube_tension_bc,
            ls_full_delta=True)
    if "cfg1_cantilever" in which:
        run("cfg1_cantilever", (50, 10, 10, dict(L=(1.0, 0.2, 0.2))), cantilever_bc)

"""Mixed precision's fallback to FP64 sweeps (the stall test of the FP64
defect loop under FP32 sweeps).  The fallback is forced from the host: with
zero outer tolerances the loop cannot stop before nCorrectors, |initRes|
settles at rounding level, and a short stall window hands the rest of the
loop to FP64 sweeps from the state it has reached.  The result is the FP64
fixed point; an ordinary mixed solve does not fall back; batches report
their PCG sweeps per instance."""
import numpy as np
import pytest

from tests._cases import E_STEEL, NU_STEEL, cube_tension_bc

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ctx():
    from solids4foam_amd import Context
    c = Context(0)
    yield c
    c.close()


def _cube(ctx, n=(8, 7, 6), seed=3):
    from solids4foam_amd import PolyMesh, SolverMesh
    pm = PolyMesh.box(*n, L=(1, 1, 1), jitter=0.15, seed=seed)
    sl = [pm.patch_faces(p) for p in range(6)]
    pt, bv = cube_tension_bc(pm.nBoundaryFaces, sl)
    m = SolverMesh(ctx, pm, pt)
    m.set_material(E_STEEL, NU_STEEL)
    for p in range(6):
        m.set_bc(p, bv[sl[p]])
    return pm, m


def test_forced_fallback_reaches_the_fp64_fixed_point(ctx):
    print("[fallback] forced", flush=True)
    pm, m = _cube(ctx)
    ref = m.solve(precision=0, nCorrectors=20000, solutionTolerance=1e-12, alternativeTolerance=1e-12,
                  pcgTolerance=1e-13)
    assert ref["converged"]
    out = m.solve(precision=1, nCorrectors=3000, solutionTolerance=0.0, alternativeTolerance=0.0,
                  pcgTolerance=1e-13, mixedStallWindow=5)
    print(f"  nOuter {out['nOuter']} fallback {out['mixedFallback']} err {rel_l2(out['D'], ref['D']):.2e}",
          flush=True)
    assert out["mixedFallback"] == 1
    assert out["nOuter"] == 3000 and not out["converged"]
    assert rel_l2(out["D"], ref["D"]) < 1e-10
    assert out["pcgIterations"] > 0


def test_ordinary_mixed_solve_does_not_fall_back(ctx):
    pm, m = _cube(ctx)
    out = m.solve(precision=1, nCorrectors=20000, solutionTolerance=1e-10, alternativeTolerance=1e-10)
    assert out["converged"] and out["mixedFallback"] == 0


def test_batch_reports_pcg_sweeps_per_instance(ctx):
    from solids4foam_amd.batch import Batch
    ms = [_cube(ctx, n, seed) for n, seed in (((6, 5, 4), 1), ((9, 4, 4), 2))]
    b = Batch(ctx)
    for _, m in ms:
        b.add(m)
    res = b.solve(precision=1, nCorrectors=20000, solutionTolerance=1e-9, alternativeTolerance=1e-9)
    alone = [m.solve(precision=1, nCorrectors=20000, solutionTolerance=1e-9, alternativeTolerance=1e-9)
             for _, m in ms]
    for r, a in zip(res, alone):
        assert r["converged"] and r["pcgIterations"] == a["pcgIterations"] > 0

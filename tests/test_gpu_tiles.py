"""Workgroup-tile rules (s4f_mesh_tiles_set, the autotuner's knob): another
rule changes only the reduction order (the solve agrees with the default to
rounding), a batch member keeps its own rule (batch == alone, bitwise, and a
rule change re-packs), invalid rules are refused, and the autotuner leaves a
mesh on the cheapest candidate it timed."""
import numpy as np
import pytest

from tests._cases import E_STEEL, NU_STEEL, cube_tension_bc

pytestmark = pytest.mark.gpu

OPTS = dict(nCorrectors=20000, solutionTolerance=1e-10, alternativeTolerance=1e-10, pcgTolerance=1e-12)


@pytest.fixture(scope="module")
def ctx():
    from solids4foam_amd import Context
    c = Context(0)
    yield c
    c.close()


def _cube(ctx, n, seed):
    from solids4foam_amd import PolyMesh, SolverMesh
    pm = PolyMesh.box(*n, L=(1, 1, 1), jitter=0.15, seed=seed)
    sl = [pm.patch_faces(p) for p in range(6)]
    pt, bv = cube_tension_bc(pm.nBoundaryFaces, sl)
    m = SolverMesh(ctx, pm, pt)
    m.set_material(E_STEEL, NU_STEEL)
    for p in range(6):
        m.set_bc(p, bv[sl[p]])
    return pm, m


def test_tile_rule_changes_only_rounding(ctx):
    print("[tiles] rule vs default", flush=True)
    pm, m = _cube(ctx, (20, 18, 16), 3)
    ref = m.solve(precision=0, **OPTS)
    for cap, mn in ((3, 4), (1000, 1), (7, 16)):
        m.set_tiles(cap, mn)
        out = m.solve(precision=0, **OPTS)
        err = np.linalg.norm(out["D"] - ref["D"]) / np.linalg.norm(ref["D"])
        print(f"  rule {cap}/{mn}: err {err:.2e}", flush=True)
        assert out["converged"] and err < 1e-10


def test_batch_member_keeps_its_rule(ctx):
    from solids4foam_amd import S4FError
    from solids4foam_amd.batch import Batch
    a = _cube(ctx, (9, 8, 7), 1)[1]
    b = _cube(ctx, (12, 6, 5), 2)[1]
    a.set_tiles(65536, 1)
    bt = Batch(ctx)
    bt.add(a)
    bt.add(b)
    res = bt.solve(precision=1, **OPTS)
    alone = [a.solve(precision=1, **OPTS), b.solve(precision=1, **OPTS)]
    for r, s in zip(res, alone):
        assert np.array_equal(r["D"], s["D"])
    b.set_tiles(2, 8)  # the packing follows
    res2 = bt.solve(precision=1, **OPTS)
    assert np.array_equal(res2[1]["D"], b.solve(precision=1, **OPTS)["D"])
    with pytest.raises(S4FError):
        a.set_tiles(0, 4)
    with pytest.raises(S4FError):
        a.set_tiles(10, 0)


def test_autotuner_keeps_the_cheapest_rule(ctx):
    from solids4foam_amd import autotune
    pm, m = _cube(ctx, (24, 24, 24), 2)
    r = autotune.tune(m, candidates=[(8192, 4), (65536, 1), (4, 16)], launches=4)
    costs = {(row["max_tiles"], row["min_slices"]): row["ms_per_outer"] for row in r["rows"]}
    print(f"[tiles] autotune {costs} -> {r['best']}", flush=True)
    assert r["best"] == min(costs, key=costs.get)
    assert all(v > 0 for v in costs.values())
    out = m.solve(precision=1, **OPTS)
    assert out["converged"]

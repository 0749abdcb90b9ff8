"""State changes between solves (GPU): a batch repacks when a member's
material or the stabilisation scales change, packs the steady matrix of a
mesh whose last solve was transient, the server and batches honour the
solve's Rhie-Chow scales, planeStress switched after the material takes
effect, and nu = 0.5 is refused.  Every comparison is with the same solve run
alone (bitwise, as an instance's tiles and reductions depend only on itself).
"""
import numpy as np
import pytest

from tests._cases import E_STEEL, NU_STEEL, FIXED, TRACTION, cube_tension_bc

pytestmark = pytest.mark.gpu

OPTS = dict(solutionTolerance=1e-9, alternativeTolerance=1e-9, nCorrectors=20000)


@pytest.fixture(scope="module")
def ctx():
    from solids4foam_amd import Context
    c = Context(0)
    yield c
    c.close()


def _mesh(ctx, n, seed, E=E_STEEL, nu=NU_STEEL):
    from solids4foam_amd import PolyMesh, SolverMesh
    pm = PolyMesh.box(*n, L=(1.0, 1.0, 1.0), jitter=0.1, seed=seed)
    sl = [pm.patch_faces(p) for p in range(6)]
    pt, bv = cube_tension_bc(pm.nBoundaryFaces, sl)
    m = SolverMesh(ctx, pm, pt)
    m.set_material(E, nu)
    for p in range(6):
        m.set_bc(p, bv[sl[p]])
    return pm, m, bv


def _same(r, a):
    assert r["converged"] and a["converged"]
    assert r["nOuter"] == a["nOuter"]
    assert np.array_equal(r["D"], a["D"]) and np.array_equal(r["sigma"], a["sigma"])


def test_batch_repacks_after_material_change(ctx):
    from solids4foam_amd.batch import Batch
    print("[state] batch repack on material change", flush=True)
    _, m1, _ = _mesh(ctx, (7, 6, 5), 1)
    _, m2, _ = _mesh(ctx, (6, 6, 6), 2)
    b = Batch(ctx)
    b.add(m1)
    b.add(m2)
    first = b.solve(**OPTS)
    m1.set_material(90e9, 0.35)
    second = b.solve(**OPTS)
    _, ref, _ = _mesh(ctx, (7, 6, 5), 1, E=90e9, nu=0.35)
    _same(second[0], ref.solve(**OPTS))
    assert not np.array_equal(first[0]["D"], second[0]["D"])
    _same(second[1], m2.solve(**OPTS))
    b.close()


def test_batch_after_transient_solve_is_steady(ctx):
    # the packed matrix is the steady one (no inertia diagonal of the
    # member's previous transient solve)
    from solids4foam_amd.batch import Batch
    print("[state] batch after a transient solve", flush=True)
    pm, m, _ = _mesh(ctx, (6, 5, 5), 3)
    m.set_density(7800.0)
    m.set_old_time(np.zeros((pm.nCells, 3)))
    assert m.solve(d2dt2=1, deltaT=1e-4, **OPTS)["converged"]
    b = Batch(ctx)
    b.add(m)
    res = b.solve(**OPTS)
    _, ref, _ = _mesh(ctx, (6, 5, 5), 3)
    _same(res[0], ref.solve(**OPTS))
    b.close()


def test_batch_rhie_chow_scale(ctx):
    from solids4foam_amd.batch import Batch
    print("[state] batch with rhieChowScale 0.2", flush=True)
    _, m1, _ = _mesh(ctx, (7, 6, 5), 4)
    _, m2, _ = _mesh(ctx, (5, 6, 7), 5)
    b = Batch(ctx)
    b.add(m1)
    b.add(m2)
    opts = dict(OPTS, rhieChowScale=0.2)
    d = b.solve(**OPTS)
    res = b.solve(**opts)
    _same(res[0], m1.solve(**opts))
    _same(res[1], m2.solve(**opts))
    assert not np.array_equal(res[0]["D"], d[0]["D"])
    b.close()


def test_server_rhie_chow_scale(ctx):
    import time
    from solids4foam_amd.serve import Server
    print("[state] server with rhieChowScale 0.2", flush=True)
    meshes = [_mesh(ctx, (7, 6, 5), 6), _mesh(ctx, (6, 6, 6), 7)]
    opts = dict(OPTS, rhieChowScale=0.2, pcgRelTol=0.1)
    alone = [m.solve(**opts) for _, m, _ in meshes]
    srv = Server(ctx, 128 << 20, max_slots=4, **opts)
    ids = [srv.submit(m, bv) for _, m, bv in meshes]
    got = {}
    t0 = time.time()
    while len(got) < len(ids) and time.time() - t0 < 120:
        rid = srv.poll(100)
        if rid is not None:
            got[rid] = srv.result(rid)
    srv.close()
    assert len(got) == len(ids)
    for rid, a in zip(ids, alone):
        _same(got[rid], a)


def test_plane_stress_switch_after_material(ctx):
    from solids4foam_amd import PATCH_EMPTY as EMPTY, PolyMesh, SolverMesh
    print("[state] planeStress after the material", flush=True)
    pm = PolyMesh.box(12, 4, 1, L=(1.0, 0.2, 0.05))
    pt = [FIXED, TRACTION, TRACTION, TRACTION, EMPTY, EMPTY]
    bv = np.zeros((pm.nBoundaryFaces, 3))
    bv[pm.patch_faces(1)] = (0.0, 1e6, 0.0)
    out = []
    for before in (True, False):
        m = SolverMesh(ctx, pm, pt)
        if before:
            m.set_plane_stress(True)
        m.set_material(200e9, 0.3)
        if not before:
            m.set_plane_stress(True)
        for p in range(4):
            m.set_bc(p, bv[pm.patch_faces(p)])
        out.append(m.solve(**OPTS))
    _same(out[1], out[0])


def test_incompressible_nu_is_refused(ctx):
    from solids4foam_amd import _abi
    _, m, _ = _mesh(ctx, (3, 3, 3), 8)
    with pytest.raises(_abi.S4FError, match="pressure equation") as e:
        m.set_material(E_STEEL, 0.5)
    assert e.value.status == _abi.S4F_ERR_INVALID

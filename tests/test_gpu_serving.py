"""Serving (continuous batching) on the GPU: each request's result equals the
same solve run alone (bitwise: an instance's tiles and reductions depend only
on itself), oversize requests are refused by HBM-footprint admission, repeated
meshes warm start from their resident field, and requests that arrive while
others are in flight join them."""
import time

import numpy as np
import pytest

from tests._cases import E_STEEL, NU_STEEL, cantilever_bc, cube_tension_bc

pytestmark = pytest.mark.gpu

OPTS = dict(solutionTolerance=1e-8, alternativeTolerance=1e-8, nCorrectors=20000, pcgRelTol=0.1)


@pytest.fixture(scope="module")
def ctx():
    from solids4foam_amd import Context
    c = Context(0)
    yield c
    c.close()


def _mesh(ctx, n, L=(1.0, 1.0, 1.0), jitter=0.1, seed=0, E=E_STEEL, nu=NU_STEEL, bc="cube"):
    from solids4foam_amd import PolyMesh, SolverMesh
    pm = PolyMesh.box(*n, L=L, jitter=jitter, seed=seed)
    sl = [pm.patch_faces(p) for p in range(6)]
    pt, bv = (cube_tension_bc if bc == "cube" else cantilever_bc)(pm.nBoundaryFaces, sl)
    m = SolverMesh(ctx, pm, pt)
    m.set_material(E, nu)
    for p in range(6):
        m.set_bc(p, bv[sl[p]])
    return m, bv


def _specs():
    return [((10, 8, 6), dict(seed=1)), ((12, 4, 4), dict(L=(1.0, 0.3, 0.3), bc="cantilever", jitter=0.0)),
            ((6, 6, 6), dict(seed=2, E=100e9, nu=0.25)), ((20, 9, 5), dict(seed=3)),
            ((8, 8, 8), dict(seed=4, E=300e9, nu=0.4)), ((16, 16, 3), dict(seed=5))]


@pytest.mark.parametrize("precision", [0, 1])
def test_server_results_equal_alone(ctx, precision):
    from solids4foam_amd.serve import Server
    print(f"[serving] equal-alone precision={precision}", flush=True)
    meshes = [_mesh(ctx, n, **kw) for n, kw in _specs()]
    alone = [m.solve(precision=precision, **OPTS) for m, _ in meshes]
    srv = Server(ctx, 256 << 20, max_slots=8, precision=precision, **OPTS)
    ids = [srv.submit(m, bv) for m, bv in meshes]
    got = {}
    t0 = time.time()
    while len(got) < len(ids) and time.time() - t0 < 120:
        rid = srv.poll(100)
        if rid is not None:
            got[rid] = srv.result(rid)
    srv.close()
    assert len(got) == len(ids)
    for rid, a in zip(ids, alone):
        r = got[rid]
        assert r["converged"] and r["nOuter"] == a["nOuter"]
        assert np.array_equal(r["D"], a["D"])
        assert np.array_equal(r["sigma"], a["sigma"])
        assert r["latency"] > 0


@pytest.mark.parametrize("precision", [0, 1])
def test_server_multigrid_equals_alone(ctx, precision):
    # the server packs each request's own multigrid hierarchy into one
    # block-diagonal hierarchy (shallower ones pass through to their dense
    # block): every result is bitwise the multigrid solve alone
    from solids4foam_amd import PRECOND_GAMG
    from solids4foam_amd.serve import Server
    print(f"[serving] multigrid equal-alone precision={precision}", flush=True)
    meshes = [_mesh(ctx, n, **kw) for n, kw in _specs()]
    opts = dict(OPTS, preconditioner=PRECOND_GAMG)
    alone = [m.solve(precision=precision, **opts) for m, _ in meshes]
    srv = Server(ctx, 256 << 20, max_slots=4, precision=precision, **opts)
    ids = []
    for m, bv in meshes:  # 4 slots for 6 requests: the packed set changes as they retire
        ids.append(srv.submit(m, bv))
    got = {}
    t0 = time.time()
    while len(got) < len(ids) and time.time() - t0 < 120:
        rid = srv.poll(100)
        if rid is not None:
            got[rid] = srv.result(rid)
    srv.close()
    assert len(got) == len(ids)
    for rid, a in zip(ids, alone):
        r = got[rid]
        assert r["converged"] and r["nOuter"] == a["nOuter"] and r["innerTotal"] == a["innerTotal"]
        assert np.array_equal(r["D"], a["D"]) and np.array_equal(r["sigma"], a["sigma"])


def test_server_keeps_plain_multigrid(ctx):
    # the server packs plain-aggregation hierarchies: it refuses the smoothed
    # preconditioner, and a mesh that also holds a smoothed hierarchy (from a
    # solve alone) is served with its plain one, bitwise the plain solve alone
    from solids4foam_amd import PRECOND_GAMG, PRECOND_GAMG_SMOOTHED, S4FError
    from solids4foam_amd.serve import Server
    print("[serving] smoothed multigrid refused, plain kept", flush=True)
    with pytest.raises(S4FError):
        Server(ctx, 256 << 20, max_slots=2, precision=1, **dict(OPTS, preconditioner=PRECOND_GAMG_SMOOTHED))
    m, bv = _mesh(ctx, (20, 9, 5), seed=3)
    sm = m.solve(precision=1, **dict(OPTS, preconditioner=PRECOND_GAMG_SMOOTHED))
    alone = m.solve(precision=1, **dict(OPTS, preconditioner=PRECOND_GAMG))
    assert sm["converged"] and alone["converged"]
    srv = Server(ctx, 256 << 20, max_slots=2, precision=1, **dict(OPTS, preconditioner=PRECOND_GAMG))
    rid = srv.submit(m, bv)
    r = None
    t0 = time.time()
    while r is None and time.time() - t0 < 60:
        got = srv.poll(100)
        if got is not None:
            r = srv.result(got)
    srv.close()
    assert r is not None and r["converged"] and r["nOuter"] == alone["nOuter"]
    assert np.array_equal(r["D"], alone["D"]) and np.array_equal(r["sigma"], alone["sigma"])


def test_server_aitken_equals_alone(ctx):
    # Aitken relaxation state lives per cell and boundary face of each
    # request's arena range: a served request gets the bits it gets alone
    from solids4foam_amd import _abi
    from solids4foam_amd.serve import Server
    meshes = [_mesh(ctx, n, **kw) for n, kw in _specs()[:4]]
    opts = dict(OPTS, relaxationMethod=_abi.RELAX_AITKEN)
    alone = [m.solve(**opts) for m, _ in meshes]
    srv = Server(ctx, 256 << 20, max_slots=2, **opts)
    ids = [srv.submit(m, bv) for m, bv in meshes]
    got = {}
    t0 = time.time()
    while len(got) < len(ids) and time.time() - t0 < 120:
        rid = srv.poll(100)
        if rid is not None:
            got[rid] = srv.result(rid)
    srv.close()
    assert len(got) == len(ids)
    for rid, a in zip(ids, alone):
        r = got[rid]
        assert r["converged"] and r["nOuter"] == a["nOuter"]
        assert np.array_equal(r["D"], a["D"]) and np.array_equal(r["sigma"], a["sigma"])


def test_server_admission_by_footprint(ctx):
    from solids4foam_amd import S4FError, _abi
    from solids4foam_amd.serve import Server
    srv = Server(ctx, 4 << 20, max_slots=4, **OPTS)
    st = srv.stats()
    small, bv_s = _mesh(ctx, (6, 6, 6), seed=1)
    big, bv_b = _mesh(ctx, (40, 40, 40), seed=1)
    with pytest.raises(S4FError) as e:
        srv.submit(big, bv_b)
    assert e.value.status == _abi.S4F_ERR_ADMISSION
    rid = srv.submit(small, bv_s)
    while srv.poll(200) != rid:
        pass
    assert srv.result(rid)["converged"]
    assert st["cell_capacity"] < big.nCells
    srv.close()


def test_server_warm_start_repeat(ctx):
    from solids4foam_amd.serve import Server
    m, bv = _mesh(ctx, (10, 10, 10), seed=7)
    srv = Server(ctx, 128 << 20, max_slots=4, **OPTS)
    r1 = srv.result(_wait(srv, srv.submit(m, bv)))
    # same mesh and load, warm start from the resident converged field
    r0 = srv.result(_wait(srv, srv.submit(m, bv, warm_start=True)))
    # new load (scaled): warm and cold reach the same fixed point.  The
    # reference's denominator is |D - D_old| with D_old the warm-start field,
    # so a changed load is judged relative to its increment.
    r2 = srv.result(_wait(srv, srv.submit(m, 1.5 * bv, warm_start=True)))
    r3 = srv.result(_wait(srv, srv.submit(m, 1.5 * bv)))  # cold
    srv.close()
    assert r1["converged"] and r0["converged"] and r2["converged"] and r3["converged"]
    assert r0["warm_started"] and r2["warm_started"] and not r1["warm_started"] and not r3["warm_started"]
    assert np.linalg.norm(r0["D"] - r1["D"]) / np.linalg.norm(r1["D"]) < 1e-6
    assert np.linalg.norm(r2["D"] - r3["D"]) / np.linalg.norm(r3["D"]) < 1e-6
    assert np.linalg.norm(r3["D"] - 1.5 * r1["D"]) / np.linalg.norm(r3["D"]) < 1e-6  # linearity


def _wait(srv, rid, timeout=60):
    t0 = time.time()
    while time.time() - t0 < timeout:
        got = srv.poll(100)
        if got == rid:
            return rid
    raise TimeoutError(rid)


def test_server_streaming_arrivals(ctx):
    from solids4foam_amd.serve import Server
    meshes = [_mesh(ctx, n, **kw) for n, kw in _specs()]
    alone = [m.solve(**OPTS) for m, _ in meshes]
    srv = Server(ctx, 256 << 20, max_slots=3, **OPTS)
    ids = []
    for m, bv in meshes:  # arrive one by one while earlier ones run (3 slots)
        ids.append(srv.submit(m, bv))
        time.sleep(0.005)
    got = {}
    t0 = time.time()
    while len(got) < len(ids) and time.time() - t0 < 120:
        rid = srv.poll(100)
        if rid is not None:
            got[rid] = srv.result(rid)
    srv.close()
    assert len(got) == len(ids)
    starts = sorted(got[i]["t_start"] for i in ids)
    finishes = sorted(got[i]["t_finish"] for i in ids)
    assert starts[-1] > finishes[0]  # later requests were admitted after earlier ones retired
    for rid, a in zip(ids, alone):
        assert np.array_equal(got[rid]["D"], a["D"])

"""The benchmark workloads' own generators, at sizes the oracle solves in
seconds, against the FP64 CPU oracle (called through the C ABI; converged
fields, plain FP64 to 1e-8 and mixed precision to 1e-6 relative L2):

  heterogeneous tetrahedra   jittered 6^3 cube split 6 tets per hex, seeded
                             spherical inclusions at E = 20 GPa in a 200 GPa
                             matrix, x-min fixed, x-max shear 5 MPa
  batched sweep              boxes from the sweep generator (uneven sizes,
                             log-normal E, nu up to 0.45, random tip traction,
                             field relaxation 0.6, multigrid PCG, FP32 sweeps),
                             each instance of one batch against its oracle solve
  serving stream             requests from the serving generator (new meshes
                             and repeated meshes with a new load, warm started)
                             through the continuous-batching server, each
                             result against the oracle solve of its mesh + load
  Cook's membrane            the reference tutorial's skewed 2-D mesh, field
                             against the oracle and the top-right displacement
                             against its README value
"""
import time

import numpy as np
import pytest

from tests._cases import FIXED, TRACTION

pytestmark = pytest.mark.gpu

TIGHT = dict(nCorrectors=200000, solutionTolerance=1e-12, alternativeTolerance=1e-12, pcgTolerance=1e-13)
ORACLE_TIGHT = dict(nCorr=200000, solTol=1e-12, altTol=1e-12, pcgTol=1e-13)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ctx():
    from solids4foam_amd import Context
    c = Context(0)
    yield c
    c.close()


def _bar(precision):
    return 1e-8 if precision == 0 else 1e-6


@pytest.mark.parametrize("zones", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_heterogeneous_tets_match_oracle(ctx, oracle_mod, precision, zones):
    from solids4foam_amd import cases
    n = 6
    # inclusion radii scaled to the coarse mesh so that several spheres
    # capture cell centroids (the full-size case uses 0.02-0.06 at n = 200);
    # zones: matrix and inclusions as two laws on their own sub-meshes
    pm, m, frac, k, E = cases.tet_inclusions(ctx, n=n, jitter=0.15, seed=2, inc_seed=3, rmin=0.12, rmax=0.22,
                                             zones=zones)
    print(f"[cfg tets] {pm.nCells} tets, {k} inclusions, {frac:.1%} soft cells, precision {precision}, "
          f"sub-meshes {zones}", flush=True)
    assert 0.02 < frac < 0.5
    out = m.solve(precision=precision, **TIGHT)
    om = oracle_mod.Mesh.box(n, n, n, (1.0, 1.0, 1.0), 0.15, 2, tets=True)
    assert np.array_equal(om.C, pm.cell_centres())
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(1)] = (0.0, 5e6, 0.0)
    oc = oracle_mod.Case(om, E, 0.3, [FIXED] + [TRACTION] * 5, bv, ls_full_delta=True)
    if zones:
        assert oc.set_zones((E != E.max()).astype(np.int32)) > 0
        oc.init()
    ro = oc.solve(**ORACLE_TIGHT)
    Do, so = oc.get("D").reshape(-1, 3), oc.get("sigma").reshape(-1, 6)
    eD, es = rel_l2(out["D"], Do), rel_l2(out["sigma"], so)
    from solids4foam_amd import sigma_eq
    eq = rel_l2(sigma_eq(out["sigma"]), oc.sigma_eq())
    print(f"  nOuter {out['nOuter']} (oracle {ro['nOuter']}) err D {eD:.2e} sigma {es:.2e} sigmaEq {eq:.2e}",
          flush=True)
    assert out["converged"] and ro["converged"]
    assert eD < _bar(precision) and es < _bar(precision) and eq < _bar(precision)


def _oracle_box(oracle_mod, spec, traction, relax):
    om = oracle_mod.Mesh.box(*spec["n"], spec["L"], spec["jitter"], spec["seed"])
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(1)] = traction
    oc = oracle_mod.Case(om, spec["E"], spec["nu"], [FIXED] + [TRACTION] * 5, bv)
    ro = oc.solve(relaxD=relax, **ORACLE_TIGHT)
    assert ro["converged"], spec
    return oc.get("D").reshape(-1, 3), oc.get("sigma").reshape(-1, 6)


def _sweep_specs():
    from solids4foam_amd import cases
    specs = cases.box_sweep_specs(6, 17, nmin=4, nmax=9)
    specs[0]["nu"] = 0.44  # the sweep's stiff end: the loop needs the field relaxation
    assert max(sp["nu"] for sp in specs) >= 0.4
    return specs


def test_batched_sweep_matches_oracle(ctx, oracle_mod):
    from solids4foam_amd import PRECOND_GAMG, cases
    from solids4foam_amd.batch import Batch
    specs = _sweep_specs()
    meshes = [cases.box_case(ctx, sp) for sp in specs]
    b = Batch(ctx)
    for _, m, _ in meshes:
        b.add(m)
    print(f"[cfg batch] {len(specs)} boxes {[sp['n'] for sp in specs]}", flush=True)
    res = b.solve(precision=1, preconditioner=PRECOND_GAMG, relaxD=0.6, pcgRelTol=0.1, **TIGHT)
    for sp, r in zip(specs, res):
        Do, so = _oracle_box(oracle_mod, sp, sp["traction"], 0.6)
        eD, es = rel_l2(r["D"], Do), rel_l2(r["sigma"], so)
        print(f"  {sp['n']} nu {sp['nu']:.3f} E {sp['E']:.3g}: nOuter {r['nOuter']} err D {eD:.2e} "
              f"sigma {es:.2e}", flush=True)
        assert r["converged"]
        assert eD < 1e-6 and es < 1e-6


def test_serving_stream_matches_oracle(ctx, oracle_mod):
    from solids4foam_amd import PRECOND_GAMG, cases
    from solids4foam_amd.serve import Server
    specs, stream, arrivals = cases.serve_requests(10, 23, nmin=4, nmax=9, repeat=0.3, rate=200.0)
    assert any(w for _, _, w in stream), "the stream must repeat a mesh"
    meshes = [cases.box_case(ctx, sp) for sp in specs]
    srv = Server(ctx, 256 << 20, max_slots=4, precision=1, preconditioner=PRECOND_GAMG, relaxD=0.6,
                 pcgRelTol=0.1, **TIGHT)
    loads, ids, got = [], {}, {}
    t0 = time.perf_counter()
    for k, (j, t, warm) in enumerate(stream):
        while time.perf_counter() - t0 < arrivals[k]:
            time.sleep(0.001)
        pm, m, bv = meshes[j]
        trac = specs[j]["traction"] if t is None else t
        if t is not None:
            bv = np.zeros((pm.nBoundaryFaces, 3))
            bv[pm.patch_faces(1)] = t
        loads.append(trac)
        ids[srv.submit(m, bv, warm_start=warm)] = k
    while len(got) < len(stream) and time.perf_counter() - t0 < 300:
        rid = srv.poll(100)
        if rid is not None:
            got[ids[rid]] = srv.result(rid)
    srv.close()
    assert len(got) == len(stream)
    print(f"[cfg serve] {len(stream)} requests on {len(specs)} meshes, "
          f"{sum(g['warm_started'] for g in got.values())} warm", flush=True)
    for k, (j, t, warm) in enumerate(stream):
        Do, so = _oracle_box(oracle_mod, specs[j], loads[k], 0.6)
        r = got[k]
        eD, es = rel_l2(r["D"], Do), rel_l2(r["sigma"], so)
        print(f"  request {k}: mesh {j} {specs[j]['n']} warm {warm} nOuter {r['nOuter']} err D {eD:.2e} "
              f"sigma {es:.2e}", flush=True)
        assert r["converged"]
        assert eD < 1e-6 and es < 1e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_cooks_membrane_matches_oracle_and_reference_value(ctx, oracle_mod, precision):
    # tutorials/solids/linearElasticity/cooksMembrane on its own skewed 20x20
    # mesh (extendedLeastSquares 0, 2-D): converged field against the oracle
    # and the top-right vertical displacement against the README's 32.12 mm
    from solids4foam_amd import PATCH_EMPTY, PolyMesh, SolverMesh
    from tests import _blockmesh as bm
    from tests.test_oracle_pins import cooks_membrane, top_right_uy
    a = bm.single_block(bm.COOK_CORNERS, (20, 20, 1), bm.COOK_PATCHES)
    pm = PolyMesh.from_arrays(a["points"], a["fOff"], a["fVerts"], a["owner"], a["neighbour"], a["nCells"],
                              a["patchStart"], a["patchSize"])
    # the zero-size defaultFaces patch (hex_blocks appends it) carries no values: typed empty
    m = SolverMesh(ctx, pm, [TRACTION, FIXED, TRACTION, PATCH_EMPTY, PATCH_EMPTY], ls_full_delta=True)
    m.set_material(bm.COOK_E, bm.COOK_NU)
    m.set_bc(0, (0.0, 0.0, 0.0))
    m.set_bc(1, (0.0, 0.0, 0.0))
    m.set_bc(2, bm.COOK_TRACTION)
    out = m.solve(precision=precision, **TIGHT)
    om, oc = cooks_membrane(oracle_mod, 20)
    ro = oc.solve(**ORACLE_TIGHT)
    eD = rel_l2(out["D"], oc.get("D").reshape(-1, 3))
    pD = m.point_displacement()
    p = int(np.argmin(np.linalg.norm(om.points - np.array(bm.COOK_TOP_RIGHT), axis=1)))
    uy, uyo = float(pD[p, 1]), top_right_uy(oracle_mod, om, oc)
    print(f"[cook] precision {precision}: nOuter {out['nOuter']} (oracle {ro['nOuter']}) err D {eD:.2e} "
          f"uy {uy * 1e3:.4f} mm (oracle {uyo * 1e3:.4f}, README {bm.COOK_UY_README * 1e3:.2f})", flush=True)
    assert out["converged"] and ro["converged"]
    assert eD < _bar(precision)
    assert abs(uy / uyo - 1.0) < _bar(precision)
    assert abs(uy / bm.COOK_UY_README - 1.0) < 0.01

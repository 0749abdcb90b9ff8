"""Sub-mesh (multi-law) discretisation on the GPU against the FP64 oracle.

Cases with one mechanical law per cellZone solve each law on its own
sub-mesh (mechanicalModel.C:571-615, solidSubMeshes.C:1550-2031): the faces
between zones carry an interface displacement (k_iface_pre), each side's
gradient sees it as a boundary value (k_grad's interface slots) and each
side's boundary stress feeds the next update (k_iface_post).

  history     first outer iterations of a jittered cube with a soft
              spherical inclusion follow the oracle's |initRes|, residualvf
              and PCG counts (diagonal preconditioner on both sides)
  converged   the same case converged, FP64 to 1e-8 and mixed precision to
              1e-6 relative L2 (D, sigma, pointD), normal and full
              least-squares patch delta
  column      the bi-material column's exact piecewise-linear field (uniform
              and graded cells), which the per-cell mapping misses
  guards      batches and the server refuse a mesh with zones; clearing the
              zones gives the one-material solve bit for bit
"""
import numpy as np
import pytest

from tests._cases import FIXED, TRACTION

pytestmark = pytest.mark.gpu

SYM = 3
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


def _inclusion(ctx, oracle_mod, n=8, full=False, zones=True):
    from solids4foam_amd.solid import PolyMesh, SolverMesh
    om = oracle_mod.Mesh.box(n, n, n, (1.0, 1.0, 1.0), 0.15, 3)
    z = (np.linalg.norm(om.C - 0.5, axis=1) < 0.3).astype(np.int32)
    E = np.where(z == 0, 200e9, 20e9)
    nu = np.where(z == 0, 0.3, 0.45)
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(5)] = (2e6, 0.0, 1e7)
    types = [TRACTION] * 4 + [FIXED, TRACTION]
    oc = oracle_mod.Case(om, E, nu, types, bv, ls_full_delta=full)
    pm = PolyMesh.box(n, n, n, L=(1.0, 1.0, 1.0), jitter=0.15, seed=3)
    assert np.array_equal(pm.cell_centres(), om.C)
    m = SolverMesh(ctx, pm, types, ls_full_delta=full)
    m.set_material(E, nu)
    for p in range(6):
        m.set_bc(p, bv[pm.patch_faces(p)])
    if zones:
        assert oc.set_zones(z) > 0
        oc.init()
        m.set_zones(z)
    return oc, m, pm, om


def test_submesh_history_follows_oracle(ctx, oracle_mod):
    print("[submesh history] 8^3 inclusion", flush=True)
    oc, m, pm, om = _inclusion(ctx, oracle_mod)
    ro = oc.solve(history=12, nCorr=12, precond=oracle_mod.PRECOND_DIAGONAL)
    m.solve(nCorrectors=12)
    h = m.history()
    n = min(len(h["initRes"]), len(ro["history"]["initRes"]))
    assert n >= 10
    assert np.allclose(h["initRes"][:n], ro["history"]["initRes"][:n], rtol=1e-8, atol=1e-14)
    assert np.allclose(h["rvf"][:n], ro["history"]["rvf"][:n], rtol=1e-8)
    assert np.array_equal(h["inner"][:n], ro["history"]["inner"][:n])


@pytest.mark.parametrize("full", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_submesh_converged_matches_oracle(ctx, oracle_mod, precision, full):
    print(f"[submesh converged] precision {precision} full delta {full}", flush=True)
    oc, m, pm, om = _inclusion(ctx, oracle_mod, full=full)
    ro = oc.solve(precond=oracle_mod.PRECOND_DIC, **ORACLE_TIGHT)
    out = m.solve(precision=precision, **TIGHT)
    assert ro["converged"] and out["converged"]
    Do, so = oc.get("D").reshape(-1, 3), oc.get("sigma").reshape(-1, 6)
    bar = 1e-8 if precision == 0 else 1e-6
    eD, es = rel_l2(out["D"], Do), rel_l2(out["sigma"], so)
    pD = m.point_displacement()
    pDo = oc.point_interpolate(oc.get("D"), oc.get("G")).reshape(-1, 3)
    ep = rel_l2(pD, pDo)
    print(f"  nOuter {out['nOuter']} (oracle {ro['nOuter']}) err D {eD:.2e} sigma {es:.2e} pointD {ep:.2e}",
          flush=True)
    assert eD < bar and es < bar and ep < bar
    if precision == 1:  # the multigrid preconditioner sees the base-mesh matrix: the same fixed point
        from solids4foam_amd import _abi
        _, mg, _, _ = _inclusion(ctx, oracle_mod, full=full)
        outg = mg.solve(precision=1, preconditioner=_abi.PRECOND_GAMG, pcgRelTol=0.9, **TIGHT)
        eg = rel_l2(outg["D"], Do)
        print(f"  multigrid: nOuter {outg['nOuter']} err D {eg:.2e}", flush=True)
        assert outg["converged"] and eg < 1e-6
    # the per-cell mapping of the same materials is a different solution
    oc1, m1, _, _ = _inclusion(ctx, oracle_mod, full=full, zones=False)
    out1 = m1.solve(precision=0, **TIGHT)
    assert rel_l2(out1["D"], Do) > 1e-3


def _column(ctx, oracle_mod, zi, zones):
    from solids4foam_amd.solid import PolyMesh, SolverMesh
    b = oracle_mod.Mesh.box(3, 3, 8, (1.0, 1.0, 1.0))
    p = b.points.copy()
    zz = p[:, 2]
    p[:, 2] = np.where(zz <= 0.5, zz * zi / 0.5, zi + (zz - 0.5) * (1.0 - zi) / 0.5)
    pm = PolyMesh.from_arrays(p, b.fOff, b.fVerts, b.owner, b.neighbour, b.nCells, b.patchStart, b.patchSize)
    Cz = pm.cell_centres()[:, 2]
    z = (Cz > zi).astype(np.int32)
    mats = ((200e9, 0.3), (70e9, 0.33))
    E = np.where(z == 0, mats[0][0], mats[1][0])
    nu = np.where(z == 0, mats[0][1], mats[1][1])
    m = SolverMesh(ctx, pm, [SYM, SYM, SYM, SYM, FIXED, TRACTION])
    m.set_material(E, nu)
    for q in range(5):
        m.set_bc(q, (0.0, 0.0, 0.0))
    m.set_bc(5, (0.0, 0.0, 1e7))
    if zones:
        m.set_zones(z)
    k = [e * v / ((1 + v) * (1 - 2 * v)) + e / (1 + v) for e, v in mats]
    exact = np.where(Cz <= zi, 1e7 * Cz / k[0], 1e7 * zi / k[0] + 1e7 * (Cz - zi) / k[1])
    return m, exact


@pytest.mark.parametrize("zi", [0.5, 0.4])
def test_bimaterial_column_exact(ctx, oracle_mod, zi):
    print(f"[submesh column] interface at z = {zi}", flush=True)
    # graded cells: no interface Rhie-Chow term (the exact field's kink
    # feeds it where the two sides' cells differ)
    kw = dict(TIGHT, interfaceScale=0.0) if zi != 0.5 else TIGHT
    m, exact = _column(ctx, oracle_mod, zi, True)
    out = m.solve(precision=0, **kw)
    assert out["converged"]
    err = np.abs(out["D"][:, 2] - exact).max() / np.abs(exact).max()
    assert err < 1e-9, err
    assert np.allclose(out["sigma"][:, 5], 1e7, rtol=1e-9)
    m0, _ = _column(ctx, oracle_mod, zi, False)
    out0 = m0.solve(precision=0, **kw)
    err0 = np.abs(out0["D"][:, 2] - exact).max() / np.abs(exact).max()
    print(f"  sub-mesh {err:.1e}, per-cell mapping {err0:.2%}", flush=True)
    assert err0 > 1e-2


def test_zones_guards_and_clear(ctx, oracle_mod):
    from solids4foam_amd.batch import Batch
    from solids4foam_amd.serve import Server
    print("[submesh guards]", flush=True)
    oc, m, pm, om = _inclusion(ctx, oracle_mod, n=5)
    b = Batch(ctx)
    with pytest.raises(RuntimeError, match="zones"):
        b.add(m)
    b.close()
    S = Server(ctx, 1 << 28, max_slots=4)
    bc = np.zeros((om.nBoundaryFaces, 3))
    with pytest.raises(RuntimeError, match="zones"):
        S.submit(m, bc)
    S.close()
    # zones cleared: the one-material solve, bit for bit
    m.set_zones(None)
    a = m.solve(precision=0, nCorrectors=30)
    oc2, m2, _, _ = _inclusion(ctx, oracle_mod, n=5, zones=False)
    c = m2.solve(precision=0, nCorrectors=30)
    assert np.array_equal(a["D"], c["D"]) and np.array_equal(a["sigma"], c["sigma"])


def _pair(ctx, oracle_mod, n, L, jitter, seed, z_of, E_of, nu, types, bv_of, tets=False, full=False):
    """oracle case and device mesh of the same box, zones and loads"""
    from solids4foam_amd.solid import PolyMesh, SolverMesh
    om = oracle_mod.Mesh.box(*n, L, jitter, seed, tets=tets)
    z = z_of(om.C).astype(np.int32)
    E = E_of(z)
    bv = bv_of(om)
    oc = oracle_mod.Case(om, E, nu, types, bv, ls_full_delta=full)
    assert oc.set_zones(z) > 0
    oc.init()
    pm = PolyMesh.box(*n, L=L, jitter=jitter, seed=seed, tets=tets)
    assert np.array_equal(pm.cell_centres(), om.C)
    m = SolverMesh(ctx, pm, types, ls_full_delta=full)
    m.set_material(E, nu)
    for p in range(len(types)):
        if types[p] != 2:
            m.set_bc(p, bv[pm.patch_faces(p)])
    m.set_zones(z)
    return oc, m


def test_submesh_tets_and_2d_match_oracle(ctx, oracle_mod):
    # tetrahedra (full-delta gradient) with a soft inclusion, and a 2-D
    # two-layer cantilever (empty faces: the sub-mesh LS inverses keep the
    # empty axis out)
    def shear(om):
        bv = np.zeros((om.nBoundaryFaces, 3))
        bv[om.patch_faces(1)] = (0.0, 5e6, 0.0)
        return bv

    def tip(om):
        bv = np.zeros((om.nBoundaryFaces, 3))
        bv[om.patch_faces(1)] = (0.0, -1e6, 0.0)
        return bv

    cases = [("tets", _pair(ctx, oracle_mod, (4, 4, 4), (1.0, 1.0, 1.0), 0.1, 5,
                            lambda C: np.linalg.norm(C - 0.5, axis=1) < 0.3, lambda z: np.where(z == 0, 200e9, 20e9),
                            0.3, [FIXED] + [TRACTION] * 5, shear, tets=True, full=True)),
             ("2-D", _pair(ctx, oracle_mod, (20, 6, 1), (1.0, 0.3, 0.05), 0.1, 3, lambda C: C[:, 1] > 0.15,
                           lambda z: np.where(z == 0, 200e9, 70e9), 0.3, [FIXED, TRACTION, TRACTION, TRACTION, 2, 2],
                           tip))]
    for name, (oc, m) in cases:
        print(f"[submesh {name}]", flush=True)
        ro = oc.solve(precond=oracle_mod.PRECOND_DIC, **ORACLE_TIGHT)
        out = m.solve(precision=0, **TIGHT)
        eD = rel_l2(out["D"], oc.get("D").reshape(-1, 3))
        es = rel_l2(out["sigma"], oc.get("sigma").reshape(-1, 6))
        print(f"  nOuter {out['nOuter']} (oracle {ro['nOuter']}) err D {eD:.2e} sigma {es:.2e}", flush=True)
        assert ro["converged"] and out["converged"] and eD < 1e-8 and es < 1e-8


def _incl6(ctx, oracle_mod):
    def load(om):
        bv = np.zeros((om.nBoundaryFaces, 3))
        bv[om.patch_faces(5)] = (2e6, 0.0, 1e7)
        return bv
    return _pair(ctx, oracle_mod, (6, 6, 6), (1.0, 1.0, 1.0), 0.15, 3,
                 lambda C: np.linalg.norm(C - 0.5, axis=1) < 0.3, lambda z: np.where(z == 0, 200e9, 20e9),
                 0.3, [TRACTION] * 4 + [FIXED, TRACTION], load)


@pytest.mark.parametrize("method", ["aitken", "fixed"])
def test_submesh_relaxation_follows_oracle(ctx, oracle_mod, method):
    # field relaxation with sub-meshes: the interface update sees the relaxed
    # field (Aitken's from its own buffer); the first outer iterations follow
    # the oracle and the converged field is the oracle's
    print(f"[submesh relaxation {method}]", flush=True)
    kw_o = dict(relaxAitken=1, relaxD=0.5) if method == "aitken" else dict(relaxD=0.7)
    kw_g = dict(relaxationMethod=1, relaxD=0.5) if method == "aitken" else dict(relaxD=0.7)
    oc, m = _incl6(ctx, oracle_mod)
    ro = oc.solve(history=10, nCorr=8, precond=oracle_mod.PRECOND_DIAGONAL, **kw_o)
    m.solve(nCorrectors=8, **kw_g)
    h = m.history()
    assert np.allclose(h["initRes"][:8], ro["history"]["initRes"][:8], rtol=1e-8, atol=1e-14)
    assert np.allclose(h["rvf"][:8], ro["history"]["rvf"][:8], rtol=1e-8)
    oc, m = _incl6(ctx, oracle_mod)
    ro = oc.solve(precond=oracle_mod.PRECOND_DIC, **ORACLE_TIGHT, **kw_o)
    out = m.solve(precision=0, **TIGHT, **kw_g)
    err = rel_l2(out["D"], oc.get("D").reshape(-1, 3))
    print(f"  nOuter {out['nOuter']} (oracle {ro['nOuter']}) err D {err:.2e}", flush=True)
    assert ro["converged"] and out["converged"] and err < 1e-8


def test_submesh_transient_steps_match_oracle(ctx, oracle_mod):
    # Euler d2dt2 with sub-meshes: the interface state carries over between
    # steps; every converged step matches the oracle's
    print("[submesh transient]", flush=True)

    def load(om):
        bv = np.zeros((om.nBoundaryFaces, 3))
        bv[om.patch_faces(1)] = (0.0, -2e6, 1e6)
        return bv
    oc, m = _pair(ctx, oracle_mod, (8, 4, 4), (1.0, 0.25, 0.25), 0.1, 5, lambda C: C[:, 0] > 0.5,
                  lambda z: np.where(z == 0, 200e9, 70e9), 0.3, [FIXED] + [TRACTION] * 5, load)
    oc.set("rho", np.full(oc.mesh.nCells, 7800.0))
    m.set_density(7800.0)
    kw = dict(nCorr=100000, solTol=1e-12, altTol=1e-12, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC)
    C = oc.mesh.nCells
    Dm, Dmm = np.zeros((C, 3)), np.zeros((C, 3))
    for i in range(3):
        if i:
            oc.advance_time()
        ro = oc.solve(d2dt2=1, deltaT=2e-5, deltaT0=2e-5, **kw)
        m.set_old_time(Dm, Dmm)
        out = m.solve(d2dt2=1, deltaT=2e-5, deltaT0=2e-5, nCorrectors=100000, solutionTolerance=1e-12,
                      alternativeTolerance=1e-12, pcgTolerance=1e-15)
        err = rel_l2(out["D"], oc.get("D").reshape(-1, 3))
        print(f"  step {i}: nOuter {out['nOuter']} (oracle {ro['nOuter']}) err {err:.2e}", flush=True)
        assert ro["converged"] and out["converged"] and err < 1e-8
        Dmm, Dm = Dm, out["D"]

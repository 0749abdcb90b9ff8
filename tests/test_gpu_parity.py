"""Parity of the HIP path (called through the C ABI) with the FP64 CPU oracle.

Operator level: each kernel against the oracle on the same state (FP64 to
1e-12 relative, FP32 SpMV to 1e-6).  Solve level: converged displacement and
stress against the committed oracle fixtures (tests/golden), plain FP64 to
1e-8 relative L2, mixed precision to 1e-6.  Plus determinism, batching (alone
== in a batch), and the divergence / max-iteration error paths.
"""
import os

import numpy as np
import pytest

from tests._cases import E_STEEL, NU_STEEL, FIXED, TRACTION, cantilever_bc, cube_tension_bc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TIGHT = dict(nCorrectors=200000, solutionTolerance=1e-12, alternativeTolerance=1e-12, pcgTolerance=1e-13)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ctx():
    from solids4foam_amd import Context
    c = Context(0)
    yield c
    c.close()


def _product_mesh(ctx, n, L=(1, 1, 1), jitter=0.0, seed=0, tets=False, bc="cube", ls_full_delta=False):
    from solids4foam_amd import PolyMesh, SolverMesh
    pm = PolyMesh.box(*n, L=L, jitter=jitter, seed=seed, tets=tets)
    sl = [pm.patch_faces(p) for p in range(6)]
    pt, bv = (cube_tension_bc if bc == "cube" else cantilever_bc)(pm.nBoundaryFaces, sl)
    m = SolverMesh(ctx, pm, pt, ls_full_delta=ls_full_delta)
    m.set_material(E_STEEL, NU_STEEL)
    for p in range(6):
        m.set_bc(p, bv[sl[p]])
    return pm, m, pt, bv


def _oracle_case(oracle_mod, n, L=(1, 1, 1), jitter=0.0, seed=0, tets=False, bc="cube", ls_full_delta=False):
    om = oracle_mod.Mesh.box(*n, L=L, jitter=jitter, seed=seed, tets=tets)
    sl = [om.patch_faces(p) for p in range(6)]
    pt, bv = (cube_tension_bc if bc == "cube" else cantilever_bc)(om.nBoundaryFaces, sl)
    return om, oracle_mod.Case(om, E_STEEL, NU_STEEL, pt, bv, ls_full_delta=ls_full_delta)


@pytest.mark.parametrize("tets", [False, True])
def test_operators_match_oracle(ctx, oracle_mod, tets):
    n = (5, 4, 4) if tets else (9, 7, 6)
    print(f"[operators] tets={tets} n={n}", flush=True)
    om, oc = _oracle_case(oracle_mod, n, jitter=0.15, seed=3, tets=tets, ls_full_delta=tets)
    pm, m, _, _ = _product_mesh(ctx, n, jitter=0.15, seed=3, tets=tets, ls_full_delta=tets)
    for _ in range(4):
        oc.iterate(precond=oracle_mod.PRECOND_DIAGONAL)
    C, B = om.nCells, om.nBoundaryFaces
    D, Db, snb = oc.get("D").reshape(C, 3), oc.get("Db").reshape(B, 3), oc.get("snb").reshape(B, 3)
    # gradient with the boundary correction
    G, Gb = m.grad(D, Db, snb)
    oc.op("grad")
    Go, Gbo = oc.get("G").reshape(C, 9), oc.get("Gb").reshape(B, 9)
    assert rel_l2(G, Go) < 1e-12
    assert rel_l2(Gb, Gbo) < 1e-12
    # Hooke
    oc.op("sigma")
    assert rel_l2(m.sigma(Go), oc.get("sigma").reshape(C, 6)) < 1e-13
    # residual: div(sigma) + Rhie-Chow
    R = m.residual(D, Go, oc.get("sigmab").reshape(B, 6))
    oc.op("residual")
    Ro = oc.get("R").reshape(C, 3)
    assert np.abs(R - Ro).max() < 1e-10 * np.abs(Ro).max()
    # matrix and 3-RHS SpMV
    diag, a = m.matrix()
    assert rel_l2(diag, oc.get("diag")) < 1e-14
    assert rel_l2(-a, oc.get("upper")) < 1e-14
    x = np.random.default_rng(0).normal(size=(C, 3))
    y = m.spmv3(x)
    yo = np.stack([oc.amul(np.ascontiguousarray(x[:, q])) for q in range(3)], axis=1)
    assert rel_l2(y, yo) < 1e-13
    y32 = m.spmv3(x, mixed=True)
    assert rel_l2(y32, yo) < 1e-5
    # PCG to a tight tolerance == dense solve
    M = np.diag(oc.get("diag"))
    o, nb = om.owner[:om.nInternalFaces], om.neighbour
    M[o, nb] = oc.get("upper")
    M[nb, o] = oc.get("upper")
    xs, its = m.pcg3(Ro, np.abs(Ro).sum(axis=0), pcgTolerance=1e-14, pcgRelTol=0.0)
    exact = np.linalg.solve(M, Ro)
    assert rel_l2(xs, exact) < 1e-9
    assert np.all(its > 0)


@pytest.mark.parametrize("name,n,kw", [
    ("cube16_jitter", (16, 16, 16), dict(jitter=0.15, seed=2)),
    ("tet4_jitter", (4, 4, 4), dict(jitter=0.15, seed=5, tets=True, ls_full_delta=True)),
    ("cfg1_cantilever", (50, 10, 10), dict(L=(1.0, 0.2, 0.2), bc="cantilever")),
])
def test_fp64_solve_matches_oracle_fixture(ctx, name, n, kw):
    print(f"[fp64 solve] {name}", flush=True)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    pm, m, _, _ = _product_mesh(ctx, n, **kw)
    out = m.solve(**TIGHT)
    print(f"  nOuter={out['nOuter']} (oracle {int(g['nOuter'])}) converged={out['converged']} "
          f"{out['seconds']:.2f}s err_D={rel_l2(out['D'], g['D']):.2e} "
          f"err_sigma={rel_l2(out['sigma'], g['sigma']):.2e}", flush=True)
    assert out["converged"]
    assert rel_l2(out["D"], g["D"]) < 1e-8
    assert rel_l2(out["sigma"], g["sigma"]) < 1e-8


@pytest.mark.parametrize("name,n,kw", [
    ("cube16_jitter", (16, 16, 16), dict(jitter=0.15, seed=2)),
    ("cfg1_cantilever", (50, 10, 10), dict(L=(1.0, 0.2, 0.2), bc="cantilever")),
])
def test_mixed_precision_solve_error(ctx, name, n, kw):
    from solids4foam_amd import sigma_eq
    print(f"[mixed solve] {name}", flush=True)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    pm, m, _, _ = _product_mesh(ctx, n, **kw)
    out = m.solve(precision=1, **TIGHT)
    eD, eS = rel_l2(out["D"], g["D"]), rel_l2(sigma_eq(out["sigma"]), sigma_eq(g["sigma"]))
    print(f"  nOuter={out['nOuter']} err_D={eD:.2e} err_vonMises={eS:.2e}", flush=True)
    assert out["converged"]
    assert eD < 1e-6 and eS < 1e-6


@pytest.mark.parametrize("smoothed", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name,n,kw", [
    ("cube16_jitter", (16, 16, 16), dict(jitter=0.15, seed=2)),
    ("tet4_jitter", (4, 4, 4), dict(jitter=0.15, seed=5, tets=True, ls_full_delta=True)),
    ("cfg1_cantilever", (50, 10, 10), dict(L=(1.0, 0.2, 0.2), bc="cantilever")),
])
def test_multigrid_preconditioned_solve_matches_fixture(ctx, name, n, kw, precision, smoothed):
    # the preconditioner changes the path, not the fixed point: same converged
    # field as the oracle (which uses DIC), in fewer PCG iterations; plain
    # aggregation and the smoothed first prolongation alike
    from solids4foam_amd import PRECOND_GAMG, PRECOND_GAMG_SMOOTHED, sigma_eq
    print(f"[gamg solve] {name} precision={precision} smoothed={smoothed}", flush=True)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    pm, m, _, _ = _product_mesh(ctx, n, **kw)
    out = m.solve(precision=precision, preconditioner=PRECOND_GAMG_SMOOTHED if smoothed else PRECOND_GAMG, **TIGHT)
    jac = m.solve(precision=precision, **TIGHT)
    eD, eS = rel_l2(out["D"], g["D"]), rel_l2(sigma_eq(out["sigma"]), sigma_eq(g["sigma"]))
    print(f"  nOuter={out['nOuter']} pcg={out['pcgIterations']} (diagonal {jac['pcgIterations']}) "
          f"err_D={eD:.2e} err_vonMises={eS:.2e}", flush=True)
    assert out["converged"]
    tol = 1e-8 if precision == 0 else 1e-6
    assert eD < tol and eS < tol
    assert out["pcgIterations"] < jac["pcgIterations"]


@pytest.mark.parametrize("precision", [0, 1])
def test_plane_stress_matches_oracle(ctx, oracle_mod, precision):
    # 2-D cantilever (empty z faces) with mechanicalProperties planeStress:
    # the device path's converged field equals the oracle's, and differs
    # from the plane-strain answer (the switch reaches the moduli)
    from solids4foam_amd import PolyMesh, SolverMesh
    print(f"[plane stress] precision={precision}", flush=True)
    E, nu, t = 200e9, 0.3, 1e6
    n, L = (16, 4, 1), (1.0, 0.2, 0.05)
    pt = [FIXED, TRACTION, TRACTION, TRACTION, oracle_mod.EMPTY, oracle_mod.EMPTY]
    om = oracle_mod.Mesh.box(*n, L=L)
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(1)] = (0.0, t, 0.0)
    oc = oracle_mod.Case(om, E, nu, pt, bv, plane_stress=True)
    r = oc.solve(nCorr=200000, solTol=1e-12, altTol=1e-12, pcgTol=1e-14, precond=oracle_mod.PRECOND_DIC)
    assert r["converged"]
    Do = oc.get("D").reshape(-1, 3)
    pm = PolyMesh.box(*n, L=L)
    sl = [pm.patch_faces(p) for p in range(6)]
    fields = {}
    for ps in (True, False):
        m = SolverMesh(ctx, pm, pt)
        m.set_plane_stress(ps)
        m.set_material(E, nu)
        for p in range(4):
            m.set_bc(p, bv[sl[p]])
        out = m.solve(precision=precision, **TIGHT)
        assert out["converged"]
        fields[ps] = out["D"]
    tol = 1e-8 if precision == 0 else 1e-6
    err = rel_l2(fields[True], Do)
    print(f"  err vs oracle {err:.2e}, plane strain differs by {rel_l2(fields[False], Do):.2e}", flush=True)
    assert err < tol
    assert rel_l2(fields[False], Do) > 1e-3


def test_multigrid_hierarchy_info(ctx):
    # level 0 is the mesh (two stored entries per internal face), every level
    # coarsens, the coarsest fits one dense block; q = offdiag(A) P groups a
    # row's neighbour entries per coarse row, so it never stores more
    from solids4foam_amd import PRECOND_GAMG, S4FError
    pm, m, _, _ = _product_mesh(ctx, (40, 40, 40), jitter=0.15, seed=2)
    with pytest.raises(S4FError):
        m.mg_info()
    out = m.solve(preconditioner=PRECOND_GAMG, nCorrectors=2)
    info, mi = m.mg_info(), m.info()
    print(f"[mg info] {info}", flush=True)
    rows, ents = info["rows"], info["entries"]
    assert len(rows) >= 3 and rows[0] == mi["nCells"] and ents[0] == 2 * mi["nInternalFaces"]
    assert all(a > b for a, b in zip(rows, rows[1:])) and rows[-1] <= 1024
    assert all(e > 0 for e in ents)
    assert mi["nCells"] < info["qEntries"] < ents[0]
    assert out["nOuter"] >= 1


def test_smoothed_multigrid_fewer_iterations(ctx):
    # the smoothed first prolongation (P = (I - 2/3 D^-1 A) P_agg) widens the
    # first coarse operator and its up-sweep operator, keeps the coarse rows,
    # and needs fewer PCG iterations per outer iteration than plain aggregation
    # on the same outer path (same outer count, same converged field)
    from solids4foam_amd import PRECOND_GAMG, PRECOND_GAMG_SMOOTHED
    pm, m, _, _ = _product_mesh(ctx, (40, 40, 40), jitter=0.15, seed=2)
    opts = dict(solutionTolerance=1e-8, alternativeTolerance=1e-8, nCorrectors=20000, pcgRelTol=0.9, precision=1)
    plain = m.solve(preconditioner=PRECOND_GAMG, **opts)
    ip = m.mg_info()
    sm = m.solve(preconditioner=PRECOND_GAMG_SMOOTHED, **opts)
    isd = m.mg_info()
    print(f"[smoothed mg] plain {plain['nOuter']} outer / {plain['pcgIterations']} PCG, levels {ip['rows']} "
          f"entries {ip['entries']} q {ip['qEntries']}; smoothed {sm['nOuter']} / {sm['pcgIterations']}, levels "
          f"{isd['rows']} entries {isd['entries']} G {isd['qEntries']}", flush=True)
    assert plain["converged"] and sm["converged"]
    assert isd["rows"][1] == ip["rows"][1] and isd["entries"][1] > ip["entries"][1]
    assert isd["qEntries"] > ip["qEntries"]
    assert sm["pcgIterations"] < 0.8 * plain["pcgIterations"]
    assert abs(sm["nOuter"] - plain["nOuter"]) <= 0.1 * plain["nOuter"]
    assert rel_l2(sm["D"], plain["D"]) < 1e-6


def test_multigrid_batch_equals_alone(ctx):
    # a batch is one block-diagonal matrix: each instance's hierarchy, cycle
    # and reductions are the ones it has alone (an instance already at its
    # coarsest size passes through the remaining levels untouched), with the
    # plain and the smoothed first prolongation
    from solids4foam_amd import PRECOND_GAMG, PRECOND_GAMG_SMOOTHED
    from solids4foam_amd.batch import Batch
    print("[gamg batch]", flush=True)
    specs = [((9, 7, 6), dict(jitter=0.1, seed=1)), ((30, 6, 6), dict(L=(1.0, 0.2, 0.2), bc="cantilever")),
             ((5, 5, 5), dict(jitter=0.1, seed=3)), ((20, 11, 13), dict(jitter=0.1, seed=4)),
             ((24, 24, 12), dict(jitter=0.1, seed=5))]
    meshes = [_product_mesh(ctx, n, **kw) for n, kw in specs]
    for precision, pc in ((0, PRECOND_GAMG), (1, PRECOND_GAMG), (1, PRECOND_GAMG_SMOOTHED)):
        opts = dict(solutionTolerance=1e-10, alternativeTolerance=1e-10, nCorrectors=50000,
                    preconditioner=pc, precision=precision)
        alone = [m.solve(**opts) for _, m, _, _ in meshes]
        b = Batch(ctx)
        for _, m, _, _ in meshes:
            b.add(m)
        res = b.solve(**opts)
        for a, r in zip(alone, res):
            assert r["converged"] and a["converged"]
            assert r["nOuter"] == a["nOuter"] and r["innerTotal"] == a["innerTotal"]
            assert np.array_equal(r["D"], a["D"]) and np.array_equal(r["sigma"], a["sigma"])
        b.close()


def _nu_box(ctx, oracle_mod, n, nu):
    from solids4foam_amd import PolyMesh, SolverMesh
    om = oracle_mod.Mesh.box(n, n, n, (0.1, 0.1, 0.1), 0.1, 3)
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(1)] = (1e6, 5e5, 0.0)
    oc = oracle_mod.Case(om, 100e9, nu, [FIXED] + [TRACTION] * 5, bv)
    pm = PolyMesh.box(n, n, n, L=(0.1, 0.1, 0.1), jitter=0.1, seed=3)
    m = SolverMesh(ctx, pm, [FIXED] + [TRACTION] * 5)
    m.set_material(100e9, nu)
    for p in range(6):
        m.set_bc(p, bv[pm.patch_faces(p)])
    return oc, m


def test_first_iterations_follow_oracle(ctx, oracle_mod):
    # with the oracle on the same (diagonal) preconditioner the device loop
    # takes the oracle's path: per outer iteration the same |initRes|,
    # residualvf and PCG iteration counts, for a loaded cube, a moved box (its
    # first residual sees the zero initial stress) and a tetrahedral mesh
    print("[history]", flush=True)
    cases = [("cube", _oracle_case(oracle_mod, (7, 6, 5), jitter=0.15, seed=3)[1],
              _product_mesh(ctx, (7, 6, 5), jitter=0.15, seed=3)[1]),
             ("tets", _oracle_case(oracle_mod, (3, 3, 3), jitter=0.1, seed=5, tets=True, ls_full_delta=True)[1],
              _product_mesh(ctx, (3, 3, 3), jitter=0.1, seed=5, tets=True, ls_full_delta=True)[1])]
    oc, m = _moved_box(ctx, oracle_mod, 5, 0.3)
    cases.append(("moved", oc, m))
    for name, oc, m in cases:
        ro = oc.solve(history=12, nCorr=12, precond=oracle_mod.PRECOND_DIAGONAL)
        m.solve(nCorrectors=12)
        h = m.history()
        n = min(len(h["initRes"]), len(ro["history"]["initRes"]))
        assert n >= 10, name
        assert np.allclose(h["initRes"][:n], ro["history"]["initRes"][:n], rtol=1e-8, atol=1e-14), name
        assert np.allclose(h["rvf"][:n], ro["history"]["rvf"][:n], rtol=1e-8), name
        assert np.array_equal(h["inner"][:n], ro["history"]["inner"][:n]), name


def test_field_relaxation_follows_oracle(ctx, oracle_mod):
    # fixed under-relaxation of D (relaxField): the first outer iterations
    # reproduce the oracle's history, and at nu = 0.45 (where the unrelaxed
    # loop oscillates) the relaxed solve converges to the oracle's field
    oc, m = _nu_box(ctx, oracle_mod, 6, 0.45)
    ro = oc.solve(history=10, nCorr=6, precond=oracle_mod.PRECOND_DIAGONAL, relaxD=0.7)
    out = m.solve(nCorrectors=6, relaxD=0.7)
    h = m.history()
    assert np.allclose(h["initRes"][:6], ro["history"]["initRes"][:6], rtol=1e-8, atol=1e-14)
    assert np.allclose(h["rvf"][:6], ro["history"]["rvf"][:6], rtol=1e-8)
    assert rel_l2(out["D"], oc.get("D").reshape(-1, 3)) < 1e-9
    # converge both sides tightly (a PCG tolerance of 1e-9 stops the outer loop
    # where the normalised residual drops below it, ~1e-8 apart between paths)
    oc, m = _nu_box(ctx, oracle_mod, 6, 0.45)
    kw = dict(solTol=1e-14, altTol=1e-14, nCorr=100000, pcgTol=1e-15)
    ro = oc.solve(precond=oracle_mod.PRECOND_DIC, relaxD=0.7, **kw)
    Do = oc.get("D").reshape(-1, 3)
    for precision in (0, 1):
        out = m.solve(nCorrectors=100000, solutionTolerance=1e-14, alternativeTolerance=1e-14, relaxD=0.7,
                      precision=precision, pcgTolerance=1e-15)
        err = rel_l2(out["D"], Do)
        print(f"  relaxed nu=0.45 precision={precision}: nOuter={out['nOuter']} (oracle {ro['nOuter']}) "
              f"err={err:.2e}", flush=True)
        assert ro["converged"] and out["converged"]
        assert err < (1e-8 if precision == 0 else 1e-6)


def _moved_box(ctx, oracle_mod, n, nu):
    # x-min fixed, x-max moved by a fixed displacement, other faces free
    from solids4foam_amd import PolyMesh, SolverMesh
    types = [FIXED, FIXED] + [TRACTION] * 4
    om = oracle_mod.Mesh.box(n, n, n, (0.1, 0.1, 0.1), 0.1, 3)
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(1)] = (1e-5, 5e-6, 0.0)
    oc = oracle_mod.Case(om, 100e9, nu, types, bv)
    pm = PolyMesh.box(n, n, n, L=(0.1, 0.1, 0.1), jitter=0.1, seed=3)
    m = SolverMesh(ctx, pm, types)
    m.set_material(100e9, nu)
    for p in range(6):
        m.set_bc(p, bv[pm.patch_faces(p)])
    return oc, m


def test_aitken_relaxation_follows_oracle(ctx, oracle_mod):
    # relaxationMethod Aitken (per cell and boundary face factors): the first
    # outer iterations reproduce the oracle's history for a displaced and a
    # loaded box, both converge to the oracle's field, and a batch of the two
    # gives each the bits it has alone
    from solids4foam_amd import _abi
    from solids4foam_amd.batch import Batch
    AIT = _abi.RELAX_AITKEN
    boxes = {"moved": _moved_box(ctx, oracle_mod, 6, 0.3), "loaded": _nu_box(ctx, oracle_mod, 6, 0.3)}
    for name, (oc, m) in boxes.items():
        ro = oc.solve(history=10, nCorr=8, precond=oracle_mod.PRECOND_DIAGONAL, relaxAitken=1)
        m.solve(nCorrectors=8, relaxationMethod=AIT)
        h = m.history()
        assert np.allclose(h["initRes"][:8], ro["history"]["initRes"][:8], rtol=1e-8, atol=1e-14), name
        assert np.allclose(h["rvf"][:8], ro["history"]["rvf"][:8], rtol=1e-8), name
    for name, (oc, m) in (("moved", _moved_box(ctx, oracle_mod, 6, 0.3)), ("loaded", _nu_box(ctx, oracle_mod, 6, 0.3))):
        ro = oc.solve(precond=oracle_mod.PRECOND_DIC, relaxAitken=1, solTol=1e-13, altTol=1e-13, nCorr=100000,
                      pcgTol=1e-15)
        Do = oc.get("D").reshape(-1, 3)
        for precision in (0, 1):
            out = m.solve(nCorrectors=100000, solutionTolerance=1e-13, alternativeTolerance=1e-13,
                          relaxationMethod=AIT, precision=precision, pcgTolerance=1e-15)
            err = rel_l2(out["D"], Do)
            print(f"  Aitken {name} precision={precision}: nOuter={out['nOuter']} (oracle {ro['nOuter']}) "
                  f"err={err:.2e}", flush=True)
            assert ro["converged"] and out["converged"]
            assert err < (1e-8 if precision == 0 else 1e-6)
    opts = dict(nCorrectors=100000, solutionTolerance=1e-10, alternativeTolerance=1e-10, relaxationMethod=AIT,
                precision=1)
    ms = [boxes["moved"][1], boxes["loaded"][1]]
    alone = [mm.solve(**opts) for mm in ms]
    b = Batch(ctx)
    for mm in ms:
        b.add(mm)
    res = b.solve(**opts)
    b.close()
    for a, r in zip(alone, res):
        assert r["converged"] and r["nOuter"] == a["nOuter"]
        assert np.array_equal(r["D"], a["D"]) and np.array_equal(r["sigma"], a["sigma"])


def _eps_eq(e):
    """sqrt(2/3 |dev e|^2) of symmetric tensors (n, 6: xx xy xz yy yz zz)"""
    tr = (e[:, 0] + e[:, 3] + e[:, 5]) / 3.0
    d2 = (e[:, 0] - tr) ** 2 + (e[:, 3] - tr) ** 2 + (e[:, 5] - tr) ** 2 + 2 * (e[:, 1] ** 2 + e[:, 2] ** 2 + e[:, 4] ** 2)
    return np.sqrt(2.0 / 3.0 * d2)


def test_post_loop_fields(ctx, oracle_mod):
    # pointD (enhancedVolPointInterpolation with gradient) against the oracle
    # on the device's own converged D and gradD; epsilon = symm(gradD),
    # epsilonEq and sigmaEq of solidModel::writeFields
    from solids4foam_amd import sigma_eq
    for tets in (False, True):
        n = (7, 6, 5) if not tets else (3, 3, 3)
        pm, m, pt, bv = _product_mesh(ctx, n, jitter=0.15, seed=3, tets=tets, ls_full_delta=tets)
        out = m.solve(want_grad=True, solutionTolerance=1e-10, alternativeTolerance=1e-10, nCorrectors=50000)
        om, oc = _oracle_case(oracle_mod, n, jitter=0.15, seed=3, tets=tets, ls_full_delta=tets)
        pD = m.point_displacement()
        pO = oc.point_interpolate(out["D"], out["gradD"])
        assert np.abs(pD - pO).max() <= 1e-12 * np.abs(pO).max()
        f = m.strain_fields()
        G = out["gradD"].reshape(-1, 3, 3)
        S = 0.5 * (G + G.transpose(0, 2, 1))
        eps = np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], axis=1)
        assert np.allclose(f["epsilon"], eps, rtol=1e-13, atol=1e-13 * np.abs(eps).max())
        assert np.allclose(f["epsilonEq"], _eps_eq(eps), rtol=1e-12)
        assert np.allclose(f["sigmaEq"], sigma_eq(out["sigma"]), rtol=1e-12)


def test_equivalent_strain_closed_form():
    # uniaxial strain a: epsilonEq = 2|a|/3; engineering shear g: g/sqrt(3)
    a, g = 3e-4, 2e-4
    e = np.array([[a, 0, 0, 0, 0, 0], [0, g / 2, 0, 0, 0, 0]])
    assert np.allclose(_eps_eq(e), [2 * a / 3, g / np.sqrt(3)], rtol=1e-14)


def test_reference_patch_test_2d(ctx, oracle_mod):
    # tutorials/solids/linearElasticity/patchTest (2-D: empty front/back,
    # linearSpatialDisplacement outer patch): the README's exact uniform strain
    # (exx 2e-6, eyy 6e-6, exy 4e-6) and the oracle's field, in both precisions
    from solids4foam_amd import PATCH_EMPTY, PolyMesh, SolverMesh
    from tests import _blockmesh as bm
    a = bm.hex_blocks(bm.PATCH_TEST_POINTS, bm.PATCH_TEST_HEXES, bm.PATCH_TEST_PATCHES)
    pm = PolyMesh.from_arrays(a["points"], a["fOff"], a["fVerts"], a["owner"], a["neighbour"], a["nCells"],
                              a["patchStart"], a["patchSize"])
    m = SolverMesh(ctx, pm, [FIXED, PATCH_EMPTY], ls_full_delta=True)
    m.set_material(200e9, 0.3)
    Cf = pm.geometry()["Cf"][pm.nInternalFaces:]
    nb = int(a["patchSize"][0])
    m.set_bc(0, bm.linear_spatial(Cf[:nb], bm.PATCH_TEST_A, bm.PATCH_TEST_B))
    om = oracle_mod.Mesh.from_arrays(a["points"], a["fOff"], a["fVerts"], a["owner"], a["neighbour"], a["nCells"],
                                     a["patchStart"], a["patchSize"])
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[:nb] = bm.linear_spatial(Cf[:nb], bm.PATCH_TEST_A, bm.PATCH_TEST_B)
    oc = oracle_mod.Case(om, 200e9, 0.3, [FIXED, 2], bv, ls_full_delta=True)
    oc.solve(nCorr=1000, solTol=1e-14, altTol=1e-14, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC)
    for precision in (0, 1):
        out = m.solve(precision=precision, want_grad=True, nCorrectors=1000, solutionTolerance=1e-14,
                      alternativeTolerance=1e-14, pcgTolerance=1e-15)
        assert out["converged"]
        G = out["gradD"].reshape(-1, 3, 3)
        assert np.allclose(G[:, 0, 0], 2e-6, rtol=1e-8) and np.allclose(G[:, 1, 1], 6e-6, rtol=1e-8)
        assert np.allclose(0.5 * (G[:, 0, 1] + G[:, 1, 0]), 4e-6, rtol=1e-8)
        assert np.abs(G[:, 2, :]).max() < 1e-20 and np.abs(out["D"][:, 2]).max() == 0.0
        assert rel_l2(out["D"], oc.get("D").reshape(-1, 3)) < 1e-10


def test_symmetry_planes(ctx, oracle_mod):
    # solidSymmetry on the three min faces, normal traction on z-max: the
    # orthogonal mesh reproduces the exact uniaxial-stress field; on a jittered
    # mesh the device field matches the oracle's
    from solids4foam_amd import PATCH_SYMMETRY, PolyMesh, SolverMesh
    E, nu, t = 200e9, 0.3, 1e7
    types = [PATCH_SYMMETRY, TRACTION, PATCH_SYMMETRY, TRACTION, PATCH_SYMMETRY, TRACTION]
    for jitter in (0.0, 0.15):
        pm = PolyMesh.box(6, 5, 4, L=(1.0, 1.0, 1.0), jitter=jitter, seed=3)
        m = SolverMesh(ctx, pm, types)
        m.set_material(E, nu)
        bv = np.zeros((pm.nBoundaryFaces, 3))
        bv[pm.patch_faces(5)] = (0.0, 0.0, t)
        for p in (1, 3, 5):
            m.set_bc(p, bv[pm.patch_faces(p)])
        om = oracle_mod.Mesh.box(6, 5, 4, (1.0, 1.0, 1.0), jitter, 3)
        oc = oracle_mod.Case(om, E, nu, [3, TRACTION, 3, TRACTION, 3, TRACTION], bv)
        oc.solve(nCorr=50000, solTol=1e-13, altTol=1e-13, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC)
        Do = oc.get("D").reshape(-1, 3)
        for precision in (0, 1):
            out = m.solve(precision=precision, nCorrectors=50000, solutionTolerance=1e-13, alternativeTolerance=1e-13,
                          pcgTolerance=1e-15)
            assert out["converged"]
            err = rel_l2(out["D"], Do)
            print(f"  symmetry jitter={jitter} precision={precision}: nOuter={out['nOuter']} err={err:.2e}", flush=True)
            assert err < (1e-8 if precision == 0 else 1e-6)
            if jitter == 0.0:
                X = pm.cell_centres()
                exact = np.stack([-nu * t * X[:, 0] / E, -nu * t * X[:, 1] / E, t * X[:, 2] / E], axis=1)
                assert np.abs(out["D"] - exact).max() < 1e-7 * np.abs(exact).max()


def test_zero_shear_patches(ctx, oracle_mod):
    # fixedDisplacementZeroShear rollers on the three min faces.  (i) z-max
    # moved by delta through a zero-shear patch (tangential part of its vector
    # ignored): the exact uniaxial field on a skewed mesh; (ii) an oblique
    # traction on z-max (shear reaches the rollers): the device field matches
    # the oracle's, with and without field relaxation
    from solids4foam_amd import PATCH_ZERO_SHEAR, PolyMesh, SolverMesh
    E, nu, delta = 200e9, 0.3, 1e-4
    ZS = PATCH_ZERO_SHEAR
    for case in ("displacement", "traction"):
        top = ZS if case == "displacement" else TRACTION
        types = [ZS, TRACTION, ZS, TRACTION, ZS, top]
        pm = PolyMesh.box(6, 5, 4, L=(1.0, 1.0, 1.0), jitter=0.15, seed=3)
        m = SolverMesh(ctx, pm, types, ls_full_delta=True)
        m.set_material(E, nu)
        bv = np.zeros((pm.nBoundaryFaces, 3))
        bv[pm.patch_faces(5)] = (0.3 * delta, -0.2 * delta, delta) if case == "displacement" else (2e6, -1e6, 1e7)
        for p in range(6):
            m.set_bc(p, bv[pm.patch_faces(p)])
        om = oracle_mod.Mesh.box(6, 5, 4, (1.0, 1.0, 1.0), 0.15, 3)
        oc = oracle_mod.Case(om, E, nu, types, bv, ls_full_delta=True)
        oc.solve(nCorr=50000, solTol=1e-13, altTol=1e-13, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC)
        Do = oc.get("D").reshape(-1, 3)
        for precision, relax in ((0, 0.0), (1, 0.0), (0, 0.6)):
            out = m.solve(precision=precision, nCorrectors=50000, solutionTolerance=1e-13, alternativeTolerance=1e-13,
                          pcgTolerance=1e-15, relaxD=relax)
            assert out["converged"]
            err = rel_l2(out["D"], Do)
            print(f"  zero shear {case} precision={precision} relax={relax}: nOuter={out['nOuter']} err={err:.2e}",
                  flush=True)
            assert err < (1e-8 if precision == 0 else 1e-6)
            if case == "displacement" and precision == 0:
                X = pm.cell_centres()
                exact = np.stack([-nu * delta * X[:, 0], -nu * delta * X[:, 1], delta * X[:, 2]], axis=1)
                assert np.abs(out["D"] - exact).max() < 1e-8 * delta


def test_transient_and_body_force(ctx, oracle_mod):
    # rho*d2dt2(D) (Euler) and rho*g: (i) a free body falls with
    # D^n = n (n + 1)/2 g dt^2 exactly; (ii) a loaded cantilever under gravity,
    # three time steps (the last with a longer step), matches the oracle step
    # by step; (iii) the steady self-weight field matches the oracle
    from solids4foam_amd import PolyMesh, SolverMesh
    pm = PolyMesh.box(5, 4, 3, L=(1.0, 0.8, 0.6), jitter=0.1, seed=3)
    g, dt = np.array([0.0, -9.81, 2.0]), 1e-3
    # FP64 only: a free body's rigid mode is held by the inertia diagonal
    # alone (3e-4 of the elastic diagonal here), which FP32 sweeps resolve too
    # coarsely for the 1e-12 outer tolerance
    for precision in (0,):
        m = SolverMesh(ctx, pm, [TRACTION] * 6)
        m.set_material(200e9, 0.3)
        m.set_density(7800.0)
        m.set_gravity(g)
        for p in range(6):
            m.set_bc(p, np.zeros((pm.patch_faces(p).stop - pm.patch_faces(p).start, 3)))
        Dm, Dmm = np.zeros((pm.nCells, 3)), np.zeros((pm.nCells, 3))
        for n in range(1, 5):
            m.set_old_time(Dm, Dmm)
            out = m.solve(precision=precision, d2dt2=1, deltaT=dt, deltaT0=dt, nCorrectors=200,
                          solutionTolerance=1e-12, alternativeTolerance=1e-12, pcgTolerance=1e-14, pcgRelTol=0.0)
            assert out["converged"]
            exact = n * (n + 1) / 2 * g * dt * dt
            err = np.abs(out["D"] - exact).max() / np.abs(exact).max()
            assert err < 1e-11, (n, err)
            Dmm, Dm = Dm, out["D"]
    # (ii) cantilever: x-min fixed, x-max sudden traction, gravity
    n3 = (8, 4, 4)
    types = [FIXED] + [TRACTION] * 5
    om = oracle_mod.Mesh.box(*n3, (1.0, 0.25, 0.25), 0.1, 5)
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(1)] = (0.0, -2e6, 1e6)
    oc = oracle_mod.Case(om, 200e9, 0.3, types, bv)
    oc.set("rho", np.full(om.nCells, 7800.0))
    pm = PolyMesh.box(*n3, L=(1.0, 0.25, 0.25), jitter=0.1, seed=5)
    m = SolverMesh(ctx, pm, types)
    m.set_material(200e9, 0.3)
    m.set_density(7800.0)
    grav = (0.0, 0.0, -9.81)
    m.set_gravity(grav)
    for p in range(6):
        m.set_bc(p, bv[pm.patch_faces(p)])
    kw = dict(nCorr=100000, solTol=1e-12, altTol=1e-12, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC)
    Dm, Dmm = np.zeros((pm.nCells, 3)), np.zeros((pm.nCells, 3))
    steps = [(2e-5, 2e-5), (2e-5, 2e-5), (3e-5, 2e-5)]
    for i, (dti, dt0) in enumerate(steps):
        if i:
            oc.advance_time()
        ro = oc.solve(d2dt2=1, deltaT=dti, deltaT0=dt0, g=grav, **kw)
        m.set_old_time(Dm, Dmm)
        out = m.solve(d2dt2=1, deltaT=dti, deltaT0=dt0, nCorrectors=100000, solutionTolerance=1e-12,
                      alternativeTolerance=1e-12, pcgTolerance=1e-15)
        Do = oc.get("D").reshape(-1, 3)
        err = rel_l2(out["D"], Do)
        print(f"  transient step {i}: nOuter={out['nOuter']} (oracle {ro['nOuter']}) err={err:.2e}", flush=True)
        assert ro["converged"] and out["converged"] and err < 1e-8
        Dmm, Dm = Dm, out["D"]
    # the linear predictor (D = D_old + U dt) starts each step elsewhere and
    # reaches the same fields
    from solids4foam_amd import LinGeomTotalDispSolid
    patches = [dict(type="fixedDisplacement")] + [dict(type="solidTraction", traction=bv[pm.patch_faces(p)])
                                                   for p in range(1, 6)]
    runs = {}
    for pred in (False, True):
        sm = LinGeomTotalDispSolid(ctx, pm, patches, 200e9, 0.3, rho=7800.0, g=grav, d2dt2="Euler", deltaT=2e-5,
                                   predictor=pred, solver=dict(nCorrectors=100000, solutionTolerance=1e-12,
                                                               alternativeTolerance=1e-12, pcgTolerance=1e-15))
        outer = []
        for _ in range(3):
            sm.evolve()
            outer.append(sm.report["nOuter"])
        runs[pred] = (sm.D.copy(), outer)
    print(f"  predictor outer iterations {runs[True][1]} (without {runs[False][1]})", flush=True)
    assert rel_l2(runs[True][0], runs[False][0]) < 1e-9
    # (iii) steady self weight
    oc = oracle_mod.Case(om, 200e9, 0.3, types, np.zeros_like(bv))
    oc.set("rho", np.full(om.nCells, 7800.0))
    oc.solve(g=grav, **kw)
    for p in range(6):
        m.set_bc(p, np.zeros((pm.patch_faces(p).stop - pm.patch_faces(p).start, 3)))
    for precision in (0, 1):
        out = m.solve(precision=precision, nCorrectors=100000, solutionTolerance=1e-12, alternativeTolerance=1e-12,
                      pcgTolerance=1e-15)
        err = rel_l2(out["D"], oc.get("D").reshape(-1, 3))
        print(f"  self weight precision={precision}: nOuter={out['nOuter']} err={err:.2e}", flush=True)
        assert out["converged"] and err < (1e-8 if precision == 0 else 1e-6)


def test_solve_is_bitwise_deterministic(ctx):
    pm, m, _, _ = _product_mesh(ctx, (12, 10, 9), jitter=0.15, seed=8)
    a = m.solve(solutionTolerance=1e-9, alternativeTolerance=1e-9)
    b = m.solve(solutionTolerance=1e-9, alternativeTolerance=1e-9)
    assert np.array_equal(a["D"], b["D"]) and np.array_equal(a["sigma"], b["sigma"])
    assert a["nOuter"] == b["nOuter"]


def test_graph_replay_equals_eager(ctx):
    # the per-iteration launch sequences replayed from HIP graphs give the
    # same bits as launching every kernel individually, for one mesh (each
    # preconditioner and precision) and for a batch whose launch narrows
    # (and is captured again) as its instances finish
    from solids4foam_amd import PRECOND_GAMG
    from solids4foam_amd.batch import Batch
    print("[graphs]", flush=True)
    _, m, _, _ = _product_mesh(ctx, (14, 10, 9), jitter=0.15, seed=8)
    for pc in (0, PRECOND_GAMG):
        for precision in (0, 1):
            opts = dict(solutionTolerance=1e-9, alternativeTolerance=1e-9, preconditioner=pc, precision=precision)
            e = m.solve(eagerLaunch=1, **opts)
            g1 = m.solve(**opts)
            g2 = m.solve(**opts)
            for g in (g1, g2):
                assert g["converged"] and g["nOuter"] == e["nOuter"] and g["innerTotal"] == e["innerTotal"]
                assert np.array_equal(g["D"], e["D"]) and np.array_equal(g["sigma"], e["sigma"])
    specs = [((9, 7, 6), dict(jitter=0.1, seed=1)), ((16, 4, 4), dict(L=(1.0, 0.25, 0.25), bc="cantilever")),
             ((5, 5, 5), dict(jitter=0.1, seed=3)), ((20, 11, 3), dict(jitter=0.1, seed=4))]
    meshes = [_product_mesh(ctx, n, **kw) for n, kw in specs]
    b = Batch(ctx)
    for _, bm, _, _ in meshes:
        b.add(bm)
    opts = dict(solutionTolerance=1e-10, alternativeTolerance=1e-10, nCorrectors=20000, preconditioner=PRECOND_GAMG,
                precision=1)
    e = b.solve(eagerLaunch=1, **opts)
    g = b.solve(**opts)
    assert len(set(r["nOuter"] for r in e)) > 1  # instances finish at different outer iterations
    for a, r in zip(e, g):
        assert r["converged"] and r["nOuter"] == a["nOuter"] and r["innerTotal"] == a["innerTotal"]
        assert np.array_equal(r["D"], a["D"]) and np.array_equal(r["sigma"], a["sigma"])
    b.close()


def test_batch_equals_alone(ctx):
    from solids4foam_amd.batch import Batch
    print("[batch]", flush=True)
    specs = [((9, 7, 6), dict(jitter=0.1, seed=1)), ((16, 4, 4), dict(L=(1.0, 0.25, 0.25), bc="cantilever")),
             ((5, 5, 5), dict(jitter=0.1, seed=3)), ((20, 11, 3), dict(jitter=0.1, seed=4))]
    opts = dict(solutionTolerance=1e-10, alternativeTolerance=1e-10, nCorrectors=20000)
    meshes = [_product_mesh(ctx, n, **kw) for n, kw in specs]
    alone = [m.solve(**opts) for _, m, _, _ in meshes]
    b = Batch(ctx)
    for _, m, _, _ in meshes:
        b.add(m)
    res = b.solve(**opts)
    for a, r in zip(alone, res):
        # an instance's tiles and reductions do not depend on the rest of the batch
        assert r["converged"]
        assert r["nOuter"] == a["nOuter"]
        assert np.array_equal(r["D"], a["D"]) and np.array_equal(r["sigma"], a["sigma"])


def test_divergence_and_maxiter_status(ctx):
    from solids4foam_amd import _abi
    # tets with the normal-delta leastSquares gradient: the segregated loop diverges
    pm, m, _, _ = _product_mesh(ctx, (3, 3, 3), jitter=0.0, tets=True, ls_full_delta=False)
    with pytest.raises(_abi.S4FError) as e:
        m.solve(nCorrectors=5000)
    assert e.value.status == _abi.S4F_ERR_DIVERGED
    pm, m, _, _ = _product_mesh(ctx, (6, 6, 6), jitter=0.1, seed=1)
    out = m.solve(nCorrectors=3)
    assert out["status"] == _abi.S4F_WARN_MAXITER and out["nOuter"] == 3 and not out["converged"]


def test_invalid_options_are_refused(ctx):
    # option and state errors come back as S4F_ERR_INVALID with a message,
    # before any kernel runs
    from solids4foam_amd import _abi
    from solids4foam_amd.batch import Batch
    from solids4foam_amd.serve import Server
    pm, m, _, _ = _product_mesh(ctx, (4, 4, 4), jitter=0.1, seed=1)
    for bad in (dict(relaxationMethod=2), dict(d2dt2=3), dict(eagerLaunch=2), dict(relaxD=1.5),
                dict(mgCoarseScale=3.0), dict(d2dt2=1, deltaT=1e-3)):  # last: no density / old-time fields
        with pytest.raises(_abi.S4FError) as e:
            m.solve(**bad)
        assert e.value.status == _abi.S4F_ERR_INVALID, bad
    m.set_density(7800.0)
    with pytest.raises(_abi.S4FError, match="old-time"):
        m.solve(d2dt2=1, deltaT=1e-3)
    with pytest.raises(_abi.S4FError, match="deltaT"):
        m.set_old_time(np.zeros((pm.nCells, 3)))
        m.solve(d2dt2=1, deltaT=0.0)
    with pytest.raises(_abi.S4FError):
        Server(ctx, 64 << 20, max_slots=2, d2dt2=1, deltaT=1e-3)
    # body force: single meshes only; a mesh with gravity set is refused by a batch
    m.set_gravity((0.0, 0.0, -9.81))
    _, m2, _, _ = _product_mesh(ctx, (4, 4, 4), jitter=0.1, seed=2)
    b = Batch(ctx)
    with pytest.raises(_abi.S4FError):
        b.add(m)
        b.add(m2)
        b.solve()
    b.close()
    # a mesh whose last solve was transient is refused by the server; the
    # steady solve after it is the plain steady solve
    m.set_gravity((0.0, 0.0, 0.0))
    out = m.solve(d2dt2=1, deltaT=1e-3, solutionTolerance=1e-10, alternativeTolerance=1e-10)
    assert out["converged"]
    srv = Server(ctx, 64 << 20, max_slots=2)
    with pytest.raises(_abi.S4FError, match="static"):
        srv.submit(m, np.zeros((pm.nBoundaryFaces, 3)))
    srv.close()
    a = m.solve(solutionTolerance=1e-10, alternativeTolerance=1e-10)
    _, m3, _, _ = _product_mesh(ctx, (4, 4, 4), jitter=0.1, seed=1)
    ref = m3.solve(solutionTolerance=1e-10, alternativeTolerance=1e-10)
    assert a["nOuter"] == ref["nOuter"] and np.array_equal(a["D"], ref["D"])


def test_ldu_solver_matches_oracle_pcg(ctx, oracle_mod):
    # lduMatrix::solver seam: one component of the oracle's M D = M D0 + R,
    # PCG with the diagonal preconditioner from the warm start D0, relTol 0.1:
    # same normalised initial residual and (to rounding) the same iterate
    from solids4foam_amd.solid import LduSolver
    om, oc = _oracle_case(oracle_mod, (8, 7, 6), jitter=0.15, seed=6)
    for _ in range(3):
        oc.iterate(precond=oracle_mod.PRECOND_DIAGONAL)
    D0 = oc.get("D").reshape(-1, 3).copy()
    oc.op("update_traction")
    oc.op("residual")
    R = oc.get("R").reshape(-1, 3)
    diag, upper = oc.get("diag"), oc.get("upper")
    L = LduSolver(ctx, om.nCells, om.owner[:om.nInternalFaces], om.neighbour)
    q = 1
    src = oc.amul(np.ascontiguousarray(D0[:, q])) + R[:, q]
    psi, perf = L.solve(diag, upper, src, D0[:, q], tolerance=1e-9, rel_tol=0.1)
    oc.set("D", D0)
    conv, ir, rvf, inner = oc.iterate(precond=oracle_mod.PRECOND_DIAGONAL)
    D1 = oc.get("D").reshape(-1, 3)
    assert perf["nIterations"] > 0
    assert abs(perf["finalResidual"]) < 0.1 * perf["initialResidual"]
    assert rel_l2(psi, D1[:, q]) < 1e-10
    # tight solve == dense
    M = np.diag(diag)
    o, nb = om.owner[:om.nInternalFaces], om.neighbour
    M[o, nb] = upper
    M[nb, o] = upper
    psi2, perf2 = L.solve(diag, upper, src, D0[:, q], tolerance=1e-15, rel_tol=0.0)
    assert rel_l2(psi2, np.linalg.solve(M, src)) < 1e-10
    L.close()

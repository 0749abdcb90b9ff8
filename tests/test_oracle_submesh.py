"""Pins the oracle's sub-mesh (multi-law) discretisation against closed forms.

A case with one mechanical law per cellZone is solved on one sub-mesh per
zone: the faces between zones carry an interface displacement updated to
traction balance every outer iteration (solidSubMeshes.C:1550-1950), each
side's least-squares gradient sees it as a boundary value
(mechanicalModel.C:571-615) and pointD averages the sub-meshes'
interpolations (solidSubMeshesTemplates.C:271-336).

  bi-material column   laterally confined column of two materials under end
                       traction: the exact solution is uniaxial strain,
                       piecewise linear with the kink on the interface; the
                       sub-mesh form reproduces it to round-off (gradient,
                       interface displacement, interface stress, pointD),
                       while the per-cell material mapping misses it by
                       percents -- the size of the difference between the
                       two discretisations
  invariants           zones without interface faces, and clearing zones,
                       give the one-material case bit for bit
"""
import numpy as np
import pytest

from tests._cases import FIXED, TRACTION

SYM = 3
T = 1e7
MATS = ((200e9, 0.3), (70e9, 0.33))
TIGHT = dict(nCorr=20000, solTol=1e-12, altTol=1e-12, pcgTol=1e-15)


def _graded(oracle_mod, n, nz, zi):
    """box with the plane z = 0.5 moved to z = zi: cells below are thinner
    (or thicker) than cells above, still orthogonal"""
    b = oracle_mod.Mesh.box(n, n, nz, (1.0, 1.0, 1.0))
    p = b.points.copy()
    z = p[:, 2]
    p[:, 2] = np.where(z <= 0.5, z * zi / 0.5, zi + (z - 0.5) * (1.0 - zi) / 0.5)
    return oracle_mod.Mesh.from_arrays(p, b.fOff, b.fVerts, b.owner, b.neighbour, b.nCells, b.patchStart,
                                       b.patchSize)


def _column(oracle_mod, n=3, nz=8, zones=True, full=False, zi=0.5, interface_scale=0.01):
    m = oracle_mod.Mesh.box(n, n, nz, (1.0, 1.0, 1.0)) if zi == 0.5 else _graded(oracle_mod, n, nz, zi)
    z = (m.C[:, 2] > zi).astype(np.int32)
    E = np.where(z == 0, MATS[0][0], MATS[1][0])
    nu = np.where(z == 0, MATS[0][1], MATS[1][1])
    bv = np.zeros((m.nBoundaryFaces, 3))
    bv[m.patch_faces(5)] = (0.0, 0.0, T)
    c = oracle_mod.Case(m, E, nu, [SYM, SYM, SYM, SYM, FIXED, TRACTION], bv, ls_full_delta=full,
                        interface_scale=interface_scale)
    nif = c.set_zones(z) if zones else 0
    c.init()
    return m, c, z, nif


def _impK(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    return lam + E / (1 + nu), lam


def _exact(zc, zi=0.5):
    """uniaxial strain: sigma_zz = T in both layers, eps_zz = T/impK"""
    k0, _ = _impK(*MATS[0])
    k1, _ = _impK(*MATS[1])
    return np.where(zc <= zi, T * zc / k0, T * zi / k0 + T * (zc - zi) / k1)


@pytest.mark.parametrize("full", [False, True])
def test_bimaterial_column_exact(oracle_mod, full):
    m, c, z, nif = _column(oracle_mod, full=full)
    assert nif == 9  # the 3 x 3 faces of the plane z = 0.5
    r = c.solve(precond=oracle_mod.PRECOND_DIC, **TIGHT)
    assert r["converged"]
    D = c.get("D").reshape(-1, 3)
    ex = _exact(m.C[:, 2])
    assert np.abs(D[:, 2] - ex).max() < 1e-9 * np.abs(ex).max()
    assert np.abs(D[:, :2]).max() < 1e-9 * np.abs(ex).max()
    s = c.get("sigma").reshape(-1, 6)
    lam = np.where(z == 0, _impK(*MATS[0])[1], _impK(*MATS[1])[1])
    k = np.where(z == 0, _impK(*MATS[0])[0], _impK(*MATS[1])[0])
    assert np.allclose(s[:, 5], T, rtol=1e-9)
    assert np.allclose(s[:, 0], lam * T / k, rtol=1e-9) and np.allclose(s[:, 3], lam * T / k, rtol=1e-9)
    # interface displacement: the exact value on the plane (the post-loop
    # update leaves it there); each side's boundary stress carries traction T
    DI = c.get("ifDI").reshape(-1, 3)
    assert np.allclose(DI[:, 2], _exact(np.array([0.5]))[0], rtol=1e-9)
    sI = c.get("ifSigma").reshape(-1, 2, 6)
    assert np.allclose(sI[:, :, 5], T, rtol=1e-9)
    assert np.allclose(sI[:, 0, 0], _impK(*MATS[0])[1] * T / _impK(*MATS[0])[0], rtol=1e-9)
    assert np.allclose(sI[:, 1, 0], _impK(*MATS[1])[1] * T / _impK(*MATS[1])[0], rtol=1e-9)
    # pointD: each material interpolates its own linear field exactly
    pD = c.point_interpolate(c.get("D"), c.get("G")).reshape(-1, 3)
    pex = _exact(m.points[:, 2])
    assert np.abs(pD[:, 2] - pex).max() < 1e-9 * np.abs(pex).max()


@pytest.mark.parametrize("full", [False, True])
def test_graded_bimaterial_column_exact(oracle_mod, full):
    # interface at z = 0.4 of a column whose lower cells are 0.1 and upper
    # cells 0.15 high: the two sides' patch deltaCoeffs differ (20 and 13.3).
    # Without the interface Rhie-Chow term (interfaceScaleFactor 0: the
    # exact field's kink would otherwise feed it), the exact solution is
    # again the discrete fixed point.
    zi = 0.4
    m, c, z, nif = _column(oracle_mod, zi=zi, full=full, interface_scale=0.0)
    dc = c.get("ifDc").reshape(-1, 2)
    assert nif == 9 and np.allclose(dc[:, 0], 20.0) and np.allclose(dc[:, 1], 1.0 / 0.075)
    r = c.solve(precond=oracle_mod.PRECOND_DIC, **TIGHT)
    assert r["converged"]
    D = c.get("D").reshape(-1, 3)
    ex = _exact(m.C[:, 2], zi)
    assert np.abs(D[:, 2] - ex).max() < 1e-9 * np.abs(ex).max()
    assert np.allclose(c.get("sigma").reshape(-1, 6)[:, 5], T, rtol=1e-9)
    assert np.allclose(c.get("ifDI").reshape(-1, 3)[:, 2], _exact(np.array([zi]), zi)[0], rtol=1e-9)


def _traction_jump(c):
    """normal traction of side b minus side a at each interface face (z)"""
    c.op("grad")
    c.iface_step("sigma")
    sI = c.get("ifSigma").reshape(-1, 2, 6)
    return sI[:, 1, 5] - sI[:, 0, 5]


def test_interface_update_balances_traction(oracle_mod):
    # the update is a Newton step on the interface's traction balance: from
    # the exact cell field with D_I off by delta, one update (cells fixed)
    # gives D_I that balances the normal traction of the two sides; after
    # the cells of side b move by c (the sides' previous state balanced), the
    # update moves D_I to the new balance.  Graded mesh: da != db, Ka != Kb.
    zi = 0.4
    m, c, z, nif = _column(oracle_mod, zi=zi, interface_scale=0.0)
    ex = _exact(m.C[:, 2], zi)
    D = np.zeros((m.nCells, 3))
    D[:, 2] = ex
    DI0 = _exact(np.array([zi]), zi)[0]
    c.set("D", D)
    c.set("Dprev", D)
    DI = np.zeros((nif, 3))
    DI[:, 2] = DI0 + 3e-6
    c.set("ifDI", DI)
    assert np.abs(_traction_jump(c)).min() > 1e-3 * T  # off balance
    c.iface_step("update")
    assert np.allclose(c.get("ifDI").reshape(-1, 3)[:, 2], DI0, rtol=1e-12, atol=0)
    assert np.abs(_traction_jump(c)).max() < 1e-9 * T
    # move side b's cells: the interface follows to the new balance
    D2 = D.copy()
    D2[z == 1, 2] += 2e-6
    c.set("D", D2)
    c.iface_step("update")
    assert np.abs(_traction_jump(c)).max() < 1e-9 * T
    assert np.abs(c.get("ifDI").reshape(-1, 3)[:, 2] - DI0).min() > 1e-7


def test_per_cell_mapping_misses_interface(oracle_mod):
    # the same column with the materials mapped per cell (no sub-meshes):
    # the gradient of the cells next to the interface mixes both slopes
    m, c, z, _ = _column(oracle_mod, zones=False)
    r = c.solve(precond=oracle_mod.PRECOND_DIC, **TIGHT)
    assert r["converged"]
    D = c.get("D").reshape(-1, 3)
    ex = _exact(m.C[:, 2])
    eD = np.abs(D[:, 2] - ex).max() / np.abs(ex).max()
    s = c.get("sigma").reshape(-1, 6)
    eS = np.abs(s[:, 5] - T).max() / T
    print(f"per-cell mapping on the bi-material column: max |D - exact| {eD:.2%}, max |sigma_zz - T| {eS:.2%}")
    assert eD > 0.01 and eS > 0.05


def test_zones_without_interfaces_and_cleared(oracle_mod):
    m = oracle_mod.Mesh.box(4, 3, 3, (1.0, 1.0, 1.0), jitter=0.15, seed=4)
    bv = np.zeros((m.nBoundaryFaces, 3))
    bv[m.patch_faces(1)] = (1e6, 2e6, 0.0)
    E = np.where(m.C[:, 0] > 0.5, 100e9, 200e9)

    def case():
        return oracle_mod.Case(m, E, 0.3, [FIXED] + [TRACTION] * 5, bv)

    ref = case()
    ref.solve(precond=oracle_mod.PRECOND_DIC, nCorr=40)
    a = case()
    assert a.set_zones(np.zeros(m.nCells, np.int32)) == 0
    a.init()
    a.solve(precond=oracle_mod.PRECOND_DIC, nCorr=40)
    b = case()
    assert b.set_zones((m.C[:, 0] > 0.5).astype(np.int32)) > 0
    assert b.set_zones(None) == 0
    b.init()
    b.solve(precond=oracle_mod.PRECOND_DIC, nCorr=40)
    for c in (a, b):
        for f in ("lsP", "lsN", "lsB", "D", "G", "sigma"):
            assert np.array_equal(c.get(f), ref.get(f)), f


def test_submesh_changes_heterogeneous_solution(oracle_mod):
    # a jittered cube with a soft spherical inclusion: both forms converge,
    # and the displacement fields differ by percents (the interface
    # treatment is not a detail at a 10x stiffness contrast)
    m = oracle_mod.Mesh.box(8, 8, 8, (1.0, 1.0, 1.0), jitter=0.15, seed=3)
    z = (np.linalg.norm(m.C - 0.5, axis=1) < 0.3).astype(np.int32)
    E = np.where(z == 0, 200e9, 20e9)
    nu = np.where(z == 0, 0.3, 0.45)
    bv = np.zeros((m.nBoundaryFaces, 3))
    bv[m.patch_faces(5)] = (2e6, 0.0, 1e7)
    out = []
    for zones in (False, True):
        c = oracle_mod.Case(m, E, nu, [TRACTION] * 4 + [FIXED, TRACTION], bv)
        if zones:
            assert c.set_zones(z) > 0
            c.init()
        r = c.solve(precond=oracle_mod.PRECOND_DIC, nCorr=20000, solTol=1e-9, altTol=1e-9, pcgTol=1e-12)
        assert r["converged"]
        out.append(c.get("D"))
    d = np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0])
    print(f"sub-mesh vs per-cell displacement on the 8^3 inclusion: {d:.2%} relative L2")
    assert d > 0.005

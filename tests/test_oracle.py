"""Pins the FP64 CPU oracle against closed forms, invariants and known answers.

The reference's own test suite (tutorials/Alltest) pins no numbers, so each
oracle function is checked against what the mathematics fixes:
  mesh/geometry      exact counts, closed-cell sums, volumes, orthogonal limits
  LS gradient        exact for linear fields (extendedLeastSquaresVectors.C:123-285)
  M assembly         b == M D + R against the literal fvm - fvc assembly
  Hooke / material   linearElastic.C:62-133 closed forms
  div(sigma)         exact for a quadratic displacement on a uniform mesh
  Rhie-Chow          closed-form fourth-difference value for D = x^4
  traction/fixed BC  uniaxial-stress linear field is the discrete fixed point
  whole solve        patch test (linear field exact), cantilever vs beam theory
"""
import numpy as np
import pytest

from tests._cases import E_STEEL, NU_STEEL, FIXED, TRACTION, cantilever_bc, cube_tension_bc


def _slices(m):
    return [m.patch_faces(p) for p in range(m.nPatches)]


def test_box_counts_cfg1(oracle_mod):
    m = oracle_mod.Mesh.box(50, 10, 10, (1.0, 0.2, 0.2))
    # 50x10x10: C=5000, F=13900 internal, B=2200 boundary, P=6171
    assert (m.nCells, m.nInternalFaces, m.nBoundaryFaces, m.nPoints) == (5000, 13900, 2200, 6171)
    assert np.isclose(m.V.sum(), 0.04, rtol=1e-13)
    assert np.allclose(m.V, 0.02**3, rtol=1e-12)
    # upper-triangular LDU order: owner < neighbour, faces sorted by (owner, neighbour)
    o, n = m.owner[:m.nInternalFaces], m.neighbour
    assert np.all(o < n)
    key = o.astype(np.int64) * m.nCells + n
    assert np.all(np.diff(key) > 0)
    assert list(m.patchSize) == [100, 100, 500, 500, 500, 500]


@pytest.mark.parametrize("tets", [False, True])
def test_closed_cells_and_volume(oracle_mod, tets):
    m = oracle_mod.Mesh.box(5, 4, 3, (1.0, 0.8, 0.6), jitter=0.15, seed=7, tets=tets)
    S = np.zeros((m.nCells, 3))
    np.add.at(S, m.owner, m.Sf)
    np.add.at(S, m.neighbour, -m.Sf[:m.nInternalFaces])
    assert np.abs(S).max() < 1e-14
    assert np.all(m.V > 0)
    assert np.isclose(m.V.sum(), 0.48, rtol=1e-12)
    # each face's area vector points out of its owner
    d = m.Cf - m.C[m.owner]
    assert np.all(np.einsum("ij,ij->i", d, m.Sf) > 0)
    if tets:
        assert m.nCells == 6 * 60
        assert np.all(np.diff(m.fOff) == 3)


def test_jitter_keeps_boundary(oracle_mod):
    m0 = oracle_mod.Mesh.box(6, 5, 4, (1.0, 1.0, 1.0))
    m1 = oracle_mod.Mesh.box(6, 5, 4, (1.0, 1.0, 1.0), jitter=0.15, seed=3)
    moved = np.any(m0.points != m1.points, axis=1)
    p = m0.points
    onb = np.any((p == 0.0) | (p == 1.0), axis=1)
    assert not np.any(moved & onb)
    assert np.all(moved[~onb])
    h = np.array([1 / 6, 1 / 5, 1 / 4])
    assert np.all(np.abs(m1.points - m0.points) <= 0.15 * h + 1e-15)


def _case(oracle_mod, m, pt, bv, **kw):
    return oracle_mod.Case(m, E_STEEL, NU_STEEL, pt, bv, **kw)


def test_material_closed_form(oracle_mod):
    m = oracle_mod.Mesh.box(2, 2, 2)
    c = _case(oracle_mod, m, [FIXED] * 6, np.zeros((m.nBoundaryFaces, 3)))
    mu, lam, K, impK = c.get("mu")[0], c.get("lambda")[0], c.get("K")[0], c.get("impK")[0]
    assert np.isclose(mu, 200e9 / 2.6, rtol=1e-15)
    assert np.isclose(lam, 0.3 * 200e9 / (1.3 * 0.4), rtol=1e-15)
    assert np.isclose(K, 200e9 / 1.2, rtol=1e-15)
    assert np.isclose(impK, 2 * mu + lam, rtol=1e-15)
    assert np.isclose(impK, 2.6923076923e11, rtol=1e-10)


def test_ls_grad_linear_exact(oracle_mod):
    rng = np.random.default_rng(1)
    A = rng.normal(size=(3, 3))
    a0 = rng.normal(size=3)
    for jit, full in ((0.0, False), (0.15, True), (0.15, False)):
        m = oracle_mod.Mesh.box(6, 5, 4, jitter=jit, seed=11)
        c = _case(oracle_mod, m, [FIXED] * 6, np.zeros((m.nBoundaryFaces, 3)), ls_full_delta=full)
        Dl = lambda x: a0 + x @ A  # D_j = a0_j + x_i A_ij  => grad G_ij = A_ij
        c.set("D", Dl(m.C))
        c.set("Db", Dl(m.Cf[m.nInternalFaces:]))
        c.set("snb", np.zeros(m.nBoundaryFaces * 3))
        c.op("grad")
        G = c.get("G").reshape(-1, 3, 3)
        err = np.abs(G - A).max(axis=(1, 2))
        bc = np.unique(m.owner[m.nInternalFaces:])
        interior = np.setdiff1d(np.arange(m.nCells), bc)
        assert err[interior].max() < 1e-12
        if full or jit == 0.0:
            assert err.max() < 1e-12
        else:
            # OpenFOAM leastSquares uses the normal patch delta: inexact on skewed boundary cells
            assert err[bc].max() > 1e-6


def test_matrix_properties_and_reference_source(oracle_mod):
    m = oracle_mod.Mesh.box(6, 5, 4, jitter=0.15, seed=5)
    pt, bv = cube_tension_bc(m.nBoundaryFaces, _slices(m))
    c = _case(oracle_mod, m, pt, bv)
    diag, upper = c.get("diag"), c.get("upper")
    assert np.all(upper < 0) and np.all(diag > 0)
    rowsum = diag.copy()
    np.add.at(rowsum, m.owner[:m.nInternalFaces], upper)
    np.add.at(rowsum, m.neighbour, upper)
    fixed_cells = np.unique(m.owner[m.patch_faces(4).start + m.nInternalFaces:
                                    m.patch_faces(4).stop + m.nInternalFaces])
    free = np.setdiff1d(np.arange(m.nCells), fixed_cells)
    assert np.abs(rowsum[free]).max() < 1e-9 * diag.max()
    assert np.all(rowsum[fixed_cells] > 0)
    # advance a few iterations so every field is non-trivial, then compare
    for _ in range(3):
        c.iterate()
    c.op("update_traction")
    c.op("residual")
    D = c.get("D").reshape(-1, 3)
    R = c.get("R").reshape(-1, 3)
    for q in range(3):
        b_ref = c.reference_source(q)
        b_def = c.amul(np.ascontiguousarray(D[:, q])) + R[:, q]
        assert np.abs(b_ref - b_def).max() <= 1e-9 * np.abs(b_ref).max()


def _uniform_case(oracle_mod, n=8):
    m = oracle_mod.Mesh.box(n, n, n, (1.0, 1.0, 1.0))
    c = _case(oracle_mod, m, [FIXED] * 6, np.zeros((m.nBoundaryFaces, 3)))
    return m, c


def _interior(m, n, margin):
    idx = np.arange(m.nCells)
    i, j, k = idx % n, (idx // n) % n, idx // (n * n)
    sel = (i >= margin) & (i < n - margin) & (j >= margin) & (j < n - margin) & (k >= margin) & (k < n - margin)
    return idx[sel]


def test_div_sigma_quadratic_exact(oracle_mod):
    # D = (x^2, 0, 0): G_xx = 2x, sigma linear; Gauss linear div exact:
    # R_x = V d/dx sigma_xx = V 2 impK, R_y = R_z = 0; Rhie-Chow vanishes for quadratics
    n = 8
    m, c = _uniform_case(oracle_mod, n)
    D = np.zeros((m.nCells, 3)); D[:, 0] = m.C[:, 0] ** 2
    Db = np.zeros((m.nBoundaryFaces, 3)); Db[:, 0] = m.Cf[m.nInternalFaces:, 0] ** 2
    c.set("D", D); c.set("Db", Db); c.set("snb", np.zeros(3 * m.nBoundaryFaces))
    c.op("grad"); c.op("sigma"); c.op("residual")
    R = c.get("R").reshape(-1, 3)
    inner = _interior(m, n, 2)
    impK = c.get("impK")[0]
    assert np.allclose(R[inner, 0], 2 * impK * m.V[inner], rtol=1e-9)
    assert np.abs(R[inner, 1:]).max() < 1e-9 * 2 * impK * m.V[0]


def test_rhie_chow_quartic_closed_form(oracle_mod):
    # D = (x^4, 0, 0), sigma forced to zero: R = Rhie-Chow only.
    # Uniform spacing h, LS gradient = central difference 4x^3 + 4xh^2:
    # R_x = gamma |S| Delta (-6 h^4) = -6 gamma |S| h^3 with gamma = 0.1 impK (dissipative sign).
    n = 10
    m, c = _uniform_case(oracle_mod, n)
    h = 1.0 / n
    D = np.zeros((m.nCells, 3)); D[:, 0] = m.C[:, 0] ** 4
    Db = np.zeros((m.nBoundaryFaces, 3)); Db[:, 0] = m.Cf[m.nInternalFaces:, 0] ** 4
    c.set("D", D); c.set("Db", Db); c.set("snb", np.zeros(3 * m.nBoundaryFaces))
    c.op("grad")
    c.set("sigma", np.zeros(6 * m.nCells)); c.set("sigmab", np.zeros(6 * m.nBoundaryFaces))
    c.op("residual")
    R = c.get("R").reshape(-1, 3)
    inner = _interior(m, n, 2)
    gamma = 0.1 * c.get("impK")[0]
    expect = -6.0 * gamma * (h * h) * h ** 3
    assert np.allclose(R[inner, 0], expect, rtol=1e-6)


def test_patch_test_linear_field(oracle_mod):
    # fixedDisplacement D = a + x.B on every patch: the linear field is the
    # discrete fixed point (exact strain), orthogonal mesh + default LS, and
    # jittered mesh + full-delta LS (patchTest/README.md, "Expected Results").
    a = np.array([1e-6, 4e-6, 0.0])
    B = np.array([[2e-6, 5e-6, 0.0], [3e-6, 6e-6, 0.0], [0.0, 0.0, 1e-6]])  # G_ij = dD_j/dx_i
    for jit, full in ((0.0, False), (0.12, True)):
        m = oracle_mod.Mesh.box(5, 4, 4, jitter=jit, seed=9)
        bv = a + m.Cf[m.nInternalFaces:] @ B
        c = _case(oracle_mod, m, [FIXED] * 6, bv, ls_full_delta=full)
        r = c.solve(nCorr=3000, solTol=1e-13, altTol=1e-13, pcgTol=1e-14)
        D = c.get("D").reshape(-1, 3)
        exact = a + m.C @ B
        assert np.abs(D - exact).max() < 1e-9 * np.abs(exact).max(), r
        eps = c.get("G").reshape(-1, 3, 3)
        assert np.abs(eps - B).max() < 1e-9 * np.abs(B).max()


def test_uniaxial_traction_fixed_point(oracle_mod):
    # z-min: fixedDisplacement with the uniaxial-stress field; z-max traction t;
    # sides traction-free.  The linear field D = (-nu t/E x, -nu t/E y, t/E z)
    # is the exact discrete solution; at convergence n.sigma_b = t on traction faces.
    t = 1e7
    for jit, full in ((0.0, False), (0.12, True)):
        m = oracle_mod.Mesh.box(4, 4, 5, jitter=jit, seed=4)
        Ex = lambda x: np.stack([-NU_STEEL * t / E_STEEL * x[:, 0], -NU_STEEL * t / E_STEEL * x[:, 1],
                                 t / E_STEEL * x[:, 2]], axis=1)
        bv = np.zeros((m.nBoundaryFaces, 3))
        zmin = m.patch_faces(4)
        bv[zmin] = Ex(m.Cf[m.nInternalFaces:][zmin])
        bv[m.patch_faces(5)] = (0, 0, t)
        c = _case(oracle_mod, m, [TRACTION] * 4 + [FIXED, TRACTION], bv, ls_full_delta=full)
        c.solve(nCorr=5000, solTol=1e-13, altTol=1e-13, pcgTol=1e-14)
        D = c.get("D").reshape(-1, 3)
        assert np.abs(D - Ex(m.C)).max() < 1e-8 * np.abs(Ex(m.C)).max()
        sb = c.get("sigmab").reshape(-1, 6)
        zmax = m.patch_faces(5)
        assert np.allclose(sb[zmax, 5], t, rtol=1e-8)


def test_traction_balance_on_skewed_cube(oracle_mod):
    # nonlinear field on a skewed mesh: at convergence the traction boundary
    # condition holds, n.sigma_b = t, on every traction face
    m = oracle_mod.Mesh.box(6, 6, 6, jitter=0.15, seed=2)
    pt, bv = cube_tension_bc(m.nBoundaryFaces, _slices(m))
    c = _case(oracle_mod, m, pt, bv)
    r = c.solve(nCorr=5000, solTol=1e-12, altTol=1e-12, pcgTol=1e-14)
    assert r["converged"]
    sb = c.get("sigmab").reshape(-1, 6)
    n = c.get("bN").reshape(-1, 3)
    ns = np.stack([n[:, 0] * sb[:, 0] + n[:, 1] * sb[:, 1] + n[:, 2] * sb[:, 2],
                   n[:, 0] * sb[:, 1] + n[:, 1] * sb[:, 3] + n[:, 2] * sb[:, 4],
                   n[:, 0] * sb[:, 2] + n[:, 1] * sb[:, 4] + n[:, 2] * sb[:, 5]], axis=1)
    for p in (0, 1, 2, 3, 5):
        s = m.patch_faces(p)
        assert np.abs(ns[s] - bv[s]).max() < 1e-6 * 1e7


def test_cantilever_vs_beam_theory(oracle_mod):
    # cfg1 cantilever; Timoshenko tip deflection F L^3/(3EI) + F L/(kappa G A)
    m = oracle_mod.Mesh.box(25, 5, 5, (1.0, 0.2, 0.2))
    pt, bv = cantilever_bc(m.nBoundaryFaces, _slices(m))
    c = _case(oracle_mod, m, pt, bv)
    r = c.solve(nCorr=20000, solTol=1e-6, altTol=1e-7)
    assert r["converged"]
    D = c.get("D").reshape(-1, 3)
    F, L, b, h = 1e6 * 0.04, 1.0, 0.2, 0.2
    I = b * h ** 3 / 12
    G = E_STEEL / (2 * (1 + NU_STEEL))
    tip = F * L ** 3 / (3 * E_STEEL * I) + F * L / (5.0 / 6.0 * G * b * h)
    # cell-centre maximum sits half a cell short of the tip: compare at x = L - h/2
    xc = L - 0.5 * L / 25
    expect = F * (3 * L * xc ** 2 - xc ** 3) / (6 * E_STEEL * I) + F * xc / (5.0 / 6.0 * G * b * h)
    assert expect < tip
    assert -D[:, 1].min() == pytest.approx(expect, rel=0.08)
    assert np.all(D[:, 1] <= 1e-12)


def test_pcg_matches_dense_solve(oracle_mod):
    # one outer iteration with a tight inner solve == numpy dense solve of M x = M D + R
    m = oracle_mod.Mesh.box(4, 3, 3, jitter=0.1, seed=1)
    pt, bv = cube_tension_bc(m.nBoundaryFaces, _slices(m))
    c = _case(oracle_mod, m, pt, bv)
    for _ in range(2):
        c.iterate()
    D0 = c.get("D").reshape(-1, 3).copy()
    diag, upper = c.get("diag"), c.get("upper")
    M = np.diag(diag)
    o, n = m.owner[:m.nInternalFaces], m.neighbour
    M[o, n] = upper; M[n, o] = upper
    c.op("update_traction"); c.op("residual")
    R = c.get("R").reshape(-1, 3)
    # rewind the traction update: iterate() applies it itself
    c2 = _case(oracle_mod, m, pt, bv)
    for _ in range(2):
        c2.iterate()
    for pre in (oracle_mod.PRECOND_DIC, oracle_mod.PRECOND_DIAGONAL):
        c3 = _case(oracle_mod, m, pt, bv)
        for _ in range(2):
            c3.iterate()
        c3.iterate(pcgTol=1e-15, pcgRelTol=0.0, precond=pre)
        D1 = c3.get("D").reshape(-1, 3)
        exact = D0 + np.linalg.solve(M, R)
        assert np.abs(D1 - exact).max() < 1e-10 * np.abs(exact).max()


def _box_nu(oracle_mod, n, nu, jitter=0.1, seed=3):
    om = oracle_mod.Mesh.box(n, n, n, (0.1, 0.1, 0.1), jitter, seed)
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(1)] = (1e6, 5e5, 0.0)
    return om, oracle_mod.Case(om, 100e9, nu, [FIXED] + [TRACTION] * 5, bv)


def test_field_relaxation_keeps_fixed_point(oracle_mod):
    # relaxField 'fixed' (solidModel.C:837-841) changes the path, not the
    # converged field: nu = 0.3 reaches the same D with and without it
    _, a = _box_nu(oracle_mod, 6, 0.3)
    _, b = _box_nu(oracle_mod, 6, 0.3)
    kw = dict(nCorr=20000, solTol=1e-11, altTol=1e-11, precond=oracle_mod.PRECOND_DIC)
    ra, rb = a.solve(**kw), b.solve(relaxD=0.7, **kw)
    assert ra["converged"] and rb["converged"] and rb["nOuter"] > ra["nOuter"]
    Da, Db = a.get("D"), b.get("D")
    assert np.linalg.norm(Da - Db) / np.linalg.norm(Da) < 1e-8


def test_field_relaxation_restores_convergence_high_nu(oracle_mod):
    # the segregated loop oscillates without relaxation at nu = 0.45 and
    # converges with D relaxed by 0.7
    _, a = _box_nu(oracle_mod, 6, 0.45)
    ra = a.solve(nCorr=400, solTol=1e-9, altTol=1e-9, precond=oracle_mod.PRECOND_DIC)
    assert not ra["converged"] and ra["initRes"] > 0.1
    _, b = _box_nu(oracle_mod, 6, 0.45)
    rb = b.solve(nCorr=20000, solTol=1e-9, altTol=1e-9, precond=oracle_mod.PRECOND_DIC, relaxD=0.7)
    assert rb["converged"]
    D = b.get("D").reshape(-1, 3)
    assert np.all(np.isfinite(D)) and np.abs(D).max() > 0


@pytest.mark.parametrize("tets", [False, True])
def test_point_interpolation_linear_exact(oracle_mod, tets):
    # pointD = sum w (D_c + (x_p - C_c).G_c) / sum w reproduces any linear
    # field exactly at every point (interior and boundary), whatever the weights
    rng = np.random.default_rng(4)
    A = rng.normal(size=(3, 3))
    a0 = rng.normal(size=3)
    m = oracle_mod.Mesh.box(5, 4, 3, jitter=0.15, seed=9, tets=tets)
    c = _case(oracle_mod, m, [FIXED] * 6, np.zeros((m.nBoundaryFaces, 3)))
    D = a0 + m.C @ A
    G = np.broadcast_to(A.reshape(1, 9), (m.nCells, 9))
    pD = c.point_interpolate(D, G)
    assert np.abs(pD - (a0 + m.points @ A)).max() < 1e-12


def test_point_interpolation_weights(oracle_mod):
    # inverse-distance weights: on a uniform orthogonal mesh an interior point
    # is equidistant from its 8 cells, so pointD is their plain mean (G = 0);
    # a boundary corner point has exactly one cell
    m = oracle_mod.Mesh.box(4, 4, 4, (1.0, 1.0, 1.0))
    c = _case(oracle_mod, m, [FIXED] * 6, np.zeros((m.nBoundaryFaces, 3)))
    rng = np.random.default_rng(5)
    D = rng.normal(size=(m.nCells, 3))
    pD = c.point_interpolate(D, np.zeros((m.nCells, 9)))
    # blockMesh numbering: point (i, j, k) = i + 5 j + 25 k, cell (i, j, k) = i + 4 j + 16 k
    p = 2 + 5 * 2 + 25 * 2
    cells = [i + 4 * j + 16 * k for i in (1, 2) for j in (1, 2) for k in (1, 2)]
    assert np.allclose(pD[p], D[cells].mean(axis=0), rtol=1e-13, atol=1e-15)
    assert np.allclose(pD[0], D[0], rtol=1e-14)
    # jittered: weights are 1/|x_p - C_c|, checked on one interior point
    mj = oracle_mod.Mesh.box(4, 4, 4, (1.0, 1.0, 1.0), jitter=0.2, seed=3)
    cj = _case(oracle_mod, mj, [FIXED] * 6, np.zeros((mj.nBoundaryFaces, 3)))
    pDj = cj.point_interpolate(D, np.zeros((mj.nCells, 9)))
    w = 1.0 / np.linalg.norm(mj.points[p] - mj.C[cells], axis=1)
    assert np.allclose(pDj[p], (w[:, None] * D[cells]).sum(0) / w.sum(), rtol=1e-12)


def _patch_test_case(oracle_mod, full_delta):
    from tests import _blockmesh as bm
    a = bm.hex_blocks(bm.PATCH_TEST_POINTS, bm.PATCH_TEST_HEXES, bm.PATCH_TEST_PATCHES)
    m = oracle_mod.Mesh.from_arrays(a["points"], a["fOff"], a["fVerts"], a["owner"], a["neighbour"], a["nCells"],
                                    a["patchStart"], a["patchSize"])
    bv = np.zeros((m.nBoundaryFaces, 3))
    nb = int(a["patchSize"][0])
    bv[:nb] = bm.linear_spatial(m.Cf[m.nInternalFaces:m.nInternalFaces + nb], bm.PATCH_TEST_A, bm.PATCH_TEST_B)
    return m, _case(oracle_mod, m, [FIXED, 2], bv, ls_full_delta=full_delta)


def test_reference_patch_test_2d(oracle_mod):
    # tutorials/solids/linearElasticity/patchTest: five irregular quadrilaterals,
    # linearSpatialDisplacement on the outer patch, front/back empty (2-D).  Its
    # README gives the exact answer: uniform strain exx = 2e-6, eyy = 6e-6,
    # exy = 4e-6; D at the cell centres is the prescribed linear field and the
    # out-of-plane component stays zero.  With the full-delta least-squares
    # gradient (extendedLeastSquares 0) the discretisation reproduces it.
    m, c = _patch_test_case(oracle_mod, True)
    r = c.solve(nCorr=1000, solTol=1e-14, altTol=1e-14, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC)
    assert r["converged"]
    G = c.get("G").reshape(-1, 3, 3)
    eps = 0.5 * (G + G.transpose(0, 2, 1))
    assert np.allclose(eps[:, 0, 0], 2e-6, rtol=1e-8) and np.allclose(eps[:, 1, 1], 6e-6, rtol=1e-8)
    assert np.allclose(eps[:, 0, 1], 4e-6, rtol=1e-8)
    assert np.abs(eps[:, 2, :]).max() < 1e-20
    D = c.get("D").reshape(-1, 3)
    from tests import _blockmesh as bm
    assert np.allclose(D, bm.linear_spatial(m.C, bm.PATCH_TEST_A, bm.PATCH_TEST_B), rtol=1e-8, atol=1e-20)
    # the normal-delta leastSquares of the tutorial's own fvSchemes is inexact
    # on these skewed boundary cells (the README's caveat): strains deviate
    m2, c2 = _patch_test_case(oracle_mod, False)
    r2 = c2.solve(nCorr=1000, solTol=1e-14, altTol=1e-14, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC)
    G2 = c2.get("G").reshape(-1, 3, 3)
    assert np.abs(G2[:, 0, 0] - 2e-6).max() > 1e-9


def test_symmetry_planes_uniaxial_exact(oracle_mod):
    # solidSymmetry on x-min, y-min, z-min, normal traction t on z-max, free
    # x-max / y-max: the exact solution is uniaxial stress, D = (-nu t x/E,
    # -nu t y/E, t z/E), which an orthogonal mesh reproduces exactly
    E, nu, t = 200e9, 0.3, 1e7
    m = oracle_mod.Mesh.box(6, 5, 4, (1.0, 1.0, 1.0))
    SYM = 3
    bv = np.zeros((m.nBoundaryFaces, 3))
    bv[m.patch_faces(5)] = (0.0, 0.0, t)
    c = oracle_mod.Case(m, E, nu, [SYM, TRACTION, SYM, TRACTION, SYM, TRACTION], bv)
    r = c.solve(nCorr=20000, solTol=1e-13, altTol=1e-13, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC)
    assert r["converged"]
    D = c.get("D").reshape(-1, 3)
    X = m.C
    exact = np.stack([-nu * t * X[:, 0] / E, -nu * t * X[:, 1] / E, t * X[:, 2] / E], axis=1)
    assert np.abs(D - exact).max() < 1e-8 * np.abs(exact).max()
    s = c.get("sigma").reshape(-1, 6)
    assert np.allclose(s[:, 5], t, rtol=1e-7) and np.abs(s[:, :5]).max() < 1e-6 * t
    # the symmetry value is the tangential part of the cell value: n.D_b = 0
    Db = c.get("Db").reshape(-1, 3)
    assert np.abs(Db[m.patch_faces(0), 0]).max() < 1e-20 and np.abs(Db[m.patch_faces(4), 2]).max() < 1e-20


@pytest.mark.parametrize("plane_stress", [True, False])
def test_plane_stress_2d_uniaxial_exact(oracle_mod, plane_stress):
    # 2-D plate (empty z faces), solidSymmetry rollers on x-min and y-min,
    # traction t on x-max: the discrete solution is the uniform strain with
    # sigma_xx = t, sigma_yy = 0 under linearElastic's Hooke law
    # sigma = 2 mu dev(eps) + K tr(eps) I (eps_zz = 0), whose K is
    # E/(3(1-nu)) with planeStress and E/(3(1-2nu)) without
    # (linearElastic.C:62-133, :336-338).  Solving those two equations for
    # (eps_xx, eps_yy) fixes the strain; a wrong plane-stress K or the 3-D
    # one fails the other case.
    E, nu, t = 200e9, 0.3, 1e7
    m = oracle_mod.Mesh.box(6, 5, 1, (1.0, 0.8, 0.1))
    SYM, EMPTY = 3, oracle_mod.EMPTY
    bv = np.zeros((m.nBoundaryFaces, 3))
    bv[m.patch_faces(1)] = (t, 0.0, 0.0)
    c = oracle_mod.Case(m, E, nu, [SYM, TRACTION, SYM, TRACTION, EMPTY, EMPTY], bv, plane_stress=plane_stress)
    r = c.solve(nCorr=50000, solTol=1e-13, altTol=1e-13, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC)
    assert r["converged"]
    mu = E / (2 * (1 + nu))
    K = E / (3 * (1 - nu)) if plane_stress else E / (3 * (1 - 2 * nu))
    le = K - 2 * mu / 3  # sigma = 2 mu eps + le tr(eps) I
    A = np.array([[2 * mu + le, le], [le, 2 * mu + le]])
    exx, eyy = np.linalg.solve(A, [t, 0.0])
    D = c.get("D").reshape(-1, 3)
    X = m.C
    exact = np.stack([exx * X[:, 0], eyy * X[:, 1], np.zeros(m.nCells)], axis=1)
    assert np.abs(D - exact).max() < 1e-8 * np.abs(exact).max()
    s = c.get("sigma").reshape(-1, 6)
    assert np.allclose(s[:, 0], t, rtol=1e-7) and np.abs(s[:, 3]).max() < 1e-6 * t


def _zero_shear_box(oracle_mod, jitter, top, relax=None):
    # fixedDisplacementZeroShear rollers (zero normal displacement) on x-min,
    # y-min, z-min; x-max / y-max traction free; z-max either a normal
    # traction ("traction", t) or a zero-shear patch with prescribed
    # displacement ("zs", its tangential part must not matter)
    E, nu = 200e9, 0.3
    m = oracle_mod.Mesh.box(6, 5, 4, (1.0, 1.0, 1.0), jitter, 3)
    ZS = oracle_mod.ZERO_SHEAR
    bv = np.zeros((m.nBoundaryFaces, 3))
    kind, v = top
    bv[m.patch_faces(5)] = v
    types = [ZS, TRACTION, ZS, TRACTION, ZS, TRACTION if kind == "traction" else ZS]
    # the full boundary delta in the least-squares vectors (extendedLeastSquares)
    # keeps the gradient of a linear field exact on skewed boundary cells
    c = oracle_mod.Case(m, E, nu, types, bv, ls_full_delta=jitter > 0)
    kw = dict(relaxD=relax) if relax else {}
    r = c.solve(nCorr=50000, solTol=1e-13, altTol=1e-13, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC, **kw)
    return m, c, r, E, nu


@pytest.mark.parametrize("jitter", [0.0, 0.15])
def test_zero_shear_rollers_uniaxial_exact(oracle_mod, jitter):
    # rollers + normal traction t: uniaxial stress, D = (-nu t x/E, -nu t y/E,
    # t z/E); linear fields are exact for the corrected scheme on skewed cells
    t = 1e7
    m, c, r, E, nu = _zero_shear_box(oracle_mod, jitter, ("traction", (0.0, 0.0, t)))
    assert r["converged"]
    X = m.C
    exact = np.stack([-nu * t * X[:, 0] / E, -nu * t * X[:, 1] / E, t * X[:, 2] / E], axis=1)
    assert np.abs(c.get("D").reshape(-1, 3) - exact).max() < 1e-8 * np.abs(exact).max()
    s = c.get("sigma").reshape(-1, 6)
    assert np.allclose(s[:, 5], t, rtol=1e-7) and np.abs(s[:, :5]).max() < 1e-6 * t
    # the roller value: zero normal component, the tangential part follows the
    # cell (not the prescribed zero vector, which a fixed patch would impose)
    Db = c.get("Db").reshape(-1, 3)
    f0 = m.patch_faces(0)
    assert np.abs(Db[f0, 0]).max() < 1e-20
    assert np.abs(Db[f0, 1]).max() > 1e-6 * t / E


def test_zero_shear_prescribed_normal_displacement(oracle_mod):
    # z-max moved by delta along its normal through a zero-shear patch whose
    # prescribed vector also has tangential components: only n (n.value)
    # binds, so the field is uniaxial stress with strain delta:
    # D = (-nu delta x, -nu delta y, delta z), sigma_zz = E delta
    delta = 1e-4
    m, c, r, E, nu = _zero_shear_box(oracle_mod, 0.15, ("zs", (0.3 * delta, -0.2 * delta, delta)))
    assert r["converged"]
    X = m.C
    exact = np.stack([-nu * delta * X[:, 0], -nu * delta * X[:, 1], delta * X[:, 2]], axis=1)
    assert np.abs(c.get("D").reshape(-1, 3) - exact).max() < 1e-8 * delta
    s = c.get("sigma").reshape(-1, 6)
    assert np.allclose(s[:, 5], E * delta, rtol=1e-7)
    # with the relaxation factor 0.6 the loop reaches the same field
    m2, c2, r2, _, _ = _zero_shear_box(oracle_mod, 0.15, ("zs", (0.3 * delta, -0.2 * delta, delta)), relax=0.6)
    assert r2["converged"]
    assert np.abs(c2.get("D").reshape(-1, 3) - exact).max() < 1e-8 * delta


def _aitken_case(oracle_mod, driven):
    # 6^3 jittered box, x-min fixed; x-max either moved (fixed displacement)
    # or loaded (traction); other faces traction free
    m = oracle_mod.Mesh.box(6, 6, 6, (0.1, 0.1, 0.1), 0.1, 3)
    bv = np.zeros((m.nBoundaryFaces, 3))
    if driven == "displacement":
        bv[m.patch_faces(1)] = (1e-5, 5e-6, 0.0)
        types = [FIXED, FIXED] + [TRACTION] * 4
    else:
        bv[m.patch_faces(1)] = (1e6, 5e5, 0.0)
        types = [FIXED] + [TRACTION] * 5
    return m, oracle_mod.Case(m, 100e9, 0.3, types, bv)


def test_aitken_first_iteration_extrapolates(oracle_mod):
    # relaxField Aitken subtracts alpha (D_prev - D) from the solved field:
    # on the first iteration (alpha = the D relaxation factor, or 1) the
    # relaxed field is D* + a0 (D* - D_prev), with D_prev = 0 here: (1 + a0) D*
    out = {}
    for key, kw in (("plain", {}), ("a1", dict(relaxAitken=1)), ("a05", dict(relaxAitken=1, relaxD=0.5))):
        m, c = _aitken_case(oracle_mod, "displacement")
        c.solve(nCorr=1, pcgTol=1e-14, pcgRelTol=0.0, precond=oracle_mod.PRECOND_DIC, **kw)
        out[key] = (c.get("D"), c.get("aitkenAlpha"))
    D = out["plain"][0]
    assert np.abs(D).max() > 0
    assert np.abs(out["a1"][0] - 2.0 * D).max() <= 1e-14 * np.abs(D).max()
    assert np.abs(out["a05"][0] - 1.5 * D).max() <= 1e-14 * np.abs(D).max()
    assert np.all(out["a1"][1] == 1.0) and np.all(out["a05"][1] == 0.5)


@pytest.mark.parametrize("driven", ["displacement", "traction"])
def test_aitken_reaches_the_fixed_point(oracle_mod, driven):
    # the relaxation changes the path, not the converged field; the factors
    # stay within the clip [0, 2]
    m, c = _aitken_case(oracle_mod, driven)
    r0 = c.solve(nCorr=20000, solTol=1e-11, altTol=1e-11, pcgTol=1e-13, precond=oracle_mod.PRECOND_DIC)
    m, ca = _aitken_case(oracle_mod, driven)
    ra = ca.solve(nCorr=20000, solTol=1e-11, altTol=1e-11, pcgTol=1e-13, precond=oracle_mod.PRECOND_DIC,
                  relaxAitken=1)
    assert r0["converged"] and ra["converged"]
    D, Da = c.get("D"), ca.get("D")
    assert np.linalg.norm(Da - D) / np.linalg.norm(D) < 1e-8
    a, ab = ca.get("aitkenAlpha"), ca.get("aitkenAlphab")
    assert a.min() >= 0.0 and a.max() <= 2.0 and ab.min() >= 0.0 and ab.max() <= 2.0
    if driven == "traction":
        # the first solve leaves the cells at rest (the load reaches them
        # through the boundary values), so r_prev = 0 at the second iteration
        # and the multiplicative update keeps the cell factors at zero; the
        # fixed patch's values are never relaxed (its factor is not set)
        assert np.all(a == 0.0)
        m, c1 = _aitken_case(oracle_mod, driven)
        c1.solve(nCorr=1, pcgTol=1e-13, precond=oracle_mod.PRECOND_DIC, relaxAitken=1)
        ab1 = c1.get("aitkenAlphab")
        assert np.all(ab1[m.patch_faces(0)] == 0.0) and np.all(ab1[m.patch_faces(1)] == 1.0)


def test_free_fall_follows_euler_d2dt2(oracle_mod):
    # a free body (all faces traction free) under gravity: every cell moves
    # together and the Euler d2dt2 of a start from rest gives exactly
    # D^n = n (n + 1) / 2 g dt^2 (the discrete second difference of it is g);
    # no strain, no stress
    m = oracle_mod.Mesh.box(5, 4, 3, (1.0, 0.8, 0.6), 0.1, 3)
    c = oracle_mod.Case(m, 200e9, 0.3, [TRACTION] * 6, np.zeros((m.nBoundaryFaces, 3)))
    c.set("rho", np.full(m.nCells, 7800.0))
    g, dt = np.array([0.0, -9.81, 2.0]), 1e-3
    for n in range(1, 5):
        if n > 1:
            c.advance_time()
        r = c.solve(nCorr=200, solTol=1e-12, altTol=1e-12, pcgTol=1e-14, pcgRelTol=0.0,
                    precond=oracle_mod.PRECOND_DIC, d2dt2=1, deltaT=dt, deltaT0=dt, g=tuple(g))
        assert r["converged"]
        exact = n * (n + 1) / 2 * g * dt * dt
        assert np.abs(c.get("D").reshape(-1, 3) - exact).max() < 1e-12 * np.abs(exact).max()
        assert np.abs(c.get("sigma")).max() < 1e-6 * 7800.0 * 9.81
    # a changed time step follows the variable-step coefficients: with
    # dt0 = dt the constant-step result, and D^{n+1} for dt != dt0 satisfies
    # coefft (D^{n+1} - D^n)/dt ... i.e. 2/(dt + dt0) ((D^{n+1} - D^n)/dt - (D^n - D^{n-1})/dt0) = g
    D4, D3 = c.get("D").reshape(-1, 3).copy(), c.get("Dold").reshape(-1, 3).copy()
    c.advance_time()
    dt1 = 2e-3
    c.solve(nCorr=200, solTol=1e-12, altTol=1e-12, pcgTol=1e-14, pcgRelTol=0.0,
            precond=oracle_mod.PRECOND_DIC, d2dt2=1, deltaT=dt1, deltaT0=dt, g=tuple(g))
    D5 = c.get("D").reshape(-1, 3)
    acc = 2.0 / (dt1 + dt) * ((D5 - D4) / dt1 - (D4 - D3) / dt)
    assert np.abs(acc - g).max() < 1e-9 * np.abs(g).max()


def test_self_weight_reaction_equals_weight(oracle_mod):
    # a body fixed at z-min under gravity (steady): at convergence the
    # boundary stresses carry the weight, sum_b S_b . sigma_b = -rho g V
    # (the discrete residual is conservative: interior fluxes cancel,
    # gamma_b = 0), and the traction-free faces carry none
    m = oracle_mod.Mesh.box(6, 5, 4, (0.5, 0.4, 1.0), 0.1, 3)
    c = oracle_mod.Case(m, 1e9, 0.3, [TRACTION] * 4 + [FIXED, TRACTION], np.zeros((m.nBoundaryFaces, 3)))
    rho, g = 2000.0, np.array([1.0, 0.0, -9.81])
    c.set("rho", np.full(m.nCells, rho))
    r = c.solve(nCorr=50000, solTol=1e-12, altTol=1e-12, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC,
                g=tuple(g))
    assert r["converged"]
    sb = c.get("sigmab").reshape(-1, 6)
    S = m.Sf[m.nInternalFaces:]
    T = np.array([[sb[:, 0], sb[:, 1], sb[:, 2]], [sb[:, 1], sb[:, 3], sb[:, 4]], [sb[:, 2], sb[:, 4], sb[:, 5]]])
    f = np.einsum("bi,ijb->bj", S, T)          # S . sigma per boundary face
    weight = rho * m.V.sum() * g
    bottom = m.patch_faces(4)
    assert np.allclose(f[bottom].sum(axis=0), -weight, rtol=1e-7, atol=1e-9 * np.abs(weight).max())
    bidx = np.arange(m.nBoundaryFaces)[bottom]
    others = np.setdiff1d(np.arange(m.nBoundaryFaces), bidx)
    assert np.abs(f[others]).max() < 1e-6 * np.abs(weight).max() / len(bidx)


def test_incompressible_nu_is_refused(oracle_mod):
    # linearElastic.C:101-133: nu = 0.5 needs the pressure equation; the
    # oracle (like the device path) refuses it rather than dividing by zero
    om = oracle_mod.Mesh.box(2, 2, 2)
    bv = np.zeros((om.nBoundaryFaces, 3))
    oracle_mod.Case(om, 200e9, 0.49, [0, 1, 1, 1, 1, 1], bv)
    for nu in (0.5, 0.6, -1.0):
        with pytest.raises(ValueError, match="nu"):
            oracle_mod.Case(om, 200e9, nu, [0, 1, 1, 1, 1, 1], bv)


def test_backward_d2dt2_is_exact_for_cubic_motion(oracle_mod):
    # backwardD2dt2Scheme (backwardD2dt2Scheme.C:227-300) applies backward
    # ddt to D and its two old levels: a five-level second difference exact
    # for cubics.  A free body under the body force rho (g0 + 6 j t) whose four
    # old-time levels lie on D(t) = a + v t + g0 t^2 / 2 + j t^3 moves on
    # exactly that curve, with no strain and no stress; Euler's three-level
    # difference (exact only for quadratics) leaves it
    m = oracle_mod.Mesh.box(4, 3, 3, (1.0, 0.8, 0.6), 0.1, 3)
    c = oracle_mod.Case(m, 200e9, 0.3, [TRACTION] * 6, np.zeros((m.nBoundaryFaces, 3)))
    c.set("rho", np.full(m.nCells, 7800.0))
    g0, a, v = np.array([0.0, -9.81, 2.0]), np.array([1e-3, 0.0, -2e-3]), np.array([0.5, 0.2, -0.1])
    j, dt = np.array([300.0, -150.0, 80.0]), 1e-3
    curve = lambda t: np.broadcast_to(a + v * t + 0.5 * g0 * t * t + j * t ** 3, (m.nCells, 3))  # noqa: E731
    for name, k in (("Dold", 0), ("Doo", 1), ("Dooo", 2), ("Doooo", 3)):
        c.set(name, curve(-k * dt))
    c.init(curve(0.0))
    opts = dict(nCorr=200, solTol=1e-13, altTol=1e-13, pcgTol=1e-15, pcgRelTol=0.0, precond=oracle_mod.PRECOND_DIC,
                deltaT=dt, deltaT0=dt)
    for n in range(1, 5):
        if n > 1:
            c.advance_time()
        r = c.solve(d2dt2=2, g=tuple(g0 + 6.0 * j * n * dt), **opts)
        assert r["converged"]
        exact = curve(n * dt)
        assert np.abs(c.get("D").reshape(-1, 3) - exact).max() < 1e-9 * np.abs(exact).max(), n
        assert np.abs(c.get("sigma")).max() < 1e-6 * 7800.0 * 9.81
    c.advance_time()
    c.solve(d2dt2=1, g=tuple(g0 + 6.0 * j * 5 * dt), **opts)
    exact = curve(5 * dt)
    assert np.abs(c.get("D").reshape(-1, 3) - exact).max() > 1e-7 * np.abs(exact).max()

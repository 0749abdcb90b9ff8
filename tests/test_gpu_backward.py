"""backward d2dt2 (backwardD2dt2Scheme.C:227-300) on the device against its
closed form and the oracle: (i) a free body whose four old-time levels lie on
a cubic in time, under a body force that makes the cubic the exact motion,
follows it to rounding (Euler would not: its second difference is exact only
for quadratics); (ii) a loaded cantilever under gravity started from rest,
four steps, matches the oracle step by step; (iii) a time step other than
the previous one is refused (the scheme's constant-step restriction)."""
import numpy as np
import pytest

from tests._cases import FIXED, TRACTION

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from solids4foam_amd import Context
    c = Context(0)
    yield c
    c.close()


def test_backward_free_body_follows_cubic_motion(ctx):
    from solids4foam_amd import PolyMesh, SolverMesh
    print("[backward] free body, cubic history", flush=True)
    pm = PolyMesh.box(4, 3, 3, L=(1.0, 0.8, 0.6), jitter=0.1, seed=3)
    m = SolverMesh(ctx, pm, [TRACTION] * 6)
    m.set_material(200e9, 0.3)
    m.set_density(7800.0)
    for p in range(6):
        m.set_bc(p, np.zeros((pm.patch_faces(p).stop - pm.patch_faces(p).start, 3)))
    g0, a, v = np.array([0.0, -9.81, 2.0]), np.array([1e-3, 0.0, -2e-3]), np.array([0.5, 0.2, -0.1])
    j, dt = np.array([300.0, -150.0, 80.0]), 1e-3
    curve = lambda t: np.broadcast_to(a + v * t + 0.5 * g0 * t * t + j * t ** 3, (pm.nCells, 3)).copy()  # noqa: E731
    levels = [curve(-k * dt) for k in range(4)]
    for n in range(1, 5):
        m.set_gravity(g0 + 6.0 * j * n * dt)
        m.set_old_times(levels)
        out = m.solve(D0=levels[0], precision=0, d2dt2=2, deltaT=dt, deltaT0=dt, nCorrectors=200,
                      solutionTolerance=1e-13, alternativeTolerance=1e-13, pcgTolerance=1e-15, pcgRelTol=0.0)
        assert out["converged"]
        exact = curve(n * dt)
        err = np.abs(out["D"] - exact).max() / np.abs(exact).max()
        print(f"  step {n}: err {err:.2e}", flush=True)
        assert err < 1e-9
        levels = [out["D"]] + levels[:3]


@pytest.mark.parametrize("precision", [0, 1])
def test_backward_cantilever_matches_oracle(ctx, oracle_mod, precision):
    from solids4foam_amd import PolyMesh, SolverMesh, S4FError
    print(f"[backward] cantilever under gravity, precision {precision}", flush=True)
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
    dt = 2e-5
    kw = dict(nCorr=100000, solTol=1e-12, altTol=1e-12, pcgTol=1e-15, precond=oracle_mod.PRECOND_DIC)
    levels = [np.zeros((pm.nCells, 3))] * 4
    for n in range(1, 5):
        if n > 1:
            oc.advance_time()
        ro = oc.solve(d2dt2=2, deltaT=dt, deltaT0=dt, g=grav, **kw)
        m.set_old_times(levels)
        out = m.solve(precision=precision, d2dt2=2, deltaT=dt, deltaT0=dt, nCorrectors=100000,
                      solutionTolerance=1e-12, alternativeTolerance=1e-12, pcgTolerance=1e-15)
        Do = oc.get("D").reshape(-1, 3)
        err = float(np.linalg.norm(out["D"] - Do) / np.linalg.norm(Do))
        print(f"  step {n}: nOuter {out['nOuter']} (oracle {ro['nOuter']}) err D {err:.2e}", flush=True)
        assert out["converged"] and ro["converged"]
        assert err < (1e-8 if precision == 0 else 1e-6)
        levels = [out["D"]] + levels[:3]
    with pytest.raises(S4FError, match="constant time step"):
        m.solve(precision=precision, d2dt2=2, deltaT=dt, deltaT0=2 * dt)

"""Synthetic case generators for the benchmark configurations.

Box patches follow the generator's blockMesh order: 0 x-min, 1 x-max,
2 y-min, 3 y-max, 4 z-min, 5 z-max.
"""
import numpy as np

from . import _abi
from .solid import PolyMesh, SolverMesh

E_STEEL, NU_STEEL = 200e9, 0.3
FIX, TRAC = _abi.PATCH_FIXED_DISPLACEMENT, _abi.PATCH_TRACTION


def cantilever(ctx, nx=50, ny=10, nz=10, L=(1.0, 0.2, 0.2), traction=(0.0, -1e6, 0.0), E=E_STEEL, nu=NU_STEEL,
               jitter=0.0, seed=0, tets=False):
    """3-D cantilever: x-min fixedDisplacement 0, x-max uniform traction, other faces traction-free
    (50x10x10 of a 1.0x0.2x0.2 m box is the 5,000-cell small case)."""
    pm = PolyMesh.box(nx, ny, nz, L, jitter, seed, tets)
    types = [FIX, TRAC, TRAC, TRAC, TRAC, TRAC]
    m = SolverMesh(ctx, pm, types)
    m.set_material(E, nu)
    m.set_bc(0, (0.0, 0.0, 0.0))
    m.set_bc(1, traction)
    for p in range(2, 6):
        m.set_bc(p, (0.0, 0.0, 0.0))
    return pm, m


def cube_tension(ctx, n=200, jitter=0.15, seed=2, t=1e7, E=E_STEEL, nu=NU_STEEL, tets=False):
    """unit cube, interior vertices jittered +/- jitter*h: z-min fixed, z-max normal traction t,
    sides traction-free (200^3 = 8M cells is the large case)."""
    pm = PolyMesh.box(n, n, n, (1.0, 1.0, 1.0), jitter, seed, tets)
    types = [TRAC, TRAC, TRAC, TRAC, FIX, TRAC]
    m = SolverMesh(ctx, pm, types)
    m.set_material(E, nu)
    for p in range(4):
        m.set_bc(p, (0.0, 0.0, 0.0))
    m.set_bc(4, (0.0, 0.0, 0.0))
    m.set_bc(5, (0.0, 0.0, t))
    return pm, m


SMALL_OPTS = dict(pcgTolerance=1e-9, pcgRelTol=0.1, solutionTolerance=1e-7, alternativeTolerance=1e-7,
                  nCorrectors=500)


def box_sweep_specs(count, seed, nmin=8, nmax=64, jitter=0.1, h=0.01):
    """Specifications of the batched-sweep boxes: n_x, n_y, n_z uniform in
    nmin..nmax (uneven sizes), spacing h in every direction, interior
    vertices jittered +/- jitter*h, E log-normal (median 100 GPa, sigma 0.5),
    nu uniform 0.2-0.45, x-max traction of magnitude uniform 1e5-1e7 Pa in a
    uniformly random direction; x-min fixed, other faces traction-free."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = rng.integers(nmin, nmax + 1, size=3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        specs.append(dict(n=tuple(int(v) for v in n), L=tuple(float(v * h) for v in n), jitter=jitter,
                          seed=int(rng.integers(1 << 30)), E=float(100e9 * np.exp(0.5 * rng.normal())),
                          nu=float(rng.uniform(0.2, 0.45)), traction=tuple(float(v) for v in
                                                                            rng.uniform(1e5, 1e7) * d)))
    return specs


def box_case(ctx, spec, traction=None):
    """Mesh of one sweep box and its per-boundary-face values (patch order):
    returns (PolyMesh, SolverMesh, bc_values (nBoundaryFaces, 3))."""
    pm = PolyMesh.box(*spec["n"], spec["L"], spec["jitter"], spec["seed"], False)
    types = [FIX, TRAC, TRAC, TRAC, TRAC, TRAC]
    m = SolverMesh(ctx, pm, types)
    m.set_material(spec["E"], spec["nu"])
    t = spec["traction"] if traction is None else traction
    bv = np.zeros((pm.nBoundaryFaces, 3))
    bv[pm.patch_faces(1)] = t
    for p in range(6):
        m.set_bc(p, bv[pm.patch_faces(p)])
    return pm, m, bv

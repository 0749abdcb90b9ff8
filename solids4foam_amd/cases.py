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


def inclusions(seed, volume_fraction=0.1, rmin=0.02, rmax=0.06):
    """Spherical inclusions in the unit cube: centres uniform in the cube,
    radii uniform in [rmin, rmax], added until their summed volume reaches
    the volume fraction.  Returns (centres (k, 3), radii (k,))."""
    rng = np.random.default_rng(seed)
    cs, rs, vol = [], [], 0.0
    while vol < volume_fraction:
        r = rng.uniform(rmin, rmax)
        cs.append(rng.uniform(0.0, 1.0, size=3))
        rs.append(r)
        vol += 4.0 / 3.0 * np.pi * r ** 3
    return np.array(cs), np.array(rs)


def inside_spheres_box(C, n, centres, radii, per_hex=1):
    """mask of cell centres C inside any sphere, for cells of the generated
    unit-cube box (hex (i, j, k) -> cells per_hex*(i + n j + n^2 k) + 0..per_hex-1,
    centres within one spacing of their hex): each sphere tests only the
    cells of the hexes its bounding box (plus one hex) covers"""
    out = np.zeros(len(C), bool)
    h = 1.0 / n
    sub = np.arange(per_hex)
    for c, r in zip(centres, radii):
        lo = np.maximum(np.floor((c - r) / h).astype(int) - 1, 0)
        hi = np.minimum(np.ceil((c + r) / h).astype(int) + 1, n - 1)
        i = np.arange(lo[0], hi[0] + 1)
        j = np.arange(lo[1], hi[1] + 1)
        k = np.arange(lo[2], hi[2] + 1)
        hexes = (i[None, None, :] + n * j[None, :, None] + n * n * k[:, None, None]).ravel()
        cells = (per_hex * hexes[:, None] + sub[None, :]).ravel()
        d = C[cells] - c
        out[cells[np.einsum("ij,ij->i", d, d) <= r * r]] = True
    return out


def tet_inclusions(ctx, n=200, jitter=0.15, seed=2, inc_seed=3, E_matrix=200e9, E_inc=20e9, nu=0.3,
                   traction=(0.0, 5e6, 0.0), rmin=0.02, rmax=0.06, zones=False):
    """The jittered n^3 cube split into 6 tetrahedra per hex (48M cells at
    n = 200), per-cell E: E_inc for cells whose centroid lies in a seeded
    spherical inclusion (10 % volume fraction), E_matrix elsewhere; x-min
    fixed, x-max shear traction, other faces traction-free.  Tetrahedra use
    the full-delta least-squares gradient (extendedLeastSquares 0): with the
    normal-delta leastSquares the segregated loop diverges on tets.  zones:
    matrix and inclusions as two laws on their own sub-meshes (the
    reference's route to two materials, linearElastic's E being one value
    per law); else one per-cell law.  Returns (PolyMesh, SolverMesh,
    inclusion cell fraction, inclusion count, E per cell)."""
    pm = PolyMesh.box(n, n, n, (1.0, 1.0, 1.0), jitter, seed, True)
    types = [FIX, TRAC, TRAC, TRAC, TRAC, TRAC]
    m = SolverMesh(ctx, pm, types, ls_full_delta=True)
    cen, rad = inclusions(inc_seed, rmin=rmin, rmax=rmax)
    inc = inside_spheres_box(pm.cell_centres(), n, cen, rad, per_hex=6)
    E = np.where(inc, E_inc, E_matrix)
    m.set_material(E, nu)
    if zones:
        m.set_zones(inc.astype(np.int32))
    m.set_bc(0, (0.0, 0.0, 0.0))
    m.set_bc(1, traction)
    for p in range(2, 6):
        m.set_bc(p, (0.0, 0.0, 0.0))
    return pm, m, float(inc.mean()), len(rad), E


def serve_requests(count, seed, nmin=8, nmax=128, repeat=0.2, rate=100.0):
    """The serving request stream: `count` requests whose boxes come from
    box_sweep_specs (n up to nmax); after the first, each request repeats an
    earlier mesh with probability `repeat` (a new x-max traction of magnitude
    uniform 1e5-1e7 Pa in a random direction, warm started from the mesh's
    resident field), else brings the next new mesh with its own load; Poisson
    arrival times at `rate` per second.  Returns (specs of the unique meshes,
    [(mesh index, traction or None for the mesh's own, warm start)], arrival
    times in seconds)."""
    rng = np.random.default_rng(seed + 1000)
    n_unique = max(1, int(round(count * (1.0 - repeat))))
    specs = box_sweep_specs(n_unique, seed, nmin, nmax)
    reqs, seen = [], []
    u = 0
    for _ in range(count):
        if seen and (u >= n_unique or rng.random() < repeat):
            j = int(rng.choice(seen))
            d = rng.normal(size=3)
            reqs.append((j, tuple(float(v) for v in rng.uniform(1e5, 1e7) * d / np.linalg.norm(d)), True))
        else:
            reqs.append((u, None, False))
            seen.append(u)
            u += 1
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    return specs, reqs, arrivals

"""Host-side mirror of solids4foam's solidModel interface for the
``linearGeometryTotalDisplacement`` path, over the libs4f C ABI.

Reference interface (paths relative to the solids4foam tree):
  solidModel::New / run-time table       src/solids4FoamModels/solidModels/solidModel/solidModel.C:1673-1749
  linGeomTotalDispSolid (ctor, evolve)   src/solids4FoamModels/solidModels/linGeomTotalDispSolid/linGeomTotalDispSolid.C:64-232
  solidModel::writeFields (sigmaEq)      solidModel.C:1936-1996
Unknown model types raise ValueError listing the valid types, as
solidModel::New aborts with FatalError (solidModel.C:1729-1743).
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import check, f64, i32, lib, ptr


class Context:
    """One HIP device + stream (s4f_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().s4f_ctx_create(int(device), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().s4f_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class PolyMesh:
    """polyMesh addressing + OpenFOAM geometry (host)."""

    def __init__(self, handle):
        self.h = handle
        sz = np.zeros(6, np.int64)
        check(lib().s4f_polymesh_sizes(self.h, ptr(sz)))
        self.nPoints, self.nCells, self.nFaces, self.nInternalFaces, self.nPatches, self._nfv = map(int, sz)
        self.nBoundaryFaces = self.nFaces - self.nInternalFaces
        self._addr = None
        self._geom = None
        self._patch_sizes = None  # known without the addressing for generated boxes

    @classmethod
    def box(cls, nx, ny, nz, L=(1.0, 1.0, 1.0), jitter=0.0, seed=0, tets=False):
        h = ctypes.c_void_p()
        Lv = f64(L)
        check(lib().s4f_polymesh_box(int(nx), int(ny), int(nz), ptr(Lv), float(jitter), int(seed), int(bool(tets)),
                                     ctypes.byref(h)))
        pm = cls(h)
        # blockMesh patch order x-min, x-max, y-min, y-max, z-min, z-max; a
        # tet split cuts every boundary quad into two triangles
        k = 2 if tets else 1
        pm._patch_sizes = [k * ny * nz, k * ny * nz, k * nx * nz, k * nx * nz, k * nx * ny, k * nx * ny]
        assert sum(pm._patch_sizes) == pm.nBoundaryFaces
        return pm

    @classmethod
    def from_arrays(cls, points, face_offsets, face_verts, owner, neighbour, n_cells, patch_start, patch_size):
        pts, fo, fv, ow, nb = f64(points), i32(face_offsets), i32(face_verts), i32(owner), i32(neighbour)
        ps, pz = i32(patch_start), i32(patch_size)
        h = ctypes.c_void_p()
        check(lib().s4f_polymesh_from_arrays(len(pts), ptr(pts), int(n_cells), len(ow), len(nb), ptr(fo), ptr(fv),
                                             ptr(ow), ptr(nb), len(ps), ptr(ps), ptr(pz), ctypes.byref(h)))
        return cls(h)

    def addressing(self):
        if self._addr is None:
            a = dict(points=np.zeros((self.nPoints, 3)), fOff=np.zeros(self.nFaces + 1, np.int32),
                     fVerts=np.zeros(self._nfv, np.int32), owner=np.zeros(self.nFaces, np.int32),
                     neighbour=np.zeros(self.nInternalFaces, np.int32), patchStart=np.zeros(self.nPatches, np.int32),
                     patchSize=np.zeros(self.nPatches, np.int32))
            check(lib().s4f_polymesh_addressing(self.h, ptr(a["points"]), ptr(a["fOff"]), ptr(a["fVerts"]),
                                                ptr(a["owner"]), ptr(a["neighbour"]), ptr(a["patchStart"]),
                                                ptr(a["patchSize"])))
            self._addr = a
        return self._addr

    def geometry(self):
        if self._geom is None:
            g = dict(Cf=np.zeros((self.nFaces, 3)), Sf=np.zeros((self.nFaces, 3)), C=np.zeros((self.nCells, 3)),
                     V=np.zeros(self.nCells))
            check(lib().s4f_polymesh_geometry(self.h, ptr(g["Cf"]), ptr(g["Sf"]), ptr(g["C"]), ptr(g["V"])))
            self._geom = g
        return self._geom

    def point_cells(self):
        """pointCells as CSR (offsets (nPoints+1,), cells), ascending cell order"""
        off = np.zeros(self.nPoints + 1, np.int32)
        check(lib().s4f_polymesh_point_cells(self.h, ptr(off), None))
        cells = np.zeros(int(off[-1]), np.int32)
        check(lib().s4f_polymesh_point_cells(self.h, ptr(off), ptr(cells)))
        return off, cells

    def cell_centres(self):
        """cell centres only (nCells, 3), without the face geometry"""
        if self._geom is not None:
            return self._geom["C"]
        C = np.zeros((self.nCells, 3))
        check(lib().s4f_polymesh_geometry(self.h, None, None, ptr(C), None))
        return C

    def patch_faces(self, p):
        if self._patch_sizes is not None:
            s = sum(self._patch_sizes[:p])
            return slice(s, s + self._patch_sizes[p])
        a = self.addressing()
        s = int(a["patchStart"][p]) - self.nInternalFaces
        return slice(s, s + int(a["patchSize"][p]))

    def close(self):
        if getattr(self, "h", None):
            lib().s4f_polymesh_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


class SolverMesh:
    """Device-resident mesh (s4f_mesh): addressing, geometry, LS matrices, ELL lists."""

    def __init__(self, ctx, polymesh, patch_types, ls_full_delta=False):
        self.ctx = ctx
        self.poly = polymesh
        self.patch_types = i32(patch_types)
        if len(self.patch_types) != polymesh.nPatches:
            raise ValueError("one patch type per patch")
        h = ctypes.c_void_p()
        check(lib().s4f_mesh_create_from_polymesh(ctx.h, polymesh.h, ptr(self.patch_types), int(bool(ls_full_delta)),
                                                  ctypes.byref(h)))
        self.h = h
        self.nCells = polymesh.nCells
        self.nBoundaryFaces = polymesh.nBoundaryFaces

    def info(self):
        out = np.zeros(6, np.int64)
        check(lib().s4f_mesh_info(self.h, ptr(out)))
        return dict(nCells=int(out[0]), nInternalFaces=int(out[1]), nBoundaryFaces=int(out[2]),
                    ellEntries=int(out[3]), ellSlices=int(out[4]), deviceBytes=int(out[5]))

    def mg_info(self):
        """Multigrid hierarchy of the current matrix (after a GAMG solve):
        rows and stored off-diagonal entries per level, and the stored entries
        of the level-0 grouped operator offdiag(A) P"""
        n = ctypes.c_int32(0)
        rows = np.zeros(32, np.int64)
        ents = np.zeros(32, np.int64)
        q = ctypes.c_int64(0)
        check(lib().s4f_mesh_mg_info(self.h, 32, ctypes.byref(n), ptr(rows), ptr(ents), ctypes.byref(q)))
        k = min(n.value, 32)
        return dict(rows=[int(v) for v in rows[:k]], entries=[int(v) for v in ents[:k]], qEntries=int(q.value))

    def set_material(self, E, nu):
        E = np.atleast_1d(f64(E))
        nu = np.atleast_1d(f64(nu))
        per_cell = E.size > 1 or nu.size > 1
        if per_cell:
            E = f64(np.broadcast_to(E, (self.nCells,)))
            nu = f64(np.broadcast_to(nu, (self.nCells,)))
        check(lib().s4f_material_set(self.h, ptr(E), ptr(nu), int(per_cell)))

    def set_plane_stress(self, on=True):
        """mechanicalProperties planeStress: linearElastic's 2-D moduli, used
        by the next set_material"""
        check(lib().s4f_plane_stress_set(self.h, int(bool(on))))

    def set_density(self, rho):
        """mechanicalProperties rho: one value or one per cell"""
        r = np.atleast_1d(f64(rho))
        check(lib().s4f_density_set(self.h, ptr(r), int(r.size > 1)))

    def set_gravity(self, g):
        """constant/g: the body force rho g of the momentum equation"""
        check(lib().s4f_gravity_set(self.h, ptr(f64(g, (3,)))))

    def set_old_time(self, D_old, D_old_old=None):
        """D.oldTime() and D.oldTime().oldTime() for a transient (Euler
        d2dt2) step; the step starts from D_old"""
        a = f64(D_old, (self.nCells, 3))
        b = None if D_old_old is None else f64(D_old_old, (self.nCells, 3))
        check(lib().s4f_old_time_set(self.h, ptr(a), ptr(b)))

    def set_tiles(self, max_tiles, min_slices):
        """workgroup tiles: at most max_tiles tiles of at least min_slices 64-cell slices"""
        check(lib().s4f_mesh_tiles_set(self.h, int(max_tiles), int(min_slices)))

    def set_zones(self, zone):
        """Material zone of every cell of a multi-law case (one law per
        cellZone), or None for one material: each zone is solved on its own
        sub-mesh, with interface displacements on the faces between zones
        (mechanicalModel.C:571-615, solidSubMeshes.C:1550-2031)."""
        if zone is None:
            check(lib().s4f_mesh_zones_set(self.h, None, None))
            self.zones = None
            return
        z = i32(np.asarray(zone).reshape(self.nCells))
        Cf = f64(self.poly.geometry()["Cf"])
        check(lib().s4f_mesh_zones_set(self.h, ptr(z), ptr(Cf)))
        self.zones = z

    def set_old_times(self, levels):
        """D.oldTime(), D.oldTime().oldTime(), ... (1 to 4 fields of (nCells, 3))"""
        a = f64(np.stack([np.asarray(l, np.float64).reshape(self.nCells, 3) for l in levels]))
        check(lib().s4f_old_times_set(self.h, len(levels), ptr(a)))

    def set_bc(self, patch, value, pressure=None):
        sl = self.poly.patch_faces(patch)
        n = sl.stop - sl.start
        v = f64(np.broadcast_to(np.asarray(value, np.float64), (n, 3)))
        p = None if pressure is None else f64(np.broadcast_to(np.asarray(pressure, np.float64), (n,)))
        check(lib().s4f_bc_set(self.h, int(patch), ptr(v), ptr(p)))

    def solve(self, D0=None, want_sigma=True, want_grad=False, want_Db=False, **opts):
        o = _abi.default_opts(**opts)
        D = np.zeros((self.nCells, 3)) if D0 is None else f64(D0, (self.nCells, 3)).copy()
        if D0 is not None:
            o.warmStart = 1
        sig = np.zeros((self.nCells, 6)) if want_sigma else None
        G = np.zeros((self.nCells, 9)) if want_grad else None
        Db = np.zeros((self.nBoundaryFaces, 3)) if want_Db else None
        rep = _abi.SolveReport()
        st = lib().s4f_solve(self.h, ctypes.byref(o), ptr(D), ptr(sig), ptr(G), ptr(Db), ctypes.byref(rep))
        check(st, allow=(_abi.S4F_OK, _abi.S4F_WARN_MAXITER))
        out = dict(D=D, sigma=sig, gradD=G, Db=Db, nOuter=rep.nOuter, converged=bool(rep.converged),
                   innerTotal=int(rep.innerTotal), initialResidual=rep.initialResidual, residualvf=rep.residualvf,
                   seconds=rep.seconds, pcgIterations=int(rep.pcgIterations), mixedFallback=int(rep.mixedFallback),
                   status=st)
        return out

    def strain_fields(self):
        """epsilon (C, 6), epsilonEq (C,), sigmaEq (C,) of the last solve (solidModel::writeFields)"""
        eps = np.zeros((self.nCells, 6))
        ee = np.zeros(self.nCells)
        se = np.zeros(self.nCells)
        check(lib().s4f_strain_fields(self.h, ptr(eps), ptr(ee), ptr(se)))
        return dict(epsilon=eps, epsilonEq=ee, sigmaEq=se)

    def point_displacement(self):
        """pointD (nPoints, 3) of the last solve: inverse-distance weighted,
        gradient-extrapolated cell values (mechanicalModel::interpolate)"""
        pm = self.poly
        pts = pm.addressing()["points"]
        off, cells = pm.point_cells()
        out = np.zeros((pm.nPoints, 3))
        check(lib().s4f_point_displacement(self.h, pm.nPoints, ptr(f64(pts)), ptr(off), ptr(cells), ptr(out)))
        return out

    def history(self, cap=100000):
        ir, rv = np.zeros(cap), np.zeros(cap)
        inn = np.zeros(cap, np.int32)
        cnt = ctypes.c_int32()
        check(lib().s4f_solve_history(self.h, cap, ptr(ir), ptr(rv), ptr(inn), ctypes.byref(cnt)))
        n = cnt.value
        return dict(initRes=ir[:n], rvf=rv[:n], inner=inn[:n])

    # fine-grained operators ------------------------------------------------
    def grad(self, D, Db, snb):
        G = np.zeros((self.nCells, 9))
        Gb = np.zeros((self.nBoundaryFaces, 9))
        check(lib().s4f_grad(self.h, ptr(f64(D)), ptr(f64(Db)), ptr(f64(snb)), ptr(G), ptr(Gb)))
        return G, Gb

    def sigma(self, G):
        s = np.zeros((self.nCells, 6))
        check(lib().s4f_sigma(self.h, ptr(f64(G)), ptr(s)))
        return s

    def residual(self, D, G, sigmab):
        R = np.zeros((self.nCells, 3))
        check(lib().s4f_residual(self.h, ptr(f64(D)), ptr(f64(G)), ptr(f64(sigmab)), ptr(R)))
        return R

    def spmv3(self, x, mixed=False):
        y = np.zeros((self.nCells, 3))
        check(lib().s4f_spmv3(self.h, ptr(f64(x)), ptr(y), int(bool(mixed))))
        return y

    def pcg3(self, b, norm_factor, **opts):
        o = _abi.default_opts(**opts)
        x = np.zeros((self.nCells, 3))
        it = np.zeros(3, np.int32)
        nf = f64(norm_factor)
        check(lib().s4f_pcg3(self.h, ctypes.byref(o), ptr(f64(b)), ptr(nf), ptr(x), ptr(it)))
        return x, it

    def matrix(self):
        d = np.zeros(self.nCells)
        a = np.zeros(self.poly.nInternalFaces)
        check(lib().s4f_matrix(self.h, ptr(d), ptr(a)))
        return d, a

    def close(self):
        if getattr(self, "h", None):
            lib().s4f_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def sigma_eq(sigma):
    """von Mises sqrt(3/2 |dev sigma|^2) (solidModel.C:1957-1961); sigma (n,6) xx xy xz yy yz zz"""
    s = np.asarray(sigma)
    tr3 = (s[:, 0] + s[:, 3] + s[:, 5]) / 3.0
    d0, d3, d5 = s[:, 0] - tr3, s[:, 3] - tr3, s[:, 5] - tr3
    ms = d0 * d0 + d3 * d3 + d5 * d5 + 2.0 * (s[:, 1] ** 2 + s[:, 2] ** 2 + s[:, 4] ** 2)
    return np.sqrt(1.5 * ms)


class LinGeomTotalDispSolid:
    """solidModel ``linearGeometryTotalDisplacement`` on the MI355X path.

    Built from the same inputs a case gives the reference: the mesh, the
    linearElastic law (E, nu), per-patch boundary conditions (fixedDisplacement
    value / solidTraction traction + pressure) and the solidProperties
    coefficients; ``evolve()`` runs the momentum loop
    (linGeomTotalDispSolid.C:135-209) and leaves D, gradD and sigma populated.
    """

    typeName = "linearGeometryTotalDisplacement"

    def __init__(self, ctx, polymesh, patches, E, nu, coeffs=None, solver=None, ls_full_delta=False, rho=None,
                 g=None, d2dt2="steadyState", deltaT=None, predictor=False, plane_stress=False, zones=None):
        # rho: mechanicalProperties density (needed with g or a transient
        # d2dt2); g: constant/g; d2dt2: d2dt2Schemes default (steadyState |
        # Euler | backward); deltaT: the time step of evolve() when transient;
        # predictor: linGeomTotalDispSolid::predict (:47-59), transient only;
        # plane_stress: mechanicalProperties planeStress (2-D moduli);
        # zones: the law (cellZone) of every cell when there are several
        # laws, each solved on its own sub-mesh (None: one law)
        # patches: list of dicts {type: "fixedDisplacement"|"fixedDisplacementZeroShear"|"solidTraction"|
        # "solidSymmetry"|"empty", value|traction, pressure}
        kinds = {"fixedDisplacement": _abi.PATCH_FIXED_DISPLACEMENT, "solidTraction": _abi.PATCH_TRACTION,
                 "empty": _abi.PATCH_EMPTY, "solidSymmetry": _abi.PATCH_SYMMETRY,
                 "fixedDisplacementZeroShear": _abi.PATCH_ZERO_SHEAR}
        types = []
        for p in patches:
            t = p["type"]
            if t not in kinds:
                raise ValueError(f"Unsupported boundary condition type {t!r} "
                                 "(valid: fixedDisplacement, fixedDisplacementZeroShear, solidTraction, "
                                 "solidSymmetry, empty)")
            types.append(kinds[t])
        self.mesh = SolverMesh(ctx, polymesh, types, ls_full_delta)
        if plane_stress:
            self.mesh.set_plane_stress(True)
        self.mesh.set_material(E, nu)
        if zones is not None:
            self.mesh.set_zones(zones)
        for i, p in enumerate(patches):
            if p["type"] in ("fixedDisplacement", "fixedDisplacementZeroShear"):
                self.mesh.set_bc(i, p.get("value", (0.0, 0.0, 0.0)))
            elif p["type"] == "solidTraction":
                self.mesh.set_bc(i, p.get("traction", (0.0, 0.0, 0.0)), p.get("pressure"))
        if d2dt2 not in ("steadyState", "Euler", "backward"):
            raise ValueError(f"d2dt2 scheme {d2dt2!r} is not on the device path (steadyState, Euler, backward)")
        if rho is not None:
            self.mesh.set_density(rho)
        if g is not None:
            self.mesh.set_gravity(g)
        self.transient = d2dt2 in ("Euler", "backward")
        self.scheme = d2dt2
        if self.transient and not (deltaT and deltaT > 0):
            raise ValueError("a transient model needs deltaT > 0")
        if predictor and not self.transient:
            raise ValueError("the predictor needs a transient d2dt2 scheme")
        self.predictor = predictor
        self.deltaT = deltaT
        self.deltaT0 = deltaT
        self.time = 0.0
        self.opts = dict(coeffs or {})
        self.opts.update(solver or {})
        self.D = np.zeros((polymesh.nCells, 3))
        self.D_old_old = self.D.copy()
        # backward d2dt2: D.oldTime().oldTime() and the two levels below it
        self.D_olds = [self.D.copy(), self.D.copy(), self.D.copy()]
        self.U = np.zeros((polymesh.nCells, 3))  # fvc::ddt(D) after each step (Euler)
        self.sigma = None
        self.gradD = None
        self.report = None

    def evolve(self, warm_start=False):
        """one steady solve, or one time step of deltaT when transient (the
        converged field becomes D.oldTime() of the next step)"""
        opts = dict(self.opts)
        if self.transient:
            if self.scheme == "backward":
                self.mesh.set_old_times([self.D] + self.D_olds)
                opts.update(d2dt2=_abi.D2DT2_BACKWARD, deltaT=self.deltaT, deltaT0=self.deltaT)
            else:
                self.mesh.set_old_time(self.D, self.D_old_old)
                opts.update(d2dt2=_abi.D2DT2_EULER, deltaT=self.deltaT, deltaT0=self.deltaT0)
            # predict(): D = D.oldTime() + U deltaT (gradD and sigma follow D)
            D0 = self.D + self.U * self.deltaT if self.predictor else None
            out = self.mesh.solve(D0=D0, want_sigma=True, want_grad=True, **opts)
            self.D_olds = [self.D] + self.D_olds[:2]
            self.D_old_old = self.D
            self.deltaT0 = self.deltaT
            self.time += self.deltaT
            self.U = (out["D"] - self.D_old_old) / self.deltaT  # U = fvc::ddt(D), Euler
        else:
            out = self.mesh.solve(D0=self.D if warm_start else None, want_sigma=True, want_grad=True, **opts)
        self.D, self.sigma, self.gradD = out["D"], out["sigma"], out["gradD"]
        self.report = {k: v for k, v in out.items() if k not in ("D", "sigma", "gradD", "Db")}
        return True

    def sigma_eq(self):
        return sigma_eq(self.sigma)

    def write_fields(self):
        """solidModel::writeFields (solidModel.C:1936-1962) plus pointD
        (linGeomTotalDispSolid.C:212-214): epsilon, epsilonEq, sigmaEq and
        the point displacement of the last evolve(), computed on the device"""
        f = self.mesh.strain_fields()
        f["pointD"] = self.mesh.point_displacement()
        return f


SOLID_MODELS = {LinGeomTotalDispSolid.typeName: LinGeomTotalDispSolid}


def solid_model_new(solid_model_type, *args, **kw):
    """Run-time selection by the solidProperties ``solidModel`` entry."""
    try:
        cls = SOLID_MODELS[solid_model_type]
    except KeyError:
        raise ValueError(f"Unknown solidModel type {solid_model_type!r}; valid solidModel types are: "
                         f"{sorted(SOLID_MODELS)}") from None
    return cls(*args, **kw)


class LduSolver:
    """PCG on a caller's symmetric LDU matrix (OpenFOAM lduMatrix addressing):
    the lduMatrix::solver seam (s4f_ldu_* in include/s4f_abi.h)."""

    def __init__(self, ctx, n_cells, lower_addr, upper_addr):
        self.ctx = ctx
        self.nCells = int(n_cells)
        lo, up = i32(lower_addr), i32(upper_addr)
        h = ctypes.c_void_p()
        check(lib().s4f_ldu_create(ctx.h, self.nCells, len(lo), ptr(lo), ptr(up), ctypes.byref(h)))
        self.h = h

    def solve(self, diag, upper, source, psi, tolerance=1e-9, rel_tol=0.0, max_iter=1000, mixed=False):
        psi = f64(psi).copy()
        ir, fr = ctypes.c_double(), ctypes.c_double()
        it = ctypes.c_int32()
        check(lib().s4f_ldu_solve(self.h, ptr(f64(diag)), ptr(f64(upper)), ptr(f64(source)), ptr(psi),
                                  float(tolerance), float(rel_tol), int(max_iter), int(bool(mixed)),
                                  ctypes.byref(ir), ctypes.byref(fr), ctypes.byref(it)))
        return psi, dict(initialResidual=ir.value, finalResidual=fr.value, nIterations=it.value)

    def close(self):
        if getattr(self, "h", None):
            lib().s4f_ldu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

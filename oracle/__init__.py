"""ctypes front-end of the FP64 CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, as the checker.  The product
package ``solids4foam_amd`` never imports this module.

The C sources restate solids4foam's ``linGeomTotalDispSolid`` outer loop
(``src/solids4FoamModels/solidModels/linGeomTotalDispSolid/linGeomTotalDispSolid.C:111-271``)
and the OpenFOAM finite-volume operators it calls; see the header of
``s4f_oracle.c`` for the full list of restated files.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FIXED, TRACTION, EMPTY, SYMMETRY, ZERO_SHEAR = 0, 1, 2, 3, 4
PRECOND_DIC, PRECOND_DIAGONAL, PRECOND_NONE = 0, 1, 2


def build(force=False):
    """Compile the oracle with its Makefile (gcc): when forced, missing, or
    older than its source."""
    src = os.path.join(_HERE, "s4f_oracle.c")
    stale = os.path.exists(_LIB_PATH) and os.path.getmtime(src) > os.path.getmtime(_LIB_PATH)
    if force or stale or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.or_box_mesh.restype = vp
        L.or_box_mesh.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [ctypes.c_uint64, ctypes.c_int]
        L.or_mesh_from_arrays.restype = vp
        L.or_mesh_from_arrays.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ip, ip, ip, ip, ctypes.c_int, ip, ip]
        L.or_mesh_free.argtypes = [vp]
        L.or_mesh_info.argtypes = [vp, ip]
        L.or_mesh_get.argtypes = [vp, dp, ip, ip, ip, ip, ip, ip]
        L.or_mesh_geometry.argtypes = [vp, dp, dp, dp, dp]
        L.or_case_create.restype = vp
        L.or_case_create.argtypes = [vp, dp, dp, ip, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int]
        L.or_case_free.argtypes = [vp]
        L.or_point_interpolate.argtypes = [vp, dp, dp, dp]
        L.or_field_size.restype = ctypes.c_long
        L.or_field_size.argtypes = [vp, ctypes.c_char_p]
        L.or_get_field.argtypes = [vp, ctypes.c_char_p, dp]
        L.or_set_field.argtypes = [vp, ctypes.c_char_p, dp]
        for name in ("or_grad", "or_sigma", "or_update_traction", "or_correct_boundary", "or_residual"):
            getattr(L, name).argtypes = [vp]
        L.or_init.argtypes = [vp, dp]
        L.or_iterate.restype = ctypes.c_int
        L.or_iterate.argtypes = [vp, vp, dp, dp, ip]
        L.or_solve.restype = ctypes.c_int
        L.or_solve.argtypes = [vp, vp, vp]
        L.or_time_iterations.restype = ctypes.c_double
        L.or_time_iterations.argtypes = [vp, vp, ctypes.c_int]
        L.or_advance_time.argtypes = [vp]
        L.or_omp_create.restype = vp
        L.or_omp_create.argtypes = [vp, ctypes.c_int]
        L.or_omp_free.argtypes = [vp]
        L.or_bump_icorr.argtypes = [vp]
        L.or_iterate_omp.restype = ctypes.c_int
        L.or_iterate_omp.argtypes = [vp, vp, vp, dp, dp, ip]
        L.or_time_iterations_omp.restype = ctypes.c_double
        L.or_time_iterations_omp.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
        L.or_reference_source.argtypes = [vp, ctypes.c_int, dp]
        L.or_amul.argtypes = [vp, dp, dp]
        L.or_sigma_eq.argtypes = [vp, dp]
        L.or_iface_sigma.argtypes = [vp]
        L.or_iface_update.argtypes = [vp]
        L.or_case_set_zones.restype = ctypes.c_int
        L.or_case_set_zones.argtypes = [vp, ip]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Opts(ctypes.Structure):
    _fields_ = [("nCorr", ctypes.c_int), ("solTol", ctypes.c_double), ("altTol", ctypes.c_double),
                ("precond", ctypes.c_int), ("pcgTol", ctypes.c_double), ("pcgRelTol", ctypes.c_double),
                ("pcgMaxIter", ctypes.c_int), ("rcScale", ctypes.c_double), ("interfaceScale", ctypes.c_double),
                ("historyCap", ctypes.c_int), ("histInitRes", ctypes.POINTER(ctypes.c_double)),
                ("histRvf", ctypes.POINTER(ctypes.c_double)), ("histInner", ctypes.POINTER(ctypes.c_int)),
                ("relaxD", ctypes.c_double), ("relaxAitken", ctypes.c_int), ("d2dt2", ctypes.c_int),
                ("deltaT", ctypes.c_double), ("deltaT0", ctypes.c_double), ("g", ctypes.c_double * 3)]


class Report(ctypes.Structure):
    _fields_ = [("nOuter", ctypes.c_int), ("converged", ctypes.c_int), ("innerTotal", ctypes.c_int),
                ("maxIterReached", ctypes.c_int), ("initRes", ctypes.c_double), ("rvf", ctypes.c_double)]


class Mesh:
    """polyMesh arrays + OpenFOAM geometry, built by the oracle."""

    def __init__(self, handle):
        self.h = handle
        info = np.zeros(6, np.int32)
        lib().or_mesh_info(self.h, _ip(info))
        self.nPoints, self.nCells, self.nFaces, self.nInternalFaces, self.nPatches, nfv = map(int, info)
        self.points = np.zeros((self.nPoints, 3))
        self.fOff = np.zeros(self.nFaces + 1, np.int32)
        self.fVerts = np.zeros(nfv, np.int32)
        self.owner = np.zeros(self.nFaces, np.int32)
        self.neighbour = np.zeros(self.nInternalFaces, np.int32)
        self.patchStart = np.zeros(self.nPatches, np.int32)
        self.patchSize = np.zeros(self.nPatches, np.int32)
        lib().or_mesh_get(self.h, _dp(self.points), _ip(self.fOff), _ip(self.fVerts), _ip(self.owner),
                          _ip(self.neighbour), _ip(self.patchStart), _ip(self.patchSize))
        self.Cf = np.zeros((self.nFaces, 3))
        self.Sf = np.zeros((self.nFaces, 3))
        self.C = np.zeros((self.nCells, 3))
        self.V = np.zeros(self.nCells)
        lib().or_mesh_geometry(self.h, _dp(self.Cf), _dp(self.Sf), _dp(self.C), _dp(self.V))
        self.nBoundaryFaces = self.nFaces - self.nInternalFaces

    @classmethod
    def box(cls, nx, ny, nz, L=(1.0, 1.0, 1.0), jitter=0.0, seed=0, tets=False):
        h = lib().or_box_mesh(nx, ny, nz, float(L[0]), float(L[1]), float(L[2]), float(jitter), int(seed), int(tets))
        if not h:
            raise ValueError("invalid box mesh")
        return cls(h)

    @classmethod
    def from_arrays(cls, points, fOff, fVerts, owner, neighbour, nCells, patchStart, patchSize):
        points = np.ascontiguousarray(points, np.float64)
        fOff = np.ascontiguousarray(fOff, np.int32)
        fVerts = np.ascontiguousarray(fVerts, np.int32)
        owner = np.ascontiguousarray(owner, np.int32)
        neighbour = np.ascontiguousarray(neighbour, np.int32)
        ps = np.ascontiguousarray(patchStart, np.int32)
        pz = np.ascontiguousarray(patchSize, np.int32)
        h = lib().or_mesh_from_arrays(len(points), _dp(points), int(nCells), len(owner), len(neighbour),
                                      _ip(fOff), _ip(fVerts), _ip(owner), _ip(neighbour), len(ps), _ip(ps), _ip(pz))
        return cls(h)

    def patch_faces(self, p):
        """boundary-face (0-based within the boundary) slice of patch p"""
        s = int(self.patchStart[p]) - self.nInternalFaces
        return slice(s, s + int(self.patchSize[p]))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mesh_free(self.h)
            self.h = None


def default_opts(**kw):
    o = dict(nCorr=10000, solTol=1e-6, altTol=1e-7, precond=PRECOND_DIC, pcgTol=1e-9, pcgRelTol=0.1,
             pcgMaxIter=1000, rcScale=0.1, interfaceScale=0.01, relaxD=0.0, relaxAitken=0, d2dt2=0,
             deltaT=0.0, deltaT0=0.0, g=(0.0, 0.0, 0.0))
    o.update(kw)
    return o


class Case:
    """A solid case: mesh + linearElastic material + per-patch BCs + state."""

    def __init__(self, mesh, E, nu, patch_types, b_value, b_pressure=None, ls_full_delta=False,
                 rc_scale=0.1, interface_scale=0.01, plane_stress=False):
        self.mesh = mesh
        C, B = mesh.nCells, mesh.nBoundaryFaces
        self.E = np.ascontiguousarray(np.broadcast_to(np.asarray(E, np.float64), (C,)))
        self.nu = np.ascontiguousarray(np.broadcast_to(np.asarray(nu, np.float64), (C,)))
        pt = np.ascontiguousarray(patch_types, np.int32)
        bv = np.ascontiguousarray(np.asarray(b_value, np.float64).reshape(B, 3))
        bp = np.zeros(B) if b_pressure is None else np.ascontiguousarray(b_pressure, np.float64)
        self.h = lib().or_case_create(mesh.h, _dp(self.E), _dp(self.nu), _ip(pt), _dp(bv), _dp(bp),
                                      int(ls_full_delta), float(rc_scale), float(interface_scale),
                                      int(bool(plane_stress)))
        if not self.h:
            raise ValueError("linearElastic needs E > 0 and -1 < nu < 0.5 (nu = 0.5 needs the pressure equation)")
        self.init()

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_case_free(self.h)
            self.h = None

    def get(self, name):
        n = lib().or_field_size(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        lib().or_get_field(self.h, name.encode(), _dp(out))
        return out

    def set(self, name, value):
        n = lib().or_field_size(self.h, name.encode())
        a = np.ascontiguousarray(np.asarray(value, np.float64).reshape(-1))
        assert a.size == n, (name, a.size, n)
        lib().or_set_field(self.h, name.encode(), _dp(a))

    def set_zones(self, zone):
        """Material zone per cell of a multi-law case (None: one material):
        each zone is solved on its own sub-mesh with interface displacements
        between them (solidSubMeshes).  Resets the interface state; returns
        the number of interface faces."""
        if zone is None:
            self._zone = None
            return lib().or_case_set_zones(self.h, None)
        self._zone = np.ascontiguousarray(np.asarray(zone, np.int32).reshape(self.mesh.nCells))
        return lib().or_case_set_zones(self.h, _ip(self._zone))

    def iface_step(self, name):
        """one interface step on the current state: "sigma" (the sub-meshes'
        interface stress) or "update" (D_I from D.prevIter() to D)"""
        getattr(lib(), "or_iface_" + name)(self.h)

    def init(self, D0=None):
        if D0 is None:
            lib().or_init(self.h, None)
        else:
            d = np.ascontiguousarray(np.asarray(D0, np.float64).reshape(-1))
            lib().or_init(self.h, _dp(d))

    def op(self, name):
        getattr(lib(), "or_" + name)(self.h)

    @staticmethod
    def _opts(o, hist=None):
        st = Opts()
        for k, v in o.items():
            setattr(st, k, v)
        if hist is not None:
            st.historyCap = len(hist[0])
            st.histInitRes = _dp(hist[0])
            st.histRvf = _dp(hist[1])
            st.histInner = _ip(hist[2])
        return st

    def iterate(self, **kw):
        o = self._opts(default_opts(**kw))
        ir, rv, inn = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        conv = lib().or_iterate(self.h, ctypes.byref(o), ctypes.byref(ir), ctypes.byref(rv), ctypes.byref(inn))
        return conv, ir.value, rv.value, inn.value

    def solve(self, history=0, **kw):
        hist = None
        if history:
            hist = (np.zeros(history), np.zeros(history), np.zeros(history, np.int32))
        o = self._opts(default_opts(**kw), hist)
        rep = Report()
        lib().or_solve(self.h, ctypes.byref(o), ctypes.byref(rep))
        out = dict(nOuter=rep.nOuter, converged=bool(rep.converged), innerTotal=rep.innerTotal,
                   maxIterReached=bool(rep.maxIterReached), initRes=rep.initRes, rvf=rep.rvf)
        if hist is not None:
            n = min(rep.nOuter, history)
            out["history"] = dict(initRes=hist[0][:n].copy(), rvf=hist[1][:n].copy(), inner=hist[2][:n].copy())
        return out

    def advance_time(self):
        """new time step: D -> D.oldTime() -> D.oldTime().oldTime()"""
        lib().or_advance_time(self.h)

    def time_iterations(self, n, **kw):
        o = self._opts(default_opts(**kw))
        return lib().or_time_iterations(self.h, ctypes.byref(o), int(n))

    def solve_omp(self, threads, **kw):
        """The outer loop on `threads` OpenMP threads (or_iterate_omp: block
        DIC per thread, the reference's four gradient evaluations per
        iteration) until solidModel::converged or nCorr"""
        o = self._opts(default_opts(**kw))
        p = lib().or_omp_create(self.h, int(threads))
        if not p:
            raise ValueError("the OpenMP path covers fixedDisplacement, solidTraction and empty patches")
        ir, rv, inn = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        n, inner, conv = 0, 0, 0
        try:
            while True:
                conv = lib().or_iterate_omp(self.h, p, ctypes.byref(o), ctypes.byref(ir), ctypes.byref(rv),
                                            ctypes.byref(inn))
                if conv < 0:
                    raise ValueError("options not covered by the OpenMP path")
                n += 1
                inner += inn.value
                if conv or n >= o.nCorr or not np.isfinite(ir.value):
                    break
                self._bump()
        finally:
            lib().or_omp_free(p)
        return dict(nOuter=n, converged=bool(conv), innerTotal=inner, initRes=ir.value, rvf=rv.value)

    def _bump(self):
        lib().or_bump_icorr(self.h)

    def time_iterations_omp(self, n, threads, **kw):
        """wall seconds of n outer iterations on `threads` OpenMP threads"""
        o = self._opts(default_opts(**kw))
        return lib().or_time_iterations_omp(self.h, ctypes.byref(o), int(n), int(threads))

    def point_interpolate(self, D, G):
        """pointD from cell D (C, 3) and gradD (C, 9)"""
        D = np.ascontiguousarray(D, np.float64).reshape(-1)
        G = np.ascontiguousarray(G, np.float64).reshape(-1)
        out = np.zeros(self.mesh.nPoints * 3)
        lib().or_point_interpolate(self.h, _dp(D), _dp(G), _dp(out))
        return out.reshape(-1, 3)

    def reference_source(self, q):
        out = np.zeros(self.mesh.nCells)
        lib().or_reference_source(self.h, int(q), _dp(out))
        return out

    def amul(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        lib().or_amul(self.h, _dp(x), _dp(y))
        return y

    def sigma_eq(self):
        out = np.zeros(self.mesh.nCells)
        lib().or_sigma_eq(self.h, _dp(out))
        return out

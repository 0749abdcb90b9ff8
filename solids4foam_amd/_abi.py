"""ctypes bindings of libs4f (include/s4f_abi.h).

Loads the in-tree ``solids4foam_amd/libs4f.so`` and nothing else: there is no
CPU fallback.  A missing or unloadable library raises immediately.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libs4f.so")

S4F_OK = 0
S4F_ERR_INVALID = 1
S4F_ERR_HIP = 2
S4F_ERR_OOM = 3
S4F_ERR_DIVERGED = 4
S4F_WARN_MAXITER = 5
S4F_ERR_ADMISSION = 6
S4F_ERR_NOT_READY = 7

PATCH_FIXED_DISPLACEMENT = 0
PATCH_TRACTION = 1
PATCH_EMPTY = 2
PATCH_SYMMETRY = 3
PATCH_ZERO_SHEAR = 4
RELAX_FIXED, RELAX_AITKEN = 0, 1
D2DT2_STEADY, D2DT2_EULER, D2DT2_BACKWARD = 0, 1, 2
PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECOND_DIAGONAL = 0
PRECOND_GAMG = 2
PRECOND_GAMG_SMOOTHED = 3

_STATUS_NAMES = {0: "OK", 1: "ERR_INVALID", 2: "ERR_HIP", 3: "ERR_OOM", 4: "ERR_DIVERGED", 5: "WARN_MAXITER",
                 6: "ERR_ADMISSION", 7: "ERR_NOT_READY"}


class S4FError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class SolveOpts(ctypes.Structure):
    _fields_ = [("nCorrectors", ctypes.c_int32), ("solutionTolerance", ctypes.c_double),
                ("alternativeTolerance", ctypes.c_double), ("rhieChowScale", ctypes.c_double),
                ("interfaceScale", ctypes.c_double), ("pcgTolerance", ctypes.c_double),
                ("pcgRelTol", ctypes.c_double), ("pcgMaxIter", ctypes.c_int32), ("preconditioner", ctypes.c_int32),
                ("eagerLaunch", ctypes.c_int32), ("precision", ctypes.c_int32), ("warmStart", ctypes.c_int32),
                ("mgCoarseScale", ctypes.c_double), ("relaxD", ctypes.c_double),
                ("relaxationMethod", ctypes.c_int32), ("mixedStallWindow", ctypes.c_int32),
                ("d2dt2", ctypes.c_int32), ("progressSeconds", ctypes.c_int32), ("deltaT", ctypes.c_double),
                ("deltaT0", ctypes.c_double)]


class SolveReport(ctypes.Structure):
    _fields_ = [("nOuter", ctypes.c_int32), ("converged", ctypes.c_int32), ("innerTotal", ctypes.c_int64),
                ("initialResidual", ctypes.c_double), ("residualvf", ctypes.c_double), ("seconds", ctypes.c_double),
                ("mixedFallback", ctypes.c_int32), ("warmStarted", ctypes.c_int32), ("pcgIterations", ctypes.c_int64)]


class MeshDesc(ctypes.Structure):
    _fields_ = [("nCells", ctypes.c_int32), ("nFaces", ctypes.c_int32), ("nInternalFaces", ctypes.c_int32),
                ("owner", ctypes.c_void_p), ("neighbour", ctypes.c_void_p), ("C", ctypes.c_void_p),
                ("V", ctypes.c_void_p), ("Sf", ctypes.c_void_p), ("Cf", ctypes.c_void_p),
                ("nPatches", ctypes.c_int32), ("patchStart", ctypes.c_void_p), ("patchSize", ctypes.c_void_p),
                ("patchType", ctypes.c_void_p), ("lsFullDelta", ctypes.c_int32)]


_lib = None

# every symbol include/s4f_abi.h declares
EXPORTED = [
    "s4f_ctx_create", "s4f_ctx_destroy", "s4f_last_error", "s4f_version",
    "s4f_polymesh_box", "s4f_polymesh_from_arrays", "s4f_polymesh_free", "s4f_polymesh_sizes",
    "s4f_polymesh_addressing", "s4f_polymesh_geometry",
    "s4f_mesh_create", "s4f_mesh_create_from_polymesh", "s4f_mesh_destroy", "s4f_mesh_info",
    "s4f_material_set", "s4f_bc_set", "s4f_solve_opts_default", "s4f_solve", "s4f_solve_history",
    "s4f_grad", "s4f_sigma", "s4f_residual", "s4f_spmv3", "s4f_pcg3", "s4f_matrix",
    "s4f_batch_create", "s4f_batch_add", "s4f_batch_solve", "s4f_batch_destroy",
    "s4f_kernel_time", "s4f_calibrate_copy",
    "s4f_server_create", "s4f_server_submit", "s4f_server_poll", "s4f_server_result", "s4f_server_stats",
    "s4f_server_destroy", "s4f_ldu_create", "s4f_ldu_solve", "s4f_ldu_destroy",
    "s4f_polymesh_point_cells", "s4f_strain_fields", "s4f_point_displacement",
    "s4f_density_set", "s4f_gravity_set", "s4f_old_time_set", "s4f_mesh_mg_info", "s4f_plane_stress_set",
    "s4f_pcg3_noalloc", "s4f_old_times_set", "s4f_mesh_tiles_set", "s4f_mesh_zones_set",
]


class Request(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_void_p), ("bcValue", ctypes.c_void_p), ("bcPressure", ctypes.c_void_p),
                ("warmStart", ctypes.c_int32)]

KERNELS = {"residual": 0, "spmv": 1, "pcg_update": 2, "finalize": 3, "grad": 4}
MG_KERNELS = {"spmv": 8, "mg0_down": 5, "mg0_up": 6, "vcycle": 7}  # spmv: the multigrid PCG's variant


def lib():
    """The loaded libs4f (raises if it is missing: no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                              "(hipcc --offload-arch=gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p
        st = ctypes.c_int
        L.s4f_last_error.restype = ctypes.c_char_p
        L.s4f_version.restype = ctypes.c_char_p
        L.s4f_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.s4f_ctx_destroy.argtypes = [vp]
        L.s4f_polymesh_box.argtypes = [ctypes.c_int] * 3 + [dp, ctypes.c_double, ctypes.c_uint64, ctypes.c_int,
                                                            ctypes.POINTER(vp)]
        L.s4f_polymesh_from_arrays.argtypes = [ctypes.c_int32, dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ip, ip, ip, ip, ctypes.c_int32, ip, ip, ctypes.POINTER(vp)]
        L.s4f_polymesh_free.argtypes = [vp]
        L.s4f_polymesh_sizes.argtypes = [vp, vp]
        L.s4f_polymesh_addressing.argtypes = [vp, dp, ip, ip, ip, ip, ip, ip]
        L.s4f_polymesh_geometry.argtypes = [vp, dp, dp, dp, dp]
        L.s4f_mesh_create.argtypes = [vp, ctypes.POINTER(MeshDesc), ctypes.POINTER(vp)]
        L.s4f_mesh_create_from_polymesh.argtypes = [vp, vp, ip, ctypes.c_int32, ctypes.POINTER(vp)]
        L.s4f_mesh_destroy.argtypes = [vp]
        L.s4f_mesh_info.argtypes = [vp, vp]
        L.s4f_mesh_mg_info.argtypes = [vp, ctypes.c_int32, vp, vp, vp, vp]
        L.s4f_material_set.argtypes = [vp, dp, dp, ctypes.c_int32]
        L.s4f_bc_set.argtypes = [vp, ctypes.c_int32, dp, dp]
        L.s4f_density_set.argtypes = [vp, dp, ctypes.c_int32]
        L.s4f_plane_stress_set.argtypes = [vp, ctypes.c_int32]
        L.s4f_gravity_set.argtypes = [vp, dp]
        L.s4f_old_time_set.argtypes = [vp, dp, dp]
        L.s4f_old_times_set.argtypes = [vp, ctypes.c_int32, dp]
        L.s4f_mesh_tiles_set.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
        L.s4f_mesh_zones_set.argtypes = [vp, vp, dp]
        L.s4f_solve_opts_default.argtypes = [ctypes.POINTER(SolveOpts)]
        L.s4f_solve_opts_default.restype = None
        L.s4f_solve.argtypes = [vp, ctypes.POINTER(SolveOpts), dp, dp, dp, dp, ctypes.POINTER(SolveReport)]
        L.s4f_solve_history.argtypes = [vp, ctypes.c_int32, dp, dp, ip, ip]
        L.s4f_grad.argtypes = [vp, dp, dp, dp, dp, dp]
        L.s4f_sigma.argtypes = [vp, dp, dp]
        L.s4f_residual.argtypes = [vp, dp, dp, dp, dp]
        L.s4f_spmv3.argtypes = [vp, dp, dp, ctypes.c_int32]
        L.s4f_pcg3.argtypes = [vp, ctypes.POINTER(SolveOpts), dp, dp, dp, ip]
        L.s4f_pcg3_noalloc.argtypes = [vp, ctypes.POINTER(SolveOpts), vp, dp, vp, ip]
        L.s4f_matrix.argtypes = [vp, dp, dp]
        L.s4f_batch_create.argtypes = [vp, ctypes.POINTER(vp)]
        L.s4f_batch_add.argtypes = [vp, vp, ip]
        L.s4f_batch_solve.argtypes = [vp, ctypes.POINTER(SolveOpts), vp, vp, ctypes.POINTER(SolveReport)]
        L.s4f_batch_destroy.argtypes = [vp]
        L.s4f_kernel_time.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp]
        L.s4f_calibrate_copy.argtypes = [vp, ctypes.c_int64, vp]
        L.s4f_server_create.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(SolveOpts),
                                        ctypes.POINTER(vp)]
        L.s4f_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_int64)]
        L.s4f_server_poll.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]
        L.s4f_server_result.argtypes = [vp, ctypes.c_int64, dp, dp, ctypes.POINTER(SolveReport), dp]
        L.s4f_server_stats.argtypes = [vp, vp]
        L.s4f_server_destroy.argtypes = [vp]
        L.s4f_ldu_create.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ip, ip, ctypes.POINTER(vp)]
        L.s4f_ldu_solve.argtypes = [vp, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                    ctypes.c_int32, vp, vp, vp]
        L.s4f_ldu_destroy.argtypes = [vp]
        for name in EXPORTED:
            f = getattr(L, name)
            if name not in ("s4f_last_error", "s4f_version", "s4f_solve_opts_default"):
                f.restype = st
        _lib = L
    return _lib


def check(status, allow=(S4F_OK,)):
    if status not in allow:
        raise S4FError(status, lib().s4f_last_error().decode(errors="replace"))
    return status


def ptr(a):
    """data pointer of a C-contiguous numpy array (None passes NULL)"""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def default_opts(**kw):
    o = SolveOpts()
    lib().s4f_solve_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o

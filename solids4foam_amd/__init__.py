"""MI355X implementation of solids4foam's segregated small-strain solid solver
(solidModel ``linearGeometryTotalDisplacement``).

The compute path is libs4f.so (hand-written HIP for gfx950 behind the C ABI in
include/s4f_abi.h); this package is its host-side mirror of the reference's
solidModel interface.  There is no CPU fallback: importing the solver classes
loads libs4f and fails loudly if it is missing.
"""
from ._abi import (PATCH_EMPTY, PATCH_FIXED_DISPLACEMENT, PATCH_SYMMETRY, PATCH_TRACTION, PATCH_ZERO_SHEAR,
                   PRECISION_FP64, PRECISION_MIXED, PRECOND_DIAGONAL, PRECOND_GAMG, PRECOND_GAMG_SMOOTHED, S4FError,
                   lib)
from .solid import (Context, PolyMesh, SolverMesh, LinGeomTotalDispSolid, SOLID_MODELS, solid_model_new,
                    sigma_eq)

__all__ = ["PATCH_EMPTY", "PATCH_FIXED_DISPLACEMENT", "PATCH_SYMMETRY", "PATCH_TRACTION", "PATCH_ZERO_SHEAR",
           "PRECISION_FP64", "PRECISION_MIXED", "PRECOND_DIAGONAL", "PRECOND_GAMG", "PRECOND_GAMG_SMOOTHED",
           "S4FError", "lib", "Context",
           "PolyMesh", "SolverMesh", "LinGeomTotalDispSolid", "SOLID_MODELS", "solid_model_new", "sigma_eq"]

"""Case builders shared by the tests (oracle and product sides)."""
import numpy as np

E_STEEL, NU_STEEL = 200e9, 0.3
# patch order of the box generator: x-min, x-max, y-min, y-max, z-min, z-max
FIXED, TRACTION = 0, 1


def cantilever_bc(mesh_nb, patch_slices, traction=(0.0, -1e6, 0.0)):
    """cantilever box: x-min fixed (0), x-max uniform traction, others traction-free"""
    bv = np.zeros((mesh_nb, 3))
    bv[patch_slices[1]] = traction
    return [FIXED, TRACTION, TRACTION, TRACTION, TRACTION, TRACTION], bv


def cube_tension_bc(mesh_nb, patch_slices, t=1e7):
    """unit cube: z-min fixed (0), z-max normal traction t, others traction-free"""
    bv = np.zeros((mesh_nb, 3))
    bv[patch_slices[5]] = (0.0, 0.0, t)
    return [TRACTION, TRACTION, TRACTION, TRACTION, FIXED, TRACTION], bv

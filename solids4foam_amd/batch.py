"""Batched solves: many meshes packed into one sliced-ELL structure and
advanced by shared kernel launches (s4f_batch_* in include/s4f_abi.h)."""
import ctypes

import numpy as np

from . import _abi
from ._abi import check, lib


class Batch:
    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().s4f_batch_create(ctx.h, ctypes.byref(h)))
        self.h = h
        self.meshes = []

    def add(self, mesh):
        idx = ctypes.c_int32()
        check(lib().s4f_batch_add(self.h, mesh.h, ctypes.byref(idx)))
        self.meshes.append(mesh)
        return idx.value

    def solve(self, D0=None, want_sigma=True, **opts):
        o = _abi.default_opts(**opts)
        n = len(self.meshes)
        Ds = [np.zeros((m.nCells, 3)) for m in self.meshes]
        if D0 is not None:
            o.warmStart = 1
            for i, d in enumerate(D0):
                if d is not None:
                    Ds[i][:] = d
        sig = [np.zeros((m.nCells, 6)) for m in self.meshes] if want_sigma else None
        Dp = (ctypes.c_void_p * n)(*[d.ctypes.data for d in Ds])
        Sp = (ctypes.c_void_p * n)(*[s.ctypes.data for s in sig]) if want_sigma else None
        reps = (_abi.SolveReport * n)()
        st = lib().s4f_batch_solve(self.h, ctypes.byref(o), Dp, Sp, reps)
        check(st, allow=(_abi.S4F_OK, _abi.S4F_WARN_MAXITER, _abi.S4F_ERR_DIVERGED))
        out = []
        for i in range(n):
            r = reps[i]
            out.append(dict(D=Ds[i], sigma=sig[i] if sig else None, nOuter=r.nOuter, converged=bool(r.converged),
                            innerTotal=int(r.innerTotal), initialResidual=r.initialResidual,
                            residualvf=r.residualvf, seconds=r.seconds, pcgIterations=int(r.pcgIterations),
                            mixedFallback=int(r.mixedFallback)))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().s4f_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

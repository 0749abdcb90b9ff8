"""Continuous-batching solve server (s4f_server_* in include/s4f_abi.h).

Requests (mesh + per-boundary-face loads) are admitted by HBM footprint while
others are in flight; all in-flight solves advance together and retire on
convergence.  Results carry submit / start / finish timestamps for latency.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import check, f64, lib, ptr


class Server:
    def __init__(self, ctx, budget_bytes, max_slots=64, **opts):
        self.ctx = ctx
        o = _abi.default_opts(**opts)
        h = ctypes.c_void_p()
        check(lib().s4f_server_create(ctx.h, int(budget_bytes), int(max_slots), ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        self._meshes = {}   # id -> mesh (kept alive while in flight)
        self._keep = {}     # id -> host arrays (kept alive until copied)

    def submit(self, mesh, bc_values, pressure=None, warm_start=False):
        """bc_values: (nBoundaryFaces, 3) per boundary face in patch order"""
        bv = f64(bc_values, (mesh.nBoundaryFaces, 3))
        bp = None if pressure is None else f64(pressure, (mesh.nBoundaryFaces,))
        rq = _abi.Request(mesh.h.value, bv.ctypes.data, None if bp is None else bp.ctypes.data, int(bool(warm_start)))
        rid = ctypes.c_int64()
        check(lib().s4f_server_submit(self.h, ctypes.byref(rq), ctypes.byref(rid)))
        self._meshes[rid.value] = mesh
        return rid.value

    def poll(self, wait_ms=0):
        rid = ctypes.c_int64()
        st = lib().s4f_server_poll(self.h, int(wait_ms), ctypes.byref(rid))
        if st == _abi.S4F_ERR_NOT_READY:
            return None
        check(st)
        return rid.value

    def result(self, rid):
        mesh = self._meshes.pop(rid)
        D = np.zeros((mesh.nCells, 3))
        sig = np.zeros((mesh.nCells, 6))
        rep = _abi.SolveReport()
        tm = np.zeros(3)
        st = lib().s4f_server_result(self.h, int(rid), ptr(D), ptr(sig), ctypes.byref(rep), ptr(tm))
        check(st, allow=(_abi.S4F_OK, _abi.S4F_WARN_MAXITER, _abi.S4F_ERR_DIVERGED))
        return dict(id=rid, D=D, sigma=sig, nOuter=rep.nOuter, converged=bool(rep.converged),
                    innerTotal=int(rep.innerTotal), warm_started=bool(rep.warmStarted), status=st, t_submit=tm[0], t_start=tm[1], t_finish=tm[2],
                    latency=tm[2] - tm[0])

    def stats(self):
        out = np.zeros(6, np.int64)
        check(lib().s4f_server_stats(self.h, ptr(out)))
        return dict(queued=int(out[0]), finished=int(out[1]), bytes_resident=int(out[2]), budget=int(out[3]),
                    cell_capacity=int(out[4]), slots=int(out[5]))

    def close(self):
        if getattr(self, "h", None):
            lib().s4f_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

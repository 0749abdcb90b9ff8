"""The allocation-free PCG entry point s4f_pcg3_noalloc, called through the C
ABI with the caller's own device arrays, against s4f_pcg3 on the same right-
hand sides: bitwise the same solution and iteration counts, on two calls in a
row (the second on a smaller mesh, so the shared scratch array holds the
previous call's larger contents), in both precisions, and the size limit
refused with a clear error."""
import ctypes

import numpy as np
import pytest

import solids4foam_amd as s4f
from solids4foam_amd import cases
from solids4foam_amd._abi import S4F_ERR_INVALID, S4FError, check, default_opts, f64, lib, ptr

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    """The HIP runtime libs4f is linked against (already loaded with it)."""
    lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


class DevBuf:
    """a caller-owned device array of n doubles"""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), 8 * n) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64).reshape(-1)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, 8 * self.n, _H2D) == 0

    def get(self):
        out = np.zeros(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 8 * self.n, _D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(hip, m, b, nf, **opts):
    """s4f_pcg3_noalloc on b (C, 3): the caller's device arrays are component-major"""
    C = b.shape[0]
    bd, xd = DevBuf(hip, 3 * C), DevBuf(hip, 3 * C)
    try:
        bd.put(b.T.copy())
        xd.put(np.full(3 * C, np.nan))  # every entry must be written
        o = default_opts(**opts)
        it = np.zeros(3, np.int32)
        nfa = f64(nf)
        check(lib().s4f_pcg3_noalloc(m.h, ctypes.byref(o), bd.p, ptr(nfa), xd.p, ptr(it)))
        return xd.get().reshape(3, C).T.copy(), it
    finally:
        bd.free()
        xd.free()


@pytest.mark.parametrize("mixed", [False, True])
def test_noalloc_pcg_equals_pcg3_twice(mixed):
    hip = _hip()
    ctx = s4f.Context(0)
    try:
        opts = dict(pcgTolerance=1e-12, pcgRelTol=0.0, pcgMaxIter=400,
                    precision=1 if mixed else 0)
        # first call: a jittered 12^3 cantilever; second: a smaller 7x6x5 one
        for (n, seed) in (((12, 12, 12), 3), ((7, 6, 5), 4)):
            print(f"[noalloc] mixed={mixed} mesh {n} seed {seed}", flush=True)
            pm, m = cases.cantilever(ctx, *n, L=tuple(0.1 * v for v in n), jitter=0.1, seed=seed)
            rng = np.random.default_rng(seed)
            b = rng.standard_normal((m.nCells, 3)) * 1e3
            nf = np.abs(b).sum(axis=0)
            x_ref, it_ref = m.pcg3(b, nf, **opts)
            x_na, it_na = _noalloc(hip, m, b, nf, **opts)
            print(f"  iters {list(it_ref)} vs {list(it_na)} max|dx|={np.abs(x_ref - x_na).max():.3e}", flush=True)
            assert np.all(it_ref > 0)
            assert np.array_equal(it_ref, it_na)
            assert np.array_equal(x_ref, x_na)
            m.close()
            pm.close()
    finally:
        ctx.close()


def test_noalloc_refuses_meshes_beyond_its_scratch():
    hip = _hip()
    ctx = s4f.Context(0)
    try:
        pm, m = cases.cantilever(ctx, 65, 64, 64, L=(0.65, 0.64, 0.64))
        assert m.nCells > 262144
        with pytest.raises(S4FError) as e:
            _noalloc(hip, m, np.zeros((m.nCells, 3)), np.ones(3))
        assert e.value.status == S4F_ERR_INVALID
        assert "scratch limit" in str(e.value)
        m.close()
        pm.close()
    finally:
        ctx.close()

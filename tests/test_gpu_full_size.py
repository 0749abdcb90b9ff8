"""The 200^3 jittered cube (8M cells, the headline mesh) against the FP64
oracle on the same state.  At this size every instance is cut into 7,813
workgroup tiles of 16 slices (the capped-tile path smaller meshes never take)
and the finalize kernels reduce 7,813 partial rows.  Checked: the first four
outer iterations' history (|initRes|, residualvf, PCG counts) with the
diagonal preconditioner on both sides, then on the oracle's state after them
the least-squares gradient with its boundary correction, the residual
(div sigma + Rhie-Chow) and the 3-RHS SpMV in FP64 and FP32."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 200


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_full_size_history_and_operators_match_oracle(oracle_mod):
    from solids4foam_amd import Context, PRECOND_DIAGONAL, cases
    print(f"[full size] oracle mesh {N}^3", flush=True)
    om = oracle_mod.Mesh.box(N, N, N, (1.0, 1.0, 1.0), 0.15, 2)
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(5)] = (0.0, 0.0, 1e7)
    oc = oracle_mod.Case(om, 200e9, 0.3, [1, 1, 1, 1, 0, 1], bv)
    print("[full size] oracle: 4 outer iterations (diagonal PCG, maxIter 20)", flush=True)
    ro = oc.solve(history=4, nCorr=4, precond=oracle_mod.PRECOND_DIAGONAL, pcgMaxIter=20)
    ctx = Context(0)
    try:
        print("[full size] device mesh", flush=True)
        pm, m = cases.cube_tension(ctx, n=N, jitter=0.15, seed=2)
        info = m.info()
        assert info["nCells"] == om.nCells and info["ellSlices"] == N ** 3 // 64
        print("[full size] device: 4 outer iterations", flush=True)
        m.solve(nCorrectors=4, precision=0, preconditioner=PRECOND_DIAGONAL, pcgMaxIter=20, want_sigma=False)
        h, ho = m.history(), ro["history"]
        print(f"  initRes {h['initRes']} vs {ho['initRes']}\n  rvf {h['rvf']} vs {ho['rvf']}\n"
              f"  inner {h['inner']} vs {ho['inner']}", flush=True)
        assert len(h["initRes"]) == 4 and len(ho["initRes"]) == 4
        assert np.allclose(h["initRes"], ho["initRes"], rtol=1e-8, atol=1e-14)
        assert np.allclose(h["rvf"], ho["rvf"], rtol=1e-8)
        assert np.array_equal(h["inner"], ho["inner"])

        C, B = om.nCells, om.nBoundaryFaces
        D, Db, snb = oc.get("D").reshape(C, 3), oc.get("Db").reshape(B, 3), oc.get("snb").reshape(B, 3)
        print("[full size] gradient", flush=True)
        G, Gb = m.grad(D, Db, snb)
        oc.op("grad")
        Go, Gbo = oc.get("G").reshape(C, 9), oc.get("Gb").reshape(B, 9)
        eG, eGb = rel_l2(G, Go), rel_l2(Gb, Gbo)
        del G, Gb
        print(f"  err gradD {eG:.2e} gradD_b {eGb:.2e}", flush=True)
        assert eG < 1e-12 and eGb < 1e-12
        print("[full size] residual", flush=True)
        oc.op("sigma")
        R = m.residual(D, Go, oc.get("sigmab").reshape(B, 6))
        oc.op("residual")
        Ro = oc.get("R").reshape(C, 3)
        eR = float(np.abs(R - Ro).max() / np.abs(Ro).max())
        print(f"  max|R - R_oracle| / max|R_oracle| {eR:.2e}", flush=True)
        assert eR < 1e-10
        del R, Ro, Go
        print("[full size] SpMV", flush=True)
        x = np.random.default_rng(0).normal(size=(C, 3))
        yo = np.stack([oc.amul(np.ascontiguousarray(x[:, q])) for q in range(3)], axis=1)
        y = m.spmv3(x)
        e64 = rel_l2(y, yo)
        y32 = m.spmv3(x, mixed=True)
        e32 = rel_l2(y32, yo)
        print(f"  err spmv FP64 {e64:.2e} FP32 {e32:.2e}", flush=True)
        assert e64 < 1e-13 and e32 < 1e-5
        m.close()
        pm.close()
    finally:
        ctx.close()

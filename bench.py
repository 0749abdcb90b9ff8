#!/usr/bin/env python3
"""Benchmark: converged solves/s of the segregated small-strain solid solver on
one MI355X, on the large configuration: unit cube 200^3 = 8M hex cells,
interior vertices jittered +/-15% (seed 2), E=200 GPa, nu=0.3, z-min fixed,
z-max normal traction 1e7 Pa, other faces traction-free.

A step is one converged solve (outer defect-correction loop of
linearGeometryTotalDisplacement to solution/alternative tolerance 1e-8; inner
PCG with the aggregation-multigrid V-cycle preconditioner, smoothed first
prolongation, tolerance 1e-9, relTol 0.9).  The converged field is the fixed point of the FP64 residual and
does not depend on the inner solver settings.  `value` is the mixed-precision
path (FP32 PCG and V-cycle inside the FP64 defect loop); the plain FP64 path
is timed beside it and serves as the FP64 reference for the displacement /
von Mises error.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def algorithmic_bytes(kernel, C, F, B, T, q_entries=None, smoothed=False):
    """Compulsory HBM bytes per launch: each array a kernel touches counted
    once; neighbour gathers assumed cache-resident.  C cells, F internal faces,
    B boundary faces, T = bytes of the PCG scalar (8 FP64, 4 FP32),
    q_entries = stored entries of the multigrid's level-0 up-sweep operator
    (q = offdiag(A) P, or G on a smoothed hierarchy; given: the
    multigrid-preconditioned PCG's kernels)."""
    if kernel == "spmv" and q_entries is not None:  # z, p old/new, q, diag per cell; a + col per entry
        return C * 13 * T + F * 2 * (T + 4)
    if kernel == "spmv":            # r, 1/diag, p old/new, q, diag per cell; a + col per ELL entry (2 per face)
        return C * 14 * T + F * 2 * (T + 4)
    if kernel == "pcg_update":      # x r/w, p, r r/w, q, 1/diag
        return C * 19 * T
    if kernel == "finalize":        # one workgroup: partial rows of 3 values
        return 0
    if kernel == "mg0_down":        # b, 1/diag, diag; r write; a + col per ELL entry
        return C * 8 * T + F * 2 * (T + 4)
    if kernel == "mg0_up":          # b, r, 1/diag, agg (not with G); z write; q or G + column per entry
        return C * (10 * T + (0 if smoothed else 4)) + q_entries * (T + 4)
    if kernel == "residual":        # D, gradD, C, diag + r, p writes; per entry col, S, w, a, c; boundary S, sigma_b
        return C * (24 + 72 + 24 + 8 + 6 * T) + F * 2 * (4 + 24 + 8 + 8 + 8) + B * (4 + 24 + 48)
    if kernel == "grad":            # D, x, C, Minv, gradD w, D' w; col per entry; boundary data r/w
        return C * (24 + 3 * T + 24 + 48 + 72 + 24) + F * 2 * 4 + B * (4 + 24 + 8 + 24 + 24 + 24 + 8 + 24 * 6 + 72 + 48)
    raise KeyError(kernel)


def precond_id(name):
    from solids4foam_amd import _abi
    return {"diagonal": _abi.PRECOND_DIAGONAL, "gamg": _abi.PRECOND_GAMG,
            "gamg-sa": _abi.PRECOND_GAMG_SMOOTHED}[name]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cpu_threads():
    """host threads for the CPU baseline: OMP_NUM_THREADS when set (the
    process's CPU share), else every core"""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_baseline(n, jitter, seed, nOuter, iters, reltol, tol, threads):
    """The oracle's OpenMP outer iteration (the reference's work per
    iteration: four least-squares gradients, the literal fvm/fvc face fluxes,
    PCG with DIC per thread block, relTol 0.1 as the reference's fvSolution)
    on `threads` host threads, same mesh: seconds per outer iteration over
    `iters` iterations after two warm-up ones, scaled by the GPU FP64
    outer-iteration count to solves/s (an estimate: the CPU loop is not run
    to convergence)."""
    import oracle
    t0 = time.time()
    om = oracle.Mesh.box(n, n, n, (1.0, 1.0, 1.0), jitter, seed)
    bv = np.zeros((om.nBoundaryFaces, 3))
    bv[om.patch_faces(5)] = (0.0, 0.0, 1e7)
    oc = oracle.Case(om, 200e9, 0.3, [1, 1, 1, 1, 0, 1], bv)
    setup = time.time() - t0
    kw = dict(pcgRelTol=0.1, pcgTol=1e-9, solTol=tol, altTol=tol)
    oc.time_iterations_omp(2, threads, **kw)  # iteration 0 (zero stress, zero residual) and one more
    sec = oc.time_iterations_omp(iters, threads, **kw)
    if sec < 0:
        raise RuntimeError("the oracle's OpenMP path refused the case")
    per_iter = sec / iters
    cell_iter_s = om.nCells / per_iter
    return dict(value=1.0 / (per_iter * nOuter), unit="solves/s", cores=threads, kind="port",
                cell_iter_per_s=round(cell_iter_s, 1), seconds_per_outer_iteration=round(per_iter, 3),
                sample=f"oracle C restatement on {threads} OpenMP threads (host nproc {os.cpu_count()}): the "
                       f"reference's per-iteration work (4 LS gradients, literal fvm/fvc face fluxes, PCG + DIC per "
                       f"thread block, relTol 0.1) on the same {n}^3 mesh, {iters} outer iterations timed "
                       f"({sec:.1f} s) after 2 warm-up ones, setup {setup:.1f} s untimed; solves/s = "
                       f"1/(s_per_outer_iter x {nOuter} outer iterations of the GPU FP64 solve), an estimate")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--n", type=int, default=200)
    ap.add_argument("--jitter", type=float, default=0.15)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--reltol", type=float, default=0.9)
    ap.add_argument("--pcg-maxiter", type=int, default=1000, help="PCG iteration cap per outer iteration")
    ap.add_argument("--mg-scale", type=float, default=0.0, help="GAMG coarse-correction scale (0: library default)")
    ap.add_argument("--precond", choices=["diagonal", "gamg", "gamg-sa"], default=None,
                    help="PCG preconditioner of the device solves (gamg-sa: aggregation multigrid with a smoothed "
                         "first prolongation; default gamg-sa, gamg for --config tets; the serving stream keeps gamg)")
    ap.add_argument("--replay", type=int, default=40, help="back-to-back launches per kernel for the roofline")
    ap.add_argument("--cpu-iters", type=int, default=5, help="CPU baseline: timed outer iterations")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: OMP_NUM_THREADS or nproc)")
    ap.add_argument("--ref-tol", type=float, default=1e-10,
                    help="outer tolerance of the untimed FP64 reference solve the errors are measured against")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default workgroup tiles instead of the autotuned table "
                         "(solids4foam_amd/tuning/mi355x_tiles.json)")
    ap.add_argument("--no-batch", action="store_true", help="default run: skip the batched-sweep sub-object")
    ap.add_argument("--no-serve", action="store_true", help="default run: skip the serving sub-object")
    ap.add_argument("--serve-requests", type=int, default=64,
                    help="default run: requests of the serving sub-object (a bounded sample of the stream)")
    ap.add_argument("--serve-nmax", type=int, default=128, help="default run: serving boxes' largest n")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--eager", action="store_true", help="launch every kernel individually (no HIP graphs)")
    ap.add_argument("--only", choices=["both", "mixed", "fp64"], default="both")
    ap.add_argument("--kernels-only", action="store_true",
                    help="no timed solves: one outer iteration for state, then only the kernel replays "
                         "(the run that rocprofv3 profiles)")
    ap.add_argument("--calib-bytes", type=int, default=1 << 30,
                    help="size of the plain copy kernel run once before the timed region (counter calibration)")
    ap.add_argument("--pmc-summary", nargs="+", metavar="DB",
                    help="summarise rocprofv3 --pmc SQLite outputs (no GPU work) and exit")
    ap.add_argument("--pmc-json", help="with --pmc-summary: also write the per-kernel figures as JSON")
    ap.add_argument("--trace-summary", metavar="CSV",
                    help="print a markdown summary of a rocprofv3 --kernel-trace CSV (per-kernel totals and the "
                         "averages of back-to-back replay runs) and exit")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "pmc_cube200_mixed.json"),
                    help="committed rocprof counter summary that fills roofline.traffic")
    ap.add_argument("--config", choices=["cube", "batch", "serve", "tets"], default="cube",
                    help="cube: the 200^3 solve (default); batch: the batched sweep of uneven boxes; "
                         "serve: a Poisson request stream through the continuous-batching server; "
                         "tets: the n^3 x 6 tetrahedra cube with soft spherical inclusions")
    ap.add_argument("--count", type=int, default=256, help="batch: number of boxes")
    ap.add_argument("--nmax", type=int, default=64, help="batch/serve: largest cells per direction")
    ap.add_argument("--sample", type=int, default=8, help="batch/serve: boxes solved one at a time for the baseline")
    ap.add_argument("--requests", type=int, default=200, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=100.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--tets-zones", action="store_true",
                    help="tets: matrix and inclusions as two laws on sub-meshes instead of one law with E per cell")
    ap.add_argument("--repeat", type=float, default=0.2, help="serve: fraction repeating an earlier mesh")
    ap.add_argument("--slots", type=int, default=64, help="serve: in-flight solves")
    ap.add_argument("--budget-gib", type=float, default=96.0, help="serve: HBM arena")
    ap.add_argument("--sweep-tol", type=float, default=1e-9, help="batch/serve: outer tolerance")
    ap.add_argument("--relax-method", choices=["fixed", "aitken"], default="fixed",
                    help="batch/serve: relaxationMethod of the field relaxation (aitken: --relax is its first factor)")
    ap.add_argument("--relax", type=float, default=0.6,
                    help="batch/serve: fixed under-relaxation of D (the sweep reaches nu = 0.45, where the "
                         "unrelaxed segregated loop oscillates)")
    args = ap.parse_args()
    if args.precond is None:
        args.precond = "gamg" if args.config == "tets" else "gamg-sa"
    if args.trace_summary:
        trace_summary(args.trace_summary)
        return
    if args.pmc_summary:
        pmc_summary(args.pmc_summary, args.calib_bytes, args.pmc_json,
                    dict(workload="cube 200^3 hex jitter 0.15 seed 2", precision=args.only))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding of this path is "
              "a design, not a measured configuration", file=sys.stderr)
        sys.exit(2)

    import solids4foam_amd as s4f
    from solids4foam_amd import _abi, cases

    ctx = s4f.Context(0)
    if args.config == "batch":
        return run_batch(args, ctx)
    if args.config == "serve":
        return run_serve(args, ctx)
    if args.config == "tets":
        return run_tets(args, ctx)
    t0 = time.time()
    pm, mesh = cases.cube_tension(ctx, n=args.n, jitter=args.jitter, seed=args.seed)
    from solids4foam_amd import autotune
    tiles = None if args.no_tune else autotune.apply(mesh)
    setup_s = time.time() - t0
    info = mesh.info()
    C, B = info["nCells"], info["nBoundaryFaces"]
    F = info["nInternalFaces"]
    print(f"[bench] mesh {args.n}^3: C={C} F={F} B={B} ELL entries={info['ellEntries']} "
          f"device={info['deviceBytes'] / 2**30:.2f} GiB setup={setup_s:.1f}s", file=sys.stderr, flush=True)
    opts = dict(solutionTolerance=args.tol, alternativeTolerance=args.tol, pcgTolerance=1e-9, pcgRelTol=args.reltol,
                nCorrectors=100000, pcgMaxIter=args.pcg_maxiter, mgCoarseScale=args.mg_scale, eagerLaunch=int(args.eager),
                preconditioner=precond_id(args.precond))
    L = _abi.lib()

    def run(precision, steps, profile):
        # lat: the solve's device-stream time (events recorded after the input
        # field is in HBM and before any result is copied back); wall: the
        # whole call, host<->device copies of D and sigma included
        lat, wall, outs = [], [], []
        for _ in range(steps):
            t = time.perf_counter()
            out = mesh.solve(want_sigma=True, precision=precision, **opts)
            wall.append(time.perf_counter() - t)
            lat.append(out["seconds"])
            outs.append(out)
            print(f"[bench] precision={precision} nOuter={out['nOuter']} inner={out['innerTotal']} "
                  f"converged={out['converged']} device {lat[-1]:.3f}s wall {wall[-1]:.3f}s", file=sys.stderr,
                  flush=True)
        prof = replay_all(precision) if profile else {}
        return lat, wall, outs, prof

    def replay_all(precision):
        # each kernel replayed back to back on the mesh's current state
        prof = {}
        if args.replay <= 0:
            return prof
        ks = dict(_abi.KERNELS)
        if args.precond != "diagonal":
            ks.update(_abi.MG_KERNELS)
        for k, kid in ks.items():
            n = args.replay if k not in ("residual", "grad") else max(4, args.replay // 4)
            prof[k] = dict(avg_ms=ctypes_kernel_time(L, mesh, kid, precision, n), launches=n)
            print(f"[bench] replay {k}: {prof[k]['avg_ms']:.4f} ms x {n}", file=sys.stderr, flush=True)
        return prof

    if args.calib_bytes:
        cms = ctypes_copy(L, ctx, args.calib_bytes)
        print(f"[bench] k_copy {args.calib_bytes} B read + {args.calib_bytes} B written: {cms:.4f} ms "
              f"({2 * args.calib_bytes / cms / 1e6:.0f} GB/s)", file=sys.stderr, flush=True)
    if args.kernels_only:
        precision = 0 if args.only == "fp64" else 1
        mesh.solve(want_sigma=False, precision=precision, **dict(opts, nCorrectors=1, pcgMaxIter=2))
        prof = replay_all(precision)
        T = 8 if precision == 0 else 4
        qn = mesh.mg_info()["qEntries"] if args.precond != "diagonal" else None
        out = {k: dict(avg_ms=round(p["avg_ms"], 5), launches=p["launches"],
                       gbps=(round(algorithmic_bytes(k, C, F, B, T, qn, args.precond == "gamg-sa") /
                                   (p["avg_ms"] * 1e-3) / 1e9, 1)
                             if k not in ("finalize", "vcycle") else None)) for k, p in prof.items()}
        print(json.dumps(dict(mode="kernels-only", cells=C, precision="fp64" if precision == 0 else "mixed",
                              kernels=out)))
        return

    amg_setup_s = None
    if args.precond != "diagonal":
        # the multigrid hierarchy is built once per matrix (like OpenFOAM's
        # cached GAMG agglomeration): build it outside the timed region
        t = time.perf_counter()
        mesh.solve(want_sigma=False, precision=1, **dict(opts, nCorrectors=1))
        amg_setup_s = round(time.perf_counter() - t, 3)
        print(f"[bench] multigrid hierarchy + 1 outer iteration: {amg_setup_s:.2f}s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        mesh.solve(want_sigma=False, precision=1, **opts)
    res = {}
    if args.only in ("both", "mixed"):
        res["mixed"] = run(1, args.steps, True)
    if args.only in ("both", "fp64"):
        res["fp64"] = run(0, args.steps, args.only == "fp64")

    head = "mixed" if "mixed" in res else "fp64"
    lat, wall, outs, prof = res[head]
    T = 4 if head == "mixed" else 8
    total = sum(lat)
    value = args.steps / total
    nOuter = outs[-1]["nOuter"]
    kern = {}
    pcg_it = outs[-1]["pcgIterations"]
    mg = args.precond != "diagonal"
    per_solve = {"spmv": pcg_it, "pcg_update": pcg_it, "finalize": (3 if mg else 2) * pcg_it, "residual": nOuter,
                 "grad": nOuter, "mg0_down": pcg_it + nOuter, "mg0_up": pcg_it + nOuter, "vcycle": pcg_it + nOuter}
    mgi = mesh.mg_info() if mg else None
    for k, p in prof.items():
        e = dict(avg_ms=round(p["avg_ms"], 5), replay_launches=p["launches"], launches_per_solve=per_solve[k],
                 ms_per_solve_est=round(p["avg_ms"] * per_solve[k], 1))
        if k not in ("finalize", "vcycle"):
            b = algorithmic_bytes(k, C, F, B, T, mgi["qEntries"] if mgi else None, args.precond == "gamg-sa")
            e.update(alg_bytes=int(b), gbps=round(b / (p["avg_ms"] * 1e-3) / 1e9, 1))
        kern[k] = e
    dom = max((k for k in kern if k not in ("finalize", "vcycle")), key=lambda k: kern[k]["ms_per_solve_est"]) \
        if kern else None
    roofline = None
    if dom:
        a = kern[dom]["gbps"]
        traffic = None
        if args.n == 200 and args.jitter == 0.15 and args.seed == 2:
            traffic = load_traffic(args.traffic_json, dom, head)
        roofline = dict(bound="hbm", achieved=a, peak=HBM_PEAK_GBPS, unit="GB/s", frac=round(a / HBM_PEAK_GBPS, 4),
                        traffic=traffic, kernel=dom, alg_bytes_per_launch=kern[dom]["alg_bytes"])
        if traffic:
            roofline["traffic_source"] = os.path.relpath(args.traffic_json, ROOT)
            roofline["traffic_kind"] = ("L2-egress bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated): "
                                        "the memory side includes the Infinity Cache, an upper bound of HBM bytes")
    extra = {}
    if "mixed" in res and "fp64" in res:
        lat64, wall64, outs64, _ = res["fp64"]
        Dm, D64 = outs[-1]["D"], outs64[-1]["D"]
        extra = dict(fp64_solves_per_s=round(args.steps / sum(lat64), 5),
                     mixed_over_fp64=round((args.steps / total) / (args.steps / sum(lat64)), 3),
                     err_D_vs_fp64=rel_l2(Dm, D64),
                     err_vonMises_vs_fp64=rel_l2(s4f.sigma_eq(outs[-1]["sigma"]), s4f.sigma_eq(outs64[-1]["sigma"])),
                     fp64_nOuter=outs64[-1]["nOuter"], fp64_ms_per_step=round(1e3 * sum(lat64) / args.steps, 1),
                     fp64_wall_ms_per_step_with_pcie=round(1e3 * sum(wall64) / args.steps, 1))
    if "fp64" in res:
        # error reference: FP64 converged to a tight outer tolerance (untimed)
        t = time.perf_counter()
        ref = mesh.solve(want_sigma=True, precision=0, **dict(opts, solutionTolerance=args.ref_tol,
                                                                alternativeTolerance=args.ref_tol))
        Dr, qr = ref["D"], s4f.sigma_eq(ref["sigma"])
        print(f"[bench] FP64 reference to {args.ref_tol:g}: nOuter={ref['nOuter']} converged={ref['converged']} "
              f"{time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
        errs = {}
        for name, (_, _, o_, _) in res.items():
            errs[name] = dict(err_D=rel_l2(o_[-1]["D"], Dr), err_vonMises=rel_l2(s4f.sigma_eq(o_[-1]["sigma"]), qr))
        extra["error_reference"] = dict(precision="fp64", outer_tolerance=args.ref_tol, nOuter=ref["nOuter"],
                                        converged=ref["converged"], vs_reference=errs)
        del ref
    sub = {}
    if args.config == "cube" and args.n == 200 and not args.kernels_only:
        mesh.close()
        pm.close()
        if not args.no_batch:
            print("[bench] batched sweep: 256 boxes ...", file=sys.stderr, flush=True)
            sub["batch"] = batch_run(args, ctx, 256, 64)
        if not args.no_serve:
            print(f"[bench] serving: {args.serve_requests} requests ...", file=sys.stderr, flush=True)
            sub["serve"] = serve_run(args, ctx, args.serve_requests, args.serve_nmax)
    cpu = None
    if not args.no_cpu_baseline:
        n64 = extra.get("fp64_nOuter", nOuter)
        print("[bench] CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.n, args.jitter, args.seed, n64, args.cpu_iters, args.reltol, args.tol,
                           args.cpu_threads or cpu_threads())
    line = {
        "metric": "converged solves/s & cell-iter/s at 1e-6 err vs FP64; p50/p99 solve latency",
        "value": round(value, 6),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"cube {args.n}^3 hex, jitter {args.jitter} seed {args.seed}, z-min fixed, "
                               f"z-max traction 1e7 Pa, linearElastic E=200GPa nu=0.3",
                   "cells": C, "precision": head, "outer_tolerance": args.tol, "pcg": f"{args.precond}, relTol {args.reltol}, maxIter {args.pcg_maxiter}",
                   "tiles": list(tiles) if tiles else "library default (8192 tiles of >= 4 slices)"},
        "cell_iter_per_s": round(C * nOuter * args.steps / total, 1),
        "nOuter": nOuter,
        "pcg_iterations": pcg_it,
        "p50_latency_s": sub["serve"]["p50_latency_s"] if "serve" in sub else round(float(np.percentile(lat, 50)), 4),
        "p99_latency_s": sub["serve"]["p99_latency_s"] if "serve" in sub else round(float(np.percentile(lat, 99)), 4),
        "latency_source": "serving stream (see serve)" if "serve" in sub else "back-to-back solves of this mesh",
        "timing": "device stream, inputs resident in HBM (solve report: events after the input upload and "
                  "before the result copies)",
        "wall_ms_per_step_with_pcie": round(1e3 * sum(wall) / args.steps, 2),
        "roofline": roofline,
        "kernels": kern,
        "cpu_baseline": cpu,
    }
    if amg_setup_s is not None:
        line["multigrid_setup_s"] = amg_setup_s
    if mgi:
        line["multigrid_levels"] = dict(rows=mgi["rows"], entries=mgi["entries"], q_entries=mgi["qEntries"])
    line.update(extra)
    line.update(sub)
    print(json.dumps(line))


# bench kernel name -> kernel tag in the rocprof summaries
PMC_TAGS = {"residual": ("k_residual",), "spmv": ("k_spmvp",), "pcg_update": ("k_update",),
            "grad": ("k_grad_bnd_pre", "k_grad", "k_grad_bnd_post"),  # the gradient step's three launches
            "mg0_down": ("k_mg0_down",), "mg0_up": (("k_mg0_up_g", "k_mg0_up"),)}  # G form on a smoothed hierarchy


def load_traffic(path, kernel, precision):
    """Bytes per launch between L2 and memory (FETCH_SIZE + WRITE_SIZE,
    copy-kernel scaled; the memory side includes the Infinity Cache, so an
    upper bound of the HBM bytes) of `kernel` from a committed rocprof
    summary (JSON written by --pmc-json), or None."""
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    tags = PMC_TAGS.get(kernel)
    if not tags:
        return None
    suffix = "<f>" if precision == "mixed" else "<d>"
    total = 0.0
    for tag in tags:
        alts = tag if isinstance(tag, tuple) else (tag,)
        names = [a + sfx for a in alts for sfx in (suffix, "")]
        for name in names:
            e = d.get("kernels", {}).get(name)
            if e and e.get("FETCH_SIZE") is not None and e.get("WRITE_SIZE") is not None:
                total += e["FETCH_SIZE"] + e["WRITE_SIZE"]
                break
        else:
            return None
    return int(round(total * 1e6))


def trace_summary(path, min_run=5):
    """Markdown tables from a rocprofv3 kernel-trace CSV: per kernel the
    calls, average and total duration; and each run of >= min_run
    consecutive dispatches of one kernel (the replays of --kernels-only),
    whose average is what the HIP-event replay timing measures."""
    import collections
    import csv
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    name = lambda r: r["Kernel_Name"].split("(")[0].replace("void ", "")
    dur = lambda r: (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0
    tot = collections.defaultdict(lambda: [0, 0.0])
    for r in rows:
        t = tot[name(r)]
        t[0] += 1
        t[1] += dur(r)
    allt = sum(v[1] for v in tot.values())
    print("| kernel | calls | avg µs | total ms | % |\n|---|---:|---:|---:|---:|")
    for k, (n, d) in sorted(tot.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{k}` | {n} | {d / n:.1f} | {d / 1000:.2f} | {100 * d / allt:.1f} |")
    runs, cur = [], None
    for r in rows:
        if cur and cur[0] == name(r):
            cur[1].append(dur(r))
        else:
            if cur:
                runs.append(cur)
            cur = [name(r), [dur(r)]]
    if cur:
        runs.append(cur)
    runs = [(k, d) for k, d in runs if len(d) >= min_run and not k.startswith("__amd")]
    if runs:
        print("\n| replay run | launches | avg µs |\n|---|---:|---:|")
        for k, d in runs:
            print(f"| `{k}` | {len(d)} | {sum(d) / len(d):.1f} |")


def pmc_summary(paths, copy_bytes, json_out=None, meta=None):
    """Per-kernel averages of rocprofv3 --pmc counters (FETCH_SIZE / WRITE_SIZE,
    KiB per dispatch) from its SQLite output, each scaled by what the same
    counter reads for the plain copy kernel k_copy of known size (copy_bytes
    read + copy_bytes written).  Dispatches shorter than a tenth of the
    kernel's longest are left out: they are launches past convergence that
    exit at entry.  Prints a text table; reads only the given files."""
    import collections
    import glob
    import sqlite3
    rows = collections.defaultdict(dict)
    for pat in paths:
        for path in sorted(glob.glob(pat)):
            db = sqlite3.connect(path)
            cur = db.cursor()
            cols = [r[1] for r in cur.execute("pragma table_info(rocpd_info_kernel_symbol)")]
            ncol = "kernel_name" if "kernel_name" in cols else "display_name"
            names = {r[0]: r[1] for r in cur.execute(f"select id, {ncol} from rocpd_info_kernel_symbol")}
            pmc = {r[0]: r[1] for r in cur.execute("select id, name from rocpd_info_pmc")}
            ev = collections.defaultdict(list)
            q = ("select d.kernel_id, p.pmc_id, p.value, d.end - d.start from rocpd_pmc_event p "
                 "join rocpd_kernel_dispatch d on d.event_id = p.event_id")
            for kid, pid, v, dur in cur.execute(q):
                ev[(names.get(kid, str(kid)), pmc.get(pid, str(pid)))].append((v, dur))
            acc = {}
            for key, lst in ev.items():
                top = max(dur for _, dur in lst)
                kept = [(v, dur) for v, dur in lst if dur >= 0.1 * top]
                acc[key] = (sum(v for v, _ in kept), len(kept), sum(dur for _, dur in kept))
            for (k, c), (sv, n, d) in acc.items():
                short = k.split("(")[0]
                for tag in ("k_copy", "k_spmvp", "k_update", "k_residual", "k_grad_bnd_pre", "k_grad_bnd_post",
                            "k_aitken", "k_grad", "k_mg0_down", "k_mg0_up_g", "k_mg0_up", "k_mg_down", "k_mg_up",
                            "k_mg_restrict_w", "k_mg_restrict", "k_mg_dense", "k_fin"):
                    if tag in short:
                        short = tag + ("<f>" if "IfE" in k else "<d>" if "IdE" in k else "")
                        break
                rows[short[:40]][c] = (sv / n, n, d / n / 1e3)
    scale = {}
    if "k_copy" in rows:
        for c, (v, _, _) in rows["k_copy"].items():
            scale[c] = copy_bytes / 1024.0 / v if v > 0 else None
    print(f"copy-kernel scale factors (true KiB / counter KiB, {copy_bytes} B read + written): "
          + ", ".join(f"{c}={f:.3f}" for c, f in scale.items() if f))
    print(f"{'kernel':22s} {'counter':11s} {'avg KiB':>12s} {'scaled MB':>10s} {'dispatches':>10s} {'avg us':>9s}")
    out = {}
    for k in sorted(rows):
        for c, (v, n, d) in sorted(rows[k].items()):
            sc = v * scale[c] * 1024 / 1e6 if scale.get(c) else None
            print(f"{k:22s} {c:11s} {v:12.1f} {sc if sc is not None else float('nan'):10.1f} {n:10d} {d:9.2f}")
            out.setdefault(k, {})[c] = round(sc, 2) if sc is not None else None
            out[k]["avg_us"] = round(d, 2)
    if json_out:
        with open(json_out, "w") as f:
            json.dump(dict(meta or {}, scale=scale, unit="MB per dispatch (copy-kernel scaled)", kernels=out), f,
                      indent=1, sort_keys=True)


def _sequential(ctx, cases_, specs, opts, label, tune=True, pc=None):
    """Plain FP64 solves of `specs`, one at a time (the multigrid
    preconditioner of the run, the fastest single-mesh path): per-solve
    seconds and fields."""
    from solids4foam_amd import _abi, autotune
    pc = _abi.PRECOND_GAMG if pc is None else pc
    table = autotune.load_table() if tune else {}
    secs, Ds = [], []
    for sp in specs:
        _, m, _ = cases_.box_case(ctx, sp)
        autotune.apply(m, table)
        m.solve(want_sigma=False, precision=0, preconditioner=pc, **dict(opts, nCorrectors=1))
        t = time.perf_counter()
        out = m.solve(want_sigma=False, precision=0, preconditioner=pc, **opts)
        secs.append(time.perf_counter() - t)
        Ds.append(out["D"])
        print(f"[bench] {label} {sp['n']}: nOuter={out['nOuter']} converged={out['converged']} {secs[-1]:.2f}s",
              file=sys.stderr, flush=True)
    return secs, Ds


def sweep_opts(args):
    """outer/inner settings of the batched sweep and the serving stream"""
    return dict(solutionTolerance=args.sweep_tol, alternativeTolerance=args.sweep_tol, pcgTolerance=1e-12,
                pcgRelTol=args.reltol, nCorrectors=200000, relaxD=args.relax,
                relaxationMethod=1 if args.relax_method == "aitken" else 0, eagerLaunch=int(args.eager))


def run_batch(args, ctx):
    print(json.dumps(batch_run(args, ctx, args.count, args.nmax)))


def batch_run(args, ctx, count, nmax):
    """Batched sweep: `count` uneven boxes packed into one structure and solved
    together (mixed precision, PCG with the block-diagonal multigrid V-cycle,
    each instance stopping on its own convergence), against plain FP64 solves
    of a sample of the same boxes one at a time."""
    from solids4foam_amd import cases
    from solids4foam_amd.batch import Batch
    specs = cases.box_sweep_specs(count, args.seed, 8, nmax)
    opts = sweep_opts(args)
    t0 = time.time()
    # batch members keep the library's tile rule: the tuned table is for a
    # mesh solved alone, where small meshes under-fill the chip; packed
    # together they fill it, and smaller tiles would only idle waves
    meshes = [cases.box_case(ctx, sp) for sp in specs]
    b = Batch(ctx)
    for _, m, _ in meshes:
        b.add(m)
    cells = sum(m.nCells for _, m, _ in meshes)
    print(f"[bench] batch of {len(specs)} boxes, {cells} cells, set-up {time.time() - t0:.1f}s", file=sys.stderr,
          flush=True)
    from solids4foam_amd import _abi
    bopts = dict(opts, preconditioner=precond_id(args.precond))
    # pack, multigrid hierarchy and first launch outside the clock
    b.solve(want_sigma=False, precision=1, **dict(bopts, nCorrectors=1))
    t = time.perf_counter()
    res = b.solve(want_sigma=False, precision=1, **bopts)
    T = time.perf_counter() - t
    outer = np.array([r["nOuter"] for r in res])
    nconv = sum(r["converged"] for r in res)
    print(f"[bench] batched mixed: {T:.2f}s, converged {nconv}/{len(res)}, outer median {np.median(outer):.0f} "
          f"max {outer.max()}", file=sys.stderr, flush=True)
    idx = list(range(min(args.sample, len(specs))))
    seq_pc = None if args.precond != "gamg-sa" else precond_id(args.precond)
    secs, Ds = _sequential(ctx, cases, [specs[i] for i in idx], opts, "sequential fp64", not args.no_tune, seq_pc)
    errs = [rel_l2(res[i]["D"], D) for i, D in zip(idx, Ds)]
    seq_rate = len(secs) / sum(secs)
    value = len(res) / T
    line = {
        "metric": "batched converged solves/s (sweep of uneven boxes) vs sequential plain FP64 solves",
        "value": round(value, 4), "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{len(specs)} boxes n in 8..{nmax} per direction, jitter 0.1, E log-normal "
                               f"(median 100 GPa, sigma 0.5), nu U(0.2,0.45), x-max traction U(1e5,1e7) Pa random "
                               f"direction, seed {args.seed}", "cells": int(cells), "outer_tolerance": args.sweep_tol,
                   "pcg": f"{args.precond}, relTol {args.reltol} (batched); "
                          f"{'gamg' if args.precond == 'diagonal' else args.precond} (sequential)",
                   "relaxD": args.relax, "relaxation": args.relax_method},
        "batch_seconds": round(T, 3), "converged": int(nconv), "outer_median": float(np.median(outer)),
        "outer_max": int(outer.max()),
        "outer_percentiles": {q: float(np.percentile(outer, q)) for q in (10, 50, 90, 99, 100)},
        "slowest": [dict(n=specs[i]["n"], nu=round(specs[i]["nu"], 4), nOuter=int(outer[i]))
                    for i in np.argsort(outer)[-3:][::-1]], "cell_iter_per_s": round(float(sum(m.nCells * r["nOuter"] for (_, m, _), r in
                                                                         zip(meshes, res)) / T), 1),
        "sequential_fp64_solves_per_s": round(seq_rate, 4), "sequential_sample": len(secs),
        "sequential_sample_seconds": [round(x, 3) for x in secs],
        "speedup_vs_sequential_fp64": round(value / seq_rate, 2),
        "max_err_D_vs_fp64": max(errs) if errs else None,
    }
    return line


def run_serve(args, ctx):
    print(json.dumps(serve_run(args, ctx, args.requests, args.nmax)))


def serve_run(args, ctx, requests, nmax):
    """Serving: a Poisson stream of solve requests (boxes from the sweep
    generator, n up to nmax, a fraction repeating an earlier mesh with a new
    load and warm start) through the continuous-batching server."""
    from solids4foam_amd import cases
    from solids4foam_amd.serve import Server
    specs, stream, arrivals = cases.serve_requests(requests, args.seed, 8, nmax, args.repeat, args.rate)
    nUnique = len(specs)
    t0 = time.time()
    meshes = [cases.box_case(ctx, sp) for sp in specs]  # served together: the library's tile rule
    print(f"[bench] serve: {nUnique} meshes, {sum(m.nCells for _, m, _ in meshes)} cells, set-up "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    reqs = []
    for j, t, warm in stream:
        pm, m, bv = meshes[j]
        if t is not None:
            bv = np.zeros((pm.nBoundaryFaces, 3))
            bv[pm.patch_faces(1)] = t
        reqs.append((j, bv, warm))
    opts = sweep_opts(args)
    from solids4foam_amd import _abi
    srv = Server(ctx, int(args.budget_gib * (1 << 30)), max_slots=args.slots, precision=1,
                 preconditioner=_abi.PRECOND_DIAGONAL if args.precond == "diagonal" else _abi.PRECOND_GAMG, **opts)
    got, ids = {}, {}
    start = time.perf_counter()
    k = 0
    shown = 0.0
    while len(got) < len(reqs):
        now = time.perf_counter() - start
        while k < len(reqs) and arrivals[k] <= now:
            j, bv, warm = reqs[k]
            ids[srv.submit(meshes[j][1], bv, warm_start=warm)] = k
            k += 1
        rid = srv.poll(1 if k < len(reqs) else 50)
        while rid is not None:
            got[ids[rid]] = srv.result(rid)
            rid = srv.poll(0)
        if now - shown >= 20.0:  # progress of a long stream
            shown = now
            print(f"[bench] serve {now:.0f}s: {k} submitted, {len(got)} done", file=sys.stderr, flush=True)
    wall = time.perf_counter() - start
    srv.close()
    lat = np.array([got[i]["latency"] for i in range(len(reqs))])
    conv = sum(got[i]["converged"] for i in range(len(reqs)))
    for i in range(len(reqs)):
        if not got[i]["converged"]:
            j, _, w = reqs[i]
            print(f"[bench] request {i} (mesh {j} {specs[j]['n']} nu {specs[j]['nu']:.3f}, warm {w}) not converged: "
                  f"status {got[i]['status']} nOuter {got[i]['nOuter']}", file=sys.stderr, flush=True)
    warm = sum(got[i]["warm_started"] for i in range(len(reqs)))
    first = min(got[i]["t_submit"] for i in range(len(reqs)))
    last = max(got[i]["t_finish"] for i in range(len(reqs)))
    value = conv / (last - first)  # converged solves per second of the stream
    print(f"[bench] served {len(reqs)} in {wall:.1f}s: {value:.2f} solves/s p50 {np.percentile(lat, 50):.2f}s "
          f"p99 {np.percentile(lat, 99):.2f}s converged {conv} warm {warm}", file=sys.stderr, flush=True)
    idx = [i for i in range(len(reqs)) if not reqs[i][2]][:args.sample]
    seq_pc = None if args.precond != "gamg-sa" else precond_id(args.precond)
    secs, Ds = _sequential(ctx, cases, [specs[reqs[i][0]] for i in idx], opts, "sequential fp64", not args.no_tune,
                           seq_pc)
    errs = [rel_l2(got[i]["D"], D) for i, D in zip(idx, Ds)]
    seq_rate = len(secs) / sum(secs)
    line = {
        "metric": "served converged solves/s and p50/p99 latency (Poisson arrivals) vs sequential plain FP64",
        "value": round(value, 4), "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{len(reqs)} requests, Poisson rate {args.rate}/s, boxes n in 8..{nmax}, "
                               f"{args.repeat:.0%} repeat an earlier mesh with a new load (warm start), seed "
                               f"{args.seed}", "slots": args.slots, "budget_GiB": args.budget_gib,
                   "outer_tolerance": args.sweep_tol,
                   "pcg": f"{'diagonal' if args.precond == 'diagonal' else 'gamg'} (served), relTol {args.reltol}; "
                          f"sequential {'gamg' if args.precond == 'diagonal' else args.precond}",
                   "relaxD": args.relax, "relaxation": args.relax_method},
        "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
        "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
        "converged": int(conv), "warm_started": int(warm),
        "sequential_fp64_solves_per_s": round(seq_rate, 4), "sequential_sample": len(secs),
        "speedup_vs_sequential_fp64": round(value / seq_rate, 2),
        "max_err_D_vs_fp64": max(errs) if errs else None,
    }
    return line


def run_tets(args, ctx):
    """Heterogeneous tetrahedral case: mixed-precision solve (value) and FP64
    solve of the same settings, displacement and von Mises errors."""
    import solids4foam_amd as s4f
    from solids4foam_amd import _abi, cases
    t0 = time.time()
    pm, mesh, frac, nInc, _ = cases.tet_inclusions(ctx, n=args.n, jitter=args.jitter, seed=args.seed,
                                                   zones=args.tets_zones)
    C = mesh.nCells
    print(f"[bench] tets {args.n}^3 x 6: C={C}, {nInc} inclusions, {frac:.3f} of cells soft, set-up "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    opts = dict(solutionTolerance=args.tol, alternativeTolerance=args.tol, pcgTolerance=1e-9, pcgRelTol=args.reltol,
                nCorrectors=200000, pcgMaxIter=args.pcg_maxiter, mgCoarseScale=args.mg_scale, eagerLaunch=int(args.eager),
                preconditioner=precond_id(args.precond),
                progressSeconds=30)  # minutes-long solves report their progress
    t = time.perf_counter()
    mesh.solve(want_sigma=False, precision=1, **dict(opts, nCorrectors=1))
    print(f"[bench] hierarchy + 1 outer iteration {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    res = {}
    for name, prec in (("mixed", 1), ("fp64", 0)):
        t = time.perf_counter()
        out = mesh.solve(want_sigma=True, precision=prec, **opts)
        wall = time.perf_counter() - t
        res[name] = (out["seconds"], out, wall)  # device-stream time, inputs resident in HBM
        print(f"[bench] {name}: nOuter={out['nOuter']} converged={out['converged']} device {res[name][0]:.2f}s "
              f"wall {wall:.2f}s", file=sys.stderr, flush=True)
    (tm, om, wm), (tf, of, wf) = res["mixed"], res["fp64"]
    line = {
        "metric": "converged solves/s & cell-iter/s at 1e-6 err vs FP64 (heterogeneous tetrahedra)",
        "value": round(1.0 / tm, 6), "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"cube {args.n}^3 hex jitter {args.jitter} seed {args.seed} split into 6 tets; "
                               f"E 200 GPa with {nInc} spherical inclusions (r 0.02-0.06, 10% volume) at 20 GPa; "
                               f"nu 0.3; x-min fixed, x-max traction (0,5e6,0) Pa", "cells": int(C),
                   "outer_tolerance": args.tol, "pcg": f"{args.precond}, relTol {args.reltol}",
                   "gradient": "extendedLeastSquares 0 (full delta)",
                   "materials": ("two laws (matrix, inclusions) on their own sub-meshes" if args.tets_zones else
                                 "one law, E per cell")},
        "ms_per_step": round(1e3 * tm, 1), "nOuter": om["nOuter"], "pcg_iterations": om["pcgIterations"],
        "cell_iter_per_s": round(C * om["nOuter"] / tm, 1),
        "fp64_solves_per_s": round(1.0 / tf, 6), "fp64_nOuter": of["nOuter"],
        "mixed_over_fp64": round(tf / tm, 3),
        "err_D_vs_fp64": rel_l2(om["D"], of["D"]),
        "err_vonMises_vs_fp64": rel_l2(s4f.sigma_eq(om["sigma"]), s4f.sigma_eq(of["sigma"])),
        "soft_cell_fraction": round(frac, 4),
        "timing": "device stream, inputs resident in HBM (solve report)",
        "wall_ms_per_step_with_pcie": round(1e3 * wm, 1), "fp64_wall_ms_per_step_with_pcie": round(1e3 * wf, 1),
    }
    print(json.dumps(line))


def ctypes_copy(L, ctx, nbytes):
    import ctypes
    ms = ctypes.c_double()
    L.s4f_calibrate_copy(ctx.h, int(nbytes), ctypes.byref(ms))
    return ms.value


def ctypes_kernel_time(L, mesh, kid, precision, n):
    import ctypes
    from solids4foam_amd import _abi
    ms = ctypes.c_double()
    _abi.check(L.s4f_kernel_time(mesh.h, int(kid), int(precision), int(n), ctypes.byref(ms)))
    return ms.value


if __name__ == "__main__":
    main()

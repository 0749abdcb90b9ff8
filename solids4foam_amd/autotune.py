"""Workgroup-tile autotuner.

Every kernel over a mesh's cells runs one workgroup per tile (a run of
64-cell slices) and sums its reductions per tile, so the tile rule -- at most
`max_tiles` tiles of at least `min_slices` slices (s4f_mesh_tiles_set) --
sets occupancy, L2 locality and the finalize kernels' work.  `tune` times the
solver's kernels on a live mesh for each candidate rule (back-to-back
replays, s4f_kernel_time), weights them by their launches per outer
iteration of a multigrid-preconditioned PCG solve and keeps the cheapest;
`main` tunes a ladder of cube sizes and writes the table that `apply` reads
(keyed by cell count, nearest power of two).  The table is for a mesh
solved alone: packed into a batch or served with others, small meshes fill
the chip together and keep the library's default rule.  No counterpart in
the reference, whose loops have no launch geometry.

    python -m solids4foam_amd.autotune --sizes 16 32 64 100 128 200 --out solids4foam_amd/tuning/mi355x_tiles.json
"""
import argparse
import ctypes
import json
import math
import os

from . import _abi
from ._abi import check, lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning", "mi355x_tiles.json")

# candidate rules: (max_tiles, min_slices)
CANDIDATES = [(2048, 4), (4096, 4), (8192, 4), (16384, 4), (32768, 4), (65536, 4), (8192, 8), (8192, 16),
              (65536, 1), (65536, 2)]

# kernel ids of s4f_kernel_time and their launches per outer iteration of a
# multigrid-preconditioned PCG solve with `k` PCG iterations per outer one
# (the V-cycle replay covers every level's sweeps)
KERNELS = {"residual": 0, "grad": 4, "spmv": 8, "pcg_update": 2, "vcycle": 7, "finalize": 3}


def weights(k):
    return {"residual": 1.0, "grad": 1.0, "spmv": k, "pcg_update": k, "vcycle": k + 1.0, "finalize": 3.0 * k + 2.0}


def kernel_ms(mesh, kernel, precision, launches):
    ms = ctypes.c_double()
    check(lib().s4f_kernel_time(mesh.h, int(kernel), int(precision), int(launches), ctypes.byref(ms)))
    return ms.value


def bucket(n_cells):
    """table key of a mesh: its cell count rounded to the nearest power of two"""
    return str(2 ** int(round(math.log2(max(1, n_cells)))))


def tune(mesh, precision=_abi.PRECISION_MIXED, candidates=None, launches=10, pcg_per_outer=7.7):
    """Times the solver kernels on `mesh` (material and boundary values set)
    under each tile rule and leaves the mesh on the cheapest.  Returns
    {"best": (max_tiles, min_slices), "rows": [...]} with the weighted
    milliseconds per outer iteration of each rule."""
    cands = list(candidates or CANDIDATES)
    # state for the replays: a short multigrid solve (builds the hierarchy)
    mesh.solve(want_sigma=False, precision=precision, preconditioner=_abi.PRECOND_GAMG, nCorrectors=2)
    w = weights(pcg_per_outer)
    rows = []
    for cap, mn in cands:
        mesh.set_tiles(cap, mn)
        t = {k: kernel_ms(mesh, kid, precision, launches if k not in ("residual", "grad") else max(3, launches // 3))
             for k, kid in KERNELS.items()}
        rows.append(dict(max_tiles=cap, min_slices=mn, kernels_ms=t, ms_per_outer=sum(w[k] * t[k] for k in t)))
    best = min(rows, key=lambda r: r["ms_per_outer"])
    mesh.set_tiles(best["max_tiles"], best["min_slices"])
    return dict(best=(best["max_tiles"], best["min_slices"]), rows=rows)


def load_table(path=DEFAULT_TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("tiles", {})


def apply(mesh, table=None):
    """Sets the tuned tile rule of the mesh's size bucket, if the table has
    one; returns the rule or None (the library default stays)."""
    t = load_table() if table is None else table
    rule = t.get(bucket(mesh.nCells))
    if rule is None:
        return None
    mesh.set_tiles(rule["max_tiles"], rule["min_slices"])
    return rule["max_tiles"], rule["min_slices"]


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--sizes", type=int, nargs="+", default=[16, 32, 64, 100, 128, 200])
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--launches", type=int, default=10)
    args = ap.parse_args()
    import sys
    from .solid import Context
    from . import cases
    ctx = Context(0)
    table, detail = {}, {}
    for n in args.sizes:
        pm, m = cases.cube_tension(ctx, n=n, jitter=0.15, seed=2)
        r = tune(m, launches=args.launches)
        key = bucket(m.nCells)
        table[key] = dict(max_tiles=r["best"][0], min_slices=r["best"][1], cells=m.nCells)
        detail[key] = r["rows"]
        print(f"[autotune] {n}^3 ({m.nCells} cells): best {r['best']} "
              + " ".join(f"{row['max_tiles']}/{row['min_slices']}:{row['ms_per_outer']:.3f}" for row in r["rows"]),
              file=sys.stderr, flush=True)
        m.close()
        pm.close()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(dict(device="MI355X (gfx950)", workload="jittered cube per size, mixed precision, multigrid PCG",
                       weights=weights(7.7), tiles=table, rows=detail), f, indent=1)
    print(json.dumps(table))


if __name__ == "__main__":
    main()

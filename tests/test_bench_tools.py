"""Measurement tooling of bench.py, on the CPU: the algorithmic-byte model the
roofline divides by, and the counter summary (copy-kernel scaling, launches
that exit at entry left out, the committed JSON and its lookup) run on a
synthetic rocprofv3 SQLite database with the tables the summary reads."""
import json
import sqlite3

import pytest

import bench


def test_algorithmic_bytes_model():
    C, F, B = 1000, 2900, 600
    # multigrid SpMV: z, p old and new, q, diag per cell; value + column per entry (two per face)
    assert bench.algorithmic_bytes("spmv", C, F, B, 4, q_entries=10) == C * 13 * 4 + 2 * F * 8
    # diagonal preconditioner: r and 1/diag instead of z (FP64 entry: 8-byte value, 4-byte column)
    assert bench.algorithmic_bytes("spmv", C, F, B, 8) == C * 14 * 8 + 2 * F * 12
    # level-0 up sweep: b, r, 1/diag, agg, z per cell; q value + column per stored entry
    assert bench.algorithmic_bytes("mg0_up", C, F, B, 4, q_entries=3650) == C * (10 * 4 + 4) + 3650 * 8
    assert bench.algorithmic_bytes("mg0_down", C, F, B, 4) == C * 8 * 4 + 2 * F * 8
    assert bench.algorithmic_bytes("pcg_update", C, F, B, 8) == C * 19 * 8
    assert bench.algorithmic_bytes("finalize", C, F, B, 8) == 0
    with pytest.raises(KeyError):
        bench.algorithmic_bytes("no_such_kernel", C, F, B, 4)


def _rocpd(path, dispatches):
    """dispatches: (kernel symbol, counter, KiB, duration in ns)"""
    db = sqlite3.connect(path)
    cur = db.cursor()
    cur.execute("create table rocpd_info_kernel_symbol (id integer, kernel_name text)")
    cur.execute("create table rocpd_info_pmc (id integer, name text)")
    cur.execute("create table rocpd_pmc_event (event_id integer, pmc_id integer, value real)")
    cur.execute('create table rocpd_kernel_dispatch (event_id integer, kernel_id integer, start integer, "end" integer)')
    syms, pmcs = {}, {}
    t = 0
    for ev, (name, counter, kib, dur) in enumerate(dispatches):
        kid = syms.setdefault(name, len(syms))
        pid = pmcs.setdefault(counter, len(pmcs))
        cur.execute("insert into rocpd_kernel_dispatch values (?, ?, ?, ?)", (ev, kid, t, t + dur))
        cur.execute("insert into rocpd_pmc_event values (?, ?, ?)", (ev, pid, kib))
        t += dur + 1000
    cur.executemany("insert into rocpd_info_kernel_symbol values (?, ?)", [(i, n) for n, i in syms.items()])
    cur.executemany("insert into rocpd_info_pmc values (?, ?)", [(i, n) for n, i in pmcs.items()])
    db.commit()
    db.close()


def test_counter_summary_scaling_and_exit_filter(tmp_path):
    copy_bytes = 1 << 20  # 1024 KiB read and 1024 KiB written by k_copy
    copy = "_ZN3s4f6k_copyEPK15HIP_vector_typeIdLj4EEPS1_m"
    update = "_ZN3s4f8k_updateIfEEvNS_7DevMeshENS_8BlockMapE"
    fetch = [(copy, "FETCH_SIZE", 512.0, 400_000)]  # the counter sees half the bytes read: scale 2
    fetch += [(update, "FETCH_SIZE", 1000.0, 100_000)] * 3
    fetch += [(update, "FETCH_SIZE", 1.0, 2_000)]  # a launch past convergence: exits at entry
    write = [(copy, "WRITE_SIZE", 1024.0, 400_000)]  # scale 1
    write += [(update, "WRITE_SIZE", 300.0, 100_000)] * 3 + [(update, "WRITE_SIZE", 0.0, 2_000)]
    _rocpd(str(tmp_path / "fetch.db"), fetch)
    _rocpd(str(tmp_path / "write.db"), write)
    out = tmp_path / "pmc.json"
    bench.pmc_summary([str(tmp_path / "fetch.db"), str(tmp_path / "write.db")], copy_bytes, str(out),
                      meta=dict(workload="synthetic"))
    d = json.loads(out.read_text())
    assert d["scale"]["FETCH_SIZE"] == pytest.approx(2.0)
    assert d["scale"]["WRITE_SIZE"] == pytest.approx(1.0)
    k = d["kernels"]["k_update<f>"]
    assert k["FETCH_SIZE"] == pytest.approx(1000.0 * 2.0 * 1024 / 1e6, abs=0.01)  # the no-op launch left out
    assert k["WRITE_SIZE"] == pytest.approx(300.0 * 1024 / 1e6, abs=0.01)
    assert k["avg_us"] == pytest.approx(100.0)
    assert d["kernels"]["k_copy"]["FETCH_SIZE"] == pytest.approx(copy_bytes / 1e6, abs=0.01)
    # the roofline's traffic lookup sums both counters of the kernel's tag
    # (the JSON keeps MB to two decimals)
    assert bench.load_traffic(str(out), "pcg_update", "mixed") == pytest.approx((2.048 + 0.3072) * 1e6, abs=1e4)
    assert bench.load_traffic(str(out), "pcg_update", "fp64") is None

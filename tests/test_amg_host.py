"""The host half of the multigrid set-up (no GPU): aggregation, the Galerkin
product and the per-aggregate grouping of the level-0 off-diagonals, built
from the library's own source and checked against a map-based restatement
(tests/amg_host_check.hip).  Sizes that leave a partial last 64-row slice
and several threads' ranges are covered."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("amg") / "amg_host_check")
    cmd = ["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "amg_host_check.hip"), os.path.join(ROOT, "solids4foam_amd", "csrc", "s4f_polymesh.cpp"),
           "-o", out]
    subprocess.check_call(cmd, timeout=600)
    return out


@pytest.mark.parametrize("n", [23, 70])
def test_host_multigrid_build(harness, n):
    r = subprocess.run([harness, str(n)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr

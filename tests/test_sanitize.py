"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

Builds the product's polyMesh generator and geometry
(solids4foam_amd/csrc/s4f_polymesh.cpp) and the FP64 oracle
(oracle/s4f_oracle.c) with -fsanitize=address,undefined (g++/gcc, host code
only: no HIP source is built here) into a driver that runs small hex, tet,
2-D and transient cases (tests/sanitize/driver.cpp), and fails on any
sanitizer report or leak."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc/g++")
def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
    obj = str(tmp_path / "oracle.o")
    subprocess.check_call(["gcc", "-std=c11", "-c", *san, os.path.join(ROOT, "oracle", "s4f_oracle.c"), "-o", obj])
    exe = str(tmp_path / "driver")
    subprocess.check_call(["g++", "-std=c++17", *san, os.path.join(ROOT, "tests", "sanitize", "driver.cpp"),
                           os.path.join(ROOT, "solids4foam_amd", "csrc", "s4f_polymesh.cpp"), obj, "-o", exe, "-lm"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], env=env, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "rc=0" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr

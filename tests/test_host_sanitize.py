"""Host-side plan logic (the quarter transform's window schedule and its copy of C, the
SYRK tile-pair decode, the per-XCD cut and tile enumeration of grouped launches, the
workspace layout) built host-only with AddressSanitizer and UndefinedBehaviorSanitizer
and run on the CPU: tests/host/plan_check.cpp."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    pytest.skip("hipcc not available")


def test_host_plan_checks_under_asan_ubsan(tmp_path):
    exe = tmp_path / "plan_check"
    cmd = [hipcc(), "--cuda-host-only", "-O1", "-g", "-std=c++17",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all",
           f"-I{ROOT}/include", f"-I{ROOT}/pynof_amd/csrc", f"{ROOT}/tests/host/plan_check.cpp", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "all host plan checks passed" in r.stdout, r.stdout + r.stderr

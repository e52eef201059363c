"""The host-side ring-table build at the 1024^2 FEM workload's settings
(tests/host/groups_check.cpp: 1.18 M pixels, the multi-threaded stable sort,
two builds at once) compiled with g++ under AddressSanitizer + UBSan and
under ThreadSanitizer, host code only: no report, and the tables' invariants
hold (every listed pixel inside the compact plane, monotone offsets, every
group pixel listed once in the segment tiles)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "host", "groups_check.cpp"),
       os.path.join(ROOT, "ms-stem-fem_amd", "csrc", "msf_host.cpp")]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_ring_tables_at_1024_under_sanitizers(tmp_path, san):
    exe = str(tmp_path / "groups_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", f"-fsanitize={san}",
           "-fno-sanitize-recover=all", "-ffp-contract=off", "-pthread", *SRC, "-o", exe]
    subprocess.run(cmd, check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout, r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "checks 0 0 0" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "WARNING: ThreadSanitizer" not in r.stderr

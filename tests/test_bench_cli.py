"""bench.py's command line on the CPU: the one-GPU refusal of --gpus N > 1
(exit status 2 with a message, before any device work) and the workload
choices."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_workload_choices():
    r = subprocess.run([sys.executable, BENCH, "--help"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    for w in ("cuzr512", "cuzr512x4", "si256", "uneven48", "si1024"):
        assert w in r.stdout
    r = subprocess.run([sys.executable, BENCH, "--workload", "nope"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "invalid choice" in r.stderr

"""bench.py's command line on the CPU: the one-GPU build refuses --gpus N > 1
with a message and exit status 2, before it touches a device, and every
workload and partition option parses."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in out["error"]


def test_options_parse():
    r = subprocess.run([sys.executable, BENCH, "--help"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for opt in ("--workload", "--tiles", "--sweep", "--materials", "--copy-probe", "--mx", "--mz"):
        assert opt in r.stdout, opt
    for choice in ("roofline", "batched", "learned", "auto", "fused", "cells", "tune"):
        assert choice in r.stdout, choice

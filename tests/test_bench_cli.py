"""bench.py's command line on the CPU: the multi-GPU request is refused
with a message, before anything touches a device."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_help_lists_the_workloads():
    r = run("--help")
    assert r.returncode == 0
    for word in ("--workload", "mix", "serve", "--lanes", "--serve-cus"):
        assert word in r.stdout

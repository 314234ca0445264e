"""bench.py's command-line contract on the CPU: more than one GPU is refused with a
message, before anything touches a device."""
import os
import subprocess
import sys

from conftest import ROOT


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""

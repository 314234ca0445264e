"""spacelab_collider's command line on the CPU: argument errors and unreadable run files
end the program with a message before anything touches a GPU (exit 2 for usage, 1 for a
file it cannot use)."""
import os
import subprocess

import pytest

import spacelab_amd as sl

from conftest import GOLDEN


def _run(*args):
    prog = sl.collider_program()
    if not os.path.exists(prog):
        pytest.skip("spacelab_collider not built")
    return subprocess.run([prog, "--quiet", *args], capture_output=True, text=True, timeout=60)


def test_usage_errors():
    for args in ((), ("--bogus",), ("--generate", "10", "1", "--continue", "x/", "y_"),
                 ("--two-cluster", "x/", "p_"), ("--generate", "10", "1", "--order", "fastest")):
        r = _run(*args)
        assert r.returncode == 2, args
        assert "usage: spacelab_collider" in r.stderr


def test_missing_run_files(tmp_path):
    r = _run("--continue", str(tmp_path) + "/", "none_")
    assert r.returncode == 1 and "cannot read" in r.stderr
    d = os.path.join(GOLDEN, "two_cluster") + os.sep
    r = _run("--two-cluster", d, "proj_", "missing_")
    assert r.returncode == 1 and "missing_simData.csv" in r.stderr


def test_bad_constants_file(tmp_path):
    d = str(tmp_path) + "/"
    with open(d + "bad_simData.csv", "w") as f:
        f.write(sl.csv_sim_header(2) + "\n" + "0,0,0,0,0,0,0,0,0,0,0,1,0,0,0,0,0,0,0,0,0,0")
    with open(d + "bad_constants.csv", "w") as f:
        f.write("1e-5,1e-14\n")                 # a ball short of its moment of inertia
    r = _run("--continue", d, "bad_")
    assert r.returncode == 1 and "bad constants file" in r.stderr

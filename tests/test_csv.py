"""The CSV layer against the reference's own files (no GPU).

tests/golden/csv_cluster16/expected/ holds the files the reference's sim_init_write
and sim_looper wrote for a 60-step run of a 16-sphere cluster (written by
tests/golden/make_golden.py through the reference harness), and the matching .npz
holds the reference's own state and energies of the same run.  Formatting those
values must give the reference's bytes: headers, the 6-digit state rows, the first
energy row at 6 digits and the loop's energy rows at 12, the leading newlines and the
missing final newline."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden

import spacelab_amd as sl

CASE = os.path.join(GOLDEN, "csv_cluster16")


def _read(*p):
    with open(os.path.join(CASE, *p), "rb") as f:
        return f.read().decode()


@pytest.fixture(scope="module")
def g():
    return load_golden("csv_cluster16_s60_k3")


def test_constants_bytes(g):
    txt = sl.csv_constants(g["R"], g["m"], g["moi"])
    assert txt == _read("expected", "constants.csv") == _read("input", "constants.csv")
    assert txt.endswith("\n") and txt.count("\n") == 16


def test_sim_header_and_rows(g):
    txt = _read("expected", "simData.csv")
    assert not txt.endswith("\n")
    lines = txt.split("\n")
    n_steps, skip = int(g["n_steps"]), int(g["skip"])
    assert len(lines) == 2 + (n_steps - 1) // skip          # header, initial row, one per write step
    assert lines[0] == sl.csv_sim_header(16)
    assert lines[1] == sl.csv_sim_row(g["pos"], g["w"], g["vel"])
    # the last write step is the final step: the reference's final state gives its row
    assert lines[-1] == sl.csv_sim_row(g["out_pos"], g["out_w"], g["out_vel"])
    assert sl.csv_last_line(os.path.join(CASE, "expected", "simData.csv")) == lines[-1]
    # the input file of a continued run: header and the initial row
    assert _read("input", "simData.csv") == lines[0] + "\n" + lines[1]


def test_energy_rows(g):
    lines = _read("expected", "energy.csv").split("\n")
    dt, skip = float(g["params"][0]), int(g["skip"])
    assert lines[0] == "Time,PE,KE,E,p,L"
    e = g["e_init"]
    assert lines[1] == sl.csv_energy_row(0.0, e[0], e[1], e[2:5], e[5:8], digits=6)
    t = 0.0
    ew = g["e_writes"]
    assert len(lines) == 2 + len(ew)
    for k, e in enumerate(ew):
        t += dt * skip                                     # simTimeElapsed += dt * skip
        assert lines[2 + k] == sl.csv_energy_row(t, e[0], e[1], e[2:5], e[5:8], digits=12), k


def test_parse_round_trip(g):
    last = sl.csv_last_line(os.path.join(CASE, "expected", "simData.csv"))
    pos, w, vel = sl.csv_parse_sim_row(last)
    r6 = lambda a: np.array([float("%.6g" % x) for x in a.ravel()]).reshape(a.shape)
    assert np.array_equal(pos, r6(g["out_pos"]))
    assert np.array_equal(w, r6(g["out_w"]))
    assert np.array_equal(vel, r6(g["out_vel"]))
    R, m, moi = sl.csv_parse_constants(_read("input", "constants.csv"), 16)
    assert np.array_equal(R, g["R"]) and np.array_equal(m, g["m"]) and np.array_equal(moi, g["moi"])
    # the input state is exactly what a continued run starts from
    pos0, w0, vel0 = sl.csv_parse_sim_row(_read("input", "simData.csv").split("\n")[1])
    assert np.array_equal(pos0, g["pos"]) and np.array_equal(w0, g["w"]) and np.array_equal(vel0, g["vel"])


def test_last_line_edges(tmp_path):
    f = tmp_path / "a.csv"
    f.write_bytes(b"head\nrow1\nrow2")
    assert sl.csv_last_line(str(f)) == "row2"
    f.write_bytes(b"only")
    assert sl.csv_last_line(str(f)) == "only"
    f.write_bytes(b"x\n")
    assert sl.csv_last_line(str(f)) == ""
    # getLastLine's backward scan stops at the second byte: a file whose only newline is
    # its first byte reads as its (empty) first line
    f.write_bytes(b"\nrow")
    assert sl.csv_last_line(str(f)) == ""
    f.write_bytes(b"\n\nrow")
    assert sl.csv_last_line(str(f)) == "row"
    big = b"h\n" + b"1," * 10000 + b"9"                     # longer than one read block
    f.write_bytes(big)
    assert sl.csv_last_line(str(f)) == big.decode().split("\n")[-1]
    with pytest.raises(sl.SpacelabError):
        sl.csv_last_line(str(tmp_path / "missing.csv"))


def test_parse_errors():
    with pytest.raises(sl.SpacelabError):
        sl.csv_parse_sim_row("1,2,3,4,5,6,7,8,9")           # a ball short of columns
    with pytest.raises(sl.SpacelabError):
        sl.csv_parse_sim_row("1,2,x,4,5,6,7,8,9,10,0")
    with pytest.raises(sl.SpacelabError):
        sl.csv_parse_constants("1,2,3\n4,5\n", 2)
    assert sl.csv_sim_row(np.zeros((1, 3)), np.array([[3.0, 4.0, 0.0]]), np.zeros((1, 3))) == "0,0,0,3,4,0,5,0,0,0,0"


def test_formatting_special_values():
    # C++ streams print a double at 6 significant digits in the %g style
    row = sl.csv_energy_row(1e-9, -0.0, 123456789.0, [3.0, 4.0, 12.0], [0.0, 0.0, 0.0], digits=6)
    assert row == "1e-09,-0,1.23457e+08,1.23457e+08,13,0"

"""The C++ driver (spacelab_collider) end to end on the GPU: it continues a run from
the reference's input files and writes constants.csv and simData.csv byte for byte
as the reference's own sim_init_write + sim_looper do on the same input (the ordered
path is bitwise the reference's arithmetic).  energy.csv has the same layout and
times; its sums are taken in another order than the reference's running sums, so
PE, KE, E and |L| agree to 1e-9 relative and |p| (a sum that cancels to rounding
noise) to the momentum scale."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, load_golden

import spacelab_amd as sl

pytestmark = pytest.mark.gpu

CASE = os.path.join(GOLDEN, "csv_cluster16")


def _run(tmp, extra, cwd=None, out=None):
    g = load_golden("csv_cluster16_s60_k3")
    for f in ("constants.csv", "simData.csv"):
        shutil.copy(os.path.join(CASE, "input", f), os.path.join(tmp, "in_" + f))
    dt, kin, kout = (repr(float(x)) for x in g["params"])
    cmd = [sl.collider_program(), "--continue", str(tmp) + "/", "in_",
           "--out", out if out is not None else str(tmp) + "/out_",
           "--no-center", "--dt", dt, "--kin", kin, "--kout", kout,
           "--steps", str(int(g["n_steps"])), "--skip", str(int(g["skip"])),
           "--preset", "reference", "--quiet"] + extra
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=cwd)
    assert r.returncode == 0, r.stderr
    return g


def _energy(path):
    lines = open(path).read().split("\n")
    return lines[0], [l.split(",") for l in lines[1:]]


def test_driver_matches_reference_files(tmp_path):
    g = _run(tmp_path, ["--order", "ordered"])
    for f in ("constants.csv", "simData.csv"):
        got = open(tmp_path / ("out_" + f), "rb").read()
        exp = open(os.path.join(CASE, "expected", f), "rb").read()
        assert got == exp, f
    h1, a = _energy(tmp_path / "out_energy.csv")
    h2, b = _energy(os.path.join(CASE, "expected", "energy.csv"))
    assert h1 == h2 and len(a) == len(b) == 21
    pscale = float(np.sum(g["m"] * np.linalg.norm(g["vel"], axis=1)))
    for ra, rb in zip(a, b):
        assert ra[0] == rb[0]                                   # time column, same text
        x, y = np.array(ra[1:], float), np.array(rb[1:], float)
        esc = abs(y[0]) + abs(y[1])
        assert abs(x[0] - y[0]) <= 1e-9 * esc and abs(x[1] - y[1]) <= 1e-9 * esc   # PE, KE
        assert abs(x[2] - y[2]) <= 1e-9 * esc                   # E
        assert abs(x[3] - y[3]) <= 1e-12 * pscale               # |p|
        assert abs(x[4] - y[4]) <= 1e-9 * abs(y[4])             # |L|


def test_driver_fast_mode_close(tmp_path):
    _run(tmp_path, ["--order", "fast"])
    got = open(tmp_path / "out_simData.csv").read().split("\n")
    exp = open(os.path.join(CASE, "expected", "simData.csv")).read().split("\n")
    assert got[0] == exp[0] and len(got) == len(exp)
    for la, lb in zip(got[1:], exp[1:]):
        pa, wa, va = sl.csv_parse_sim_row(la)
        pb, wb, vb = sl.csv_parse_sim_row(lb)
        # the fast path reorders sums; printed at 6 digits the states agree to that
        assert np.allclose(pa, pb, rtol=2e-5, atol=1e-5 * np.abs(pb).max())
        assert np.allclose(va, vb, rtol=2e-5, atol=1e-5 * np.abs(vb).max())


def test_driver_file_counter(tmp_path):
    """A second run with the same output prefix gets a "1_" prefix, as sim_init_write does."""
    _run(tmp_path, ["--order", "ordered"], cwd=tmp_path, out="out_")
    _run(tmp_path, ["--order", "ordered"], cwd=tmp_path, out="out_")
    a = open(tmp_path / "out_simData.csv", "rb").read()
    b = open(tmp_path / "1_out_simData.csv", "rb").read()
    assert a == b


def test_driver_rejects_bad_input(tmp_path):
    r = subprocess.run([sl.collider_program(), "--continue", str(tmp_path) + "/", "nothing_"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "cannot" in r.stderr


def test_driver_generate_mode(tmp_path):
    """--generate builds a seeded aggregate and runs it; the files have the reference's
    layout: a header and one row per write step plus the initial row."""
    r = subprocess.run([sl.collider_program(), "--generate", "100", "0x5EED", "--out", str(tmp_path) + "/gen_",
                        "--steps", "31", "--skip", "10", "--preset", "config", "--order", "fast", "--quiet"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    sim = open(tmp_path / "gen_simData.csv").read()
    assert not sim.endswith("\n")
    lines = sim.split("\n")
    assert lines[0] == sl.csv_sim_header(100) and len(lines) == 2 + 3
    for row in lines[1:]:
        pos, w, vel = sl.csv_parse_sim_row(row)
        assert pos.shape == (100, 3) and np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))
    en = open(tmp_path / "gen_energy.csv").read().split("\n")
    assert en[0] == "Time,PE,KE,E,p,L" and len(en) == 2 + 3
    # the generated group is the library's seeded aggregate, centred on its centre of mass
    g = sl.generate_bpca(100, seed=0x5EED)
    R, m, moi = sl.csv_parse_constants(open(tmp_path / "gen_constants.csv").read(), 100)
    assert np.allclose(R, g.R, rtol=1e-5) and np.allclose(m, g.m, rtol=1e-5)


def test_driver_dynamic_time(tmp_path):
    """--dynamic-time recalibrates at write steps by the velocity rule.  A generated
    aggregate starts calibrated at v_custom (0.36 cm/s); at the first write step the rule
    takes its balls' largest infall speed (about 3e-3 cm/s), dt grows, and calibrate_dt's
    integer rescaling (int)(dt_old / dt) = 0 ends the run at that step, as the reference's
    loop would (the same calls on the CPU give steps = 10)."""
    r = subprocess.run([sl.collider_program(), "--generate", "50", "0x5EED", "--out", str(tmp_path) + "/dyn_",
                        "--steps", "41", "--skip", "10", "--dynamic-time", "--order", "ordered", "--quiet"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    en = open(tmp_path / "dyn_energy.csv").read().split("\n")
    assert len(en) == 3                      # header, initial row, the row of step 10
    sim = open(tmp_path / "dyn_simData.csv").read().split("\n")
    assert len(sim) == 3


def test_driver_two_cluster(tmp_path):
    """--two-cluster loads a projectile and a target run and sets up their collision as
    the reference's collision constructor does (ColliderSingleCore.cpp:32,
    ball_group.hpp:100-112).  The merged, calibrated and centred group is pinned bitwise
    to the reference's own constructor (tests/golden/two_cluster/agg_vcustom.npz), so the
    driver's constants file and first simData row are that group's bytes, and its dt and
    k are the reference's."""
    d = os.path.join(GOLDEN, "two_cluster") + os.sep
    z = np.load(d + "agg_vcustom.npz")
    cmd = [sl.collider_program(), "--two-cluster", d, "proj_", "targ_", "--out", str(tmp_path) + "/tc_",
           "--steps", "21", "--skip", "10", "--order", "ordered"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    run = dict(zip([str(f) for f in z["fields"]], z["run"]))
    assert f"dt: {run['dt']:.6g}\n" in r.stderr and f"k: {run['kin']:.6g}\n" in r.stderr
    assert open(tmp_path / "tc_constants.csv").read() == sl.csv_constants(z["R"], z["m"], z["moi"])
    sim = open(tmp_path / "tc_simData.csv").read().split("\n")
    assert sim[0] == sl.csv_sim_header(72) and len(sim) == 4
    assert sim[1] == sl.csv_sim_row(z["pos"], z["w"], z["vel"])
    for row in sim[2:]:
        pos, w, vel = sl.csv_parse_sim_row(row)
        assert np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))
    e = sl.csv_read_energy(str(tmp_path / "tc_energy.csv"))
    assert e.shape == (3, 6) and np.all(np.isfinite(e))

"""The reference's own data files as run input (no GPU).

tests/golden/generator/ holds the 1000-sphere files the reference's Generator wrote
(Generator/1000-...-dt0.04_{simData,Constants,Energy}.csv): the Generator names its
constants file "Constants.csv" and its energy file "Energy.csv" with the columns
Time,PE,KE,E,p,L,Bound,Unbound,m (Generator.cpp:168-170,195), where ball_group writes
"constants.csv" and Time,PE,KE,E,p,L.  The loader reads both.  What the files fix:
three radii (1481, 2962, 4443 cm), density 1.7 and moi = 0.4 m R^2 to the 6 printed
digits, E = PE + KE and the total mass in the m column."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

import spacelab_amd as sl

DIR = os.path.join(GOLDEN, "generator") + os.sep
PREFIX = "1000-R29019.1-k1e+18-cor0.64-mu0.3-rho1.70-dt0.04_"


def _lines(name):
    with open(DIR + PREFIX + name, "rb") as f:
        return f.read().decode().split("\n")


def test_continue_from_generator_files():
    g = sl.BallGroup.from_csv(DIR, PREFIX)
    assert g.num_particles == 1000
    assert sorted(set(g.R.tolist())) == [1481.0, 2962.0, 4443.0]
    assert np.allclose(g.moi, 0.4 * g.m * g.R ** 2, rtol=1e-5)
    assert np.allclose(g.m / (4 / 3 * np.pi * g.R ** 3), 1.7, rtol=1e-5)
    # the last line, split independently: x,y,z,wx,wy,wz,wmag,vx,vy,vz,comp per ball
    cols = np.array(_lines("simData.csv")[-1].split(","), dtype=np.float64).reshape(1000, 11)
    assert np.array_equal(g.pos, cols[:, 0:3])
    assert np.array_equal(g.w, cols[:, 3:6])
    assert np.array_equal(g.vel, cols[:, 7:10])


def test_every_row_parses():
    lines = _lines("simData.csv")
    assert sl.csv_sim_header(1000).count(",") == lines[0].count(",")
    for row in lines[1:]:
        pos, w, vel = sl.csv_parse_sim_row(row)
        assert pos.shape == (1000, 3) and np.isfinite(pos).all()


def test_both_energy_layouts():
    e = sl.csv_read_energy(DIR + PREFIX + "Energy.csv")
    assert e.shape == (4, 9)
    scale = np.abs(e[:, 1]) + np.abs(e[:, 2])
    assert np.all(np.abs(e[:, 3] - (e[:, 1] + e[:, 2])) <= 1e-5 * scale)
    R, m, _ = sl.csv_parse_constants("\n".join(_lines("Constants.csv")), 1000)
    assert np.allclose(e[:, 8], m.sum(), rtol=1e-5)
    # the six-column layout of the reference's own energy.csv
    own = os.path.join(GOLDEN, "csv_cluster16", "expected", "energy.csv")
    z = np.load(os.path.join(GOLDEN, "csv_cluster16_s60_k3.npz"))
    e6 = sl.csv_read_energy(own)
    assert e6.shape[1] == 6 and len(e6) == 1 + len(z["e_writes"])
    assert np.allclose(e6[1:, 1], z["e_writes"][:, 0], rtol=1e-11)
    assert np.allclose(e6[1:, 2], z["e_writes"][:, 1], rtol=1e-11)


def test_energy_layout_errors():
    assert sl.csv_energy_columns("Time,PE,KE,E,p,L\r\n") == 6
    for bad in ("Time,PE,KE", "time,PE,KE,E,p,L", ""):
        with pytest.raises(sl.SpacelabError):
            sl.csv_energy_columns(bad)
    import ctypes as C
    out = np.zeros(9)
    dp = out.ctypes.data_as(C.POINTER(C.c_double))
    assert sl.lib().sl_csv_parse_energy_row(b"1,2,3,5,4,6", 6, dp) == 0
    assert list(out[:6]) == [1, 2, 3, 5, 4, 6]
    assert sl.lib().sl_csv_parse_energy_row(b"1,2,3,5,4", 6, dp) < 0          # short
    assert sl.lib().sl_csv_parse_energy_row(b"1,2,3,5,4,6,7", 6, dp) < 0      # long
    assert sl.lib().sl_csv_parse_energy_row(b"1,2,3,5,4,6", 9, dp) < 0


def test_constants_extra_columns_ignored():
    # loadConsts reads three fields of each line and ignores the rest
    R, m, moi = sl.csv_parse_constants("1,2,3,9\n4,5,6\n", 2)
    assert list(R) == [1, 4] and list(m) == [2, 5] and list(moi) == [3, 6]


def test_oracle_pe_matches_generator_energy(oracle_mod):
    """The Generator's first energy row (t = 0.04) is the potential energy of the state
    in its first simData row.  At these sizes it is all gravitational (-G m_A m_B / d per
    pair; van der Waals is ~1e-9 of it), so the oracle's PE of that state with gravity on
    every pair must agree to what the printed digits allow: positions are printed to 6
    digits (a relative change of ~1e-6 in PE, measured 1.6e-6) and so is the PE."""
    g = sl.BallGroup.from_csv(DIR, PREFIX)
    pos, w, vel = sl.csv_parse_sim_row(_lines("simData.csv")[1])
    p = sl.preset("config")
    P = oracle_mod.Params(G=p.G, Ha=p.Ha, h_min=p.h_min, u_s=p.u_s, u_r=p.u_r, kin=1.0, kout=1.0, dt=0.0,
                          far_gravity=1)
    _, e0, _ = oracle_mod.run(oracle_mod.State(pos, vel, w, g.R, g.m, g.moi), P, 1, 0, mode="gather")
    pe_file = sl.csv_read_energy(DIR + PREFIX + "Energy.csv")[0, 1]
    assert abs(e0[0] - pe_file) <= 5e-6 * abs(pe_file)
    # without gravity the same state has almost no potential energy
    P0 = oracle_mod.Params(G=p.G, Ha=p.Ha, h_min=p.h_min, u_s=p.u_s, u_r=p.u_r, kin=1.0, kout=1.0, dt=0.0,
                           far_gravity=0)
    _, e1, _ = oracle_mod.run(oracle_mod.State(pos, vel, w, g.R, g.m, g.moi), P0, 1, 0, mode="gather")
    assert abs(e1[0]) <= 1e-6 * abs(pe_file)

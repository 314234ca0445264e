"""calibrate_dt's run control against the reference's own (no GPU).

tests/golden/calib/ holds, for four groups, the reference's calc_helpfuls,
calc_v_collapse and three calibrate_dt calls (ball_group.hpp:188-238, 302-350,
1313-1329), run by the reference harness: step 0 with the velocity rule, step 100 with
the velocities halved, step 200 at a custom speed.  The library's RunControl must give
the same numbers bit for bit, including the reference's integer step rescaling."""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

import spacelab_amd as sl

CASES = sorted(f[:-4] for f in os.listdir(os.path.join(GOLDEN, "calib")) if f.endswith(".npz"))
FIELDS = ("kin", "kout", "dt", "steps", "skip", "v_collapse", "v_max", "v_max_prev", "soc")


def _group(d):
    pos, vel, R, m = d["pos"], d["vel"], d["R"], d["m"]
    return sl.BallGroup(pos.copy(), vel.copy(), np.zeros_like(pos), R, m, 0.4 * m * R * R)


def _row(st):
    return np.array([float(getattr(st, f)) for f in FIELDS])


@pytest.mark.parametrize("case", CASES)
def test_run_control_matches_reference(case):
    d = dict(np.load(os.path.join(GOLDEN, "calib", case + ".npz")))
    rec = d["records"]
    g = _group(d)
    rc = sl.RunControl(g, sl.preset("reference"), cor=0.4)
    got = [_row(rc.calibrate(g, 0))]
    g.vel = g.vel * 0.5
    got.append(_row(rc.calibrate(g, 100)))
    got.append(_row(rc.calibrate(g, 200, float(d["custom_speed"]))))
    for k in range(3):
        for f, a, b in zip(FIELDS, got[k], rec[k]):
            assert a == b or (math.isinf(a) and math.isinf(b)), (case, k, f, a, b)


def test_resting_group_uses_collapse_speed():
    d = dict(np.load(os.path.join(GOLDEN, "calib", "resting16.npz")))
    g = _group(d)
    rc = sl.RunControl(g, sl.preset("reference"))
    st = rc.calibrate(g, 0)
    assert st.v_max == st.v_collapse > 0


def test_group_stats_closed_form():
    # two equal spheres 4 apart: centre of mass halfway, radius 2, soc = 4 r_max + 2
    import ctypes as C
    pos = np.array([[0.0, 0, 0], [4.0, 0, 0]])
    R = np.array([0.5, 0.5])
    m = np.array([1.0, 1.0])
    out = np.zeros(5)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert sl.lib().sl_group_stats(2, dp(pos), dp(R), dp(m), dp(out)) == 0
    assert list(out) == [0.5, 0.5, 2.0, 2.0, 4.0]


def test_run_length_rule():
    import ctypes as C
    st, sk = C.c_int(), C.c_int()
    assert sl.lib().sl_run_length(1e-9, -1.0, 0, 0, 0.5e-3, 1e-5, C.byref(st), C.byref(sk)) == 0
    assert (st.value, sk.value) == (500000, 10000)
    # later calls rescale the remaining steps by the integer part of dt_old / dt
    assert sl.lib().sl_run_length(0.4e-9, 1e-9, 100, 500000, 0.5e-3, 1e-5, C.byref(st), C.byref(sk)) == 0
    assert st.value == 2 * (500000 - 100) + 100

"""The two-cluster collision set-up against the reference's own constructor (no GPU).

tests/golden/two_cluster/ holds run files in the reference's layout (a 60-sphere
target, a 12-sphere projectile and a one-sphere dart, with their own motion and spins)
and, per case, what the reference's Ball_group(path, projectile, target, customVel)
built from them (ball_group.hpp:100-112, sim_init_two_cluster :1368-1442, run by the
reference harness): the merged group after calibrate_dt and to_origin, and its run
control.  Loading the same files, sl_two_cluster, RunControl.calibrate and to_origin
must give the same numbers bit for bit."""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

import spacelab_amd as sl

DIR = os.path.join(GOLDEN, "two_cluster") + os.sep
CASES = sorted(f[:-4] for f in os.listdir(DIR) if f.endswith(".npz"))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("case", CASES)
def test_two_cluster_matches_reference(case):
    d = dict(np.load(DIR + case + ".npz"))
    target = sl.BallGroup.from_csv(DIR, str(d["target"]))
    projectile = sl.BallGroup.from_csv(DIR, str(d["projectile"]))
    g, (v_small, v_big) = sl.two_cluster(target, projectile)
    assert v_small == -sl.V_CUSTOM
    fields = [str(f) for f in d["fields"]]
    ref = dict(zip(fields, d["run"]))
    # calc_helpfuls of the merged group, before centring
    st = np.zeros(5)
    import ctypes as C
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert sl.lib().sl_group_stats(g.num_particles, dp(g.pos), dp(g.R), dp(g.m), dp(st)) == 0
    assert (st[1], st[2], st[3]) == (ref["r_max"], ref["m_total"], ref["initial_radius"])
    rc = sl.RunControl(g, sl.preset("reference"), cor=0.4)
    rs = rc.calibrate(g, 0, float(d["custom_speed"]))
    for f in ("kin", "kout", "dt", "steps", "skip", "v_collapse", "v_max", "v_max_prev", "soc"):
        a, b = float(getattr(rs, f)), float(ref[f])
        assert a == b or (math.isinf(a) and math.isinf(b)), (case, f, a, b)
    g.to_origin()
    for k in ("pos", "vel", "w", "R", "m", "moi"):
        assert np.array_equal(bits(getattr(g, k)), bits(d[k])), (case, k)


def test_impact_angle_offset():
    """The projectile moves by (its radius + the target's radius + 2 r_max of the target)
    along (cos a, sin a, 0); the target does not move; momentum along x balances."""
    target = sl.BallGroup.from_csv(DIR, "targ_")
    projectile = sl.BallGroup.from_csv(DIR, "proj_")
    st = {}
    for name, grp in (("t", target), ("p", projectile)):
        c = (grp.m[:, None] * grp.pos).sum(0) / grp.m.sum()
        st[name] = (np.linalg.norm(grp.pos - c, axis=1).max(), grp.R.max())
    shift = st["p"][0] + st["t"][0] + 2 * st["t"][1]
    for a in (0.0, 0.3, -2.0):
        g, (vs, vb) = sl.two_cluster(target, projectile, v_impact=0.7, impact_angle=a)
        nt = target.num_particles
        assert np.array_equal(g.pos[:nt], target.pos)
        d = g.pos[nt:] - projectile.pos
        assert np.allclose(d, [shift * math.cos(a), shift * math.sin(a), 0.0], rtol=1e-12, atol=1e-18)
        assert vs == -0.7
        p = (g.m[:, None] * (g.vel - np.vstack([target.vel, projectile.vel]))).sum(0)
        assert abs(p[0]) <= 1e-12 * projectile.m.sum() * 0.7 and p[1] == 0 and p[2] == 0


def test_two_cluster_errors():
    t = sl.BallGroup.from_csv(DIR, "targ_")
    massless = t.copy()
    massless.m = np.zeros_like(massless.m)
    with pytest.raises(sl.SpacelabError):
        sl.two_cluster(massless, t)
    with pytest.raises(sl.SpacelabError):
        sl.two_cluster(t, t, v_impact=float("nan"))

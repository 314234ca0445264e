"""Collision set-up, outcome labelling and descriptors (host code, no GPU).

The set-up follows the two-cluster arrangement of ball_group.hpp:1368-1442 with the
impact parameter as a distance; the outcome rules and descriptors have no reference
counterpart and are pinned here by constructed cases whose answer is known.
"""
import numpy as np
import pytest

import spacelab_amd as sl

R0 = 1e-5


def test_collision_geometry_and_momentum():
    s = sl.generate_collision(300, 40, seed=11, speed=50.0, b_frac=0.4, gap=2.0)
    g = s.group
    t, p = slice(0, 300), slice(300, 340)
    mom = (g.m[:, None] * g.vel).sum(0)
    assert np.abs(mom).max() <= 1e-12 * (g.m * np.abs(g.vel[:, 0])).sum()
    # relative velocity of the bodies is the impact speed along +x
    assert np.allclose(g.vel[p, 0] - g.vel[t, 0][0], 50.0, rtol=1e-12)
    assert np.all(g.vel[:, 1:] == 0)
    com_t = (g.m[t, None] * g.pos[t]).sum(0) / g.m[t].sum()
    com_p = (g.m[p, None] * g.pos[p]).sum(0) / g.m[p].sum()
    assert np.allclose(com_t, 0, atol=1e-18)
    assert com_p[0] == pytest.approx(-(s.R_t + s.R_p + 2.0 * s.r_max), rel=1e-12)
    assert com_p[1] == pytest.approx(0.4 * (s.R_t + s.R_p), rel=1e-12)
    # the bodies do not touch at the start: surface gap at least 2 r_max
    d = np.linalg.norm(g.pos[t, None, :] - g.pos[None, p, :], axis=2) - g.R[t, None] - g.R[None, p]
    assert d.min() >= 2.0 * s.r_max * (1 - 1e-12)
    # outer radii as defined
    assert s.R_t == pytest.approx((np.linalg.norm(g.pos[t] - com_t, axis=1) + g.R[t]).max(), rel=1e-12)


def test_collision_deterministic_and_seeded():
    a = sl.generate_collision(50, 10, seed=3, speed=10.0, b_frac=0.0)
    b = sl.generate_collision(50, 10, seed=3, speed=10.0, b_frac=0.0)
    c = sl.generate_collision(50, 10, seed=4, speed=10.0, b_frac=0.0)
    assert np.array_equal(a.group.pos, b.group.pos)
    assert not np.array_equal(a.group.pos, c.group.pos)


def _group(pos, vel, R):
    R = np.asarray(R, dtype=float)
    m = 2.25 * 4 / 3 * np.pi * R ** 3
    pos = np.asarray(pos, dtype=float)
    return sl.BallGroup(pos, np.asarray(vel, dtype=float), np.zeros_like(pos), R, m, 0.4 * m * R * R)


def test_outcome_stick_bounce_fragment():
    link = 0.1 * R0
    # a chain of 4 touching spheres: target = first 2, projectile = last 2 -> stick
    x = np.arange(4) * 2 * R0
    pos = np.stack([x, 0 * x, 0 * x], 1)
    g = _group(pos, np.zeros((4, 3)), [R0] * 4)
    o = sl.collision_outcome(2, g, link)
    assert o.label == sl.OUTCOME_STICK and o.n_components == 1
    assert o.mass_transfer == pytest.approx(1.0)
    # the same with the pairs 10 radii apart, moving apart -> bounce
    pos2 = pos.copy()
    pos2[2:, 0] += 10 * R0
    vel = np.zeros((4, 3))
    vel[:2, 0] = -1.0
    vel[2:, 0] = 1.0
    o = sl.collision_outcome(2, _group(pos2, vel, [R0] * 4), link)
    assert o.label == sl.OUTCOME_BOUNCE and o.n_components == 2
    assert o.mass_transfer == pytest.approx(0.0)
    assert abs(o.v_largest[0]) == pytest.approx(1.0) and abs(o.v_second[0]) == pytest.approx(1.0)
    # every sphere alone -> fragment
    pos3 = pos.copy()
    pos3[:, 0] = np.arange(4) * 10 * R0
    o = sl.collision_outcome(2, _group(pos3, np.zeros((4, 3)), [R0] * 4), link)
    assert o.label == sl.OUTCOME_FRAGMENT and o.n_components == 4


def test_outcome_link_distance():
    """Spheres with a surface gap below the link distance are bound."""
    pos = np.array([[0, 0, 0], [2.05 * R0, 0, 0]])
    g = _group(pos, np.zeros((2, 3)), [R0, R0])
    assert sl.collision_outcome(1, g, 0.1 * R0).n_components == 1
    assert sl.collision_outcome(1, g, 0.01 * R0).n_components == 2


def test_descriptors_closed_forms():
    link = 0.1 * R0
    # a straight chain of n touching spheres: coordination 2 (n - 1) / n, and the
    # radius of gyration of n equal point masses spaced 2r is 2r sqrt((n^2 - 1) / 12)
    n = 9
    x = np.arange(n) * 2 * R0
    pos = np.vstack([np.stack([x, 0 * x, 0 * x], 1), [[0, 50 * R0, 0]]])
    R = np.full(n + 1, R0)
    m = 2.25 * 4 / 3 * np.pi * R ** 3
    setup = sl.CollisionSetup(sl.BallGroup(pos, np.zeros_like(pos), np.zeros_like(pos), R, m, 0.4 * m * R * R),
                              n, 1, 30.0, 0.25, 0, 0, 0, 0, 0, R0)
    d = sl.collision_descriptors(setup, link)
    assert d[0] == pytest.approx(np.log(n))
    assert d[1] == pytest.approx(2 * np.sqrt((n * n - 1) / 12), rel=1e-12)
    assert d[2] == pytest.approx(np.log(n) / np.log(d[1]), rel=1e-12)
    assert d[3] == pytest.approx(2 * (n - 1) / n)
    assert d[4] == pytest.approx(0.0) and d[6] == pytest.approx(3.0)   # single sphere
    assert d[8] == pytest.approx(np.log10(30.0)) and d[9] == pytest.approx(0.25)
    assert d[10] == pytest.approx(1 / (n + 1))


def test_approach_two_spheres_closed_form():
    # projectile sphere 10 R0 behind and R0 to the side: it touches after 10 - sqrt(3) radii
    pos = np.array([[0.0, 0, 0], [-10 * R0, R0, 0]])
    g = _group(pos, np.zeros((2, 3)), [R0, R0])
    s = sl.CollisionSetup(g, 1, 1, 1.0, 0.5, R0, R0, R0, 0, 0, R0)
    met, shift = sl.collision_approach(s, 0.01 * R0)
    assert met
    assert shift == pytest.approx((10 - np.sqrt(3) - 0.01) * R0, rel=1e-12)
    assert g.pos[1, 0] == pytest.approx(-(np.sqrt(3) + 0.01) * R0, rel=1e-12)
    assert g.pos[0, 0] == 0.0
    # 3 R0 to the side: a miss, nothing moves
    g2 = _group(np.array([[0.0, 0, 0], [-10 * R0, 3 * R0, 0]]), np.zeros((2, 3)), [R0, R0])
    s2 = sl.CollisionSetup(g2, 1, 1, 1.0, 1.0, R0, R0, 3 * R0, 0, 0, R0)
    assert sl.collision_approach(s2, 0.0) == (False, 0.0)
    assert g2.pos[1, 0] == -10 * R0


def test_approach_aggregates_touch_after_gap():
    """After approach with gap 0 the bodies touch without overlap; any further travel overlaps."""
    s = sl.generate_collision(120, 30, seed=17, speed=20.0, b_frac=0.3, gap=3.0)
    met, shift = sl.collision_approach(s, 0.0)
    assert met and shift > 0
    g = s.group
    t, p = g.pos[:120], g.pos[120:]
    S = g.R[:120, None] + g.R[None, 120:]

    def min_gap(dx):
        d = np.linalg.norm(t[:, None, :] - (p[None, :, :] + [dx, 0, 0]), axis=2)
        return (d - S).min()

    r = g.R.max()
    assert abs(min_gap(0.0)) <= 1e-9 * r
    assert min_gap(1e-3 * r) < 0
    assert min_gap(-1e-3 * r) > 0


def test_descriptors_impact_geometry():
    link = 0.1 * R0

    def desc(py):
        pos = np.array([[0.0, 0, 0], [-10 * R0, py, 0]])
        g = _group(pos, np.zeros((2, 3)), [R0, R0])
        s = sl.CollisionSetup(g, 1, 1, 30.0, py / (2 * R0), R0, R0, py, g.m[0], g.m[1], R0)
        return sl.collision_descriptors(s, link)

    d = desc(0.0)                       # head-on
    assert d[12] == 1 and d[13] == 1 and d[14] == 1
    assert d[15] == pytest.approx(np.log(2.0))
    assert d[16] == pytest.approx(1.0)
    assert d[17] == pytest.approx(1.0) and d[18] == pytest.approx(1.0)   # contact on both surfaces
    assert d[19] == 0 and d[20] == 0
    d = desc(R0)                        # offset by one radius: normal at 30 degrees to x
    assert d[12] == 1 and d[16] == pytest.approx(np.sqrt(3) / 2, rel=1e-12)
    assert d[17] == pytest.approx(1.0) and d[18] == pytest.approx(1.0)
    d = desc(3 * R0)                    # miss
    assert d[12] == 0 and d[13] == 0 and d[14] == 0 and d[15] == 0
    assert np.all(d[16:] == 0)
    # the chain case: neighbour counts of the touching balls
    n = 5
    x = np.arange(n) * 2 * R0
    pos = np.vstack([np.stack([x, 0 * x, 0 * x], 1), [[-5 * R0, 0, 0]]])
    g = _group(pos, np.zeros((n + 1, 3)), [R0] * (n + 1))
    s = sl.CollisionSetup(g, n, 1, 30.0, 0.0, 0, 0, 0, 0, 0, R0)
    d = sl.collision_descriptors(s, link)
    assert d[12] == 1 and d[16] == pytest.approx(1.0)
    assert d[19] == 1 and d[20] == 0    # the end ball of the chain touches one neighbour
    assert d[15] == pytest.approx(np.log(1 + n))


def test_descriptors_batch_equals_one_at_a_time():
    """The threaded batch entry gives every collision's descriptors exactly as the
    single call does, whatever the thread count (sizes from a lone ball to hundreds)."""
    setups = []
    for k, (nt, npj) in enumerate([(1, 1), (300, 40), (32, 1), (120, 64), (7, 3), (512, 2)]):
        s = sl.generate_collision(nt, npj, seed=100 + k, speed=50.0 * (k + 1), b_frac=0.1 * k)
        sl.collision_approach(s, 5e-7)
        setups.append(s)
    link = 1e-6
    one = np.stack([sl.collision_descriptors(s, link) for s in setups])
    for threads in (0, 1, 3, 16):
        assert np.array_equal(sl.collision_descriptors_batch(setups, link, threads), one), threads
    assert sl.collision_descriptors_batch([], link).shape == (0, sl.N_DESCRIPTORS)
    with pytest.raises(sl.SpacelabError):
        sl.collision_descriptors_batch(setups, -1.0)

"""Closed-form and invariant checks that pin the CPU oracle beyond the goldens.

These are the checks the mathematics fixes, chosen so that a dropped term, a
wrong sign or a swapped operand in the oracle fails one of them:

* free flight: x(t) = x0 + v t with no pair forces;
* vdW force magnitude = dU/dh (central difference of the potential);
* head-on restitution of the hysteretic spring: e = sqrt(kout/kin) and contact
  time (pi/2)(sqrt(m/kin) + sqrt(m/kout)) for two equal spheres;
* Kepler two-body orbit with all-pairs gravity (the far_gravity extension, which
  the reference leaves commented out): period 2 pi sqrt(a^3 / G M), energy and
  angular momentum conserved;
* Newton's third law: sum_i m_i a_i = 0; total angular momentum conserved when
  rolling friction is off.
"""
import math

import numpy as np
import pytest

from conftest import load_golden


def _state(oracle, pos, vel, w, R, m):
    R = np.asarray(R, dtype=np.float64)
    m = np.asarray(m, dtype=np.float64)
    return oracle.State(np.asarray(pos, np.float64), np.asarray(vel, np.float64),
                        np.asarray(w, np.float64), R, m, 0.4 * m * R * R)


def test_free_flight(oracle_mod):
    # two spheres far apart, no cohesion or gravity
    P = oracle_mod.Params(G=0.0, Ha=0.0, kin=1.0, kout=1.0, dt=1e-3)
    st = _state(oracle_mod, [[0, 0, 0], [10, 0, 0]], [[1, 2, -3], [-0.5, 0.25, 4]],
                np.zeros((2, 3)), [1e-2, 1e-2], [1.0, 2.0])
    s, _, _ = oracle_mod.run(st, P, 1001, 0, mode="tri")
    expect = st.pos + st.vel * (1000 * 1e-3)
    assert np.allclose(s.pos, expect, rtol=1e-12, atol=1e-12)
    assert np.array_equal(s.vel, st.vel)


@pytest.mark.parametrize("h", [2e-6, 1e-5, 5e-5])
def test_vdw_force_is_potential_gradient(oracle_mod, h):
    L = oracle_mod.lib()
    Ha, Ra, Rb = 4.7e-12, 1e-5, 1.3e-5
    eps = h * 1e-4
    dU = (L.orc_vdw_potential(Ha, Ra, Rb, h + eps) - L.orc_vdw_potential(Ha, Ra, Rb, h - eps)) / (2 * eps)
    F = L.orc_vdw_force_mag(Ha, Ra, Rb, h)
    assert F > 0 and dU > 0
    assert abs(F - dU) / F < 1e-6
    # symmetric in the two spheres
    assert abs(L.orc_vdw_force_mag(Ha, Rb, Ra, h) - F) / F < 1e-14
    # potential vanishes far away
    assert abs(L.orc_vdw_potential(Ha, Ra, Rb, 1.0)) < 1e-20


def test_headon_restitution_and_contact_time(oracle_mod):
    R, m = 1.0, 1.0
    kin, cor = 100.0, 0.64
    kout = cor * kin
    tc = math.pi / 2 * (math.sqrt(m / kin) + math.sqrt(m / kout))
    dt = tc / 4000
    v = 1.0
    P = oracle_mod.Params(G=0.0, Ha=0.0, u_s=0.0, u_r=0.0, kin=kin, kout=kout, dt=dt)
    gap = 2 * v * dt * 10
    st = _state(oracle_mod, [[-(R + gap / 2), 0, 0], [R + gap / 2, 0, 0]], [[v, 0, 0], [-v, 0, 0]],
                np.zeros((2, 3)), [R, R], [m, m])
    s, _, _ = oracle_mod.run(st, P, 1, 1, mode="tri")
    in_contact = False
    t_in = t_out = None
    for k in range(12000):
        s.pos_prev = None
        s, _, _ = oracle_mod.run(s, P, 2, 0, mode="gather", do_init=False)
        d = s.pos[1, 0] - s.pos[0, 0]
        if not in_contact and d < 2 * R:
            in_contact, t_in = True, k
        elif in_contact and d >= 2 * R:
            t_out = k
            break
    assert t_in is not None and t_out is not None
    e = (s.vel[1, 0] - s.vel[0, 0]) / (2 * v)
    assert abs(e - math.sqrt(cor)) < 2e-3
    assert abs((t_out - t_in) * dt - tc) / tc < 2e-3


def test_kepler_orbit_far_gravity(oracle_mod):
    G = 6.67e-8
    m1, m2 = 3e5, 1e5
    a = 100.0
    M = m1 + m2
    T = 2 * math.pi * math.sqrt(a ** 3 / (G * M))
    vrel = math.sqrt(G * M / a)
    r1, r2 = a * m2 / M, a * m1 / M
    pos = [[-r1, 0, 0], [r2, 0, 0]]
    vel = [[0, -vrel * m2 / M, 0], [0, vrel * m1 / M, 0]]
    nsteps = 4000
    P = oracle_mod.Params(G=G, Ha=0.0, u_s=0.0, u_r=0.0, kin=0.0, kout=0.0, dt=T / nsteps, far_gravity=1)
    st = _state(oracle_mod, pos, vel, np.zeros((2, 3)), [1.0, 1.0], [m1, m2])
    s, e0, ew = oracle_mod.run(st, P, nsteps + 1, nsteps // 8, mode="tri")
    # back where it started after one period
    assert np.allclose(s.pos, st.pos, atol=1e-3 * a)
    # PE = -G m1 m2 / a; energy and angular momentum conserved
    assert abs(e0[0] - (-G * m1 * m2 / a)) / (G * m1 * m2 / a) < 1e-12
    E0 = e0[0] + e0[1]
    for e in ew:
        assert abs((e[0] + e[1]) - E0) / abs(E0) < 1e-5
        assert abs(e[7] - e0[7]) / abs(e0[7]) < 1e-10
    # the same orbit with far gravity off does not close (sanity of the flag)
    P0 = oracle_mod.Params(G=G, Ha=0.0, u_s=0.0, u_r=0.0, kin=0.0, kout=0.0, dt=T / nsteps, far_gravity=0)
    s0, _, _ = oracle_mod.run(st, P0, nsteps + 1, 0, mode="tri")
    assert not np.allclose(s0.pos, st.pos, atol=1e-3 * a)


def test_newton_third_law_and_angular_momentum(oracle_mod):
    z = load_golden("cluster64_s100_k25")
    dt, kin, kout = z["params"]
    st = oracle_mod.State(z["pos"], z["vel"], z["w"], z["R"], z["m"], z["moi"])
    # linear momentum: forces are equal and opposite
    P = oracle_mod.Params(kin=float(kin), kout=float(kout), dt=float(dt))
    s, _, _ = oracle_mod.run(st, P, 2, 1, mode="tri")
    ma = z["m"][:, None] * s.acc
    assert np.linalg.norm(ma.sum(0)) < 1e-12 * np.abs(ma).sum()
    # angular momentum: conserved by contact, cohesion and sliding when u_r = 0
    P0 = oracle_mod.Params(kin=float(kin), kout=float(kout), dt=float(dt), u_r=0.0)
    s, e0, ew = oracle_mod.run(st, P0, 101, 10, mode="tri")
    Lscale = float(np.sum(z["m"] * np.linalg.norm(np.cross(z["pos"], z["vel"]), axis=1)
                          + z["moi"] * np.linalg.norm(z["w"], axis=1)))
    drift = max(np.linalg.norm(e[5:8] - e0[5:8]) for e in ew) / Lscale
    assert drift < 1e-6
    # with rolling friction on and unequal radii, the reference's rolling term
    # is not torque-free, so the same check must see it (guards the sign/term)
    P1 = oracle_mod.Params(kin=float(kin), kout=float(kout), dt=float(dt), u_r=0.5)
    s1, e1, ew1 = oracle_mod.run(st, P1, 101, 10, mode="tri")
    drift1 = max(np.linalg.norm(e[5:8] - e1[5:8]) for e in ew1) / Lscale
    assert drift1 > 10 * drift

"""GPU parity of the fast FP64 path and the mixed-precision path.

The fast path evaluates each pair once and reassociates the far-field algebra
(s = d^2 - S^2 instead of h = d - S), so it agrees with the reference's
single-core arithmetic (the oracle) to rounding, not bit for bit.  In these
contact-dominated aggregates a rounding-level difference grows with an
e-folding time of about five steps, so the bars are: forces to 1e-13 after one
step, positions to 1e-12 and velocities to 1e-10 after 100 steps, and energy
conservation as good as the reference arithmetic's over 10k steps of
conservative dynamics.  The mixed path's bar is an energy error within 1e-6 of
FP64 over the first 100 steps and the same energy conservation over 10k steps.
Relative errors of arrays are norm-wise: max |a - b| / max |b|.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, bits_equal, golden_names, load_golden

import spacelab_amd as sl

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
V_CUSTOM = 0.36301555459799423


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _oracle_params(oracle, p):
    return oracle.Params(G=p.G, Ha=p.Ha, h_min=p.h_min, u_s=p.u_s, u_r=p.u_r, kin=p.kin,
                         kout=p.kout, dt=p.dt, far_gravity=p.far_gravity)


def test_rotation_lane_layout():
    """The tile kernel's travelling reaction sums rely on this lane rotation."""
    out = sl.probe_rotation()
    lanes = np.arange(64)
    assert np.array_equal(out[0], (lanes + 1) % 64)
    assert np.array_equal(out[1], (lanes + 1) % 64)
    assert np.array_equal(out[2], (lanes + 31) % 64)


def _run(g, p, steps, write_every=0):
    with sl.Collider(g.num_particles, p) as c:
        c.upload(g)
        e0 = c.init_conditions()
        ew = c.step(steps, write_every)
        return c.download(), e0, ew


@pytest.mark.parametrize("far_gravity", [0, 1])
@pytest.mark.parametrize("v_sigma", [1e-3, 5.0])
def test_fast_1000_vs_oracle(oracle_mod, far_gravity, v_sigma):
    """Force-level agreement after one step, and the trajectory after 100 steps.

    Measured on MI355X: acc agrees to ~3e-15 and velocities to ~1e-15 over the
    first 20 steps; afterwards any rounding difference grows with an e-folding
    time of about five steps (the loading/unloading spring switch at maximum
    compression is discontinuous), reaching ~1e-14 in positions and ~1e-11 in
    velocities at step 100 and O(1) by step ~500.  Spins, driven by the friction
    terms whose direction switches discontinuously, drift fastest (~5e-10 at step
    100), so their bar there is 1e-9; the one-step bars on forces and torques are
    what pin the fast arithmetic itself.  Only the ordered path, which reproduces
    the reference's rounding exactly, stays on the reference trajectory beyond that."""
    n = 1000
    g = sl.generate_bpca(n, seed=0x5EED0000, v_sigma=v_sigma)
    base = sl.preset("config" if far_gravity else "reference")
    p = sl.calibrate(base, g.R, g.m, velocity=V_CUSTOM, cor=0.64).replace(order=sl.ORDER_FAST)
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    P = _oracle_params(oracle_mod, p)
    out1, e0, _ = _run(g, p, 1)
    s1, oe0, _ = oracle_mod.run(st, P, 2, 0, mode="gather", nthreads=NTHREADS)
    assert relerr(out1.acc, s1.acc) <= 1e-13
    assert relerr(out1.aacc, s1.aacc) <= 1e-13
    assert relerr(out1.vel, s1.vel) <= 1e-14
    assert relerr(out1.w, s1.w) <= 1e-13
    assert relerr(out1.pos, s1.pos) <= 1e-15
    assert abs(e0.PE - oe0[0]) <= 1e-12 * abs(oe0[0])
    out, _, ew = _run(g, p, 100, 100)
    s, _, oew = oracle_mod.run(st, P, 101, 100, mode="gather", nthreads=NTHREADS)
    assert relerr(out.pos, s.pos) <= 1e-12
    assert relerr(out.vel, s.vel) <= 1e-10
    assert relerr(out.w, s.w) <= 1e-9 or np.max(np.abs(s.w)) == 0
    scale = abs(oew[-1][0]) + oew[-1][1]
    assert abs(ew[-1].E - (oew[-1][0] + oew[-1][1])) <= 1e-12 * scale


@pytest.mark.parametrize("name", golden_names())
def test_fast_goldens(name):
    """Small contact-dominated cases: every pair here is near, so the fast path
    evaluates them in the reference arithmetic; far pairs differ by rounding."""
    z = load_golden(name)
    dt, kin, kout = z["params"]
    g = sl.BallGroup(z["pos"], z["vel"], z["w"], z["R"], z["m"], z["moi"])
    p = sl.preset("reference").replace(kin=float(kin), kout=float(kout), dt=float(dt),
                                       order=sl.ORDER_FAST)
    with sl.Collider(len(z["R"]), p) as c:
        c.upload(g)
        c.init_conditions()
        c.sim_looper(int(z["n_steps"]), int(z["skip"]))
        out = c.download()
    for k in ("pos", "vel", "w"):
        assert relerr(getattr(out, k), z["out_" + k]) <= 1e-12, k


def _gas(n, seed):
    """Spheres in a loose cloud with elastic, frictionless, cohesionless contacts and
    gravity on all pairs: conservative dynamics, so total energy is an invariant."""
    g = sl.generate_bpca(n, seed=seed, v_sigma=5.0)
    g.pos *= 1.6   # open the aggregate up so spheres fly and collide
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=1.0)
    p = p.replace(Ha=0.0, u_s=0.0, u_r=0.0, kout=p.kin)
    return g, p


def _max_drift(ew, e0):
    scale = abs(e0.PE) + e0.KE
    return max(abs(e.E - e0.E) for e in ew) / scale


def test_fast_energy_conservation_10k():
    """10k steps of conservative dynamics: the fast path conserves energy as well as
    the reference arithmetic (ordered path) does.  Trajectories themselves diverge
    (chaos), so this compares the invariant, not the states."""
    g, p = _gas(400, 0x5EED0003)
    _, e0o, ewo = _run(g, p, 10000, 500)
    _, e0f, ewf = _run(g, p.replace(order=sl.ORDER_FAST), 10000, 500)
    do, df = _max_drift(ewo, e0o), _max_drift(ewf, e0f)
    assert e0o.contacts == 0 or do < 1e-2
    assert df <= 2 * do + 1e-9


def test_mixed_energy_within_1e6_of_fp64():
    """Mixed precision against FP64 from the same state: energy within 1e-6 over the
    first 100 steps (measured ~1e-8; beyond that the trajectories part by chaos,
    as any two non-bitwise evaluations do), and the same energy conservation over
    10k steps of conservative dynamics."""
    n = 2000
    g = sl.generate_bpca(n, seed=0x5EED0002, v_sigma=1e-3)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64)
    ref, e0r, ewr = _run(g, p.replace(order=sl.ORDER_FAST), 100, 10)
    mix, e0m, ewm = _run(g, p.replace(order=sl.ORDER_FAST, precision=sl.PREC_MIXED), 100, 10)
    assert abs(e0m.E - e0r.E) <= 1e-6 * (abs(e0r.PE) + e0r.KE)
    for a, b in zip(ewm, ewr):
        assert abs(a.E - b.E) <= 1e-6 * (abs(b.PE) + b.KE)
    assert relerr(mix.pos, ref.pos) <= 1e-5
    gg, pg = _gas(400, 0x5EED0004)
    _, e0f, ewf = _run(gg, pg.replace(order=sl.ORDER_FAST), 10000, 500)
    _, e0x, ewx = _run(gg, pg.replace(order=sl.ORDER_FAST, precision=sl.PREC_MIXED), 10000, 500)
    assert _max_drift(ewx, e0x) <= 2 * _max_drift(ewf, e0f) + 1e-6


@pytest.mark.parametrize("split", [1, 2, 4])
def test_fast_reproducible(split):
    """Two runs at the same tile split are bitwise equal.  The split is forced: the
    autotuner times the splits, and two that time alike may be picked in turn, which
    changes the order of the tile's partial sums."""
    n = 700
    g = sl.generate_bpca(n, seed=5, v_sigma=3.0)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64)
    for prec in (sl.PREC_FP64, sl.PREC_MIXED):
        q = p.replace(order=sl.ORDER_FAST, precision=prec)
        outs = []
        for _ in range(2):
            with sl.Collider(n, q) as c:
                c.set_option("tile_split", split)
                c.upload(g)
                c.init_conditions()
                c.step(60)
                assert c.get_option("tile_split") == split
                outs.append(c.download())
        for k in ("pos", "vel", "w", "acc", "aacc"):
            assert bits_equal(getattr(outs[0], k), getattr(outs[1], k)), (prec, k)


def test_mode_switch_continues_state(oracle_mod):
    """ordered -> fast -> ordered on one context keeps the trajectory."""
    n = 300
    g = sl.generate_bpca(n, seed=77, v_sigma=2.0)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64)
    with sl.Collider(n, p) as c:
        c.upload(g)
        c.init_conditions()
        c.step(20)
        c.set_params(p.replace(order=sl.ORDER_FAST))
        c.step(20)
        c.set_params(p)
        c.step(20)
        out = c.download()
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    s, _, _ = oracle_mod.run(st, _oracle_params(oracle_mod, p), 61, 0, mode="gather",
                             nthreads=NTHREADS)
    assert relerr(out.pos, s.pos) <= 1e-12
    assert relerr(out.vel, s.vel) <= 1e-12


def test_many_partners_per_sphere(oracle_mod):
    """A large sphere pressed against by 150 small ones: more near partners than one
    gather chunk of the per-ball kernel (64), all of them touching."""
    rng = np.random.default_rng(3)
    k = 150
    Rb, Rs = 2e-4, 1e-5
    d = rng.standard_normal((k, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    # spread the small spheres over the surface so they do not overlap each other
    for _ in range(300):
        diff = d[:, None, :] - d[None, :, :]
        dist = np.linalg.norm(diff, axis=2) + np.eye(k)
        d += 0.002 * (diff / dist[:, :, None] ** 3).sum(1)
        d /= np.linalg.norm(d, axis=1)[:, None]
    ang = np.arccos(np.clip(d @ d.T - 2 * np.eye(k), -1, 1))
    assert ang.min() * (Rb + Rs) > 2.2 * Rs
    pos = np.vstack([[0, 0, 0], d * (Rb + Rs) * (1 - 1e-3)])
    R = np.array([Rb] + [Rs] * k)
    m = 2.25 * 4 / 3 * np.pi * R ** 3
    vel = rng.standard_normal(pos.shape) * 1e-3
    g = sl.BallGroup(pos, vel, np.zeros_like(pos), R, m, 0.4 * m * R * R)
    p = sl.calibrate(sl.preset("reference"), R, m, velocity=V_CUSTOM, cor=0.64).replace(
        order=sl.ORDER_FAST)
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    P = _oracle_params(oracle_mod, p)
    out1, e0, _ = _run(g, p, 1)
    s1, oe0, _ = oracle_mod.run(st, P, 2, 0, mode="gather", nthreads=NTHREADS)
    assert e0.contacts == k
    assert relerr(out1.acc, s1.acc) <= 1e-13
    assert relerr(out1.vel, s1.vel) <= 1e-14
    assert abs(e0.PE - oe0[0]) <= 1e-12 * abs(oe0[0])
    out, _, _ = _run(g, p, 20)
    s, _, _ = oracle_mod.run(st, P, 21, 0, mode="gather", nthreads=NTHREADS)
    assert relerr(out.pos, s.pos) <= 1e-12
    assert relerr(out.vel, s.vel) <= 1e-10


def test_fast_40000_vs_oracle(oracle_mod):
    """A 40,000-sphere aggregate: 625 ball blocks, so a wavefront of the per-ball kernel
    visits more than 64 partial slots and numbers its near entries by rescanning every
    slot's partner masks rather than from its 64-bit list of slots with partners.  The
    step-0 pass and one step against the oracle.  With this many contacts some touch by
    only a few 1e-6 of their radii, where the overlap S - d cancels: a one-ulp change of a
    drifted position (from the rounding of the forces that drifted it) changes the
    elastic force, and the friction torque proportional to it, by some 1e-11.  So each
    array's bar is the one-step bar above or ten times the oracle's own change when only
    the numbering of the spheres (and so the order of every sum) is permuted, whichever
    is larger.  Spins are zero in this step's pair pass, so renumbering, which swaps the
    roles A and B of some pairs, changes nothing but rounding.  Measured: 1.9e-14 acc,
    1.4e-14 vel, 6.5e-12 aacc and w; the GPU's differences are the same quantized
    changes, of the same size."""
    n = 40000
    assert (n + 63) // 64 > 64 * 8
    g = sl.generate_bpca(n, seed=0x5EED0040, v_sigma=1e-3)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64).replace(
        order=sl.ORDER_FAST)
    P = _oracle_params(oracle_mod, p)
    out1, e0, _ = _run(g, p, 1)
    s1, oe0, _ = oracle_mod.run(oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi), P, 2, 0,
                                mode="gather", nthreads=NTHREADS)
    perm = np.random.default_rng(0).permutation(n)
    sp, _, _ = oracle_mod.run(oracle_mod.State(g.pos[perm], g.vel[perm], g.w[perm], g.R[perm], g.m[perm],
                                               g.moi[perm]), P, 2, 0, mode="gather", nthreads=NTHREADS)
    inv = np.argsort(perm)
    assert e0.contacts > 0
    for k, bar in (("acc", 1e-13), ("vel", 1e-14), ("aacc", 1e-13), ("w", 1e-13), ("pos", 1e-15)):
        cond = relerr(getattr(sp, k)[inv], getattr(s1, k))
        assert relerr(getattr(out1, k), getattr(s1, k)) <= max(bar, 10 * cond), (k, cond)
    assert abs(e0.PE - oe0[0]) <= 1e-12 * abs(oe0[0])


def test_generator_aggregate_fast_and_mixed(oracle_mod):
    """The 1000-sphere aggregate the reference's Generator wrote (the first row of its
    simData file, tests/golden/generator/): three radii of 1481-4443 cm, masses up to
    6e11 g, every pair apart by at least 1e-4 of its radii -- a far-field input at
    astronomical scale.  Fast path: the step-0 pass and one step against the oracle to the
    one-step bars.  Mixed precision (FP32 far field in scaled units): energy within 1e-6
    of FP64 over 100 steps."""
    d = os.path.join(GOLDEN, "generator") + os.sep
    prefix = "1000-R29019.1-k1e+18-cor0.64-mu0.3-rho1.70-dt0.04_"
    consts = sl.BallGroup.from_csv(d, prefix)
    with open(d + prefix + "simData.csv", "rb") as f:
        pos, w, vel = sl.csv_parse_sim_row(f.read().decode().split("\n")[1])
    g = sl.BallGroup(pos, vel, w, consts.R, consts.m, consts.moi)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64).replace(
        order=sl.ORDER_FAST)
    out1, e0, _ = _run(g, p, 1)
    s1, oe0, _ = oracle_mod.run(oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi),
                                _oracle_params(oracle_mod, p), 2, 0, mode="gather", nthreads=NTHREADS)
    assert relerr(out1.acc, s1.acc) <= 1e-13
    assert relerr(out1.vel, s1.vel) <= 1e-14
    assert relerr(out1.pos, s1.pos) <= 1e-15
    assert abs(e0.PE - oe0[0]) <= 1e-12 * abs(oe0[0])
    _, e0f, ewf = _run(g, p, 100, 10)
    _, e0m, ewm = _run(g, p.replace(precision=sl.PREC_MIXED), 100, 10)
    for a, b in zip([e0m] + ewm, [e0f] + ewf):
        assert abs(a.E - b.E) <= 1e-6 * (abs(b.PE) + b.KE)


def test_sphere_with_4000_partners(oracle_mod):
    """A sphere pressed against by 4000 small ones spread over its whole surface (a
    Fibonacci lattice, neighbours 2.4e-8 cm apart, inside the h_min band).  In the
    engine's space-filling-curve order its partners lie in all 63 ball blocks, so each
    wavefront of the per-ball kernel meets them in more slots than it keeps partner
    masks for and numbers them by reading the masks again, over several rounds of 512
    evaluations in the big sphere's block.  The step-0 pass and one step against the
    oracle, to the one-step bars; spins are nonzero, so the friction terms that are not
    symmetric in A and B are exercised with the caller's orientation.  After one step the
    velocities are the kicks of the compressed contacts (about 0.7 cm/s against 1e-3 cm/s
    at the start), so their bar is the forces' (1e-13) rather than 1e-14."""
    k = 4000
    rs, Rb = 1e-5, 4e-4
    t = np.arange(k) + 0.5
    phi, th = np.arccos(1 - 2 * t / k), np.pi * (1 + 5 ** 0.5) * t
    d = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    pos = np.vstack([[0.0, 0.0, 0.0], d * (Rb + rs) * (1 - 1e-3)])
    R = np.array([Rb] + [rs] * k)
    m = 2.25 * 4 / 3 * np.pi * R ** 3
    rng = np.random.default_rng(11)
    vel = rng.standard_normal(pos.shape) * 1e-3
    w = rng.standard_normal(pos.shape) * 1e-2
    g = sl.BallGroup(pos, vel, w, R, m, 0.4 * m * R * R)
    p = sl.calibrate(sl.preset("config"), R, m, velocity=V_CUSTOM, cor=0.64).replace(order=sl.ORDER_FAST)
    out1, e0, _ = _run(g, p, 1)
    s1, oe0, _ = oracle_mod.run(oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi),
                                _oracle_params(oracle_mod, p), 2, 0, mode="gather", nthreads=NTHREADS)
    assert e0.contacts == k
    assert relerr(out1.acc, s1.acc) <= 1e-13
    assert relerr(out1.aacc, s1.aacc) <= 1e-13
    assert relerr(out1.vel, s1.vel) <= 1e-13
    assert relerr(out1.w, s1.w) <= 1e-13
    assert relerr(out1.pos, s1.pos) <= 1e-15
    assert abs(e0.PE - oe0[0]) <= 1e-12 * abs(oe0[0])


def test_benchmark_aggregate_5000_fast_and_mixed(oracle_mod):
    """The benchmark's own workload (the 5000-sphere aggregate, seed 0x5EED0001, config
    physics, dt = t_c/50) through the fast path: the step-0 pass and one step against the
    oracle to the one-step bars, and 100 steps to the 100-step bars; the mixed path's
    energy within 1e-6 of the fast FP64 path's over the same 100 steps."""
    n = 5000
    g = sl.generate_bpca(n, seed=0x5EED0001, v_sigma=1e-3)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64, dt_rule="tc50").replace(
        order=sl.ORDER_FAST)
    P = _oracle_params(oracle_mod, p)
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    out1, e0, _ = _run(g, p, 1)
    s1, oe0, _ = oracle_mod.run(st, P, 2, 0, mode="gather", nthreads=NTHREADS)
    assert relerr(out1.acc, s1.acc) <= 1e-13
    assert relerr(out1.aacc, s1.aacc) <= 1e-13
    assert relerr(out1.vel, s1.vel) <= 1e-14
    assert relerr(out1.pos, s1.pos) <= 1e-15
    assert abs(e0.PE - oe0[0]) <= 1e-12 * abs(oe0[0])
    out, _, ew = _run(g, p, 100, 50)
    s, _, oew = oracle_mod.run(st, P, 101, 50, mode="gather", nthreads=NTHREADS)
    assert relerr(out.pos, s.pos) <= 1e-12
    assert relerr(out.vel, s.vel) <= 1e-10
    for e, r in zip(ew, oew):
        assert abs(e.E - (r[0] + r[1])) <= 1e-12 * (abs(r[0]) + r[1])
    _, e0m, ewm = _run(g, p.replace(precision=sl.PREC_MIXED), 100, 50)
    for a, b in zip([e0m] + ewm, [e0] + ew):
        assert abs(a.E - b.E) <= 1e-6 * (abs(b.PE) + b.KE)


def test_mixed_against_oracle_5000(oracle_mod):
    """The mixed path against the oracle (the reference's single-core arithmetic) on the
    benchmark's own workload, the 5000-sphere aggregate: after one step the
    accelerations to 1e-6 (binary32 far field, binary64 contacts; measured 1.6e-7, spins
    3.6e-8), after 100 steps the positions to 1e-6 (measured 1.1e-7) and velocities to
    2e-3 (7.5e-4: the one-step difference grows with the aggregate's e-folding time of
    about five steps) of their largest values, and the energies at the write steps to
    1e-6 of the energy scale (1.5e-10 and 4.9e-8)."""
    n = 5000
    g = sl.generate_bpca(n, seed=0x5EED0001, v_sigma=1e-3)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64, dt_rule="tc50").replace(
        order=sl.ORDER_FAST, precision=sl.PREC_MIXED)
    P = _oracle_params(oracle_mod, p)
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    out1, e0, _ = _run(g, p, 1)
    s1, oe0, _ = oracle_mod.run(st, P, 2, 0, mode="gather", nthreads=NTHREADS)
    print("one step: acc", relerr(out1.acc, s1.acc), "aacc", relerr(out1.aacc, s1.aacc))
    assert relerr(out1.acc, s1.acc) <= 1e-6
    assert relerr(out1.aacc, s1.aacc) <= 1e-6
    assert abs(e0.PE - oe0[0]) <= 1e-6 * abs(oe0[0])
    out, _, ew = _run(g, p, 100, 50)
    s, _, oew = oracle_mod.run(st, P, 101, 50, mode="gather", nthreads=NTHREADS)
    print("100 steps: pos", relerr(out.pos, s.pos), "vel", relerr(out.vel, s.vel),
          "E", [abs(e.E - (r[0] + r[1])) / (abs(r[0]) + r[1]) for e, r in zip(ew, oew)])
    assert relerr(out.pos, s.pos) <= 1e-6
    assert relerr(out.vel, s.vel) <= 2e-3
    for e, r in zip(ew, oew):
        assert abs(e.E - (r[0] + r[1])) <= 1e-6 * (abs(r[0]) + r[1])


@pytest.mark.parametrize("split", [1, 2, 4])
def test_mixed_one_step_every_split(oracle_mod, split):
    """The mixed mode's tile partials are binary32: the SPLIT wavefronts sweeping a tile
    add their sums in binary32, so each split rounds differently.  At every split the
    step-0 forces and one step's accelerations stay within 1e-6 of the oracle's (measured
    1.6e-7 at split 4 on the 5000-sphere aggregate), with gravity on all pairs."""
    n = 1500
    g = sl.generate_bpca(n, seed=0x5EED000B, v_sigma=1e-3)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64, dt_rule="tc50").replace(
        order=sl.ORDER_FAST, precision=sl.PREC_MIXED)
    with sl.Collider(n, p) as c:
        c.set_option("tile_split", split)
        c.upload(g)
        c.init_conditions()
        a0 = c.download()
        c.step(1)
        assert c.get_option("tile_split") == split
        a1 = c.download()
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    s0, _, _ = oracle_mod.run(st, _oracle_params(oracle_mod, p), 1, 0, mode="gather", nthreads=NTHREADS)
    s1, _, _ = oracle_mod.run(st, _oracle_params(oracle_mod, p), 2, 0, mode="gather", nthreads=NTHREADS)
    assert relerr(a0.acc, s0.acc) <= 1e-6
    assert relerr(a1.acc, s1.acc) <= 1e-6
    assert relerr(a1.aacc, s1.aacc) <= 1e-6
    assert relerr(a1.pos, s1.pos) <= 1e-10


_REPEAT_JOB = r'''
import hashlib, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import spacelab_amd as sl
g = sl.generate_bpca(700, seed=5, v_sigma=3.0)
p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=0.36301555459799423, cor=0.64).replace(
    order=sl.ORDER_FAST)
with sl.Collider(700, p) as c:
    c.upload(g)
    c.init_conditions()
    c.step(50, 0)
    st = c.download()
    split = c.get_option("tile_split")
h = hashlib.sha256(np.concatenate([st.pos.ravel(), st.vel.ravel(), st.w.ravel()]).tobytes()).hexdigest()
print(split, h)
'''


def test_default_runs_repeat_bitwise_through_tune_cache(tmp_path):
    """Two processes run the same fast-path job with the tile split left to the
    autotuner: the first tunes and stores its choice in the tune cache, the second reads
    it back, so the two repeat bit for bit without forcing a split."""
    import json
    import subprocess
    import sys
    env = dict(os.environ, SPACELAB_TUNE_CACHE=str(tmp_path / "tune.json"))
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    outs = []
    for _ in range(2):
        r = subprocess.run([sys.executable, "-c", _REPEAT_JOB, repo], env=env, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        outs.append(r.stdout.split())
    assert outs[0] == outs[1]
    cache = json.loads((tmp_path / "tune.json").read_text())
    # the stored choice: the split, plus 8 when the tiles come from the work queue
    tiles = [v["split"] & 7 for k, v in cache.items() if k.startswith("tiles:prec0:")]
    assert tiles == [int(outs[0][0])]


@pytest.mark.parametrize("prec", [sl.PREC_FP64, sl.PREC_MIXED])
def test_tile_queue_launch_same_bits(prec):
    """The work-queue launch of the tile kernel (as many workgroups as the device holds,
    pulling tiles) gives the bits of the one-workgroup-per-tile launch, at every split,
    over steps with and without energies; and two queue runs in a row (the queue rewinds
    itself after each launch) agree too."""
    n = 1500
    g = sl.generate_bpca(n, seed=21, v_sigma=3.0)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64).replace(
        order=sl.ORDER_FAST, precision=prec)
    for split in (1, 2, 4):
        outs = []
        for queue in (1, 2, 2):
            with sl.Collider(n, p) as c:
                c.set_option("tile_split", split)
                c.set_option("tile_queue", queue)
                c.upload(g)
                e0 = c.init_conditions()
                ew = c.step(40, 10)
                assert c.get_option("tile_queue") == queue
                st = c.download()
            outs.append((st, [e0] + ew))
        for st, ee in outs[1:]:
            for k in ("pos", "vel", "w", "acc", "aacc"):
                assert bits_equal(getattr(st, k), getattr(outs[0][0], k)), (split, k)
            for a, b in zip(ee, outs[0][1]):
                assert a.PE == b.PE and a.KE == b.KE and a.contacts == b.contacts

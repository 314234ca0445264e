"""GPU parity of the ordered (bitwise) path, through the C ABI.

Bar: the ordered FP64 path reproduces the reference's single-core arithmetic bit
for bit in every state array (positions, velocities, half-step velocities,
spins, accelerations).  Energies agree to rounding (1e-13 relative; the GPU
sums per block in a fixed order, the reference in one running sum).
"""
import os

import numpy as np
import pytest

from conftest import (angmom_scale, bits_equal, golden_names, load_golden,
                      momentum_scale)

import spacelab_amd as sl

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)


def _params_ref(kin, kout, dt, **kw):
    return sl.preset("reference").replace(kin=float(kin), kout=float(kout), dt=float(dt), **kw)


def _oracle_params(oracle, p):
    return oracle.Params(G=p.G, Ha=p.Ha, h_min=p.h_min, u_s=p.u_s, u_r=p.u_r, kin=p.kin,
                         kout=p.kout, dt=p.dt, far_gravity=p.far_gravity)


def _energy_close(got, ref, m, moi, pos, vel, w, rtol=1e-13):
    assert abs(got.PE - ref[0]) <= rtol * abs(ref[0]) + 1e-300
    assert abs(got.KE - ref[1]) <= rtol * abs(ref[1]) + 1e-300
    ps = momentum_scale(m, vel)
    Ls = angmom_scale(m, moi, pos, vel, w)
    assert np.max(np.abs(got.p - ref[2:5])) <= rtol * ps + 1e-300
    assert np.max(np.abs(got.L - ref[5:8])) <= rtol * Ls + 1e-300


def test_device_division_sqrt_correctly_rounded():
    """Bitwise parity needs IEEE-correct binary64 a/b and sqrt(a) on the device."""
    rng = np.random.default_rng(11)
    n = 1 << 20
    a = np.exp(rng.uniform(-230, 230, n)) * rng.choice([-1.0, 1.0], n)
    b = np.exp(rng.uniform(-230, 230, n)) * rng.choice([-1.0, 1.0], n)
    a_pos = np.abs(a)
    q, s, lg = sl.probe_divsqrt(a_pos, b)
    assert bits_equal(q, a_pos / b)
    assert bits_equal(s, np.sqrt(a_pos))
    # log is only used for the vdW energy; it need not be correctly rounded
    assert np.max(np.abs(lg - np.log(a_pos)) / np.abs(np.log(a_pos))) < 1e-15


def _adversarial_divisors(rng, n):
    """Divisors whose significand is all ones or nearly (the hardest reciprocals), and
    quotients that land on or next to exact values."""
    e = rng.integers(-200, 200, n)
    sig = (np.uint64((1 << 52) - 1) - rng.integers(0, 256, n).astype(np.uint64))
    bits = ((e + 1023).astype(np.uint64) << np.uint64(52)) | sig
    return bits.view(np.float64) * rng.choice([-1.0, 1.0], n)


def test_device_division_by_reciprocal_correctly_rounded():
    """The ordered kernels divide by a distance or a mass through the correctly rounded
    reciprocal and two FMA residual corrections; bitwise parity needs the result to be
    the correctly rounded quotient.  Random operands over 2^+-330, divisors with
    all-ones significands, and quotients that are exact or one ulp off exact."""
    rng = np.random.default_rng(12)
    n = 1 << 20
    a = np.exp(rng.uniform(-230, 230, n)) * rng.choice([-1.0, 1.0], n)
    b = np.exp(rng.uniform(-230, 230, n)) * rng.choice([-1.0, 1.0], n)
    assert bits_equal(sl.probe_div_rcp(a, b), a / b)
    b2 = _adversarial_divisors(rng, n)
    a2 = np.exp(rng.uniform(-200, 200, n)) * rng.choice([-1.0, 1.0], n)
    assert bits_equal(sl.probe_div_rcp(a2, b2), a2 / b2)
    # exact quotients and their neighbours
    qx = np.round(rng.uniform(-1e6, 1e6, n)) * 2.0 ** rng.integers(-40, 40, n)
    a3 = qx * b
    a3 = np.where(rng.random(n) < 0.5, a3, np.nextafter(a3, np.inf))
    assert bits_equal(sl.probe_div_rcp(a3, b), a3 / b)


@pytest.mark.parametrize("name", golden_names())
def test_golden_bitwise(name):
    z = load_golden(name)
    dt, kin, kout = z["params"]
    g = sl.BallGroup(z["pos"], z["vel"], z["w"], z["R"], z["m"], z["moi"])
    with sl.Collider(len(z["R"]), _params_ref(kin, kout, dt)) as c:
        c.upload(g)
        e0 = c.init_conditions()
        ew = c.sim_looper(int(z["n_steps"]), int(z["skip"]))
        out = c.download()
    for k in ("pos", "vel", "velh", "w", "wh", "acc", "aacc"):
        assert bits_equal(getattr(out, k), z["out_" + k]), f"{name}: {k}"
    _energy_close(e0, z["e_init"], z["m"], z["moi"], z["pos"], z["vel"], z["w"])
    assert len(ew) == len(z["e_writes"])
    for e, r in zip(ew, z["e_writes"]):
        _energy_close(e, r, z["m"], z["moi"], out.pos, out.vel, out.w)


@pytest.mark.parametrize("far_gravity", [0, 1])
@pytest.mark.parametrize("v_sigma", [1e-3, 5.0])
def test_bpca_1000_bitwise_100_steps(oracle_mod, far_gravity, v_sigma):
    """Aggregate of 1000 spheres (lognormal radii), 100 steps: bitwise vs the oracle."""
    n = 1000
    g = sl.generate_bpca(n, seed=0x5EED0000, v_sigma=v_sigma)
    base = sl.preset("config" if far_gravity else "reference")
    p = sl.calibrate(base, g.R, g.m, velocity=0.36301555459799423, cor=0.64)
    with sl.Collider(n, p) as c:
        c.upload(g)
        e0 = c.init_conditions()
        ew = c.step(100, 50)
        out = c.download()
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    s, oe0, oew = oracle_mod.run(st, _oracle_params(oracle_mod, p), 101, 50, mode="gather",
                                 nthreads=NTHREADS)
    for k in ("pos", "vel", "velh", "w", "wh", "acc", "aacc"):
        assert bits_equal(getattr(out, k), getattr(s, k)), k
    _energy_close(e0, oe0, g.m, g.moi, g.pos, g.vel, g.w, rtol=1e-12)
    for e, r in zip(ew, oew):
        _energy_close(e, r, g.m, g.moi, out.pos, out.vel, out.w, rtol=1e-12)
    assert e0.contacts >= 0


def test_benchmark_aggregate_5000_bitwise_100_steps(oracle_mod):
    """The benchmark's own workload: the 5000-sphere aggregate (seed 0x5EED0001, config
    physics with gravity on every pair, dt = t_c/50), 100 steps through the ordered path:
    every state array bitwise the oracle's, energies at the write steps to 1e-12."""
    n = 5000
    g = sl.generate_bpca(n, seed=0x5EED0001, v_sigma=1e-3)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=0.36301555459799423, cor=0.64, dt_rule="tc50")
    with sl.Collider(n, p) as c:
        c.upload(g)
        e0 = c.init_conditions()
        ew = c.step(100, 50)
        out = c.download()
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    s, oe0, oew = oracle_mod.run(st, _oracle_params(oracle_mod, p), 101, 50, mode="gather",
                                 nthreads=NTHREADS)
    for k in ("pos", "vel", "velh", "w", "wh", "acc", "aacc"):
        assert bits_equal(getattr(out, k), getattr(s, k)), k
    _energy_close(e0, oe0, g.m, g.moi, g.pos, g.vel, g.w, rtol=1e-12)
    assert len(ew) == 2
    for e, r in zip(ew, oew):
        _energy_close(e, r, g.m, g.moi, out.pos, out.vel, out.w, rtol=1e-12)


def test_energy_10k_steps(oracle_mod):
    """10k steps of a 200-sphere aggregate: states bitwise, E/p/L within 1e-9."""
    n = 200
    g = sl.generate_bpca(n, seed=0x5EED0001, v_sigma=2.0)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=0.36301555459799423, cor=0.64)
    with sl.Collider(n, p) as c:
        c.upload(g)
        c.init_conditions()
        ew = c.step(10000, 1000)
        out = c.download()
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    s, _, oew = oracle_mod.run(st, _oracle_params(oracle_mod, p), 10001, 1000, mode="gather",
                               nthreads=NTHREADS)
    for k in ("pos", "vel", "w"):
        assert bits_equal(getattr(out, k), getattr(s, k)), k
    assert len(ew) == 10
    for e, r in zip(ew, oew):
        assert abs(e.E - (r[0] + r[1])) <= 1e-9 * (abs(r[0]) + r[1])
        _energy_close(e, r, g.m, g.moi, out.pos, out.vel, out.w, rtol=1e-9)


def test_reproducible_run_to_run():
    n = 500
    g = sl.generate_bpca(n, seed=99, v_sigma=3.0)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=0.36301555459799423, cor=0.64)
    outs = []
    for _ in range(2):
        with sl.Collider(n, p) as c:
            c.upload(g)
            c.init_conditions()
            c.step(50, 0)
            outs.append(c.download())
    for k in ("pos", "vel", "w", "acc", "aacc"):
        assert bits_equal(getattr(outs[0], k), getattr(outs[1], k))


def test_errors_are_codes_not_crashes():
    p = _params_ref(1.0, 0.4, 1e-9)
    with sl.Collider(8, p) as c:
        with pytest.raises(sl.SpacelabError) as ei:
            c.step(1)
        assert ei.value.code == -3                       # before upload
        g = sl.generate_bpca(8, seed=1)
        bad = g.copy()
        bad.R[3] = 0.0
        with pytest.raises(sl.SpacelabError) as ei:
            c.upload(bad)                                # R <= 0
        assert ei.value.code == -1
        big = sl.generate_bpca(9, seed=1)
        with pytest.raises(sl.SpacelabError) as ei:
            c.upload(big)                                # n > n_max
        assert ei.value.code == -1
        c.upload(g)
        with pytest.raises(sl.SpacelabError) as ei:
            c.step(1)                                    # no forces yet
        assert ei.value.code == -3


@pytest.mark.parametrize("order", [sl.ORDER_ORDERED, sl.ORDER_FAST])
def test_coincident_spheres_flag_nonfinite(order):
    """dist = 0 gives NaN forces in the reference too; the library reports it, in both
    paths (no index depends on the values, so NaN only propagates through the sums)."""
    R = np.array([1e-5, 1e-5, 1e-5])
    m = 2.25 * 4 / 3 * np.pi * R ** 3
    pos = np.array([[0.0, 0, 0], [0.0, 0, 0], [5e-5, 0, 0]])
    g = sl.BallGroup(pos, np.zeros((3, 3)), np.zeros((3, 3)), R, m, 0.4 * m * R * R)
    p = _params_ref(1.0, 0.4, 1e-9, order=order)
    with sl.Collider(3, p) as c:
        c.upload(g)
        c.init_conditions()
        e = c.step(1, 1)[0]
    assert e.nonfinite


def test_contact_count_at_write_steps():
    """The write-step energies count the touching pairs: 150 small spheres pressed onto a
    large one (each overlapping it by 1e-3 of the radii sum) give 150 contacts at the
    step-0 pass and at the next write steps, in the ordered path and in the fast path."""
    rng = np.random.default_rng(3)
    k = 150
    Rb, Rs = 2e-4, 1e-5
    d = rng.standard_normal((k, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    for _ in range(300):
        diff = d[:, None, :] - d[None, :, :]
        dist = np.linalg.norm(diff, axis=2) + np.eye(k)
        d += 0.002 * (diff / dist[:, :, None] ** 3).sum(1)
        d /= np.linalg.norm(d, axis=1)[:, None]
    pos = np.vstack([[0, 0, 0], d * (Rb + Rs) * (1 - 1e-3)])
    R = np.array([Rb] + [Rs] * k)
    m = 2.25 * 4 / 3 * np.pi * R ** 3
    g = sl.BallGroup(pos, np.zeros_like(pos), np.zeros_like(pos), R, m, 0.4 * m * R * R)
    p = sl.calibrate(sl.preset("reference"), R, m, velocity=0.36301555459799423, cor=0.64)
    for order in (sl.ORDER_ORDERED, sl.ORDER_FAST):
        with sl.Collider(k + 1, p.replace(order=order)) as c:
            c.upload(g)
            e0 = c.init_conditions()
            ew = c.step(3, 1)
        assert e0.contacts == k, order
        assert len(ew) == 3 and all(e.contacts == k for e in ew), (order, [e.contacts for e in ew])


def _shape_count():
    """Number of launch shapes of the ordered pass: the largest value the option accepts."""
    g = sl.generate_bpca(8, seed=3)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=0.36301555459799423, cor=0.64)
    with sl.Collider(8, p) as c:
        k = 1
        while True:
            try:
                c.set_option("ordered_shape", k + 1)
            except sl.SpacelabError:
                return k
            k += 1
            assert k < 64


def test_ordered_shapes_bitwise(oracle_mod):
    """Every launch shape of the ordered pass (balls per workgroup, chunk width, contact
    arm inline or not, register cap) gives the oracle's bits: 700 spheres (a multiple of
    none of the shapes' ball counts, several partial chunks), contacts from v_sigma = 5,
    gravity on every pair, 30 steps with energies every 10."""
    n = 700
    g = sl.generate_bpca(n, seed=0x5EED0007, v_sigma=5.0)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=0.36301555459799423, cor=0.64)
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    s, _, oew = oracle_mod.run(st, _oracle_params(oracle_mod, p), 31, 10, mode="gather",
                               nthreads=NTHREADS)
    shapes = _shape_count()
    assert shapes >= 2
    ref_e = None
    for shape in range(1, shapes + 1):
        with sl.Collider(n, p) as c:
            c.set_option("ordered_shape", shape)
            c.upload(g)
            c.init_conditions()
            ew = c.step(30, 10)
            out = c.download()
            assert c.get_option("ordered_shape") == shape
        for k in ("pos", "vel", "velh", "w", "wh", "acc", "aacc"):
            assert bits_equal(getattr(out, k), getattr(s, k)), (shape, k)
        assert len(ew) == 3
        e = np.array([[x.PE, x.KE, x.contacts] for x in ew])
        if ref_e is None:
            ref_e = e
        # the energies are per-ball chains too: the same bits for every shape
        assert np.array_equal(e.view(np.uint64), ref_e.view(np.uint64)), shape
        for x, r in zip(ew, oew):
            _energy_close(x, r, g.m, g.moi, out.pos, out.vel, out.w, rtol=1e-12)
        assert ew[-1].contacts > 0


def test_ordered_graphs_same_bits():
    """Runs of ordered steps between write steps replay captured graphs, in which each
    closing kick but the run's last is fused with the next step's drift; with graphs off
    every kernel is launched directly.  Both give the same bits in every state array and
    the same energies, over runs that end on and between write steps and across calls of
    different lengths (so several captured lengths are replayed)."""
    n = 300
    g = sl.generate_bpca(n, seed=77, v_sigma=5.0)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=0.36301555459799423, cor=0.64)
    outs = []
    for graphs in (0, 1):
        with sl.Collider(n, p) as c:
            c.set_option("graphs", graphs)
            c.upload(g)
            c.init_conditions()
            ew = c.step(130, 40) + c.step(75, 40) + c.step(3, 0) + c.step(200, 100)
            out = c.download()
        outs.append((out, ew))
    for k in ("pos", "vel", "velh", "w", "wh", "acc", "aacc"):
        assert bits_equal(getattr(outs[0][0], k), getattr(outs[1][0], k)), k
    assert len(outs[0][1]) == len(outs[1][1]) == 7
    for a, b in zip(outs[0][1], outs[1][1]):
        assert a.PE == b.PE and a.KE == b.KE and a.contacts == b.contacts
        assert np.array_equal(a.p, b.p) and np.array_equal(a.L, b.L)

"""Sizes at the edges of the engines' blocking, against the oracle.

- One sphere (no pair at all), two spheres, one short of a 64-sphere block, exactly one
  block, one over, and two blocks plus one: the fast path pads every system to whole
  64-sphere blocks, so these put padded lanes in the diagonal tile, in a full tile and
  in a tile of one real sphere.
- 7,519, 7,520 and 15,040 spheres: for the ordered pair pass's launch shapes (2, 4 or 8
  spheres per workgroup) a ragged last workgroup (7,519) and whole ones, with the last
  partner chunk cut short; the first step times every shape on them.
- The 20-step cases replay captured graphs of 18 ordered steps, in which each closing kick
  but the last is fused with the next step's drift.

The ordered path is bitwise the oracle at every size.  The fast path meets its one-step
bars (forces and torques to 1e-13 relative, velocities to 1e-14, positions to 1e-15) and,
after 20 steps, positions to 1e-12 and velocities to 1e-10; with one sphere it has no pair
to evaluate and is bitwise the oracle.
"""
import os

import numpy as np
import pytest

from conftest import bits_equal

import spacelab_amd as sl

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
V_CUSTOM = 0.36301555459799423
FIELDS = ("pos", "vel", "velh", "w", "wh", "acc", "aacc")


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _oracle_params(oracle, p):
    return oracle.Params(G=p.G, Ha=p.Ha, h_min=p.h_min, u_s=p.u_s, u_r=p.u_r, kin=p.kin,
                         kout=p.kout, dt=p.dt, far_gravity=p.far_gravity)


def _case(n):
    g = sl.generate_bpca(n, seed=0x5EED0200 + n, v_sigma=1.0)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64)
    return g, p


def _run(g, p, steps, write_every):
    with sl.Collider(g.num_particles, p) as c:
        c.upload(g)
        e0 = c.init_conditions()
        ew = c.step(steps, write_every)
        return c.download(), e0, ew


def _pe_close(e, ref_pe, rtol):
    assert abs(e.PE - ref_pe) <= rtol * abs(ref_pe) + 1e-300, (e.PE, ref_pe)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 129, 7519, 7520, 15040])
def test_ordered_sizes_bitwise(oracle_mod, n):
    g, p = _case(n)
    steps = 20 if n < 1000 else 3 if n < 10000 else 2
    print(f"ordered n={n} steps={steps}", flush=True)
    out, e0, ew = _run(g, p, steps, steps)
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    s, oe0, oew = oracle_mod.run(st, _oracle_params(oracle_mod, p), steps + 1, steps,
                                 mode="gather", nthreads=NTHREADS)
    for k in FIELDS:
        assert bits_equal(getattr(out, k), getattr(s, k)), (n, k)
    _pe_close(e0, oe0[0], 1e-12)
    assert len(ew) == len(oew) == 1
    _pe_close(ew[0], oew[0][0], 1e-12)
    if n == 1:
        assert e0.PE == 0.0 and e0.contacts == 0
        assert np.all(out.acc == 0) and np.all(out.aacc == 0)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 129])
def test_fast_sizes_vs_oracle(oracle_mod, n):
    g, p = _case(n)
    q = p.replace(order=sl.ORDER_FAST)
    print(f"fast n={n}", flush=True)
    st = oracle_mod.State(g.pos, g.vel, g.w, g.R, g.m, g.moi)
    P = _oracle_params(oracle_mod, q)
    out1, e0, _ = _run(g, q, 1, 0)
    s1, oe0, _ = oracle_mod.run(st, P, 2, 0, mode="gather", nthreads=NTHREADS)
    if n == 1:
        for k in FIELDS:
            assert bits_equal(getattr(out1, k), getattr(s1, k)), k
        assert e0.PE == 0.0
        return
    assert relerr(out1.acc, s1.acc) <= 1e-13
    assert relerr(out1.aacc, s1.aacc) <= 1e-13 or np.max(np.abs(s1.aacc)) == 0
    assert relerr(out1.vel, s1.vel) <= 1e-14
    assert relerr(out1.pos, s1.pos) <= 1e-15
    _pe_close(e0, oe0[0], 1e-12)
    out, _, _ = _run(g, q, 20, 0)
    s, _, _ = oracle_mod.run(st, P, 21, 0, mode="gather", nthreads=NTHREADS)
    assert relerr(out.pos, s.pos) <= 1e-12
    assert relerr(out.vel, s.vel) <= 1e-10

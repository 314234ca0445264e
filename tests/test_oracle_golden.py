"""The CPU oracle against the reference's own outputs (golden vectors).

The fixtures under tests/golden were written by the reference's sim_one_step
(ColliderSingleCore/ColliderSingleCore.cpp:78-288) compiled from its sources
(tests/golden/make_golden.py).  The oracle must reproduce every state array
bit for bit, in both its triangular form (the reference's loop order) and its
ordered-gather form (the GPU's), and the write-step energies.
"""
import numpy as np
import pytest

from conftest import bits_equal, golden_names, load_golden, momentum_scale

NAMES = golden_names()


def _run(oracle, z, mode):
    st = oracle.State(z["pos"], z["vel"], z["w"], z["R"], z["m"], z["moi"])
    dt, kin, kout = z["params"]
    P = oracle.Params(kin=float(kin), kout=float(kout), dt=float(dt))
    return oracle.run(st, P, int(z["n_steps"]), int(z["skip"]), mode=mode)


def test_fixture_set_complete():
    # head-on, oblique with spin, unequal radii, 16- and 64-sphere clusters
    assert len(NAMES) >= 9
    for stem in ("two_headon", "two_oblique_spin", "three_unequal", "cluster16", "cluster64"):
        assert any(n.startswith(stem) for n in NAMES)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("mode", ["tri", "gather"])
def test_state_bitwise(oracle_mod, name, mode):
    z = load_golden(name)
    s, e_init, ew = _run(oracle_mod, z, mode)
    for k in ("pos", "vel", "velh", "w", "wh", "acc", "aacc"):
        assert bits_equal(getattr(s, k), z["out_" + k]), f"{name} {mode}: {k} differs"


@pytest.mark.parametrize("name", NAMES)
def test_energies_tri_bitwise(oracle_mod, name):
    z = load_golden(name)
    _, e_init, ew = _run(oracle_mod, z, "tri")
    assert bits_equal(e_init, z["e_init"])
    assert bits_equal(ew, z["e_writes"])


@pytest.mark.parametrize("name", NAMES)
def test_energies_gather_close(oracle_mod, name):
    # The gather sums PE per row and then over rows instead of one running sum.
    z = load_golden(name)
    _, e_init, ew = _run(oracle_mod, z, "gather")
    ref = np.vstack([z["e_init"][None, :], z["e_writes"]])
    got = np.vstack([e_init[None, :], ew])
    assert np.allclose(got[:, 0], ref[:, 0], rtol=1e-13, atol=0)
    assert bits_equal(got[:, 1:], ref[:, 1:])  # KE, p, L use the same per-ball order


def test_goldens_cover_contact_hysteresis(oracle_mod):
    """The head-on case must load (kin) and unload (kout) a contact."""
    z = load_golden("two_headon_s200_k10")
    st = oracle_mod.State(z["pos"], z["vel"], z["w"], z["R"], z["m"], z["moi"])
    dt, kin, kout = z["params"]
    P = oracle_mod.Params(kin=float(kin), kout=float(kout), dt=float(dt))
    s, _, _ = oracle_mod.run(st, P, 1, 1, mode="gather")
    prev = None
    loads = unloads = 0
    for _ in range(200):
        d = float(np.linalg.norm(s.pos[1] - s.pos[0]))
        if d < z["R"].sum() and prev is not None:
            loads += d < prev
            unloads += d >= prev
        prev = d
        s.pos_prev = None
        s, _, _ = oracle_mod.run(s, P, 2, 1, mode="gather", do_init=False)
    assert loads > 10 and unloads > 10
    p_scale = momentum_scale(z["m"], z["vel"])
    assert p_scale > 0

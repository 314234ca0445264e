"""Batched ensembles and captured step graphs of the fast engine.

A system inside a batch is laid out on its own 64-aligned slots and its pair tiles,
partial sums and energy blocks are the same as when it runs alone, so at equal tile
split its trajectory and energies are bitwise those of a single-system run.  Replaying
captured step graphs launches the same kernels as direct launches, so both agree
bitwise too.  The mixed path scales its FP32 far field by the median radius and mass
of all balls in the context, which differs between a batch and its parts, so it is
compared with FP64 to the mixed path's tolerance instead.
"""
import numpy as np
import pytest

from conftest import bits_equal

import spacelab_amd as sl

pytestmark = pytest.mark.gpu

V_CUSTOM = 0.36301555459799423
FIELDS = ("pos", "vel", "w", "acc", "aacc")


def _params(g, velocity=V_CUSTOM, precision=sl.PREC_FP64, preset="config"):
    p = sl.calibrate(sl.preset(preset), g.R, g.m, velocity=velocity, cor=0.64)
    return p.replace(order=sl.ORDER_FAST, precision=precision)


def _single(g, p, steps, write_every, split, graphs=1):
    with sl.Collider(g.num_particles, p) as c:
        c.set_option("tile_split", split)
        c.set_option("graphs", graphs)
        c.upload(g)
        e0 = c.init_conditions()
        ew = c.step(steps, write_every)
        return c.download(), e0, ew, c.get_option("tile_split")


def _ebytes(e):
    return e.as_array().tobytes() + bytes([e.nonfinite]) + int(e.contacts).to_bytes(8, "little")


@pytest.mark.parametrize("split", [1, 2, 4])
def test_graphs_match_direct_launches(split):
    g = sl.generate_bpca(700, seed=0x5EED0101, v_sigma=2.0)
    p = _params(g)
    a, ea0, eaw, sa = _single(g, p, 300, 64, split, graphs=1)
    b, eb0, ebw, sb = _single(g, p, 300, 64, split, graphs=0)
    assert sa == sb == split
    for k in FIELDS:
        assert bits_equal(getattr(a, k), getattr(b, k)), k
    assert _ebytes(ea0) == _ebytes(eb0)
    assert len(eaw) == len(ebw) == 300 // 64
    for x, y in zip(eaw, ebw):
        assert _ebytes(x) == _ebytes(y)


def test_graph_segments_across_calls():
    """Calls of different lengths (odd and even, with and without write steps) give
    the same state as one long call."""
    g = sl.generate_bpca(400, seed=0x5EED0102, v_sigma=2.0)
    p = _params(g)
    one, _, ew1, _ = _single(g, p, 517, 100, 2)
    with sl.Collider(g.num_particles, p) as c:
        c.set_option("tile_split", 2)
        c.upload(g)
        c.init_conditions()
        ew = []
        for n in (1, 2, 3, 130, 7, 257, 117):
            ew += c.step(n, 100)
        assert c.step_count == 517
        out = c.download()
    for k in FIELDS:
        assert bits_equal(getattr(out, k), getattr(one, k)), k
    assert [_ebytes(e) for e in ew] == [_ebytes(e) for e in ew1]


def test_graph_cache_eviction():
    """More distinct segment lengths than the graph cache holds (16): the cache is
    emptied while earlier segments may still run, and the state equals a run with graphs
    off, bit for bit."""
    g = sl.generate_bpca(300, seed=0x5EED0103, v_sigma=2.0)
    p = _params(g)
    lengths = list(range(3, 43, 2))          # 20 lengths, each a graph of its own
    outs = []
    for graphs in (1, 0):
        with sl.Collider(g.num_particles, p) as c:
            c.set_option("tile_split", 2)
            c.set_option("graphs", graphs)
            c.upload(g)
            c.init_conditions()
            for n in lengths + lengths[::-1]:
                c.step(n, 0)
            outs.append(c.download())
    for k in FIELDS:
        assert bits_equal(getattr(outs[0], k), getattr(outs[1], k)), k


def _ensemble(sizes, seed0):
    groups = [sl.generate_bpca(n, seed=seed0 + k, v_sigma=1.0 + k) for k, n in enumerate(sizes)]
    params = [_params(g, velocity=V_CUSTOM * (1 + 0.5 * k)) for k, g in enumerate(groups)]
    return groups, params


@pytest.mark.parametrize("sizes", [[65, 700, 1, 333, 128], [64, 2, 1000]])
def test_batch_equals_single_runs(sizes):
    groups, params = _ensemble(sizes, 0x5EED0200)
    steps, we, split = 150, 50, 2
    with sl.Batch(len(sizes), sum(sizes), params[0]) as b:
        b.set_option("tile_split", split)
        b.upload(groups, kin=[p.kin for p in params], kout=[p.kout for p in params],
                 dt=[p.dt for p in params])
        e0 = b.init_conditions()
        ew = b.step(steps, we)
        assert b.step_count == steps
        out = b.download()
    assert len(ew) == steps // we and all(len(row) == len(sizes) for row in ew)
    for k, (g, p) in enumerate(zip(groups, params)):
        s, se0, sew, _ = _single(g, p, steps, we, split)
        for f in FIELDS:
            assert bits_equal(getattr(out[k], f), getattr(s, f)), (k, f)
        assert _ebytes(e0[k]) == _ebytes(se0), k
        for w in range(len(sew)):
            assert _ebytes(ew[w][k]) == _ebytes(sew[w]), (k, w)


def test_batch_reupload_and_system_params():
    """A second upload with another layout, and new per-system dt mid-run."""
    groups, params = _ensemble([300, 90], 0x5EED0300)
    with sl.Batch(4, 2000, params[0]) as b:
        b.set_option("tile_split", 1)
        b.upload([sl.generate_bpca(1500, seed=1)])
        b.init_conditions()
        b.step(5)
        b.upload(groups, kin=[p.kin for p in params], kout=[p.kout for p in params],
                 dt=[p.dt for p in params])
        b.init_conditions()
        b.step(20)
        b.set_system_params(dt=[p.dt * 0.5 for p in params])
        b.step(20)
        out = b.download()
    for k, (g, p) in enumerate(zip(groups, params)):
        with sl.Collider(g.num_particles, p) as c:
            c.set_option("tile_split", 1)
            c.upload(g)
            c.init_conditions()
            c.step(20)
            c.set_params(p.replace(dt=p.dt * 0.5))
            c.step(20)
            s = c.download()
        for f in FIELDS:
            assert bits_equal(getattr(out[k], f), getattr(s, f)), (k, f)


def test_batch_mixed_energy_close_to_fp64():
    groups, params = _ensemble([800, 200, 50], 0x5EED0400)
    kin = [p.kin for p in params]
    kout = [p.kout for p in params]
    dt = [p.dt for p in params]
    res = {}
    for prec in (sl.PREC_FP64, sl.PREC_MIXED):
        with sl.Batch(3, 1050, params[0].replace(precision=prec)) as b:
            b.upload(groups, kin=kin, kout=kout, dt=dt)
            res[prec] = (b.init_conditions(), b.step(100, 10))
    (e0d, ewd), (e0m, ewm) = res[sl.PREC_FP64], res[sl.PREC_MIXED]
    for k in range(3):
        scale = abs(e0d[k].PE) + e0d[k].KE
        assert abs(e0m[k].E - e0d[k].E) <= 1e-6 * scale
        for a, b_ in zip(ewm, ewd):
            assert abs(a[k].E - b_[k].E) <= 1e-6 * (abs(b_[k].PE) + b_[k].KE)


def test_autotuner_picks_a_split():
    groups, params = _ensemble([500, 250], 0x5EED0500)
    with sl.Batch(2, 750, params[0]) as b:
        assert b.get_option("tile_split") == 0
        b.upload(groups)
        b.init_conditions()
        b.step(10)
        assert b.get_option("tile_split") in (1, 2, 4)


def test_batch_errors():
    g = sl.generate_bpca(100, seed=3)
    p = _params(g)
    with pytest.raises(sl.SpacelabError) as ei:
        sl.Batch(2, 100, p.replace(order=sl.ORDER_ORDERED))
    assert ei.value.code == -1
    with sl.Batch(2, 150, p) as b:
        with pytest.raises(sl.SpacelabError) as ei:
            b.upload([g, g, g])
        assert ei.value.code == -1
        with pytest.raises(sl.SpacelabError) as ei:
            b.upload([g, g])            # 200 balls > 150
        assert ei.value.code == -1
        b.upload([g])
        with pytest.raises(sl.SpacelabError) as ei:
            b.step(3)                   # no init_conditions yet
        assert ei.value.code == -3
        with pytest.raises(sl.SpacelabError) as ei:
            b.set_system_params(dt=[-1.0])
        assert ei.value.code == -1
        with pytest.raises(sl.SpacelabError) as ei:
            b.set_option("tile_split", 3)
        assert ei.value.code == -1
        with pytest.raises(sl.SpacelabError) as ei:
            b.set_option("ball_waves", 6)
        assert ei.value.code == -1


def _batch_run(groups, params, steps, we, ball_waves):
    with sl.Batch(len(groups), sum(g.num_particles for g in groups), params[0]) as b:
        b.set_option("tile_split", 2)
        b.set_option("ball_waves", ball_waves)
        b.upload(groups, kin=[p.kin for p in params], kout=[p.kout for p in params],
                 dt=[p.dt for p in params])
        e0 = b.init_conditions()
        ew = b.step(steps, we)
        return b.download(), e0, ew, b.get_option("ball_waves")


def test_ball_kernel_shapes_agree_bitwise():
    """The per-ball kernel in workgroups of 4 wavefronts (two slot classes each) and of
    8 (one each) gives the same bits: the classes fix every sum's order.  A batch of 300
    systems (more 64-ball blocks than compute units: 4 by default) against the same
    batch forced to 8, and a lone system (8 by default) forced to 4, with write steps."""
    sizes = [40 + (37 * k) % 160 for k in range(300)]
    groups, params = _ensemble(sizes, 0x5EED0600)
    a, ea0, eaw, wa = _batch_run(groups, params, 60, 30, 0)
    b, eb0, ebw, wb = _batch_run(groups, params, 60, 30, 8)
    assert (wa, wb) == (4, 8)
    for k in range(len(sizes)):
        for f in FIELDS:
            assert bits_equal(getattr(a[k], f), getattr(b[k], f)), (k, f)
        assert _ebytes(ea0[k]) == _ebytes(eb0[k]), k
        for w in range(len(eaw)):
            assert _ebytes(eaw[w][k]) == _ebytes(ebw[w][k]), (k, w)
    g = sl.generate_bpca(2000, seed=0x5EED0601, v_sigma=2.0)
    p = _params(g)
    outs = []
    for waves in (0, 4):
        with sl.Collider(g.num_particles, p) as c:
            c.set_option("tile_split", 2)
            c.set_option("ball_waves", waves)
            c.upload(g)
            e0 = c.init_conditions()
            ew = c.step(60, 30)
            outs.append((c.download(), e0, ew, c.get_option("ball_waves")))
    assert outs[0][3] == 8 and outs[1][3] == 4
    for f in FIELDS:
        assert bits_equal(getattr(outs[0][0], f), getattr(outs[1][0], f)), f
    assert _ebytes(outs[0][1]) == _ebytes(outs[1][1])
    assert [_ebytes(e) for e in outs[0][2]] == [_ebytes(e) for e in outs[1][2]]


def test_ball_kernel_shapes_agree_with_many_partners():
    """One sphere pressed by 4000 small ones (a Fibonacci lattice over its surface): its
    block numbers more than 4,000 near entries, reading partner masks again, and
    evaluates them over at least 8 rounds of 512 or 16 of 256.  Both kernel shapes give
    the same bits."""
    k = 4000
    rs, Rb = 1e-5, 4e-4
    t = np.arange(k) + 0.5
    phi, th = np.arccos(1 - 2 * t / k), np.pi * (1 + 5 ** 0.5) * t
    d = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    pos = np.vstack([[0.0, 0.0, 0.0], d * (Rb + rs) * (1 - 1e-3)])
    R = np.array([Rb] + [rs] * k)
    m = 2.25 * 4 / 3 * np.pi * R ** 3
    rng = np.random.default_rng(12)
    g = sl.BallGroup(pos, rng.standard_normal(pos.shape) * 1e-3, rng.standard_normal(pos.shape) * 1e-2,
                     R, m, 0.4 * m * R * R)
    p = sl.calibrate(sl.preset("config"), R, m, velocity=V_CUSTOM, cor=0.64).replace(order=sl.ORDER_FAST)
    outs = []
    for waves in (8, 4):
        with sl.Collider(k + 1, p) as c:
            c.set_option("tile_split", 2)
            c.set_option("ball_waves", waves)
            c.upload(g)
            e0 = c.init_conditions()
            ew = c.step(4, 2)
            outs.append((c.download(), e0, ew))
    assert outs[0][1].contacts == k
    for f in FIELDS:
        assert bits_equal(getattr(outs[0][0], f), getattr(outs[1][0], f)), f
    assert _ebytes(outs[0][1]) == _ebytes(outs[1][1])
    assert [_ebytes(e) for e in outs[0][2]] == [_ebytes(e) for e in outs[1][2]]

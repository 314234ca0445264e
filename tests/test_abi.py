"""The C-ABI library: it loads, exports every entry point its headers declare, and
its host-only functions (presets, calibration, the aggregate generator) are right.
No compute is called here: this container has no GPU."""
import ctypes as C
import glob
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import spacelab_amd as sl

HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "spacelab", "*.h")))


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:int|void|long long|char\s*\*|const char\s*\*)\s*\**\s*(sl_\w+)\s*\(",
                             txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_present():
    assert HEADERS, "include/spacelab/*.h missing"
    assert "sl_step" in declared_functions()


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    L = sl.lib()
    assert hasattr(L, name), f"{name} declared in include/ but not exported"


def test_presets():
    p = sl.preset("reference")
    assert (p.G, p.Ha, p.u_s, p.u_r, p.far_gravity) == (6.67e-8, 4.7e-12, 0.1, 1e-5, 0)
    assert p.h_min == 1e-5 * .1
    q = sl.preset("config")
    assert (q.u_s, q.far_gravity) == (0.3, 1)


def test_calibrate_closed_forms():
    R = np.array([1e-5, 2e-5, 0.8e-5])
    m = 2.25 * 4 / 3 * math.pi * R ** 3
    p = sl.calibrate(sl.preset("config"), R, m, velocity=0.36301555459799423, cor=0.64, dt_rule="tc50")
    kconsts = 3e3 * (4 / 3 * math.pi * 2.25) / (0.1 * 0.1)
    assert math.isclose(p.kin, kconsts * R.max() * 0.36301555459799423 ** 2, rel_tol=1e-15)
    assert math.isclose(p.kout, 0.64 * p.kin, rel_tol=1e-15)
    assert math.isclose(p.dt * 50, math.pi * math.sqrt(m.min() / p.kin), rel_tol=1e-14)
    # reference rule: dt = .01 sqrt(m/k_eff) when the elastic term dominates,
    # which is t_c / (100 pi) of the same spring
    r = sl.calibrate(sl.preset("reference"), R, m, velocity=10.0, cor=0.4, dt_rule="reference")
    four_thirds_pi_rho = 4 / 3 * math.pi * 2.25
    expect = 0.01 * math.sqrt(four_thirds_pi_rho / (r.kin * 0.1 * R.min() / (0.1 * R.min())) * R.min() ** 3)
    assert math.isclose(r.dt, expect, rel_tol=1e-12)


def test_bpca_generator_properties():
    g = sl.generate_bpca(300, seed=12345)
    g2 = sl.generate_bpca(300, seed=12345)
    g3 = sl.generate_bpca(300, seed=12346)
    assert np.array_equal(g.pos, g2.pos) and np.array_equal(g.R, g2.R)
    assert not np.array_equal(g.pos, g3.pos)
    n = g.num_particles
    d = np.linalg.norm(g.pos[:, None, :] - g.pos[None, :, :], axis=2)
    S = g.R[:, None] + g.R[None, :]
    gap = d - S
    np.fill_diagonal(gap, np.inf)
    # no overlaps beyond rounding, and every sphere touches at least one other
    assert gap.min() > -1e-12 * g.R.max()
    assert np.all(np.abs(gap).min(axis=1) < 1e-12 * g.R.max())
    assert np.allclose((g.m[:, None] * g.pos).sum(0) / g.m.sum(), 0, atol=1e-18)
    assert np.abs((g.m[:, None] * g.vel).sum(0)).max() < 1e-12 * (g.m * np.abs(g.vel).sum(1)).sum()
    assert np.allclose(g.m, 2.25 * 4 / 3 * math.pi * g.R ** 3, rtol=1e-14)
    assert np.allclose(g.moi, 0.4 * g.m * g.R ** 2, rtol=1e-14)
    assert np.all(g.w == 0)


def test_bpca_radius_distribution():
    g = sl.generate_bpca(4000, seed=7)
    lr = np.log(g.R / 1e-5)
    assert abs(np.median(lr)) < 0.02
    assert abs(lr.std() - 0.2) < 0.01
    assert abs(g.vel.std() - 1e-3) < 5e-5


def test_create_without_gpu_reports_nogpu():
    p = sl.preset("reference").replace(kin=1.0, kout=0.4, dt=1e-9)
    try:
        c = sl.Collider(16, p)
    except sl.SpacelabError as e:
        assert e.code == -5
        return
    c.close()
    pytest.skip("a GPU is present; covered by the gpu tests")


def test_bad_params_rejected_before_device():
    p = sl.preset("reference").replace(kin=1.0, kout=0.4, dt=-1.0)
    with pytest.raises(sl.SpacelabError) as ei:
        sl.Collider(16, p)
    assert ei.value.code in (-1, -5)


def test_binding_constants_match_header():
    """The Python binding's codes are the header's #defines (precision, order, options,
    error codes, outcome classes, descriptor count)."""
    txt = open(os.path.join(ROOT, "include", "spacelab", "sl_api.h")).read()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r"^#define (SL_\w+) (-?\d+)", txt, flags=re.M)}
    assert defs["SL_PREC_FP64"] == sl.PREC_FP64 and defs["SL_PREC_MIXED"] == sl.PREC_MIXED
    assert defs["SL_ORDER_ORDERED"] == sl.ORDER_ORDERED and defs["SL_ORDER_FAST"] == sl.ORDER_FAST
    assert defs["SL_OPT_GRAPHS"] == sl.OPT_GRAPHS
    assert defs["SL_OPT_TILE_SPLIT"] == sl.OPT_TILE_SPLIT
    assert defs["SL_OPT_BALL_WAVES"] == sl.OPT_BALL_WAVES
    assert defs["SL_OPT_ORDERED_SHAPE"] == sl.OPT_ORDERED_SHAPE
    assert defs["SL_OPT_TILE_QUEUE"] == sl.OPT_TILE_QUEUE
    for code, name in sl.ERRORS.items():
        assert defs[name] == code, name
    assert (defs["SL_OUTCOME_STICK"], defs["SL_OUTCOME_BOUNCE"], defs["SL_OUTCOME_FRAGMENT"]) == (
        sl.OUTCOME_STICK, sl.OUTCOME_BOUNCE, sl.OUTCOME_FRAGMENT)
    assert defs["SL_N_DESCRIPTORS"] == sl.N_DESCRIPTORS


def test_shipped_tune_cache_is_valid():
    """The tune cache shipped beside the library: a flat JSON object whose entries name a
    tile split (1, 2 or 4, plus 8 for the work-queue launch) or an ordered launch shape
    (1 + index, at most 64), with the timings they were chosen from."""
    import json
    path = os.path.join(os.path.dirname(sl.__file__), "data", "tune.json")
    data = json.load(open(path))
    assert data
    for key, v in data.items():
        kind = key.split(":")[0]
        assert kind in ("tiles", "ordered"), key
        if kind == "tiles":
            assert v["split"] in (1, 2, 4, 9, 10, 12), key
            assert len(v["ms"]) in (3, 6)
        else:
            assert 1 <= v["split"] <= 64, key
            assert len(v["ms"]) >= 2

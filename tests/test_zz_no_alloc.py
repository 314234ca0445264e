"""The allocation-free entry point, sl_step_ordered_static, against sl_step.

The caller (this test) owns the device memory: it allocates the state and energy
arrays through the HIP runtime the library itself is linked against.  The entry point
steps that state with the ordered kernels, its scratch in the library's fixed
per-device buffer.  Two calls in a row, the second on a smaller system, must each give
bitwise the state and energies of a context stepped by sl_step from the same state."""
import ctypes as C

import numpy as np
import pytest

from conftest import bits_equal

import spacelab_amd as sl

pytestmark = pytest.mark.gpu

NFIELD = 27
F = {"pos": 0, "vel": 3, "w": 6, "acc": 9, "aacc": 12, "velh": 15, "wh": 18, "pos_prev": 21}
F_R, F_M, F_MOI = 24, 25, 26
E_STRIDE = 10
V_CUSTOM = 0.36301555459799423


class _Hip:
    """hipMalloc/hipMemcpy/hipFree of the runtime libspacelab.so is linked against."""

    def __init__(self):
        sl.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, a):
        assert self.L.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.L.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self, p):
        assert self.L.hipFree(p) == 0


def _context_run(g, p, steps, write_every):
    """The state after sl_init_forces, and after `steps` more steps of sl_step."""
    with sl.Collider(g.num_particles, p) as c:
        c.upload(g)
        c.init_conditions()
        s0 = c.download()
        ew = c.step(steps, write_every)
        s1 = c.download()
    return s0, s1, ew


def _static_run(hip, g, s0, p, steps, write_every):
    n = g.num_particles
    stride = (n + 63) // 64 * 64 + 64
    host = np.zeros((NFIELD, stride))
    for name, arr in (("pos", s0.pos), ("vel", s0.vel), ("w", s0.w), ("acc", s0.acc), ("aacc", s0.aacc)):
        host[F[name]:F[name] + 3, :n] = arr.T
    # fields the step overwrites get values it must not depend on
    host[F["velh"]:F["velh"] + 3, :n] = 1e300
    host[F["wh"]:F["wh"] + 3, :n] = -1e300
    host[F["pos_prev"]:F["pos_prev"] + 3, :n] = 3.0
    host[F_R, :n], host[F_M, :n], host[F_MOI, :n] = g.R, g.m, g.moi
    n_w = steps // write_every
    energy = np.full((max(n_w, 1), E_STRIDE), np.nan)
    d_state = hip.malloc(host.nbytes)
    d_e = hip.malloc(energy.nbytes)
    try:
        hip.h2d(d_state, host)
        hip.h2d(d_e, energy)
        L = sl.lib()
        L.sl_step_ordered_static.argtypes = [C.POINTER(sl.Params), C.c_int, C.c_int, C.c_void_p, C.c_longlong,
                                             C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        rc = L.sl_step_ordered_static(C.byref(p), n, stride, d_state, 0, steps, write_every, d_e, None)
        hip.sync()
        assert rc == n_w
        out = np.empty_like(host)
        hip.d2h(out, d_state)
        hip.d2h(energy, d_e)
    finally:
        hip.free(d_state)
        hip.free(d_e)
    st = {name: np.ascontiguousarray(out[F[name]:F[name] + 3, :n].T) for name in ("pos", "vel", "w", "acc", "aacc")}
    return st, energy[:n_w]


def _check(g, p, steps, write_every, hip):
    s0, s1, ew = _context_run(g, p, steps, write_every)
    st, energy = _static_run(hip, g, s0, p, steps, write_every)
    for k in ("pos", "vel", "w", "acc", "aacc"):
        assert bits_equal(st[k], getattr(s1, k)), k
    assert len(ew) == len(energy) > 0
    for e, row in zip(ew, energy):
        want = np.array([e.PE, e.KE, *e.p, *e.L], dtype=np.float64)
        assert bits_equal(row[:8], want)
        assert row[8] == e.contacts
        assert row[9] == 0.0 and not e.nonfinite


def test_static_entry_matches_sl_step_twice():
    """2000 spheres, then 300: the two sizes take different default launch shapes of the
    ordered pass (1500 spheres and above, and below)."""
    hip = _Hip()
    g1 = sl.generate_bpca(2000, seed=0x5EED0A01, v_sigma=2.0)
    p1 = sl.calibrate(sl.preset("config"), g1.R, g1.m, velocity=V_CUSTOM, cor=0.64)
    _check(g1, p1, 60, 20, hip)
    # a smaller system next: the buffer still holds the first call's partials
    g2 = sl.generate_bpca(300, seed=0x5EED0A02, v_sigma=2.0)
    p2 = sl.calibrate(sl.preset("reference"), g2.R, g2.m, velocity=V_CUSTOM, cor=0.64)
    _check(g2, p2, 40, 10, hip)


def test_static_entry_rejects_oversize_and_bad_arguments():
    L = sl.lib()
    L.sl_step_ordered_static.argtypes = [C.POINTER(sl.Params), C.c_int, C.c_int, C.c_void_p, C.c_longlong,
                                         C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    p = sl.preset("config").replace(kin=1.0, kout=0.64, dt=1e-12)
    fake = C.c_void_p(1)   # never dereferenced: each call fails its argument checks first
    assert L.sl_step_ordered_static(C.byref(p), 65537, 65600, fake, 0, 1, 0, None, None) == -4
    assert L.sl_step_ordered_static(C.byref(p), 10, 5, fake, 0, 1, 0, None, None) == -1
    assert L.sl_step_ordered_static(C.byref(p.replace(precision=sl.PREC_MIXED)), 10, 64, fake, 0, 1, 0, None,
                                    None) == -1
    assert L.sl_step_ordered_static(None, 10, 64, fake, 0, 1, 0, None, None) == -1


def test_static_entry_at_capacity():
    """The largest system the entry point takes, SL_STATIC_MAX_BALLS = 65,536 spheres (its
    scratch buffer full): two steps, the second a write step, bitwise sl_step's."""
    hip = _Hip()
    g = sl.generate_bpca(65536, seed=0x5EED0A03, v_sigma=2.0)
    p = sl.calibrate(sl.preset("config"), g.R, g.m, velocity=V_CUSTOM, cor=0.64)
    _check(g, p, 2, 2, hip)

"""spacelab_amd -- MI355X (gfx950) implementation of SpaceLab's DEM step.

Python binding over the C ABI in ``include/spacelab/sl_api.h`` (library
``spacelab_amd/lib/libspacelab.so``).  The binding only moves arrays and calls
the library; every step runs in the HIP kernels.  There is no CPU fallback:
if the library is missing or no gfx950 GPU is present the calls raise.

The object model mirrors the reference's (ball_group.hpp, ColliderSingleCore.cpp):

* :class:`BallGroup` holds the per-ball arrays of ``Ball_group`` (pos, vel, velh,
  acc, w, wh, aacc, R, m, moi) as float64 arrays of shape (n, 3) / (n,);
* :class:`Collider` owns a device context and provides ``init_conditions()``
  (Ball_group::init_conditions), ``sim_one_step(write_step)`` and
  ``sim_looper(steps, skip)`` (ColliderSingleCore.cpp:78-331), with the energies
  of write steps returned as :class:`Energy` records (O.PE, O.KE, O.mom, O.ang_mom);
* :class:`Batch` steps many independent groups (a collision ensemble) together.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libspacelab.so")

PREC_FP64, PREC_MIXED = 0, 1
ORDER_ORDERED, ORDER_FAST = 0, 1
OPT_GRAPHS, OPT_TILE_SPLIT, OPT_BALL_WAVES = 1, 2, 3
OPT_ORDERED_SHAPE = 4
OPT_TILE_QUEUE = 5
_OPTS = {"graphs": OPT_GRAPHS, "tile_split": OPT_TILE_SPLIT, "ball_waves": OPT_BALL_WAVES,
         "ordered_shape": OPT_ORDERED_SHAPE, "tile_queue": OPT_TILE_QUEUE}

ERRORS = {-1: "SL_E_ARG", -2: "SL_E_HIP", -3: "SL_E_STATE", -4: "SL_E_CAPACITY", -5: "SL_E_NOGPU"}


class SpacelabError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    """sl_params: the constants and globals the step reads (dust_const.hpp:21-53)."""
    _fields_ = [("G", C.c_double), ("Ha", C.c_double), ("h_min", C.c_double),
                ("u_s", C.c_double), ("u_r", C.c_double), ("kin", C.c_double),
                ("kout", C.c_double), ("dt", C.c_double), ("far_gravity", C.c_int),
                ("precision", C.c_int), ("order", C.c_int)]

    def copy(self) -> "Params":
        p = Params()
        C.pointer(p)[0] = self
        return p

    def replace(self, **kw) -> "Params":
        p = self.copy()
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class RunState(C.Structure):
    """sl_run_state: calibrate_dt's run-control state."""
    _fields_ = [("kin", C.c_double), ("kout", C.c_double), ("dt", C.c_double), ("steps", C.c_int),
                ("skip", C.c_int), ("v_collapse", C.c_double), ("v_max", C.c_double),
                ("v_max_prev", C.c_double), ("soc", C.c_double)]


class _Energy(C.Structure):
    _fields_ = [("PE", C.c_double), ("KE", C.c_double), ("p", C.c_double * 3),
                ("L", C.c_double * 3), ("nonfinite", C.c_int), ("contacts", C.c_longlong)]


class Outcome(C.Structure):
    """sl_outcome: contact-graph outcome of a finished collision."""
    _fields_ = [("label", C.c_int), ("n_components", C.c_int), ("largest_mass_frac", C.c_double),
                ("second_mass_frac", C.c_double), ("mass_transfer", C.c_double),
                ("v_largest", C.c_double * 3), ("v_second", C.c_double * 3)]


OUTCOME_STICK, OUTCOME_BOUNCE, OUTCOME_FRAGMENT = 0, 1, 2
N_DESCRIPTORS = 21


class _Profile(C.Structure):
    _fields_ = [("pair_ms", C.c_double), ("pair_launches", C.c_longlong),
                ("ball_ms", C.c_double), ("ball_launches", C.c_longlong),
                ("contacts", C.c_longlong), ("near_pairs", C.c_longlong)]


@dataclass
class Energy:
    """Energies of one write step: O.PE, O.KE, O.mom, O.ang_mom."""
    PE: float
    KE: float
    p: np.ndarray
    L: np.ndarray
    nonfinite: bool = False
    contacts: int = 0

    @property
    def E(self) -> float:
        return self.PE + self.KE

    def as_array(self) -> np.ndarray:
        return np.array([self.PE, self.KE, *self.p, *self.L], dtype=np.float64)

    @staticmethod
    def _from(e: _Energy) -> "Energy":
        return Energy(e.PE, e.KE, np.array(e.p[:]), np.array(e.L[:]), bool(e.nonfinite),
                      int(e.contacts))


_lib = None


def lib():
    """Load libspacelab.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m spacelab_amd.build` "
                          "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.sl_params_preset.argtypes = [C.POINTER(Params), C.c_int]
    L.sl_calibrate.argtypes = [C.POINTER(Params), C.c_int, dp, dp, C.c_double, C.c_double, C.c_int]
    L.sl_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(Params)]
    L.sl_set_params.argtypes = [vp, C.POINTER(Params)]
    L.sl_upload.argtypes = [vp, C.c_int] + [dp] * 9
    L.sl_init_forces.argtypes = [vp, C.POINTER(_Energy)]
    L.sl_step.argtypes = [vp, C.c_int, C.c_int, C.POINTER(_Energy), C.c_int]
    L.sl_download.argtypes = [vp] + [dp] * 5
    L.sl_download_half.argtypes = [vp, dp, dp]
    L.sl_synchronize.argtypes = [vp]
    L.sl_step_count.argtypes = [vp]
    L.sl_step_count.restype = C.c_longlong
    L.sl_last_error.argtypes = [vp]
    L.sl_last_error.restype = C.c_char_p
    L.sl_destroy.argtypes = [vp]
    L.sl_destroy.restype = None
    L.sl_profile_enable.argtypes = [vp, C.c_int, C.c_int]
    L.sl_profile_read.argtypes = [vp, C.POINTER(_Profile)]
    L.sl_set_option.argtypes = [vp, C.c_int, C.c_int]
    L.sl_get_option.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    ip = C.POINTER(C.c_int)
    L.sl_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.POINTER(Params)]
    L.sl_batch_upload.argtypes = [vp, C.c_int, ip] + [dp] * 9
    L.sl_batch_set_system_params.argtypes = [vp, dp, dp, dp]
    L.sl_batch_init_forces.argtypes = [vp, C.POINTER(_Energy)]
    L.sl_batch_step.argtypes = [vp, C.c_int, C.c_int, C.POINTER(_Energy), C.c_int]
    L.sl_batch_download.argtypes = [vp] + [dp] * 5
    L.sl_batch_set_option.argtypes = [vp, C.c_int, C.c_int]
    L.sl_batch_get_option.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.sl_batch_step_count.argtypes = [vp]
    L.sl_batch_step_count.restype = C.c_longlong
    L.sl_batch_last_error.argtypes = [vp]
    L.sl_batch_last_error.restype = C.c_char_p
    L.sl_batch_destroy.argtypes = [vp]
    L.sl_batch_destroy.restype = None
    L.sl_gen_bpca.argtypes = [C.c_int, C.c_ulonglong, C.c_double, C.c_double, C.c_double,
                              C.c_double] + [dp] * 6
    L.sl_gen_collision.argtypes = [C.c_int, C.c_int, C.c_ulonglong, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double] + [dp] * 7
    L.sl_collision_outcome.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, C.c_double, C.POINTER(Outcome)]
    L.sl_collision_approach.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, dp]
    L.sl_group_stats.argtypes = [C.c_int, dp, dp, dp, dp]
    L.sl_v_collapse.argtypes = [C.c_double, C.c_double, C.c_double, dp]
    L.sl_vel_max.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_double, dp, ip]
    L.sl_run_length.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, ip, ip]
    L.sl_run_init.argtypes = [C.POINTER(RunState), C.c_int, dp, dp, dp, C.c_double]
    L.sl_calibrate_dt.argtypes = [C.POINTER(RunState), C.POINTER(Params), C.c_int, dp, dp, dp, dp, C.c_int,
                                  C.c_double, C.c_double]
    L.sl_to_origin.argtypes = [C.c_int, dp, dp]
    L.sl_two_cluster.argtypes = [C.c_int] + [dp] * 6 + [C.c_int] + [dp] * 6 + [C.c_double, C.c_double] + [dp] * 7
    cp, ll = C.c_char_p, C.c_longlong
    for name, args in (("sl_csv_sim_header", [C.c_int, cp, ll]),
                       ("sl_csv_sim_row", [C.c_int, dp, dp, dp, cp, ll]),
                       ("sl_csv_energy_row", [C.c_double, C.c_double, C.c_double, dp, dp, C.c_int, cp, ll]),
                       ("sl_csv_constants", [C.c_int, dp, dp, dp, cp, ll]),
                       ("sl_csv_last_line", [cp, cp, ll])):
        getattr(L, name).argtypes = args
        getattr(L, name).restype = ll
    L.sl_csv_count_balls.argtypes = [cp]
    L.sl_csv_parse_sim_row.argtypes = [cp, C.c_int, dp, dp, dp]
    L.sl_csv_parse_constants.argtypes = [cp, C.c_int, dp, dp, dp]
    L.sl_csv_energy_columns.argtypes = [cp]
    L.sl_csv_parse_energy_row.argtypes = [cp, C.c_int, dp]
    L.sl_collision_descriptors.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_double, C.c_double,
                                           C.c_double, dp]
    L.sl_collision_descriptors_batch.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                 C.POINTER(C.c_longlong), dp, dp, dp, dp, dp, C.c_double, dp,
                                                 C.c_int]
    fp = C.POINTER(C.c_float)
    L.sl_mlp_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong]
    L.sl_mlp_n_params.argtypes = [vp]
    L.sl_mlp_get_params.argtypes = [vp, fp]
    L.sl_mlp_set_params.argtypes = [vp, fp]
    L.sl_mlp_forward.argtypes = [vp, C.c_int, fp, fp, fp, fp]
    L.sl_mlp_gradients.argtypes = [vp, C.c_int, fp, ip, fp, fp, fp]
    L.sl_mlp_train_step.argtypes = [vp, C.c_int, fp, ip, fp, C.c_float, fp]
    L.sl_mlp_last_error.argtypes = [vp]
    L.sl_mlp_last_error.restype = C.c_char_p
    L.sl_mlp_destroy.argtypes = [vp]
    L.sl_mlp_destroy.restype = None
    L.sl_probe_divsqrt.argtypes = [C.c_int] + [dp] * 5
    L.sl_probe_div_rcp.argtypes = [C.c_int] + [dp] * 3
    L.sl_probe_rotation.argtypes = [C.POINTER(C.c_int)]
    L.sl_probe_copy.argtypes = [C.c_ulonglong, C.c_int, C.POINTER(C.c_double)]
    _lib = L
    return L


def _dp(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a, shape) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def preset(name: str = "reference") -> Params:
    """'reference': dust_const.hpp verbatim; 'config': cor 0.64, u_s 0.3, all-pairs gravity."""
    p = Params()
    rc = lib().sl_params_preset(C.byref(p), {"reference": 0, "config": 1}[name])
    if rc:
        raise SpacelabError(rc, "sl_params_preset")
    return p


class RunControl:
    """calibrate_dt's run control (ball_group.hpp:197-238): built from a group as the
    continue constructor does (calc_helpfuls, calc_v_collapse), then calibrate(step,
    custom_speed) with the velocity rule (custom_speed <= 0) or a custom speed."""

    def __init__(self, g: BallGroup, base: Params | None = None, cor: float = 0.4):
        n = g.num_particles
        self.base = (base or preset("reference")).copy()
        self.cor = cor
        self.state = RunState()
        rc = lib().sl_run_init(C.byref(self.state), n, _dp(_f64(g.pos, (n, 3))), _dp(_f64(g.R, (n,))),
                               _dp(_f64(g.m, (n,))), self.base.G)
        if rc:
            raise SpacelabError(rc, "sl_run_init")

    def calibrate(self, g: BallGroup, step: int, custom_speed: float = 0.0) -> RunState:
        n = g.num_particles
        rc = lib().sl_calibrate_dt(C.byref(self.state), C.byref(self.base), n, _dp(_f64(g.pos, (n, 3))),
                                   _dp(_f64(g.vel, (n, 3))), _dp(_f64(g.R, (n,))), _dp(_f64(g.m, (n,))),
                                   int(step), float(custom_speed), self.cor)
        if rc:
            raise SpacelabError(rc, "sl_calibrate_dt")
        return self.state


def calibrate(p: Params, R: np.ndarray, m: np.ndarray, velocity: float, cor: float,
              dt_rule: str = "tc50") -> Params:
    """kin, kout and dt (updateDTK, ball_group.hpp:1313-1329, or dt = t_c/50)."""
    q = p.copy()
    R = np.ascontiguousarray(R, dtype=np.float64)
    m = np.ascontiguousarray(m, dtype=np.float64)
    rc = lib().sl_calibrate(C.byref(q), len(R), _dp(R), _dp(m), velocity, cor,
                            {"reference": 0, "tc50": 1}[dt_rule])
    if rc:
        raise SpacelabError(rc, "sl_calibrate")
    return q


@dataclass
class BallGroup:
    """The per-ball arrays of the reference's Ball_group (ball_group.hpp:21-53)."""
    pos: np.ndarray
    vel: np.ndarray
    w: np.ndarray
    R: np.ndarray
    m: np.ndarray
    moi: np.ndarray
    acc: np.ndarray | None = None
    aacc: np.ndarray | None = None
    velh: np.ndarray | None = None
    wh: np.ndarray | None = None

    @property
    def num_particles(self) -> int:
        return int(self.R.shape[0])

    def copy(self) -> "BallGroup":
        cp = lambda a: None if a is None else np.array(a, dtype=np.float64, copy=True)
        return BallGroup(cp(self.pos), cp(self.vel), cp(self.w), cp(self.R), cp(self.m),
                         cp(self.moi), cp(self.acc), cp(self.aacc), cp(self.velh), cp(self.wh))

    @classmethod
    def from_csv(cls, path: str, prefix: str) -> "BallGroup":
        """Continue a run: the last row of <path><prefix>simData.csv and the constants file
        (loadSim, ball_group.hpp:1236-1247).  The constants file is <prefix>constants.csv,
        or the Generator's <prefix>Constants.csv where only that exists."""
        pos, w, vel = csv_parse_sim_row(csv_last_line(path + prefix + "simData.csv"))
        consts = path + prefix + "constants.csv"
        if not os.path.exists(consts) and os.path.exists(path + prefix + "Constants.csv"):
            consts = path + prefix + "Constants.csv"
        with open(consts, "rb") as f:
            R, m, moi = csv_parse_constants(f.read().decode(), len(pos))
        return cls(pos, vel, w, R, m, moi)

    def to_origin(self) -> "BallGroup":
        """Moves the centre of mass to the origin, in place (to_origin, sl_to_origin)."""
        self.pos = _f64(self.pos, (self.num_particles, 3)).copy()
        rc = lib().sl_to_origin(self.num_particles, _dp(self.pos), _dp(_f64(self.m, (self.num_particles,))))
        if rc:
            raise SpacelabError(rc, "sl_to_origin")
        return self


V_CUSTOM = 0.36301555459799423     # dust_const.hpp:34


def two_cluster(target: BallGroup, projectile: BallGroup, v_impact: float = V_CUSTOM,
                impact_angle: float = 0.0) -> tuple[BallGroup, tuple[float, float]]:
    """The reference's two-cluster set-up of two loaded groups (sl_two_cluster): the
    merged group (target first) and the (projectile, target) speeds along x."""
    nt, npj = target.num_particles, projectile.num_particles
    n = nt + npj
    out = [np.zeros((n, 3)) for _ in range(3)] + [np.zeros(n) for _ in range(3)]
    v2 = np.zeros(2)

    def arrays(g, k):
        return [_f64(g.pos, (k, 3)), _f64(g.vel, (k, 3)), _f64(g.w, (k, 3)), _f64(g.R, (k,)),
                _f64(g.m, (k,)), _f64(g.moi, (k,))]
    ta, pa = arrays(target, nt), arrays(projectile, npj)
    rc = lib().sl_two_cluster(nt, *[_dp(a) for a in ta], npj, *[_dp(a) for a in pa], float(v_impact),
                              float(impact_angle), *[_dp(a) for a in out], _dp(v2))
    if rc:
        raise SpacelabError(rc, "sl_two_cluster")
    return BallGroup(*out), (float(v2[0]), float(v2[1]))


def generate_bpca(n: int, seed: int, r_med: float = 1e-5, sigma: float = 0.2,
                  rho: float = 2.25, v_sigma: float = 1e-3) -> BallGroup:
    """Seeded ballistic particle-cluster aggregate (see sl_gen_bpca)."""
    pos = np.zeros((n, 3)); vel = np.zeros((n, 3)); w = np.zeros((n, 3))
    R = np.zeros(n); m = np.zeros(n); moi = np.zeros(n)
    rc = lib().sl_gen_bpca(n, seed, r_med, sigma, rho, v_sigma, _dp(pos), _dp(vel), _dp(w),
                           _dp(R), _dp(m), _dp(moi))
    if rc:
        raise SpacelabError(rc, "sl_gen_bpca")
    return BallGroup(pos, vel, w, R, m, moi)


@dataclass
class CollisionSetup:
    """A two-aggregate collision (sl_gen_collision): the combined group (target
    first) and its geometry."""
    group: BallGroup
    n_t: int
    n_p: int
    speed: float
    b_frac: float
    R_t: float
    R_p: float
    b: float
    M_t: float
    M_p: float
    r_max: float


def generate_collision(n_t: int, n_p: int, seed: int, speed: float, b_frac: float, gap: float = 2.0,
                       r_med: float = 1e-5, sigma: float = 0.2, rho: float = 2.25) -> CollisionSetup:
    """Target and projectile aggregates aimed at each other (see sl_gen_collision)."""
    n = n_t + n_p
    pos = np.zeros((n, 3)); vel = np.zeros((n, 3)); w = np.zeros((n, 3))
    R = np.zeros(n); m = np.zeros(n); moi = np.zeros(n); geom = np.zeros(6)
    rc = lib().sl_gen_collision(n_t, n_p, seed, speed, b_frac, gap, r_med, sigma, rho, _dp(pos),
                                _dp(vel), _dp(w), _dp(R), _dp(m), _dp(moi), _dp(geom))
    if rc:
        raise SpacelabError(rc, "sl_gen_collision")
    return CollisionSetup(BallGroup(pos, vel, w, R, m, moi), n_t, n_p, speed, b_frac, *geom)


def collision_approach(setup: CollisionSetup, gap: float) -> tuple[bool, float]:
    """Move the projectile along +x (in place) so that its first contact with the
    target comes after ``gap`` cm of travel; (met, shift), met False on a miss."""
    g = setup.group
    n = g.num_particles
    g.pos = _f64(g.pos, (n, 3))
    sh = C.c_double()
    rc = lib().sl_collision_approach(setup.n_t, setup.n_p, _dp(g.pos), _dp(_f64(g.R, (n,))), gap, C.byref(sh))
    if rc < 0:
        raise SpacelabError(rc, "sl_collision_approach")
    return bool(rc), float(sh.value)


def collision_outcome(n_t: int, g: BallGroup, link: float) -> Outcome:
    """Contact-graph outcome (stick / bounce / fragment) of a collision's final state."""
    n = g.num_particles
    pos, vel = _f64(g.pos, (n, 3)), _f64(g.vel, (n, 3))
    R, m = _f64(g.R, (n,)), _f64(g.m, (n,))
    o = Outcome()
    rc = lib().sl_collision_outcome(n_t, n - n_t, _dp(pos), _dp(vel), _dp(R), _dp(m), link, C.byref(o))
    if rc:
        raise SpacelabError(rc, "sl_collision_outcome")
    return o


def collision_descriptors(setup: CollisionSetup, link: float) -> np.ndarray:
    """The N_DESCRIPTORS inputs of the surrogate model for a collision set-up."""
    g = setup.group
    n = g.num_particles
    out = np.zeros(N_DESCRIPTORS)
    rc = lib().sl_collision_descriptors(setup.n_t, setup.n_p, _dp(_f64(g.pos, (n, 3))), _dp(_f64(g.R, (n,))),
                                        _dp(_f64(g.m, (n,))), setup.speed, setup.b_frac, link, _dp(out))
    if rc:
        raise SpacelabError(rc, "sl_collision_descriptors")
    return out


def collision_descriptors_batch(setups: list, link: float, threads: int = 0) -> np.ndarray:
    """collision_descriptors of many set-ups at once, on ``threads`` host threads (0: one
    per core, at most 16); row k belongs to setups[k]."""
    k = len(setups)
    out = np.zeros((k, N_DESCRIPTORS))
    if k == 0:
        return out
    sizes = np.array([s.group.num_particles for s in setups], dtype=np.int64)
    first = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)[:-1]]), dtype=np.int64)
    pos = np.ascontiguousarray(np.concatenate([_f64(s.group.pos, (s.group.num_particles, 3)) for s in setups]))
    R = np.ascontiguousarray(np.concatenate([_f64(s.group.R, (s.group.num_particles,)) for s in setups]))
    m = np.ascontiguousarray(np.concatenate([_f64(s.group.m, (s.group.num_particles,)) for s in setups]))
    n_t = np.ascontiguousarray([s.n_t for s in setups], dtype=np.int32)
    n_p = np.ascontiguousarray([s.n_p for s in setups], dtype=np.int32)
    speed = np.ascontiguousarray([s.speed for s in setups], dtype=np.float64)
    b = np.ascontiguousarray([s.b_frac for s in setups], dtype=np.float64)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    rc = lib().sl_collision_descriptors_batch(k, ip(n_t), ip(n_p), first.ctypes.data_as(C.POINTER(C.c_longlong)),
                                              _dp(pos), _dp(R), _dp(m), _dp(speed), _dp(b), link, _dp(out),
                                              int(threads))
    if rc:
        raise SpacelabError(rc, "sl_collision_descriptors_batch")
    return out


class Collider:
    """One simulated Ball_group on the GPU (sim_one_step / sim_looper semantics)."""

    def __init__(self, n_max: int, params: Params):
        self._ctx = C.c_void_p()
        rc = lib().sl_create(C.byref(self._ctx), int(n_max), C.byref(params))
        if rc:
            raise SpacelabError(rc, "sl_create: no usable gfx950 device" if rc == -5
                                else "sl_create")
        self.params = params.copy()
        self.n = 0

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = lib().sl_last_error(self._ctx)
            raise SpacelabError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    def close(self):
        if self._ctx:
            lib().sl_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, params: Params):
        self._check(lib().sl_set_params(self._ctx, C.byref(params)), "sl_set_params")
        self.params = params.copy()

    def upload(self, g: BallGroup, pos_prev: np.ndarray | None = None):
        n = g.num_particles
        s3 = (n, 3)
        pos, vel, w = _f64(g.pos, s3), _f64(g.vel, s3), _f64(g.w, s3)
        acc = None if g.acc is None else _f64(g.acc, s3)
        aacc = None if g.aacc is None else _f64(g.aacc, s3)
        R, m, moi = _f64(g.R, (n,)), _f64(g.m, (n,)), _f64(g.moi, (n,))
        pp = None if pos_prev is None else _f64(pos_prev, s3)
        self._check(lib().sl_upload(self._ctx, n, _dp(pos), _dp(vel), _dp(w), _dp(acc), _dp(aacc),
                                    _dp(R), _dp(m), _dp(moi), _dp(pp)), "sl_upload")
        self.n = n
        self._static = (R.copy(), m.copy(), moi.copy())

    def init_conditions(self) -> Energy:
        e = _Energy()
        self._check(lib().sl_init_forces(self._ctx, C.byref(e)), "sl_init_forces")
        return Energy._from(e)

    def step(self, n_steps: int, write_every: int = 0) -> list[Energy]:
        """n_steps sim_one_step calls; write step when step_count % write_every == 0."""
        if write_every > 0:
            c0 = self.step_count
            cap = (c0 + n_steps) // write_every - c0 // write_every
        else:
            cap = 0
        buf = (_Energy * max(cap, 1))()
        nw = self._check(lib().sl_step(self._ctx, int(n_steps), int(write_every), buf, max(cap, 1)),
                         "sl_step")
        return [Energy._from(buf[k]) for k in range(min(nw, cap))]

    def sim_one_step(self, write_step: bool) -> Energy | None:
        """One step (ColliderSingleCore.cpp:78-288); energies when write_step."""
        if write_step:
            e = _Energy()
            # write_every = 1 makes this one step a write step
            self._check(lib().sl_step(self._ctx, 1, 1, C.byref(e), 1), "sl_step")
            return Energy._from(e)
        self._check(lib().sl_step(self._ctx, 1, 0, None, 0), "sl_step")
        return None

    def sim_looper(self, steps: int, skip: int) -> list[Energy]:
        """Steps 1..steps-1 with a write every skip steps (ColliderSingleCore.cpp:298-331)."""
        return self.step(steps - 1, skip)

    @property
    def step_count(self) -> int:
        return int(lib().sl_step_count(self._ctx))

    def download(self) -> BallGroup:
        n = self.n
        pos = np.empty((n, 3)); vel = np.empty((n, 3)); w = np.empty((n, 3))
        acc = np.empty((n, 3)); aacc = np.empty((n, 3))
        velh = np.empty((n, 3)); wh = np.empty((n, 3))
        self._check(lib().sl_download(self._ctx, _dp(pos), _dp(vel), _dp(w), _dp(acc), _dp(aacc)),
                    "sl_download")
        self._check(lib().sl_download_half(self._ctx, _dp(velh), _dp(wh)), "sl_download_half")
        R, m, moi = self._static
        return BallGroup(pos, vel, w, R.copy(), m.copy(), moi.copy(), acc, aacc, velh, wh)

    def synchronize(self):
        self._check(lib().sl_synchronize(self._ctx), "sl_synchronize")

    def set_option(self, name: str, value: int):
        """'graphs' (0/1), 'tile_split' (0 = autotune, 1, 2, 4) or 'ball_waves' (0 = by
        size, 4, 8) and 'tile_queue' (0 = autotune, 1 = one workgroup per tile group, 2 =
        work queue) of the fast engine; 'ordered_shape' (0 = autotune, 1..n) of the
        ordered pair pass."""
        self._check(lib().sl_set_option(self._ctx, _OPTS[name], int(value)), "sl_set_option")

    def get_option(self, name: str) -> int:
        v = C.c_int()
        self._check(lib().sl_get_option(self._ctx, _OPTS[name], C.byref(v)), "sl_get_option")
        return v.value

    def profile_enable(self, enable: bool = True, reset: bool = True):
        self._check(lib().sl_profile_enable(self._ctx, int(enable), int(reset)), "sl_profile_enable")

    def profile_read(self) -> dict:
        p = _Profile()
        self._check(lib().sl_profile_read(self._ctx, C.byref(p)), "sl_profile_read")
        return {k: getattr(p, k) for k, _ in p._fields_}


def _opt_f64(a, n):
    return None if a is None else _f64(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)).copy(), (n,))


class Batch:
    """Many independent Ball_groups stepped together by the fast engine.

    Each system has its own kin, kout and dt; the other parameters are common.
    ``step`` returns, per write step, one :class:`Energy` per system.
    """

    def __init__(self, max_systems: int, max_balls: int, params: Params):
        self._b = C.c_void_p()
        rc = lib().sl_batch_create(C.byref(self._b), int(max_systems), int(max_balls),
                                   C.byref(params))
        if rc:
            raise SpacelabError(rc, "sl_batch_create: no usable gfx950 device" if rc == -5
                                else "sl_batch_create")
        self.params = params.copy()
        self.sizes: list[int] = []

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = lib().sl_batch_last_error(self._b)
            raise SpacelabError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    def close(self):
        if self._b:
            lib().sl_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def n_sys(self) -> int:
        return len(self.sizes)

    def upload(self, groups: list[BallGroup], kin=None, kout=None, dt=None):
        """kin, kout, dt: one value per system (or a scalar for all; None: the params')."""
        sizes = [g.num_particles for g in groups]
        cat = lambda f, shape: np.ascontiguousarray(
            np.concatenate([_f64(getattr(g, f), shape(g.num_particles)) for g in groups]))
        pos = cat("pos", lambda n: (n, 3)); vel = cat("vel", lambda n: (n, 3))
        w = cat("w", lambda n: (n, 3))
        R = cat("R", lambda n: (n,)); m = cat("m", lambda n: (n,)); moi = cat("moi", lambda n: (n,))
        ns = len(groups)
        nps = (C.c_int * ns)(*sizes)
        ki, ko, d = _opt_f64(kin, ns), _opt_f64(kout, ns), _opt_f64(dt, ns)
        self._check(lib().sl_batch_upload(self._b, ns, nps, _dp(pos), _dp(vel), _dp(w), _dp(R),
                                          _dp(m), _dp(moi), _dp(ki), _dp(ko), _dp(d)),
                    "sl_batch_upload")
        self.sizes = sizes
        self._static = (R, m, moi)

    def set_system_params(self, kin=None, kout=None, dt=None):
        ns = self.n_sys
        self._check(lib().sl_batch_set_system_params(
            self._b, _dp(_opt_f64(kin, ns)), _dp(_opt_f64(kout, ns)), _dp(_opt_f64(dt, ns))),
            "sl_batch_set_system_params")

    def init_conditions(self) -> list[Energy]:
        buf = (_Energy * self.n_sys)()
        self._check(lib().sl_batch_init_forces(self._b, buf), "sl_batch_init_forces")
        return [Energy._from(buf[s]) for s in range(self.n_sys)]

    def step(self, n_steps: int, write_every: int = 0) -> list[list[Energy]]:
        if write_every > 0:
            c0 = self.step_count
            cap = (c0 + n_steps) // write_every - c0 // write_every
        else:
            cap = 0
        ns = self.n_sys
        buf = (_Energy * max(cap * ns, 1))()
        nw = self._check(lib().sl_batch_step(self._b, int(n_steps), int(write_every), buf,
                                             max(cap, 1) if cap else 0), "sl_batch_step")
        return [[Energy._from(buf[k * ns + s]) for s in range(ns)] for k in range(min(nw, cap))]

    @property
    def step_count(self) -> int:
        return int(lib().sl_batch_step_count(self._b))

    def download(self) -> list[BallGroup]:
        n = sum(self.sizes)
        pos = np.empty((n, 3)); vel = np.empty((n, 3)); w = np.empty((n, 3))
        acc = np.empty((n, 3)); aacc = np.empty((n, 3))
        self._check(lib().sl_batch_download(self._b, _dp(pos), _dp(vel), _dp(w), _dp(acc),
                                            _dp(aacc)), "sl_batch_download")
        R, m, moi = self._static
        out, o = [], 0
        for k in self.sizes:
            sl_ = slice(o, o + k)
            out.append(BallGroup(pos[sl_].copy(), vel[sl_].copy(), w[sl_].copy(), R[sl_].copy(),
                                 m[sl_].copy(), moi[sl_].copy(), acc[sl_].copy(), aacc[sl_].copy()))
            o += k
        return out

    def set_option(self, name: str, value: int):
        self._check(lib().sl_batch_set_option(self._b, _OPTS[name], int(value)),
                    "sl_batch_set_option")

    def get_option(self, name: str) -> int:
        v = C.c_int()
        self._check(lib().sl_batch_get_option(self._b, _OPTS[name], C.byref(v)),
                    "sl_batch_get_option")
        return v.value


def _fp(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


class MLP:
    """The surrogate perceptron on the GPU (sl_mlp_*): n_in -> width x 3 -> n_out."""

    def __init__(self, n_in: int, n_out: int, width: int = 256, max_batch: int = 4096, seed: int = 1):
        self._m = C.c_void_p()
        rc = lib().sl_mlp_create(C.byref(self._m), n_in, width, n_out, max_batch, seed)
        if rc:
            raise SpacelabError(rc, "sl_mlp_create")
        self.n_in, self.n_out, self.width, self.max_batch = n_in, n_out, width, max_batch

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = lib().sl_mlp_last_error(self._m)
            raise SpacelabError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    def close(self):
        if self._m:
            lib().sl_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def n_params(self) -> int:
        return int(lib().sl_mlp_n_params(self._m))

    def get_params(self) -> np.ndarray:
        w = np.zeros(self.n_params, dtype=np.float32)
        self._check(lib().sl_mlp_get_params(self._m, _fp(w)), "sl_mlp_get_params")
        return w

    def set_params(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.shape == (self.n_params,)
        self._check(lib().sl_mlp_set_params(self._m, _fp(w)), "sl_mlp_set_params")

    def forward(self, x: np.ndarray):
        """(outputs, class probabilities, predictive entropy) of a batch."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        B = x.shape[0]
        y = np.zeros((B, self.n_out), np.float32); p = np.zeros((B, 3), np.float32)
        e = np.zeros(B, np.float32)
        self._check(lib().sl_mlp_forward(self._m, B, _fp(x), _fp(y), _fp(p), _fp(e)), "sl_mlp_forward")
        return y, p, e

    def gradients(self, x, cls, reg):
        x = np.ascontiguousarray(x, dtype=np.float32)
        cls = np.ascontiguousarray(cls, dtype=np.int32)
        reg = np.ascontiguousarray(reg, dtype=np.float32)
        loss = C.c_float()
        g = np.zeros(self.n_params, np.float32)
        self._check(lib().sl_mlp_gradients(self._m, x.shape[0], _fp(x), _ip(cls), _fp(reg), C.byref(loss), _fp(g)),
                    "sl_mlp_gradients")
        return float(loss.value), g

    def train_step(self, x, cls, reg, lr: float) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        cls = np.ascontiguousarray(cls, dtype=np.int32)
        reg = np.ascontiguousarray(reg, dtype=np.float32)
        loss = C.c_float()
        self._check(lib().sl_mlp_train_step(self._m, x.shape[0], _fp(x), _ip(cls), _fp(reg), lr, C.byref(loss)),
                    "sl_mlp_train_step")
        return float(loss.value)


def _csv_text(fn, what: str, *args, guess: int = 256) -> str:
    """One formatting pass into `guess` bytes; a second only if the text is longer."""
    buf = C.create_string_buffer(guess + 1)
    n = fn(*args, buf, guess + 1)
    if n < 0:
        raise SpacelabError(int(n), what)
    if n > guess:
        buf = C.create_string_buffer(int(n) + 1)
        fn(*args, buf, n + 1)
    return buf.raw[:n].decode()


def csv_sim_header(n: int) -> str:
    """simData.csv header line for n balls (sl_csv_sim_header)."""
    return _csv_text(lib().sl_csv_sim_header, "sl_csv_sim_header", int(n), guess=80 * int(n) + 16)


def csv_sim_row(pos, w, vel) -> str:
    """One simData.csv row, without its leading newline (sl_csv_sim_row)."""
    n = len(pos)
    # a number takes at most 13 characters at 6 digits
    return _csv_text(lib().sl_csv_sim_row, "sl_csv_sim_row", n, _dp(_f64(pos, (n, 3))), _dp(_f64(w, (n, 3))),
                     _dp(_f64(vel, (n, 3))), guess=154 * n + 16)


def csv_energy_row(time: float, PE: float, KE: float, p, L, digits: int = 12) -> str:
    """One energy.csv row "t,PE,KE,E,|p|,|L|", without its leading newline."""
    return _csv_text(lib().sl_csv_energy_row, "sl_csv_energy_row", float(time), float(PE), float(KE),
                     _dp(_f64(p, (3,))), _dp(_f64(L, (3,))), int(digits))


def csv_constants(R, m, moi) -> str:
    """The constants.csv text (one "R,m,moi" line per ball)."""
    n = len(R)
    return _csv_text(lib().sl_csv_constants, "sl_csv_constants", n, _dp(_f64(R, (n,))), _dp(_f64(m, (n,))),
                     _dp(_f64(moi, (n,))), guess=42 * n + 16)


def csv_last_line(path: str) -> str:
    """The last line of a file as the reference's getLastLine reads it."""
    size = os.path.getsize(path) if os.path.exists(path) else 0
    return _csv_text(lib().sl_csv_last_line, f"sl_csv_last_line({path})", os.fsencode(path), guess=size)


def csv_parse_sim_row(row: str):
    """(pos, w, vel) of every ball of a simData.csv row."""
    b = row.encode()
    n = lib().sl_csv_count_balls(b)
    pos, w, vel = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    rc = lib().sl_csv_parse_sim_row(b, n, _dp(pos), _dp(w), _dp(vel))
    if rc:
        raise SpacelabError(rc, "sl_csv_parse_sim_row")
    return pos, w, vel


def csv_parse_constants(text: str, n: int):
    """(R, m, moi) of n balls from constants.csv text."""
    R, m, moi = np.zeros(n), np.zeros(n), np.zeros(n)
    rc = lib().sl_csv_parse_constants(text.encode(), n, _dp(R), _dp(m), _dp(moi))
    if rc:
        raise SpacelabError(rc, "sl_csv_parse_constants")
    return R, m, moi


def csv_energy_columns(header: str) -> int:
    """6 or 9: the columns of an energy.csv file with this header line."""
    k = lib().sl_csv_energy_columns(header.encode())
    if k < 0:
        raise SpacelabError(k, "sl_csv_energy_columns")
    return k


def csv_read_energy(path: str) -> np.ndarray:
    """Rows of an energy.csv file in either layout (Time,PE,KE,E,p,L[,Bound,Unbound,m]),
    one row per line after the header."""
    with open(path, "rb") as f:
        lines = f.read().decode().split("\n")
    k = csv_energy_columns(lines[0])
    rows = [ln for ln in lines[1:] if ln.strip()]
    out = np.zeros((len(rows), k))
    for i, ln in enumerate(rows):
        rc = lib().sl_csv_parse_energy_row(ln.encode(), k, _dp(out[i]))
        if rc:
            raise SpacelabError(rc, f"sl_csv_parse_energy_row({path}:{i + 2})")
    return out


def collider_program() -> str:
    """Path of the C++ driver (spacelab_collider) built beside the library."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "spacelab_collider")


def probe_rotation() -> np.ndarray:
    """Lane layout of the wavefront rotation used by the fast pair kernel."""
    out = (C.c_int * 192)()
    rc = lib().sl_probe_rotation(out)
    if rc:
        raise SpacelabError(rc, "sl_probe_rotation")
    return np.array(out[:], dtype=np.int64).reshape(3, 64)


def probe_copy(nbytes: int, reps: int = 1) -> float:
    """Run the plain copy kernel; returns ms per copy."""
    ms = C.c_double()
    rc = lib().sl_probe_copy(int(nbytes), int(reps), C.byref(ms))
    if rc:
        raise SpacelabError(rc, "sl_probe_copy")
    return ms.value


def probe_divsqrt(a: np.ndarray, b: np.ndarray):
    """Device a/b, sqrt(a), log(a) with the ordered kernels' compilation."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    q = np.empty_like(a); s = np.empty_like(a); lg = np.empty_like(a)
    rc = lib().sl_probe_divsqrt(len(a), _dp(a), _dp(b), _dp(q), _dp(s), _dp(lg))
    if rc:
        raise SpacelabError(rc, "sl_probe_divsqrt")
    return q, s, lg


def probe_div_rcp(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Device a/b as the ordered kernels divide by a distance or a mass (through the
    correctly rounded reciprocal and two residual corrections)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    q = np.empty_like(a)
    rc = lib().sl_probe_div_rcp(len(a), _dp(a), _dp(b), _dp(q))
    if rc:
        raise SpacelabError(rc, "sl_probe_div_rcp")
    return q

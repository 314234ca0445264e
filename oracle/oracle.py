"""ctypes front end to the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module; the product package never does.  It wraps

* ``oracle/build/liboracle.so`` -- the plain-C restatement of the reference's
  ``sim_one_step`` (ColliderSingleCore/ColliderSingleCore.cpp:78-288) and
  ``init_conditions`` (ball_group.hpp:777-947);
* ``oracle/_ref/ref_harness`` -- the reference's own ``sim_one_step`` compiled
  from its sources (built only where those sources are present).

State arrays are float64, AoS xyz of shape (n, 3), as the reference's vec3
arrays are laid out.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
import tempfile
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")
REF_MAGIC = 0x534C4731


class OrcParams(C.Structure):
    _fields_ = [("G", C.c_double), ("Ha", C.c_double), ("h_min", C.c_double),
                ("u_s", C.c_double), ("u_r", C.c_double), ("kin", C.c_double),
                ("kout", C.c_double), ("dt", C.c_double), ("far_gravity", C.c_int)]


class OrcEnergy(C.Structure):
    _fields_ = [("PE", C.c_double), ("KE", C.c_double), ("p", C.c_double * 3),
                ("L", C.c_double * 3)]

    def as_array(self) -> np.ndarray:
        return np.array([self.PE, self.KE, *self.p, *self.L], dtype=np.float64)


def build(quiet: bool = True) -> None:
    """Compile the oracle (and the reference harness where its sources exist)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.c_int, C.POINTER(OrcParams), C.c_int, C.c_int, C.c_int,
                              dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, C.c_void_p,
                              C.POINTER(OrcEnergy), C.POINTER(OrcEnergy), C.c_int, C.c_int,
                              C.c_int]
        L.orc_vdw_force_mag.restype = C.c_double
        L.orc_vdw_force_mag.argtypes = [C.c_double] * 4
        L.orc_vdw_potential.restype = C.c_double
        L.orc_vdw_potential.argtypes = [C.c_double] * 4
        _lib = L
    return _lib


@dataclass
class Params:
    """Physics parameters (dust_const.hpp:21-53).  ``h_min`` is computed as the
    reference does, ``scaleBalls * .1`` (dust_const.hpp:38-39)."""
    G: float = 6.67e-8
    Ha: float = 4.7e-12
    h_min: float = 1e-5 * .1
    u_s: float = 0.1
    u_r: float = 1e-5
    kin: float = -1.0
    kout: float = -1.0
    dt: float = -1.0
    far_gravity: int = 0

    def c(self) -> OrcParams:
        return OrcParams(self.G, self.Ha, self.h_min, self.u_s, self.u_r, self.kin,
                         self.kout, self.dt, int(self.far_gravity))


@dataclass
class State:
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
    pos_prev: np.ndarray | None = None

    @property
    def n(self) -> int:
        return int(self.R.shape[0])

    def copy(self) -> "State":
        cp = lambda a: None if a is None else np.array(a, dtype=np.float64, copy=True)
        return State(cp(self.pos), cp(self.vel), cp(self.w), cp(self.R), cp(self.m), cp(self.moi),
                     cp(self.acc), cp(self.aacc), cp(self.velh), cp(self.wh), cp(self.pos_prev))


def run(state: State, params: Params, n_steps: int, skip: int, mode: str = "tri",
        do_init: bool = True, nthreads: int = 1):
    """Step-0 pass (if ``do_init``) then steps 1..n_steps-1, as sim_looper does.

    ``mode`` "tri" is the reference's triangular loop, "gather" the ordered
    per-ball gather.  Returns (new_state, e_init[8], e_writes[k, 8])."""
    s = state.copy()
    n = s.n
    z = lambda: np.zeros((n, 3), dtype=np.float64)
    for name in ("acc", "aacc", "velh", "wh"):
        if getattr(s, name) is None:
            setattr(s, name, z())
    pp = s.pos_prev if s.pos_prev is not None else np.array(s.pos, copy=True)
    if mode == "gather":
        s.pos_prev = pp
        pp_arg = s.pos_prev.ctypes.data
    else:
        pp_arg = None if s.pos_prev is None else s.pos_prev.ctypes.data
    n_w = (n_steps - 1) // skip if skip > 0 else 0
    e_out = (OrcEnergy * max(n_w, 1))()
    e_init = OrcEnergy()
    P = params.c()
    got = lib().orc_run(n, C.byref(P), n_steps, skip, 1 if do_init else 0, s.pos, s.vel, s.velh, s.w,
                        s.wh, s.acc, s.aacc, s.R, s.m, s.moi, pp_arg, C.byref(e_init), e_out,
                        max(n_w, 1), 0 if mode == "tri" else 1, nthreads)
    ew = np.array([e_out[k].as_array() for k in range(min(got, n_w))]).reshape(-1, 8)
    return s, e_init.as_array(), ew


def ref_available() -> bool:
    return os.path.exists(REF_HARNESS)


def run_reference_csv(state: State, dt: float, kin: float, kout: float, n_steps: int, skip: int,
                      outdir: str) -> None:
    """The reference's own sim_init_write and sim_looper through ref_harness --csv:
    constants.csv, simData.csv and energy.csv written into ``outdir``."""
    if not ref_available():
        raise FileNotFoundError(REF_HARNESS)
    os.makedirs(outdir, exist_ok=True)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", REF_MAGIC, state.n, n_steps, skip))
            f.write(struct.pack("<3d", dt, kin, kout))
            for a in (state.pos, state.vel, state.w, state.R, state.m, state.moi):
                f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        r = subprocess.run([REF_HARNESS, "--csv", fin, os.path.abspath(outdir)], capture_output=True,
                           text=True, cwd=td)
        if r.returncode != 0:
            raise RuntimeError(f"ref_harness --csv failed ({r.returncode}): {r.stderr[-2000:]}")


def run_reference_calib(state: State, custom_speed: float) -> np.ndarray:
    """The reference's own calc_helpfuls, calc_v_collapse and three calibrate_dt calls
    (ref_harness --calib): rows {kin, kout, dt, steps, skip, v_collapse, v_max,
    v_max_prev, soc} after (step 0, velocity rule), (step 100, velocity rule with the
    velocities halved) and (step 200, custom_speed)."""
    if not ref_available():
        raise FileNotFoundError(REF_HARNESS)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", REF_MAGIC, state.n, 1, 1))
            f.write(struct.pack("<3d", custom_speed, 0.0, 0.0))
            for a in (state.pos, state.vel, state.w, state.R, state.m, state.moi):
                f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        r = subprocess.run([REF_HARNESS, "--calib", fin, fout], capture_output=True, text=True, cwd=td)
        if r.returncode != 0:
            raise RuntimeError(f"ref_harness --calib failed ({r.returncode}): {r.stderr[-2000:]}")
        return np.frombuffer(open(fout, "rb").read(), dtype="<f8").reshape(3, 9).copy()


TWO_CLUSTER_FIELDS = ("kin", "kout", "dt", "steps", "skip", "v_collapse", "v_max", "v_max_prev", "soc",
                      "m_total", "initial_radius", "r_max")


def run_reference_two_cluster(directory: str, projectile: str, target: str, custom_speed: float) -> dict:
    """The reference's own two-cluster constructor, Ball_group(path, projectile, target,
    customVel) (ball_group.hpp:100-112, 1368-1442), on the run files
    <directory>/<name>simData.csv and constants.csv (ref_harness --two-cluster).
    Returns the merged group after calibrate_dt and to_origin (pos, vel, w, R, m, moi)
    and its run-control values (TWO_CLUSTER_FIELDS)."""
    if not ref_available():
        raise FileNotFoundError(REF_HARNESS)
    directory = os.path.abspath(directory) + os.sep
    with tempfile.TemporaryDirectory() as td:
        fout = os.path.join(td, "out.bin")
        r = subprocess.run([REF_HARNESS, "--two-cluster", directory, projectile, target, repr(float(custom_speed)),
                            fout], capture_output=True, text=True, cwd=td)
        if r.returncode != 0:
            raise RuntimeError(f"ref_harness --two-cluster failed ({r.returncode}): {r.stderr[-2000:]}")
        raw = open(fout, "rb").read()
    n = struct.unpack("<i", raw[:4])[0]
    v = np.frombuffer(raw[4:], dtype="<f8")
    out, k = {}, 0
    for name, size in (("pos", 3), ("vel", 3), ("w", 3), ("R", 1), ("m", 1), ("moi", 1)):
        out[name] = v[k:k + size * n].reshape((n, 3) if size == 3 else (n,)).copy()
        k += size * n
    out["run"] = v[k:k + len(TWO_CLUSTER_FIELDS)].copy()
    return out


def run_reference(state: State, dt: float, kin: float, kout: float, n_steps: int, skip: int):
    """Run the reference's own sim_one_step through oracle/_ref/ref_harness.

    Physics constants are the reference's compile-time ones (dust_const.hpp).
    Returns (dict of final arrays, e_init[8], e_writes[k, 8])."""
    if not ref_available():
        raise FileNotFoundError(REF_HARNESS)
    n = state.n
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", REF_MAGIC, n, n_steps, skip))
            f.write(struct.pack("<3d", dt, kin, kout))
            for a in (state.pos, state.vel, state.w, state.R, state.m, state.moi):
                f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        r = subprocess.run([REF_HARNESS, fin, fout], capture_output=True, text=True, cwd=td)
        if r.returncode != 0:
            raise RuntimeError(f"ref_harness failed ({r.returncode}): {r.stderr[-2000:]}")
        raw = open(fout, "rb").read()
    off = 0
    out = {}
    for name in ("pos", "vel", "velh", "w", "wh", "acc", "aacc"):
        out[name] = np.frombuffer(raw, dtype="<f8", count=3 * n, offset=off).reshape(n, 3).copy()
        off += 24 * n
    e_init = np.frombuffer(raw, dtype="<f8", count=8, offset=off).copy()
    off += 64
    (nw,) = struct.unpack_from("<i", raw, off)
    off += 4
    ew = np.frombuffer(raw, dtype="<f8", count=8 * nw, offset=off).reshape(nw, 8).copy()
    return out, e_init, ew

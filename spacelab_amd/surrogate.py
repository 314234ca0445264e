"""Learned collision surrogate: DEM-labelled collision datasets from batched
ensembles, the GPU perceptron (:class:`spacelab_amd.MLP`) trained on them, and
entropy-gated inference that hands every collision whose predictive entropy is high
back to DEM.

The reference has no counterpart: it simulates every collision with DEM, one
sim_one_step per time step (ColliderSingleCore.cpp:78-288).  Here DEM is the arbiter of
the surrogate: its labels train the network and decide every collision the network
is unsure of.

Collision set-ups follow the two-cluster arrangement (ball_group.hpp:1368-1442):
target and projectile aggregates from the seeded BPCA generator, random orientations,
impact parameter b_frac * (R_t + R_p), zero total momentum.  The projectile is moved
along the line of approach until its first contact is ``gap`` median radii away, so
DEM steps are spent on the collision rather than the approach.  Springs are calibrated
at the impact speed by updateDTK's rule (ball_group.hpp:1313-1329) and divided by
``softening``.  dust_const.hpp:35's kConsts carries a factor 3e3 that the kConsts of
asteroid_const.hpp:25 and ColliderParallel/structures.hpp:1106 lack; a softening s
takes sqrt(s) times fewer steps per radius of travel (t_c = pi sqrt(m/k), dt = t_c/50).
The deepest overlap of an impact grows as sqrt(s N_p) (the projectile's N_p balls
press on the first contacts), so s stays well below 3e3: at s = 3e3 a 64-ball
projectile at 1000 cm/s drives balls through one another.
"""
from __future__ import annotations

import argparse
import json
import math
import time
from dataclasses import asdict, dataclass, field, replace

import numpy as np

from . import (N_DESCRIPTORS, OUTCOME_BOUNCE, OUTCOME_FRAGMENT, OUTCOME_STICK, ORDER_FAST,
               PREC_FP64, Batch, CollisionSetup, MLP, calibrate, collision_approach,
               collision_descriptors_batch, collision_outcome, generate_collision, preset)

CLASS_NAMES = ("stick", "bounce", "fragment")
# regression targets: mass fraction of the largest body, share of the projectile's mass
# that ends in the target's body, and the centre-of-mass velocities (x, y) of the two
# largest bodies divided by the impact speed (zero-momentum frame)
REG_NAMES = ("largest_mass_frac", "mass_transfer", "v1x", "v1y", "v2x", "v2y")
N_REG = len(REG_NAMES)


@dataclass(frozen=True)
class Spec:
    """One collision: body sizes, generator seed, impact speed (cm/s), b / (R_t + R_p)."""
    n_t: int
    n_p: int
    seed: int
    speed: float
    b_frac: float


@dataclass
class DemSettings:
    steps: int = 40000          # DEM steps after set-up (same for every collision)
    gap: float = 0.05           # travel before first contact, median radii
    link: float = 0.1           # outcome link distance, median radii
    r_med: float = 1e-5         # median ball radius, cm
    sigma: float = 0.2          # lognormal width of the radii
    softening: float = 100.0    # divisor of the calibrated spring constants
    precision: int = PREC_FP64
    max_systems: int = 512      # collisions per batched launch
    preset: str = "config"
    tile_split: int = 2         # fixed, so a collision's result does not depend on its batch


def _log_uniform_int(rng, lo, hi, n):
    return np.minimum(np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi + 1), n))), hi).astype(int)


def sample_specs(count: int, seed: int, n_t=(32, 512), n_p=(1, 64), speed=(10.0, 1000.0)) -> list[Spec]:
    """Log-uniform body sizes, uniform speed and impact parameter, per-collision seeds."""
    rng = np.random.default_rng(seed)
    nt = _log_uniform_int(rng, *n_t, count)
    npj = _log_uniform_int(rng, *n_p, count)
    v = rng.uniform(*speed, count)
    b = rng.uniform(0.0, 1.0, count)
    seeds = rng.integers(1, 2 ** 62, count)
    return [Spec(int(a), int(c), int(s), float(x), float(y)) for a, c, s, x, y in zip(nt, npj, seeds, v, b)]


@dataclass
class Prepared:
    setup: CollisionSetup
    kin: float
    kout: float
    dt: float
    met: bool


def prepare(spec: Spec, s: DemSettings) -> Prepared:
    setup = generate_collision(spec.n_t, spec.n_p, spec.seed, spec.speed, spec.b_frac, gap=2.0,
                               r_med=s.r_med, sigma=s.sigma)
    met, _ = collision_approach(setup, s.gap * s.r_med)
    g = setup.group
    p = calibrate(preset(s.preset), g.R, g.m, velocity=spec.speed, cor=0.64, dt_rule="tc50")
    kin, kout, dt = p.kin, p.kout, p.dt
    if s.softening != 1.0:
        # t_c = pi sqrt(m / k): dividing k by s multiplies t_c/50 by sqrt(s)
        kin, kout, dt = kin / s.softening, kout / s.softening, dt * math.sqrt(s.softening)
    return Prepared(setup, kin, kout, dt, met)


def descriptors(preps: list[Prepared], s: DemSettings, threads: int = 0) -> np.ndarray:
    """Descriptor rows of the set-ups, computed on ``threads`` host threads (0: all cores)."""
    if not preps:
        return np.zeros((0, N_DESCRIPTORS), dtype=np.float32)
    return collision_descriptors_batch([q.setup for q in preps], s.link * s.r_med, threads).astype(np.float32)


def targets(o, speed: float) -> np.ndarray:
    return np.array([o.largest_mass_frac, o.mass_transfer, o.v_largest[0] / speed, o.v_largest[1] / speed,
                     o.v_second[0] / speed, o.v_second[1] / speed])


class DemRunner:
    """Runs collisions to their outcome with the batched fast engine."""

    def __init__(self, s: DemSettings, max_balls: int):
        self.s = s
        p = preset(s.preset).replace(order=ORDER_FAST, precision=s.precision)
        self.batch = Batch(s.max_systems, max_balls, p)
        self.batch.set_option("tile_split", s.tile_split)

    def close(self):
        self.batch.close()

    def run(self, preps: list[Prepared], steps: int | None = None, looks: tuple = ()):
        """(class labels, regression targets, outcomes) of the collisions after
        ``steps`` (default: the settings') DEM steps.  With ``looks`` (step counts below
        ``steps``) it returns a fourth item: {step: look-ahead features} of the same
        trajectories at those steps."""
        s = self.s
        steps = s.steps if steps is None else steps
        looks = sorted(h for h in looks if 0 < h < steps)
        labels, regs, outs = [], [], []
        feats = {h: [] for h in looks}
        for c0 in range(0, len(preps), s.max_systems):
            chunk = preps[c0:c0 + s.max_systems]
            b = self.batch
            b.upload([q.setup.group for q in chunk], kin=[q.kin for q in chunk], kout=[q.kout for q in chunk],
                     dt=[q.dt for q in chunk])
            b.init_conditions()
            done = 0
            for h in looks:
                b.step(h - done)
                done = h
                for q, g in zip(chunk, b.download()):
                    feats[h].append(look_row(collision_outcome(q.setup.n_t, g, s.link * s.r_med), q.setup.speed))
            b.step(steps - done)
            for q, g in zip(chunk, b.download()):
                o = collision_outcome(q.setup.n_t, g, s.link * s.r_med)
                labels.append(o.label)
                regs.append(targets(o, q.setup.speed))
                outs.append(o)
        res = (np.array(labels, dtype=np.int32), np.array(regs, dtype=np.float32).reshape(-1, N_REG), outs)
        if looks:
            return res + ({h: np.array(v, dtype=np.float32).reshape(-1, N_LOOK) for h, v in feats.items()},)
        return res


@dataclass
class Normaliser:
    mean: np.ndarray
    std: np.ndarray

    @staticmethod
    def fit(X: np.ndarray) -> "Normaliser":
        mu = X.mean(0)
        sd = X.std(0)
        return Normaliser(mu, np.where(sd > 0, sd, 1.0))

    def __call__(self, X: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray((X - self.mean) / self.std, dtype=np.float32)


@dataclass
class Surrogate:
    """An ensemble of perceptrons (different initialisations and batch orders) with
    their input normalisation and the entropy gate.  The class distribution is the
    mean of the members' softmax outputs and the gate reads its entropy (predictive
    entropy of the ensemble)."""
    width: int = 256
    seed: int = 1
    n_models: int = 4
    max_batch: int = 4096
    n_in: int = N_DESCRIPTORS
    threshold: float = float("inf")
    norm: Normaliser | None = None
    mlps: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        if not self.mlps:
            self.mlps = [MLP(self.n_in, 3 + N_REG, self.width, self.max_batch, self.seed * 1000 + k)
                         for k in range(self.n_models)]

    def close(self):
        for m in self.mlps:
            m.close()
        self.mlps = []

    def fit(self, X, cls, reg, epochs=30, batch=256, lr=1e-3, seed=0, log=None) -> list[float]:
        """Adam on shuffled mini-batches with a cosine step-size decay; returns the
        members' mean loss of each epoch."""
        self.norm = Normaliser.fit(X)
        Xn = self.norm(X)
        cls = np.ascontiguousarray(cls, dtype=np.int32)
        reg = np.ascontiguousarray(reg, dtype=np.float32)
        n = len(Xn)
        hist = np.zeros(epochs)
        for k, m in enumerate(self.mlps):
            rng = np.random.default_rng(seed * 1000 + k)
            for ep in range(epochs):
                perm = rng.permutation(n)
                tot, cnt = 0.0, 0
                lr_ep = lr * 0.5 * (1 + math.cos(math.pi * ep / epochs))
                for b0 in range(0, n - batch + 1, batch):
                    idx = perm[b0:b0 + batch]
                    tot += m.train_step(Xn[idx], cls[idx], reg[idx], lr_ep) * len(idx)
                    cnt += len(idx)
                hist[ep] += tot / max(cnt, 1) / len(self.mlps)
        if log:
            log(f"loss by epoch: {np.round(hist, 4).tolist()}")
        return hist.tolist()

    def predict(self, X):
        """(class, regression outputs, class probabilities, predictive entropy) of
        descriptor rows."""
        Xn = self.norm(X)
        P = np.zeros((len(Xn), 3))
        Y = np.zeros((len(Xn), N_REG))
        for m in self.mlps:
            for k in range(0, len(Xn), self.max_batch):
                y, p, _ = m.forward(Xn[k:k + self.max_batch])
                P[k:k + len(p)] += p
                Y[k:k + len(y)] += y[:, 3:]
        P /= len(self.mlps)
        Y /= len(self.mlps)
        ent = -(P * np.log(np.where(P > 0, P, 1.0))).sum(1)
        return P.argmax(1).astype(np.int32), Y.astype(np.float32), P, ent

    def calibrate_gate(self, X, cls, target: float) -> dict:
        """Smallest fallback share whose gated agreement on (X, cls) reaches ``target``:
        collisions above the entropy threshold go to DEM, which is right by definition."""
        pred, _, _, ent = self.predict(X)
        order = np.argsort(-ent, kind="stable")
        wrong = (pred != cls)[order]
        n = len(cls)
        # after sending the k most uncertain to DEM, the errors left are the wrong ones among the rest
        left = wrong[::-1].cumsum()[::-1]           # left[k] = errors among order[k:]
        agree = 1.0 - np.append(left, 0) / n       # agreement when k = 0..n go to DEM
        k = int(np.argmax(agree >= target)) if np.any(agree >= target) else n
        # the gate is ent > threshold: the k most uncertain go to DEM (ties with the
        # (k+1)-th stay with the surrogate)
        self.threshold = float(ent[order[k]]) if k < n else -1.0
        fb = ent > self.threshold
        final = np.where(fb, cls, pred)
        return {"threshold": self.threshold, "fallback_share": float(fb.mean()),
                "agreement": float((final == cls).mean()), "raw_accuracy": float((pred == cls).mean())}


def run_gated(preps: list[Prepared], sur: Surrogate, dem: DemRunner, s: DemSettings):
    """Outcome classes and regression targets for every collision: the surrogate's,
    except where its entropy exceeds the gate, where DEM's.  Returns (cls, reg,
    fallback mask, timings in seconds: descriptors, inference, DEM, total)."""
    t0 = time.perf_counter()
    X = descriptors(preps, s)
    t1 = time.perf_counter()
    cls, reg, _, ent = sur.predict(X)
    fb = ent > sur.threshold
    t2 = time.perf_counter()
    idx = np.flatnonzero(fb)
    if len(idx):
        c, r, _ = dem.run([preps[i] for i in idx])
        cls[idx] = c
        reg[idx] = r
    t3 = time.perf_counter()
    return cls, reg, fb, {"descriptors": t1 - t0, "inference": t2 - t1, "dem": t3 - t2, "total": t3 - t0}


def evaluate(train: dict, heldout: dict, target: float = 0.95, z: float = 2.0, val_frac: float = 0.5,
             epochs: int = 40, lr: float = 1e-3, n_models: int = 4, seed: int = 1, log=None) -> dict:
    """Train on ``train`` with a validation share held back to set the entropy gate,
    then run the held-out collisions twice, by DEM alone and by the gated surrogate,
    with live DEM deciding the true outcome of the second.

    The gate aims above ``target`` by z standard errors of the difference between two
    binomial estimates, the validation share's and the held-out set's, so that the
    held-out agreement clears ``target`` with probability about Phi(z).  Accuracy
    grows little with the training set beyond a few thousand collisions, so half of
    them go to validation by default."""
    s = settings_of(heldout)
    n = len(train["cls"])
    perm = np.random.default_rng(seed).permutation(n)
    nv = int(n * val_frac)
    va, tr = perm[:nv], perm[nv:]
    X, C, Rg = train["X"], train["cls"], train["reg"]
    sur = Surrogate(seed=seed, n_models=n_models)
    nh = len(heldout["cls"])
    margin = z * math.sqrt(target * (1 - target) * (1.0 / max(nv, 1) + 1.0 / nh))
    dem = None
    try:
        t0 = time.perf_counter()
        hist = sur.fit(X[tr], C[tr], Rg[tr], epochs=epochs, lr=lr, seed=seed, log=log)
        t_train = time.perf_counter() - t0
        gate = sur.calibrate_gate(X[va], C[va], min(target + margin, 1.0))
        if log:
            log(f"trained {epochs} epochs in {t_train:.1f} s, loss {hist[0]:.3f} -> {hist[-1]:.3f}; gate {gate}")
        specs = specs_of(heldout)
        preps = [prepare(q, s) for q in specs]
        max_balls = s.max_systems * max(q.setup.group.num_particles for q in preps)
        dem = DemRunner(s, max_balls)
        dem.run(preps[:16], steps=256)          # first-use set-up (tile-split tuning, graphs)
        t0 = time.perf_counter()
        c_dem, r_dem, _ = dem.run(preps)
        t_dem = time.perf_counter() - t0
        c_g, r_g, fb, tg = run_gated(preps, sur, dem, s)
        raw, _, _, _ = sur.predict(descriptors(preps, s))
    finally:
        sur.close()
        if dem is not None:
            dem.close()
    nh = len(preps)
    keep = ~fb
    mse = float(((r_g[keep] - r_dem[keep]) ** 2).mean()) if keep.any() else 0.0
    mse0 = float(((Rg[tr].mean(0) - r_dem[keep]) ** 2).mean()) if keep.any() else 0.0
    c_d = t_dem / nh                                         # DEM seconds per collision, full batches
    c_s = (tg["descriptors"] + tg["inference"]) / nh          # surrogate seconds per collision
    return {
        "n_train": int(len(tr)), "n_val": int(nv), "n_heldout": nh, "epochs": epochs, "n_models": n_models,
        "train_s": t_train, "gate_target": target + margin,
        "raw_accuracy": float((raw == c_dem).mean()), "agreement": float((c_g == c_dem).mean()),
        "fallback_share": float(fb.mean()), "threshold": sur.threshold, "val_gate": gate,
        "dem_s": t_dem, "gated_s": tg["total"], "gated_parts_s": tg,
        "collisions_per_s_dem": nh / t_dem, "collisions_per_s_gated": nh / tg["total"],
        "gain": t_dem / tg["total"],
        # fallbacks of a long stream queue up into full batches: per collision, the
        # surrogate's cost plus the fallback share of a full-batch DEM collision
        "gain_streaming": c_d / (c_s + fb.mean() * c_d),
        "reg_mse": mse, "reg_mse_mean_predictor": mse0,
        "heldout_classes": np.bincount(c_dem, minlength=3).tolist(),
        "dem_steps": s.steps, "softening": s.softening,
    }


# ---- look-ahead cascade -------------------------------------------------------------
LOOK_AHEAD = 4000     # DEM steps of the cascade's look-ahead stage


@dataclass
class Cascade:
    """Two entropy gates in a row.  Stage 1 reads a set-up's descriptors.  A collision
    whose stage-1 entropy exceeds ``t1`` runs ``look`` DEM steps, and stage 2 reads the
    descriptors plus the provisional outcome at that step (:func:`look_row`).  A
    collision whose stage-2 entropy exceeds ``t2`` runs DEM to the end, from its set-up,
    so that at a fixed tile split its result is the DEM-alone one bit for bit."""
    s1: Surrogate
    s2: Surrogate
    look: int = LOOK_AHEAD
    t1: float = float("inf")
    t2: float = float("inf")

    def close(self):
        self.s1.close()
        self.s2.close()

    def calibrate(self, X, L, cls, target: float, look_cost: float) -> dict:
        """Thresholds that reach ``target`` agreement on (X, L, cls) at the least DEM
        cost per collision, look_cost x (stage-1 fallback share) + (share run to the
        end), in units of one full DEM collision.  Stage-1 thresholds tried: the 201
        quantiles of the stage-1 entropy and -1 (everything to stage 2); for each, the
        stage-2 threshold is the loosest that reaches the target."""
        p1, _, _, e1 = self.s1.predict(X)
        p2, _, _, e2 = self.s2.predict(np.hstack([X, L]))
        n = len(cls)
        need = int(math.ceil(target * n - 1e-9))     # collisions that must agree
        cands = np.unique(np.concatenate([np.quantile(e1, np.linspace(0, 1, 201)), [-1.0]]))
        best = None
        for t1 in cands:
            F1 = e1 > t1
            ok1 = int(((p1 == cls) & ~F1).sum())
            idx = np.flatnonzero(F1)
            o = idx[np.argsort(-e2[idx], kind="stable")]
            right = p2[o] == cls[o]
            # sending the k most uncertain of F1 to DEM: k right by definition, plus the
            # stage-2 predictions that are right among the rest
            right_after = np.append(right[::-1].cumsum()[::-1], 0)
            hit = np.flatnonzero(ok1 + np.arange(len(o) + 1) + right_after >= need)
            if len(hit) == 0:
                continue
            k = int(hit[0])
            # e2 > t2 takes o[:k] and any ties with o[k-1]
            t2 = float(np.nextafter(e2[o[k - 1]], -np.inf)) if k > 0 else float("inf")
            F2 = F1 & (e2 > t2)
            cost = look_cost * F1.mean() + F2.mean()
            if best is None or cost < best[0] - 1e-12:
                best = (float(cost), float(t1), t2)
        if best is None:                                   # only DEM reaches the target
            best = (1.0 + look_cost, -1.0, -1.0)
        _, self.t1, self.t2 = best
        F1 = e1 > self.t1
        F2 = F1 & (e2 > self.t2)
        final = np.where(F2, cls, np.where(F1, p2, p1))
        return {"t1": self.t1, "t2": self.t2, "stage1_share": float(F1.mean()), "dem_share": float(F2.mean()),
                "agreement": float((final == cls).mean()), "dem_cost": best[0]}


def run_cascade(preps: list[Prepared], cas: Cascade, dem: "DemRunner", s: DemSettings):
    """Outcome classes and regression targets of every collision through the cascade.
    Returns (cls, reg, stage-1 fallback mask, run-to-the-end mask, look-ahead features
    of the stage-1 fallbacks, timings in seconds)."""
    n = len(preps)
    t0 = time.perf_counter()
    X = descriptors(preps, s)
    t1 = time.perf_counter()
    cls, reg, _, e1 = cas.s1.predict(X)
    F1 = np.flatnonzero(e1 > cas.t1)
    t2 = time.perf_counter()
    m1 = np.zeros(n, dtype=bool)
    m2 = np.zeros(n, dtype=bool)
    L = np.zeros((len(F1), N_LOOK), dtype=np.float32)
    t3 = t4 = t2
    if len(F1):
        m1[F1] = True
        _, _, outs = dem.run([preps[i] for i in F1], steps=cas.look)
        L = np.array([look_row(o, preps[i].setup.speed) for o, i in zip(outs, F1)],
                     dtype=np.float32).reshape(-1, N_LOOK)
        t3 = time.perf_counter()
        c2, r2, _, e2 = cas.s2.predict(np.hstack([X[F1], L]))
        cls[F1] = c2
        reg[F1] = r2
        F2 = F1[e2 > cas.t2]
        t4 = time.perf_counter()
        if len(F2):
            m2[F2] = True
            c, r, _ = dem.run([preps[i] for i in F2])
            cls[F2] = c
            reg[F2] = r
    t5 = time.perf_counter()
    return cls, reg, m1, m2, L, {"descriptors": t1 - t0, "stage1": t2 - t1, "look_dem": t3 - t2,
                                 "stage2": t4 - t3, "dem": t5 - t4, "total": t5 - t0}


def merge_datasets(parts: list[dict]) -> dict:
    """Several datasets of the same settings as one (rows in order)."""
    out = dict(parts[0])
    for k, v in parts[0].items():
        if k in ("settings", "sample_seed"):
            continue
        out[k] = np.concatenate([p[k] for p in parts])
    for p in parts[1:]:
        if str(p["settings"]) != str(parts[0]["settings"]):
            raise ValueError("datasets of different DEM settings")
    out["sample_seed"] = np.array([int(p["sample_seed"]) for p in parts])
    return out


def evaluate_cascade(train: dict, heldout: dict, target: float = 0.95, z: float = 2.0, val_frac: float = 0.5,
                     epochs: int = 20, lr: float = 1e-3, n_models: int = 4, seed: int = 1,
                     dem_all: bool = True, check: int = 512, batch_systems: int = 1024, log=None) -> dict:
    """The cascade and the one-gate surrogate on the held-out collisions.

    Both stages train on ``train`` with a validation share held back; the thresholds of
    both gates are set on it, aiming above ``target`` by z standard errors of the
    difference of two binomial estimates (validation and held-out), as :func:`evaluate`
    does.  With ``dem_all`` every held-out collision also runs DEM alone (the reference
    labels and the DEM-alone time); otherwise the stored DEM labels are the reference,
    after the first ``check`` collisions run DEM alone and must reproduce them exactly
    (the tile split is fixed, so a collision's result does not depend on its batch), and
    the DEM-alone time is scaled from those.

    Twenty epochs, not forty: the shorter training leaves the networks less sure of
    themselves where they are wrong, and both gates' validation cost fell (DEM per
    collision 0.0986 -> 0.0888 for the cascade, fallback share 10.2% -> 9.4% for one gate).

    DEM runs (alone, look-ahead and to the end) take up to ``batch_systems`` collisions per
    batched launch; DEM alone keeps the GPU full either way, while the cascade's fallbacks
    (a few hundred collisions) then run as one batch instead of a full one and a partial
    one.  At a fixed tile split a collision's result does not depend on its batch.

    The streaming gain counts per collision the surrogate's own time (descriptors and
    both stages), the look-ahead share at look/steps of a full-batch DEM collision and
    the run-to-the-end share at one: fallbacks of a long stream fill full batches."""
    s = settings_of(heldout)
    n = len(train["cls"])
    perm = np.random.default_rng(seed).permutation(n)
    nv = int(n * val_frac)
    va, tr = perm[:nv], perm[nv:]
    look_key = f"L{LOOK_AHEAD}"
    X, C, Rg, Lf = train["X"], train["cls"], train["reg"], train[look_key]
    nh = len(heldout["cls"])
    margin = z * math.sqrt(target * (1 - target) * (1.0 / max(nv, 1) + 1.0 / nh))
    look_cost = LOOK_AHEAD / s.steps
    cas = Cascade(Surrogate(seed=seed, n_models=n_models),
                  Surrogate(seed=seed + 100, n_models=n_models, n_in=N_DESCRIPTORS + N_LOOK))
    dem = None
    try:
        t0 = time.perf_counter()
        cas.s1.fit(X[tr], C[tr], Rg[tr], epochs=epochs, lr=lr, seed=seed)
        cas.s2.fit(np.hstack([X[tr], Lf[tr]]), C[tr], Rg[tr], epochs=epochs, lr=lr, seed=seed + 100)
        t_train = time.perf_counter() - t0
        gate1 = cas.s1.calibrate_gate(X[va], C[va], min(target + margin, 1.0))
        gate2 = cas.calibrate(X[va], Lf[va], C[va], min(target + margin, 1.0), look_cost)
        if log:
            log(f"trained two stages in {t_train:.1f} s; one gate {gate1}; cascade {gate2}")
        specs = specs_of(heldout)
        preps = [prepare(q, s) for q in specs]
        s_run = replace(s, max_systems=batch_systems)
        max_balls = s_run.max_systems * max(q.setup.group.num_particles for q in preps)
        dem = DemRunner(s_run, max_balls)
        dem.run(preps[:16], steps=256)          # first-use set-up (graphs)
        ref = heldout["cls"].astype(np.int32)
        if dem_all:
            t0 = time.perf_counter()
            c_dem, r_dem, _ = dem.run(preps)
            t_dem = time.perf_counter() - t0
            n_timed = nh
        else:
            n_timed = min(check, nh)
            t0 = time.perf_counter()
            c_dem, r_dem, _ = dem.run(preps[:n_timed])
            t_dem = time.perf_counter() - t0
        stored_match = float((c_dem == ref[:len(c_dem)]).mean())
        if dem_all:
            ref = c_dem
        c_c, r_c, m1, m2, Lh, tc = run_cascade(preps, cas, dem, s_run)
        Xh = descriptors(preps, s)
        p1, _, _, e1 = cas.s1.predict(Xh)
        look_diff = float(np.abs(Lh - heldout[look_key][m1]).max()) if m1.any() else 0.0
        # provisional class: the one-hot columns of the look-ahead row
        look_cls = float((Lh[:, :3].argmax(1) == heldout[look_key][m1][:, :3].argmax(1)).mean()) if m1.any() else 1.0
    finally:
        cas.close()
        if dem is not None:
            dem.close()
    c_d = t_dem / n_timed                                     # DEM seconds per collision, full batches
    c_s = (tc["descriptors"] + tc["stage1"] + tc["stage2"]) / nh
    fb1 = e1 > cas.s1.threshold
    one = np.where(fb1, ref, p1)
    f1, f2 = float(m1.mean()), float(m2.mean())
    keep = ~m2
    out = {
        "n_train": int(len(tr)), "n_val": int(nv), "n_heldout": nh, "epochs": epochs, "n_models": n_models,
        "train_s": t_train, "gate_target": target + margin, "look_ahead_steps": LOOK_AHEAD,
        "dem_steps": s.steps, "softening": s.softening, "batch_systems": batch_systems,
        "raw_accuracy": float((p1 == ref).mean()),
        "agreement": float((c_c == ref).mean()), "stage1_share": f1, "dem_share": f2,
        "gain_streaming": c_d / (c_s + f1 * look_cost * c_d + f2 * c_d),
        "surrogate_ms_per_collision": c_s * 1e3, "dem_ms_per_collision": c_d * 1e3,
        "cascade_s": tc["total"], "cascade_parts_s": tc,
        "one_gate": {"agreement": float((one == ref).mean()), "fallback_share": float(fb1.mean()),
                     "gain_streaming": c_d / (c_s + float(fb1.mean()) * c_d)},
        "val_one_gate": gate1, "val_cascade": gate2,
        "dem_alone_collisions": n_timed, "dem_labels_match_stored": stored_match,
        "look_features_max_abs_diff_vs_stored": look_diff, "look_class_match_stored": look_cls,
        "heldout_classes": np.bincount(ref, minlength=3).tolist(),
    }
    if dem_all:
        out["dem_s"] = t_dem
        out["gain"] = t_dem / tc["total"]
        out["collisions_per_s_dem"] = nh / t_dem
        out["collisions_per_s_cascade"] = nh / tc["total"]
        out["reg_mse"] = float(((r_c[keep] - r_dem[keep]) ** 2).mean()) if keep.any() else 0.0
        out["reg_mse_mean_predictor"] = float(((Rg[tr].mean(0) - r_dem[keep]) ** 2).mean()) if keep.any() else 0.0
    return out


N_T_RANGE, N_P_RANGE, SPEED_RANGE = (32, 512), (1, 64), (10.0, 1000.0)


LOOK_STEPS = (4000, 8000)


def make_dataset(count: int, seed: int, s: DemSettings, log=None, looks=LOOK_STEPS) -> dict:
    """DEM-labelled collisions: descriptors, classes, regression targets, look-ahead
    features of the same trajectories ("L<steps>") and the specs that regenerate each
    set-up."""
    specs = sample_specs(count, seed, N_T_RANGE, N_P_RANGE, SPEED_RANGE)
    dem = DemRunner(s, s.max_systems * (N_T_RANGE[1] + N_P_RANGE[1]))
    X, C, Rg, met = [], [], [], []
    Ls = {h: [] for h in looks}
    t0 = time.perf_counter()
    try:
        for c0 in range(0, count, s.max_systems):
            preps = [prepare(q, s) for q in specs[c0:c0 + s.max_systems]]
            X.append(descriptors(preps, s))
            c, r, _, lf = dem.run(preps, looks=looks)
            for h in looks:
                Ls[h].append(lf[h])
            C.append(c); Rg.append(r); met += [q.met for q in preps]
            if log:
                log(f"{c0 + len(preps)}/{count} collisions, {time.perf_counter() - t0:.1f} s, "
                    f"classes so far {np.bincount(np.concatenate(C), minlength=3).tolist()}")
    finally:
        dem.close()
    out = {f"L{h}": np.concatenate(v) for h, v in Ls.items()}
    return out | {"X": np.concatenate(X), "cls": np.concatenate(C), "reg": np.concatenate(Rg),
            "met": np.array(met), "n_t": np.array([q.n_t for q in specs]),
            "n_p": np.array([q.n_p for q in specs]), "seed": np.array([q.seed for q in specs], dtype=np.uint64),
            "speed": np.array([q.speed for q in specs]), "b_frac": np.array([q.b_frac for q in specs]),
            "settings": np.array(json.dumps(asdict(s))), "sample_seed": np.array(seed)}


def redescribe(ds: dict) -> dict:
    """The dataset with its descriptors recomputed from the stored specs (the set-ups
    are regenerated exactly; the DEM labels are kept)."""
    s = settings_of(ds)
    out = dict(ds)
    out["X"] = np.concatenate([descriptors([prepare(q, s) for q in chunk], s)
                               for chunk in (specs_of(ds)[k:k + 512] for k in range(0, len(ds["cls"]), 512))])
    return out


def specs_of(ds: dict) -> list[Spec]:
    return [Spec(int(a), int(b), int(c), float(d), float(e))
            for a, b, c, d, e in zip(ds["n_t"], ds["n_p"], ds["seed"], ds["speed"], ds["b_frac"])]


def settings_of(ds: dict) -> DemSettings:
    return DemSettings(**json.loads(str(ds["settings"])))


# ---- look-ahead stage: the provisional outcome after a short DEM run -------------
# 12 features of the outcome at the look-ahead step: the provisional class (one-hot),
# ln(components), the two largest bodies' mass fractions, the mass transfer, the x and y
# centre-of-mass velocities of the two largest bodies over the impact speed, and their
# relative speed over the impact speed.
N_LOOK = 12


def look_row(o, speed: float) -> np.ndarray:
    v1 = np.array(o.v_largest[:]) / speed
    v2 = np.array(o.v_second[:]) / speed
    onehot = [1.0 if o.label == k else 0.0 for k in range(3)]
    return np.array(onehot + [math.log(max(o.n_components, 1)), o.largest_mass_frac, o.second_mass_frac,
                              o.mass_transfer, v1[0], v1[1], v2[0], v2[1],
                              float(np.linalg.norm(v1 - v2)) if o.n_components > 1 else 0.0])



def load_dataset(path: str) -> dict:
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


HELDOUT_FILES = ("collisions_heldout.npz", "collisions_heldout_b.npz")   # seeds 2 and 3


def load_heldout(data_dir: str) -> dict:
    """The held-out collisions: the committed held-out files of ``data_dir`` as one set
    (2,000 from sample seed 2, then 8,000 from sample seed 3)."""
    import os
    return merge_datasets([load_dataset(os.path.join(data_dir, f)) for f in HELDOUT_FILES])


def add_look_features(ds: dict, looks: tuple, s: DemSettings | None = None, log=None) -> dict:
    """The dataset with look-ahead features ("L<steps>") at more step counts, from DEM
    runs of its stored set-ups up to the largest of them (the labels are not touched)."""
    s = settings_of(ds) if s is None else s
    looks = tuple(sorted(int(h) for h in looks))
    specs = specs_of(ds)
    dem = DemRunner(s, s.max_systems * (int(ds["n_t"].max()) + int(ds["n_p"].max())))
    feats = {h: [] for h in looks}
    t0 = time.perf_counter()
    try:
        for c0 in range(0, len(specs), s.max_systems):
            preps = [prepare(q, s) for q in specs[c0:c0 + s.max_systems]]
            # run to the last look; its outcome gives that look's features
            if len(looks) > 1:
                _, _, outs, lf = dem.run(preps, steps=looks[-1], looks=looks[:-1])
            else:
                _, _, outs = dem.run(preps, steps=looks[-1])
                lf = {}
            for h in looks[:-1]:
                feats[h].append(lf[h])
            feats[looks[-1]].append(np.array([look_row(o, q.setup.speed) for o, q in zip(outs, preps)],
                                             dtype=np.float32).reshape(-1, N_LOOK))
            if log:
                log(f"{c0 + len(preps)}/{len(specs)} collisions, {time.perf_counter() - t0:.1f} s")
    finally:
        dem.close()
    out = dict(ds)
    for h in looks:
        out[f"L{h}"] = np.concatenate(feats[h])
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="DEM-labelled collision datasets for the surrogate")
    ap.add_argument("--count", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--steps", type=int, default=DemSettings.steps)
    ap.add_argument("--softening", type=float, default=DemSettings.softening)
    ap.add_argument("--look-steps", default=",".join(str(h) for h in LOOK_STEPS))
    ap.add_argument("--out", required=True)
    ap.add_argument("--add-looks", default="",
                    help="instead of a new dataset: add look-ahead features at these step counts "
                         "(comma-separated) to the dataset --in, written to --out")
    ap.add_argument("--in", dest="inp", default="")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.add_looks:
        ds = add_look_features(load_dataset(a.inp), tuple(int(x) for x in a.add_looks.split(",")), log=log)
        np.savez_compressed(a.out, **ds)
        print(f"wrote {a.out}: {sorted(k for k in ds if k.startswith('L'))}", flush=True)
        return
    s = DemSettings(steps=a.steps, softening=a.softening)
    ds = make_dataset(a.count, a.seed, s, log=log, looks=tuple(int(x) for x in a.look_steps.split(",")))
    np.savez_compressed(a.out, **ds)
    print(f"wrote {a.out}: classes {np.bincount(ds['cls'], minlength=3).tolist()}", flush=True)


if __name__ == "__main__":
    main()

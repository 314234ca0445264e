"""The surrogate's gates on synthetic predictions (host logic only, no network): the
one-gate threshold and the cascade's two thresholds must reach the agreement target on
the calibration set, send as little to DEM as any thresholds that do, and never count a
prediction as right that is not."""
import itertools

import numpy as np
import pytest

from spacelab_amd import surrogate as S


class FakeModel:
    """predict() as Surrogate has it: (class, regression, probabilities, entropy), from
    fixed rows keyed by the first input column."""

    def __init__(self, pred, ent):
        self.pred, self.ent = np.asarray(pred), np.asarray(ent, dtype=float)
        self.threshold = float("inf")

    def predict(self, X):
        idx = X[:, 0].astype(int)
        return self.pred[idx], None, None, self.ent[idx]

    calibrate_gate = S.Surrogate.calibrate_gate

    def close(self):
        pass


def _case(n, seed):
    rng = np.random.default_rng(seed)
    cls = rng.integers(0, 3, n)
    # stage 1 right on about 90%, wrong ones more uncertain on average
    ok1 = rng.random(n) < 0.9
    p1 = np.where(ok1, cls, (cls + 1 + rng.integers(0, 2, n)) % 3)
    e1 = np.where(ok1, rng.random(n) * 0.8, 0.3 + rng.random(n) * 0.8)
    ok2 = ok1 | (rng.random(n) < 0.4)
    p2 = np.where(ok2, cls, (cls + 1) % 3)
    e2 = np.where(ok2, rng.random(n) * 0.6, 0.2 + rng.random(n) * 0.8)
    X = np.arange(n, dtype=np.float32)[:, None]
    L = np.zeros((n, S.N_LOOK), dtype=np.float32)
    return X, L, cls, p1, e1, p2, e2


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_one_gate_reaches_target_with_least_fallback(seed):
    X, _, cls, p1, e1, _, _ = _case(400, seed)
    m = FakeModel(p1, e1)
    g = m.calibrate_gate(X, cls, 0.96)
    fb = e1 > m.threshold
    final = np.where(fb, cls, p1)
    assert (final == cls).mean() >= 0.96
    assert g["fallback_share"] == pytest.approx(fb.mean())
    # no smaller fallback set of the most uncertain collisions reaches the target
    order = np.argsort(-e1, kind="stable")
    for k in range(int(fb.sum())):
        send = np.zeros(len(cls), bool)
        send[order[:k]] = True
        assert (np.where(send, cls, p1) == cls).mean() < 0.96


@pytest.mark.parametrize("seed", [4, 5])
def test_cascade_reaches_target_at_least_cost(seed):
    n = 300
    X, L, cls, p1, e1, p2, e2 = _case(n, seed)
    cas = S.Cascade(FakeModel(p1, e1), FakeModel(p2, e2))
    look = 0.1
    g = cas.calibrate(X, L, cls, 0.97, look)
    F1 = e1 > cas.t1
    F2 = F1 & (e2 > cas.t2)
    final = np.where(F2, cls, np.where(F1, p2, p1))
    agree = (final == cls).mean()
    assert agree >= 0.97 and g["agreement"] == pytest.approx(agree)
    cost = look * F1.mean() + F2.mean()
    assert g["dem_cost"] == pytest.approx(cost)
    # brute force over the same stage-1 thresholds and every stage-2 cut: nothing cheaper
    cands = np.unique(np.concatenate([np.quantile(e1, np.linspace(0, 1, 201)), [-1.0]]))
    best = np.inf
    for t1 in cands:
        f1 = e1 > t1
        for t2 in np.concatenate([np.unique(e2[f1]) - 1e-12, [np.inf]]):
            f2 = f1 & (e2 > t2)
            if (np.where(f2, cls, np.where(f1, p2, p1)) == cls).mean() >= 0.97:
                best = min(best, look * f1.mean() + f2.mean())
    assert cost <= best + 1e-12


def test_cascade_all_to_dem_when_nothing_else_reaches():
    n = 50
    X, L, cls, p1, e1, p2, e2 = _case(n, 9)
    cas = S.Cascade(FakeModel((cls + 1) % 3, e1), FakeModel((cls + 1) % 3, e2))
    g = cas.calibrate(X, L, cls, 1.0, 0.1)
    assert g["agreement"] == 1.0
    assert g["dem_share"] == 1.0

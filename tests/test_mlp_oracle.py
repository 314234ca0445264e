"""The CPU oracle of the surrogate network: analytic gradients against central finite
differences (binary64, no bfloat16 rounding), and the bfloat16 rounding helper
against its definition."""
import numpy as np

from oracle import mlp as M


def _net(rng, dims):
    layers = []
    for i, j in zip(dims[:-1], dims[1:]):
        layers.append((rng.standard_normal((i, j)) * np.sqrt(2.0 / i), rng.standard_normal(j) * 0.1))
    return layers


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    dims = [5, 7, 6, 8, 6]
    layers = _net(rng, dims)
    B = 9
    x = rng.standard_normal((B, 5))
    cls = rng.integers(0, 3, B)
    reg = rng.standard_normal((B, 3))
    loss, grads = M.loss_and_grads(layers, x, cls, reg)
    flat = M.flatten(layers)
    gflat = M.flatten(grads)
    h = 1e-6
    for idx in rng.choice(len(flat), 60, replace=False):
        fp = flat.copy(); fp[idx] += h
        fm = flat.copy(); fm[idx] -= h
        lp, _ = M.loss_and_grads(M.unflatten(fp, dims), x, cls, reg)
        lm, _ = M.loss_and_grads(M.unflatten(fm, dims), x, cls, reg)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gflat[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, gflat[idx])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -3.5, 0.0], dtype=np.float32)
    r = M.bf16(x)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    assert np.array_equal(r, [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -3.5, 0.0])

"""The surrogate perceptron on the GPU against the CPU oracle (oracle/mlp.py).

The oracle rounds the operands of every matrix product to bfloat16 at the points the
device does, so device and oracle differ only by binary32 versus binary64
accumulation and, rarely, by a hidden activation that lands on the other side of a
bfloat16 rounding boundary; the tolerances below are set for that.  A wrong
matrix-core lane layout or a transposed operand gives errors of order one."""
import numpy as np
import pytest

import spacelab_amd as sl
from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _dims(n_in, width, n_out):
    return [n_in, width, width, width, n_out]


def _batch(rng, B, n_in, n_out):
    x = rng.standard_normal((B, n_in)).astype(np.float32)
    cls = rng.integers(0, 3, B).astype(np.int32)
    reg = rng.standard_normal((B, n_out - 3)).astype(np.float32)
    return x, cls, reg


@pytest.mark.parametrize("n_in,width,n_out,B", [(12, 256, 9, 300), (5, 32, 4, 1), (20, 48, 6, 130), (3, 16, 5, 64)])
def test_forward_matches_oracle(n_in, width, n_out, B):
    rng = np.random.default_rng(n_in * 1000 + B)
    with sl.MLP(n_in, n_out, width, max_batch=512, seed=7) as m:
        w = m.get_params()
        assert w.shape == (m.n_params,)
        # non-zero biases so the bias path is exercised
        w = w + np.float32(0.05) * rng.standard_normal(w.shape).astype(np.float32)
        m.set_params(w)
        assert np.array_equal(m.get_params(), w)
        x = rng.standard_normal((B, n_in)).astype(np.float32)
        y, p, e = m.forward(x)
    layers = M.unflatten(w.astype(np.float64), _dims(n_in, width, n_out))
    yo = M.forward(layers, x, emulate=True)[0][-1]
    scale = np.abs(yo).max()
    err = np.abs(y - yo)
    assert err.max() <= 2e-3 * scale, err.max() / scale
    assert np.median(err) <= 1e-5 * scale
    # against the unrounded binary64 network: the bfloat16 effect stays small
    yt = M.forward(layers, x, emulate=False)[0][-1]
    assert np.abs(y - yt).max() <= 5e-2 * scale
    # probabilities and entropy of the device's own logits
    z = y[:, :3].astype(np.float64)
    q = np.exp(z - z.max(1, keepdims=True))
    q /= q.sum(1, keepdims=True)
    assert np.abs(p - q).max() <= 1e-6
    assert np.abs(e - (-(q * np.log(q)).sum(1))).max() <= 1e-5


def test_gradients_match_oracle():
    rng = np.random.default_rng(5)
    n_in, width, n_out, B = 12, 256, 9, 200
    with sl.MLP(n_in, n_out, width, max_batch=256, seed=3) as m:
        w = m.get_params() + np.float32(0.02) * rng.standard_normal(m.n_params).astype(np.float32)
        m.set_params(w)
        x, cls, reg = _batch(rng, B, n_in, n_out)
        loss, g = m.gradients(x, cls, reg)
        loss2, g2 = m.gradients(x, cls, reg)
    assert loss == loss2 and np.array_equal(g, g2)          # fixed-order reductions
    dims = _dims(n_in, width, n_out)
    layers = M.unflatten(w.astype(np.float64), dims)
    lo, go = M.loss_and_grads(layers, x, cls, reg, emulate=True)
    assert loss == pytest.approx(lo, rel=1e-5)
    gl = M.unflatten(g.astype(np.float64), dims)
    for k, ((gW, gb), (oW, ob)) in enumerate(zip(gl, go)):
        for a, o, what in ((gW, oW, "W"), (gb, ob, "b")):
            rel = np.linalg.norm(a - o) / np.linalg.norm(o)
            assert rel <= 2e-3, (k, what, rel)
    # and against the unrounded network: the bfloat16 effect on the gradient is bounded
    _, gt = M.loss_and_grads(layers, x, cls, reg, emulate=False)
    for (gW, gb), (oW, ob) in zip(gl, gt):
        assert np.linalg.norm(gW - oW) <= 5e-2 * np.linalg.norm(oW)


def test_adam_first_step():
    """After one step from zero moments, Adam moves each weight by lr * g / (|g| + eps)."""
    rng = np.random.default_rng(9)
    with sl.MLP(6, 5, 32, max_batch=64, seed=11) as m:
        w0 = m.get_params()
        x, cls, reg = _batch(rng, 40, 6, 5)
        _, g = m.gradients(x, cls, reg)
        lr = 1e-3
        m.train_step(x, cls, reg, lr)
        w1 = m.get_params()
    g = g.astype(np.float64)
    expect = w0 - lr * g / (np.abs(g) + 1e-8)
    assert np.abs(w1 - expect).max() <= 2e-5 * lr + 2e-7 * np.abs(w0).max()


def _task(rng, B, n_in, A, Bm):
    x = rng.standard_normal((B, n_in)).astype(np.float32)
    cls = (x @ A).argmax(1).astype(np.int32)
    reg = np.tanh(x @ Bm).astype(np.float32)
    return x, cls, reg


def test_training_learns_and_is_reproducible():
    rng = np.random.default_rng(21)
    n_in, n_out = 12, 5
    A = rng.standard_normal((n_in, 3)); Bm = rng.standard_normal((n_in, 2)) * 0.5
    xt, ct, rt = _task(rng, 4096, n_in, A, Bm)
    xv, cv, rv = _task(rng, 2000, n_in, A, Bm)
    finals = []
    for rep in range(2):
        with sl.MLP(n_in, n_out, 256, max_batch=2048, seed=4) as m:
            y0, _, _ = m.forward(xv)
            mse0 = ((y0[:, 3:] - rv) ** 2).mean()
            losses = []
            for ep in range(8):
                for k in range(0, 4096, 256):
                    losses.append(m.train_step(xt[k:k + 256], ct[k:k + 256], rt[k:k + 256], 2e-3))
            y, p, e = m.forward(xv)
            finals.append(m.get_params())
        acc = (y[:, :3].argmax(1) == cv).mean()
        mse = ((y[:, 3:] - rv) ** 2).mean()
        assert acc >= 0.93, acc
        assert mse <= 0.2 * mse0, (mse, mse0)
        assert np.mean(losses[-16:]) < 0.3 * np.mean(losses[:4])
        # confident where right: wrong predictions carry more entropy on average
        right = y[:, :3].argmax(1) == cv
        assert e[~right].mean() > e[right].mean()
    assert np.array_equal(finals[0], finals[1])


def test_argument_errors():
    with sl.MLP(4, 5, 16, max_batch=8, seed=1) as m:
        x = np.zeros((9, 4), np.float32)
        with pytest.raises(sl.SpacelabError):
            m.forward(x)
        with pytest.raises(sl.SpacelabError):
            m.train_step(x[:2], np.array([0, 3], np.int32), np.zeros((2, 2), np.float32), 1e-3)
        with pytest.raises(sl.SpacelabError):
            m.train_step(x[:2], np.array([0, 1], np.int32), np.zeros((2, 2), np.float32), 0.0)
        # still usable after rejected calls
        y, p, e = m.forward(x[:3])
        assert np.all(np.isfinite(y))

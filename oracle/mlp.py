"""CPU oracle of the surrogate perceptron (test infrastructure only: the product path
is the HIP library; only tests may import this).

A binary64 restatement of the network in spacelab_amd/csrc/sl_mlp.hip: layers
n_in -> width -> width -> width -> n_out, ReLU between layers, outputs 0..2 class
logits, the rest regression targets; loss = softmax cross-entropy + mean squared
error of the regression outputs, averaged over the batch.  With emulate=True the
operands of every matrix product are rounded to bfloat16 (round to nearest even) at
exactly the points the device rounds them (inputs, weights, hidden activations,
output gradients, hidden-layer gradients), so that device and oracle differ only by
binary32 versus binary64 accumulation.  The analytic gradients are pinned by central
finite differences in the tests.  There is no reference counterpart (the surrogate
is an addition of this implementation).
"""
from __future__ import annotations

import numpy as np

NCLS = 3


def bf16(x: np.ndarray) -> np.ndarray:
    """Round binary32 values to bfloat16 (nearest even) and return them as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unflatten(params: np.ndarray, dims: list[int]):
    """Flat parameters (per layer W[in][out] then b[out]) -> [(W, b), ...]."""
    out, o = [], 0
    for k in range(len(dims) - 1):
        i, j = dims[k], dims[k + 1]
        W = np.asarray(params[o:o + i * j], dtype=np.float64).reshape(i, j); o += i * j
        b = np.asarray(params[o:o + j], dtype=np.float64); o += j
        out.append((W, b))
    assert o == len(params)
    return out


def flatten(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def forward(layers, x, emulate=False):
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    h = r(x)
    acts = [h]
    zs = []
    for k, (W, b) in enumerate(layers):
        z = h @ r(W) + b
        zs.append(z)
        if k < len(layers) - 1:
            a = np.maximum(z, 0.0)
            h = r(a)
            acts.append(h)
    return zs, acts


def loss_and_grads(layers, x, cls, reg, emulate=False):
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    zs, acts = forward(layers, x, emulate)
    y = zs[-1]
    B = y.shape[0]
    logits = y[:, :NCLS]
    mx = logits.max(1, keepdims=True)
    e = np.exp(logits - mx)
    p = e / e.sum(1, keepdims=True)
    nr = y.shape[1] - NCLS
    d = y[:, NCLS:] - reg
    loss = (np.log(e.sum(1)) + mx[:, 0] - logits[np.arange(B), cls] + (d * d).sum(1) / nr).mean()
    g = np.zeros_like(y)
    g[:, :NCLS] = p
    g[np.arange(B), cls] -= 1.0
    g[:, NCLS:] = 2.0 * d / nr
    g /= B
    grads = [None] * len(layers)
    dz = g
    for k in range(len(layers) - 1, -1, -1):
        W, b = layers[k]
        dzr = r(dz)
        grads[k] = (acts[k].T @ dzr, dz.sum(0))
        if k > 0:
            dh = dzr @ r(W).T
            dz = dh * (zs[k - 1] > 0)
    return loss, grads

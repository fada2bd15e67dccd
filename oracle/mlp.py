"""CPU restatement (float64 numpy) of the learned sizer's network, loss and
Adam step -- test infrastructure only: the tests compare the device kernels
of polars-metal_amd/csrc/pmh_mlp.hip with it, and the product never imports
it.

The network has no counterpart in the reference (which sizes nothing from
data); this file restates the kernels' own mathematics:

    h1 = relu(x W1 + b1)        W1: [F, 64]
    h2 = relu(h1 W2 + b2)       W2: [64, 64]
    o  = h2 W3 + b3             W3: [64, 2]
    loss = mean_s[(o0 - y0)^2 + lambda * [y1 >= 0] * (softplus(o1) - y1 o1)]

Parity pins (tests/test_mlp.py): `grad` against central finite differences
of `loss` on random networks (float64, so a dropped term, a transposed
operand or a wrong sign fails by orders of magnitude); `adam_step` against
the closed form of its first step (a step of exactly lr * sign(g))."""
from __future__ import annotations

import numpy as np

F, H = 16, 64
SIZES = [("W1", (F, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, 2)), ("b3", (2,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SIZES:
        if name.startswith("W"):
            d[name] = rng.standard_normal(shape) * np.sqrt(2.0 / shape[0])
        else:
            d[name] = np.zeros(shape)
    return pack(d)


def forward(p: np.ndarray, x: np.ndarray):
    w = unpack(np.asarray(p, np.float64))
    x = np.asarray(x, np.float64)
    z1 = x @ w["W1"] + w["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w["W2"] + w["b2"]
    h2 = np.maximum(z2, 0)
    o = h2 @ w["W3"] + w["b3"]
    return o, (x, z1, h1, z2, h2)


def _softplus(z):
    return np.where(z > 0, z + np.log1p(np.exp(-np.abs(z))), np.log1p(np.exp(np.minimum(z, 0))))


def loss(p: np.ndarray, x: np.ndarray, y: np.ndarray, lam: float) -> float:
    o, _ = forward(p, x)
    y = np.asarray(y, np.float64)
    m = (y[:, 1] >= 0).astype(np.float64)
    per = (o[:, 0] - y[:, 0]) ** 2 + lam * m * (_softplus(o[:, 1]) - y[:, 1] * o[:, 1])
    return float(per.mean())


def grad(p: np.ndarray, x: np.ndarray, y: np.ndarray, lam: float) -> np.ndarray:
    """Gradient of `loss` by back-propagation."""
    w = unpack(np.asarray(p, np.float64))
    o, (x, z1, h1, z2, h2) = forward(p, x)
    y = np.asarray(y, np.float64)
    n = x.shape[0]
    m = (y[:, 1] >= 0).astype(np.float64)
    sig = 1.0 / (1.0 + np.exp(-o[:, 1]))
    go = np.stack([2.0 * (o[:, 0] - y[:, 0]), lam * m * (sig - y[:, 1])], axis=1) / n
    g = {"W3": h2.T @ go, "b3": go.sum(0)}
    dz2 = (go @ w["W3"].T) * (z2 > 0)
    g["W2"] = h1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ w["W2"].T) * (z1 > 0)
    g["W1"] = x.T @ dz1
    g["b1"] = dz1.sum(0)
    return pack(g)


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba, 2015), returning (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v

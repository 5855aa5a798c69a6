"""FP64 restatement of the learned wall closure's MLP -- TEST INFRASTRUCTURE ONLY.

The closure network has no counterpart in the reference (pocketwind's wall treatment is the
Werner-Wengle wall law of src/lbm/shaders/stream_collide.wgsl:1-80); this file states the
network the library trains and evaluates, so that the HIP kernels can be checked against it:

    h1 = silu(W1 x + b1)      x: 12 features, h1: 64
    h2 = silu(W2 h1 + b2)     h2: 64
    y  = W3 h2 + b3           y: 3
    loss = mean over samples and outputs of (y - t)^2

Parameters travel as one flat float32 vector in the order W1 (64 x 12, row-major), b1 (64),
W2 (64 x 64), b2 (64), W3 (3 x 64), b3 (3).

``bf16=True`` rounds to bfloat16 (round to nearest even) exactly where the device kernels feed
their matrix cores: the inputs, the weights, the hidden activations, and the backward
deltas (dY, dZ2, dZ1) that enter a product. Sums are exact here (float64) where the device
accumulates in float32; biases, bias gradients and pre-activations stay unrounded.

Only tests/ may import this module; it is the checker, not the product.
"""
from __future__ import annotations

import numpy as np

IN, HIDDEN, OUT = 12, 64, 3
SHAPES = [("W1", (HIDDEN, IN)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
          ("W3", (OUT, HIDDEN)), ("b3", (OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    flat = np.asarray(flat, np.float64)
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = flat[k:k + n].reshape(shape)
        k += n
    return out


def pack(params) -> np.ndarray:
    return np.concatenate([np.asarray(params[name], np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """float -> bfloat16 -> float with round to nearest even (finite inputs)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def silu_prime(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def forward(flat, x, bf16_points: bool = False):
    p = unpack(flat)
    r = bf16 if bf16_points else (lambda a: np.asarray(a, np.float64))
    x = r(np.asarray(x, np.float64))
    w1, w2, w3 = r(p["W1"]), r(p["W2"]), r(p["W3"])
    z1 = x @ w1.T + p["b1"]
    h1 = r(silu(z1))
    z2 = h1 @ w2.T + p["b2"]
    h2 = r(silu(z2))
    y = h2 @ w3.T + p["b3"]
    return y, dict(x=x, w1=w1, w2=w2, w3=w3, z1=z1, h1=h1, z2=z2, h2=h2)


def loss_and_gradients(flat, x, t, bf16_points: bool = False):
    """Mean squared error over samples and outputs, and its gradient as a flat vector."""
    r = bf16 if bf16_points else (lambda a: np.asarray(a, np.float64))
    y, c = forward(flat, x, bf16_points)
    t = np.asarray(t, np.float64)
    b = x.shape[0]
    loss = float(np.mean((y - t) ** 2))
    dy = 2.0 * (y - t) / (b * OUT)
    dyr = r(dy)
    g = {"W3": dyr.T @ c["h2"], "b3": dy.sum(0)}
    dh2 = dyr @ c["w3"]
    dz2 = dh2 * silu_prime(c["z2"])
    dz2r = r(dz2)
    g["W2"] = dz2r.T @ c["h1"]
    g["b2"] = dz2.sum(0)
    dh1 = dz2r @ c["w2"]
    dz1 = dh1 * silu_prime(c["z1"])
    dz1r = r(dz1)
    g["W1"] = dz1r.T @ c["x"]
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def adam(flat, grad, m, v, step: int, lr: float, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba 2015, Algorithm 1); returns (params, m, v)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return flat - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases (the library's own initialisation is equivalent)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shape[1])
            p[name] = rng.uniform(-lim, lim, shape)
        else:
            p[name] = np.zeros(shape)
    return pack(p).astype(np.float32).astype(np.float64)

"""FP64 numpy model of the closure network (5 -> 64 -> 64 -> 1, tanh), the
checker of the GPU training kernels.  Test infrastructure only.

The model keeps the rounding points of the GPU kernels when ``bf16=True``:
the operands of the 64 x 64 products are rounded to bfloat16 (hidden layer 1
and W2 in the forward pass; dH2, H1 and W2 in the backward pass), everything
else is exact FP64.  Loss: mean squared error.  Adam: Kingma & Ba (2015).
"""
from __future__ import annotations

import numpy as np

NPARAM = 4609
OFF = {"W1": (0, (64, 5)), "b1": (320, (64,)), "W2": (384, (64, 64)), "b2": (4480, (64,)), "W3": (4544, (64,)),
       "b3": (4608, ())}


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out = {}
    for k, (o, shp) in OFF.items():
        n = int(np.prod(shp)) if shp else 1
        out[k] = p[o:o + n].reshape(shp) if shp else p[o]
    return out


def pack(d):
    p = np.zeros(NPARAM)
    for k, (o, shp) in OFF.items():
        n = int(np.prod(shp)) if shp else 1
        p[o:o + n] = np.asarray(d[k], dtype=np.float64).reshape(-1)
    return p


def bf16(x):
    """Round to the nearest bfloat16 (ties to even) through float32."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, x, bf=True):
    """y[n] and the activations."""
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    h1 = np.tanh(x @ w["W1"].T + w["b1"])
    h1r = r(h1)
    w2r = r(w["W2"])
    h2 = np.tanh(h1r @ w2r.T + w["b2"])
    y = h2 @ w["W3"] + w["b3"]
    return y, dict(h1=h1, h1r=h1r, w2r=w2r, h2=h2)


def loss_grad(p, x, t, bf=True):
    """Mean squared error and its gradient (backprop of the same rounding
    points the kernels use)."""
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n = x.shape[0]
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    y, a = forward(p, x, bf)
    e = y - t
    loss = np.mean(e * e)
    dy = 2.0 * e / n
    dh2 = (dy[:, None] * w["W3"][None, :]) * (1.0 - a["h2"] ** 2)
    dh2r = r(dh2)
    g = {"b3": dy.sum(), "W3": dy @ a["h2"], "b2": dh2.sum(axis=0), "W2": dh2r.T @ a["h1r"]}
    dh1 = (dh2r @ a["w2r"]) * (1.0 - a["h1"] ** 2)
    g["b1"] = dh1.sum(axis=0)
    g["W1"] = dh1.T @ x
    return loss, pack(g)


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v


def glorot(seed=0):
    rng = np.random.default_rng(seed)
    d = {"W1": rng.uniform(-1, 1, (64, 5)) * np.sqrt(6 / 69), "b1": np.zeros(64),
         "W2": rng.uniform(-1, 1, (64, 64)) * np.sqrt(6 / 128), "b2": np.zeros(64),
         "W3": rng.uniform(-1, 1, 64) * np.sqrt(6 / 65), "b3": 0.0}
    return pack(d)

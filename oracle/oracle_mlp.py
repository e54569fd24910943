"""CPU restatement of the densification MLP (nprush_amd/csrc/k_mlp.hip,
nml_api.cpp).  TEST INFRASTRUCTURE ONLY.

32 -> 256 -> 256 -> 256 -> 14, GELU (tanh form) on the hidden layers, MSE of
standardised outputs.  ``bf16=True`` rounds to bfloat16 (nearest even) exactly
where the GPU path does: the standardised inputs, the weight copies, each hidden
activation, the output gradient and each hidden pre-activation gradient; products
and sums are then carried in float64 (the GPU accumulates in float32, so results
agree to accumulation rounding).  ``bf16=False`` is the plain float64 network
whose gradients the finite-difference pins check.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 32, 256, 14
DIMS = [(IN, HID), (HID, HID), (HID, HID), (HID, OUT)]
NUM_PARAMS = sum(i * o + o for i, o in DIMS)
C_GELU = float(np.float32(0.7978845608028654))
A_GELU = float(np.float32(0.044715))


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def split(params):
    p = np.asarray(params, np.float64)
    out, off = [], 0
    for i, o in DIMS:
        W = p[off:off + i * o].reshape(o, i)
        off += i * o
        b = p[off:off + o]
        off += o
        out.append((W, b))
    return out


def gelu(z):
    t = np.tanh(C_GELU * (z + A_GELU * z ** 3))
    return 0.5 * z * (1.0 + t)


def gelu_d(z):
    t = np.tanh(C_GELU * (z + A_GELU * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * A_GELU * z * z)


def init_params(seed=0, out_scale=0.0):
    """He-style initial weights (N(0, 2/fan_in)), zero biases; the output layer
    scaled by out_scale (0: the network starts by proposing no change)."""
    rng = np.random.default_rng(seed)
    parts = []
    for k, (i, o) in enumerate(DIMS):
        W = rng.normal(0.0, np.sqrt(2.0 / i), (o, i))
        if k == len(DIMS) - 1:
            W *= out_scale
        parts += [W.reshape(-1), np.zeros(o)]
    return np.concatenate(parts).astype(np.float32)


def forward(params, x, norm=None, bf=True):
    """Standardised outputs (n x 14) and the cache for ``grads``."""
    r = (lambda a: bf16(np.asarray(a, np.float32))) if bf else (lambda a: np.asarray(a, np.float64))
    layers = split(params)
    x = np.asarray(x, np.float64)
    if norm is not None:
        x = (x.astype(np.float32) - np.asarray(norm[0], np.float32)) / np.asarray(norm[1], np.float32)
    h = r(x)
    cache = [h]
    zs = []
    for k, (W, b) in enumerate(layers):
        Wq = r(W)
        z = h @ Wq.T + b
        if k < len(layers) - 1:
            zs.append(z)
            h = r(gelu(z))
            cache.append(h)
        else:
            y = z
    return y, (cache, zs)


def grads(params, x, t, norm=None, bf=True):
    """(loss, flat gradient) of mean((y - t_std)^2) over n x 14 outputs."""
    r = (lambda a: bf16(np.asarray(a, np.float32))) if bf else (lambda a: np.asarray(a, np.float64))
    y, (hs, zs) = forward(params, x, norm, bf)
    t = np.asarray(t, np.float64)
    if norm is not None:
        t = (t.astype(np.float32) - np.asarray(norm[2], np.float32)) / np.asarray(norm[3], np.float32)
    n = y.shape[0]
    loss = float(np.mean((y - t) ** 2))
    d = r(2.0 * (y - t) / (n * OUT))
    layers = split(params)
    g = [None] * len(layers)
    for k in range(len(layers) - 1, -1, -1):
        W, _ = layers[k]
        g[k] = (d.T @ hs[k], d.sum(0))
        if k > 0:
            da = d @ r(W)
            d = r(da * gelu_d(zs[k - 1]))
    flat = np.concatenate([np.concatenate([gw.reshape(-1), gb]) for gw, gb in g])
    return loss, flat

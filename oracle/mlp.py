"""The warm-start MLP in numpy float64 -- TEST INFRASTRUCTURE ONLY.

The network has no counterpart in the reference: it is the build's own
learned warm start for the bounce (a per-patch regression of the converged
bounced light).  This module is the CPU oracle its HIP kernels are checked
against:

* ``features``: the 25 inputs of a patch (position scaled to the map's box,
  the normal, the log of the direct light, 8 Fourier pairs of the position);
* ``forward`` / ``loss_and_grad``: 25 -> 128 -> 128 -> 128 -> 3, GELU (tanh
  form) on the hidden layers, mean squared error over the three outputs;
  ``emulate=True`` rounds to bfloat16 exactly where the GPU path stores
  bfloat16 (inputs, weights, hidden activations, output and hidden-layer
  gradients), so the GPU's gradients can be held to a tight tolerance;
* ``adam``: one Adam step (Kingma and Ba, 2015, Algorithm 1).

The analytic backward is pinned by float64 central differences in the tests.
"""
from __future__ import annotations

import numpy as np

N_IN = 25
HIDDEN = 128
N_OUT = 3
N_FOURIER = 8
LAYERS = [(N_IN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, N_OUT)]
N_PARAMS = sum(i * o + o for i, o in LAYERS)  # 36,739
GELU_C = np.sqrt(2.0 / np.pi)


def unpack(flat):
    """The flat parameter vector (per layer: W row-major [in][out], then b)
    as a list of (W, b)."""
    flat = np.asarray(flat, np.float64)
    assert flat.shape == (N_PARAMS,)
    out, o = [], 0
    for i, n in LAYERS:
        W = flat[o:o + i * n].reshape(i, n)
        o += i * n
        b = flat[o:o + n]
        o += n
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64.
    Values are first rounded to float32, as the GPU holds them."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    nan = np.isnan(f)
    r = np.where(nan, (u.astype(np.uint32) & 0xFFFF0000) | 0x00400000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * 0.044715 * z * z)


def fourier_matrix(seed: int = 7) -> np.ndarray:
    """The 8 x 3 frequency matrix of the Fourier features (standard normal,
    seeded), float32."""
    return np.random.default_rng(seed).standard_normal((N_FOURIER, 3)).astype(np.float32)


def features(origin, normal, direct, lo, hi, freq, length, floor=1e-7):
    """(P, 25) float64: 2(p - lo)/(hi - lo) - 1, n, log(max(direct, 0) + floor),
    then sin and cos of 2 pi (F p / length) for each of the 8 rows of F."""
    p = np.asarray(origin, np.float64)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    ext = np.where(hi > lo, hi - lo, 1.0)
    x = np.empty((p.shape[0], N_IN))
    x[:, 0:3] = 2.0 * (p - lo) / ext - 1.0
    x[:, 3:6] = normal
    x[:, 6:9] = np.log(np.maximum(np.asarray(direct, np.float64), 0.0) + floor)
    ang = 2.0 * np.pi * (p @ np.asarray(freq, np.float64).T) / length
    x[:, 9:9 + N_FOURIER] = np.sin(ang)
    x[:, 9 + N_FOURIER:9 + 2 * N_FOURIER] = np.cos(ang)
    return x


def forward(flat, x, emulate=False):
    """Returns (y, cache)."""
    L = unpack(flat)
    r = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    h = r(x)
    cache = {"h": [h], "z": []}
    for li, (W, b) in enumerate(L):
        z = h @ r(W) + b
        if li < len(L) - 1:
            cache["z"].append(z)
            h = r(gelu(z))
            cache["h"].append(h)
        else:
            y = z
    return y, cache


def loss_and_grad(flat, x, t, emulate=False):
    """Mean squared error over N rows and 3 outputs, and its gradient in the
    flat layout."""
    L = unpack(flat)
    r = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    y, c = forward(flat, x, emulate)
    n = y.shape[0]
    d = y - np.asarray(t, np.float64)
    loss = float((d * d).sum() / (3 * n))
    dz = r(2.0 * d / (3 * n))
    grads = [None] * len(L)
    for li in range(len(L) - 1, -1, -1):
        W, _b = L[li]
        h = c["h"][li]
        grads[li] = (h.T @ dz, dz.sum(axis=0))
        if li > 0:
            dh = dz @ r(W).T
            dz = r(dh * gelu_grad(c["z"][li - 1]))
    return loss, pack(grads)


def adam(w, m, v, g, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step at step number t (1-based); returns (w, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v

"""Float64 CPU oracle of the openness surrogate's network -- TEST INFRASTRUCTURE.

The network has no reference counterpart (the reference has no learned
component).  This module restates its definition in NumPy float64 so that
the HIP kernels' forward pass, gradients and Adam step can be checked:

* conv3x3: "same" 3x3 correlation with zero padding on NHWC arrays;
  tap (dy, dx) reads x[y + dy - 1, x + dx - 1] with weight w[co, ci, dy, dx];
* GELU(x) = x * Phi(x) (erf form);
* layers 0..4: conv + bias + GELU; layer 5: conv + bias (linear);
* loss: sum over labelled cells of (out[..., 0] + base - label)^2, times scale.

``backward`` is the hand-derived gradient; tests/test_oracle_pins.py pins it
against central finite differences in float64.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np
from scipy.special import erf


def conv3x3(x: np.ndarray, w: np.ndarray, b: np.ndarray) -> np.ndarray:
    bsz, h, wd, cin = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((bsz, h, wd, w.shape[0]))
    for dy in range(3):
        for dx in range(3):
            out += np.einsum("bhwc,oc->bhwo", xp[:, dy:dy + h, dx:dx + wd, :], w[:, :, dy, dx])
    return out + b


def conv3x3_backward(x: np.ndarray, w: np.ndarray, dz: np.ndarray):
    """Gradients of conv3x3 (x, w, b) given dL/dz: (dx, dw, db)."""
    bsz, h, wd, cin = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for dy in range(3):
        for dx in range(3):
            dw[:, :, dy, dx] = np.einsum("bhwo,bhwc->oc", dz, xp[:, dy:dy + h, dx:dx + wd, :])
            dxp[:, dy:dy + h, dx:dx + wd, :] += np.einsum("bhwo,oc->bhwc", dz, w[:, :, dy, dx])
    return dxp[:, 1:-1, 1:-1, :], dw, dz.sum(axis=(0, 1, 2))


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def forward(params: Sequence[Tuple[np.ndarray, np.ndarray]], x: np.ndarray):
    """Returns (output [..., cout_last], cache of (input, pre-activation) per layer)."""
    cache = []
    a = x
    for li, (w, b) in enumerate(params):
        z = conv3x3(a, w, b)
        cache.append((a, z))
        a = gelu(z) if li < len(params) - 1 else z
    return a, cache


def loss(out, base, label, scale=1.0) -> float:
    m = ~np.isnan(label)
    d = out[..., 0] + base - label
    return float(scale * np.sum(np.where(m, d * d, 0.0)))


def backward(params, cache, out, base, label, scale=1.0):
    """[(dw, db)] per layer for ``loss``."""
    m = ~np.isnan(label)
    d = np.where(m, out[..., 0] + base - label, 0.0)
    dz = np.zeros_like(out)
    dz[..., 0] = scale * 2.0 * d
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        a, z = cache[li]
        if li < len(params) - 1:
            dz = dz * gelu_grad(z)
        da, dw, db = conv3x3_backward(a, params[li][0], dz)
        grads[li] = (dw, db)
        dz = da
    return grads


def adam_step(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba 2015) in float64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(rng, widths: List[int]):
    """He-initialised layers (cin -> cout) for the given channel widths."""
    params = []
    for cin, cout in zip(widths[:-1], widths[1:]):
        w = rng.normal(0.0, math.sqrt(2.0 / (9 * cin)), size=(cout, cin, 3, 3))
        params.append((w, rng.normal(0.0, 0.05, size=cout)))
    return params

"""Float64 CPU oracle of the openness surrogate's network -- TEST INFRASTRUCTURE.

The network has no reference counterpart (the reference has no learned
component).  This module restates its definition in NumPy float64 so that
the HIP kernels' forward pass, gradients and Adam step can be checked:

* conv3x3: "same" 3x3 correlation with zero padding on NHWC arrays;
  tap (dy, dx) reads x[y + dy - 1, x + dx - 1] with weight w[co, ci, dy, dx];
* GELU in its tanh form (Hendrycks & Gimpel 2016): x * sigmoid(u),
  u = 2 sqrt(2/pi) (x + 0.044715 x^3);
* layers 0..4: conv + bias + GELU; layer 5: conv + bias (linear);
* loss: sum over labelled cells of (out[..., 0] + base - label)^2 plus, with a
  detail label, sum of weight_s * (out[..., 1] - label_s)^2, times scale;
* compose: the full-resolution angle from the grid output, bilinear (pixel-
  centre aligned, clamped) of out[..., 0] + base plus out[cell, 1] * s_scale *
  (z - bilinear(zg));
* detail_label: the per-cell least-squares slope of exact - bilinear(label)
  on z - bilinear(zg).

``backward`` is the hand-derived gradient; tests/test_oracle_pins.py pins it
against central finite differences in float64, ``bilinear`` against the
exact reproduction of affine fields, and ``detail_label`` against fields built
with a known slope.
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


K0 = 2.0 * math.sqrt(2.0 / math.pi)
K1 = K0 * 0.044715


def _sig(x):
    u = x * (K0 + K1 * x * x)
    return 0.5 * (1.0 + np.tanh(0.5 * u))


def gelu(x):
    return x * _sig(x)


def gelu_grad(x):
    s = _sig(x)
    return s + x * s * (1.0 - s) * (K0 + 3.0 * K1 * x * x)


def gelu_erf(x):
    """The erf form x * Phi(x), which the tanh form approximates."""
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def forward(params: Sequence[Tuple[np.ndarray, np.ndarray]], x: np.ndarray):
    """Returns (output [..., cout_last], cache of (input, pre-activation) per layer)."""
    cache = []
    a = x
    for li, (w, b) in enumerate(params):
        z = conv3x3(a, w, b)
        cache.append((a, z))
        a = gelu(z) if li < len(params) - 1 else z
    return a, cache


def loss(out, base, label, scale=1.0, label_s=None, weight_s=None) -> float:
    m = ~np.isnan(label)
    d = out[..., 0] + base - label
    total = np.sum(np.where(m, d * d, 0.0))
    if label_s is not None:
        ms = ~np.isnan(label_s)
        e = out[..., 1] - label_s
        total += np.sum(np.where(ms, weight_s * e * e, 0.0))
    return float(scale * total)


def backward(params, cache, out, base, label, scale=1.0, label_s=None, weight_s=None):
    """[(dw, db)] per layer for ``loss``."""
    m = ~np.isnan(label)
    d = np.where(m, out[..., 0] + base - label, 0.0)
    dz = np.zeros_like(out)
    dz[..., 0] = scale * 2.0 * d
    if label_s is not None:
        ms = ~np.isnan(label_s)
        dz[..., 1] = scale * 2.0 * np.where(ms, weight_s * (out[..., 1] - label_s), 0.0)
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


def bilinear(src: np.ndarray, f: int, h: int, w: int) -> np.ndarray:
    """Pixel-centre aligned bilinear upsampling by f of a [hs, ws] grid to
    [h, w]: sample at ((i + .5) / f - .5, (j + .5) / f - .5), clamped."""
    hs, ws = src.shape
    y = np.clip((np.arange(h) + 0.5) / f - 0.5, 0, hs - 1)
    x = np.clip((np.arange(w) + 0.5) / f - 0.5, 0, ws - 1)
    y0 = np.minimum(y.astype(int), hs - 1)
    x0 = np.minimum(x.astype(int), ws - 1)
    y1, x1 = np.minimum(y0 + 1, hs - 1), np.minimum(x0 + 1, ws - 1)
    ty, tx = (y - y0)[:, None], (x - x0)[None, :]
    top = src[y0][:, x0] * (1 - tx) + src[y0][:, x1] * tx
    bot = src[y1][:, x0] * (1 - tx) + src[y1][:, x1] * tx
    return top * (1 - ty) + bot * ty


def compose(out: np.ndarray, base: np.ndarray, zg: np.ndarray, f: int, z=None,
            s_scale: float = 1.0) -> np.ndarray:
    """Full-resolution angle from the grid output out [hs, ws, cout]."""
    hs, ws = base.shape
    h, w = (hs * f, ws * f) if z is None else z.shape
    res = bilinear(out[..., 0] + base, f, h, w)
    if z is not None:
        cy = np.minimum(np.arange(h) // f, hs - 1)
        cx = np.minimum(np.arange(w) // f, ws - 1)
        res = res + s_scale * out[..., 1][cy][:, cx] * (z - bilinear(zg, f, h, w))
    return res


def detail_label(exact: np.ndarray, ey0: int, ex0: int, lab: np.ndarray, zg: np.ndarray,
                 f: int, z: np.ndarray, s_unit: float):
    """Per-cell (label_s, weight_s): the least-squares slope of exact -
    bilinear(lab) on z - bilinear(zg) over each fully labelled cell."""
    hs, ws = lab.shape
    H, W = hs * f, ws * f
    bl = bilinear(lab, f, H, W)
    r = np.full((H, W), np.nan)
    eh, ew = exact.shape
    r[ey0:ey0 + eh, ex0:ex0 + ew] = exact - bl[ey0:ey0 + eh, ex0:ex0 + ew]
    dz = z[:H, :W] - bilinear(zg, f, H, W)
    ls = np.full((hs, ws), np.nan)
    wt = np.zeros((hs, ws))
    for cy in range(hs):
        for cx in range(ws):
            rr = r[cy * f:(cy + 1) * f, cx * f:(cx + 1) * f]
            dd = dz[cy * f:(cy + 1) * f, cx * f:(cx + 1) * f]
            sdd = float(np.sum(dd * dd))
            if np.isnan(lab[cy, cx]) or np.isnan(rr).any() or sdd <= 0:
                continue
            ls[cy, cx] = s_unit * float(np.sum(rr * dd)) / sdd
            wt[cy, cx] = sdd / (f * f * s_unit * s_unit)
    return ls, wt

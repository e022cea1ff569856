"""Seeded synthetic DEMs: spectral fractional Brownian motion and NoData discs.

fBm: white Gaussian noise -> FFT -> amplitude |f|^-(H+1) (power spectrum
proportional to f^-(2H+2)), DC removed -> inverse FFT -> min-max rescale to
[0, relief] -> float32.  Generated with torch on the requested device, so the
large rasters are synthesised directly in HBM.
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np
import torch


def fbm(h: int, w: int, hurst: float, relief: float, seed: int, device="cpu",
        chunk_rows: int = 4096) -> torch.Tensor:
    """float32 (h, w) fBm surface rescaled to [0, relief]."""
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    noise = torch.randn((h, w), generator=gen, device=dev, dtype=torch.float32)
    spec = torch.fft.rfft2(noise)
    del noise
    fy = torch.fft.fftfreq(h, device=dev, dtype=torch.float32)
    fx = torch.fft.rfftfreq(w, device=dev, dtype=torch.float32)
    expo = -(float(hurst) + 1.0) / 2.0  # |f|^-(H+1) = (f^2)^expo
    for r0 in range(0, h, chunk_rows):
        r1 = min(h, r0 + chunk_rows)
        f2 = fy[r0:r1, None] ** 2 + fx[None, :] ** 2
        amp = torch.where(f2 > 0, f2.clamp_min(1e-30) ** expo, torch.zeros_like(f2))
        spec[r0:r1] *= amp
        del f2, amp
    surf = torch.fft.irfft2(spec, s=(h, w))
    del spec
    lo, hi = surf.min(), surf.max()
    span = torch.clamp(hi - lo, min=1e-30)
    surf = (surf - lo) / span * float(relief)
    return surf.to(torch.float32).contiguous()


def nodata_discs(h: int, w: int, coverage: float, seed: int, r_min: int = 5,
                 r_max: int = 50) -> List[Tuple[int, int, int]]:
    """Seeded discs (cy, cx, r) added until the union covers >= coverage of the
    raster (coverage is tracked on the host at 1/8 resolution)."""
    rng = np.random.default_rng(int(seed))
    f = 8
    ch, cw = max(1, h // f), max(1, w // f)
    cov = np.zeros((ch, cw), dtype=bool)
    target = coverage * ch * cw
    discs = []
    while cov.sum() < target:
        cy, cx = int(rng.integers(0, h)), int(rng.integers(0, w))
        r = int(rng.integers(r_min, r_max + 1))
        discs.append((cy, cx, r))
        y0, y1 = max(0, (cy - r) // f), min(ch, (cy + r) // f + 1)
        x0, x1 = max(0, (cx - r) // f), min(cw, (cx + r) // f + 1)
        yy, xx = np.ogrid[y0:y1, x0:x1]
        cov[y0:y1, x0:x1] |= ((yy * f + f // 2 - cy) ** 2 + (xx * f + f // 2 - cx) ** 2) <= r * r
    return discs


def apply_discs(dem: torch.Tensor, discs, value: float = float("nan")) -> torch.Tensor:
    """Set every disc cell of ``dem`` to ``value`` in place."""
    h, w = dem.shape
    for cy, cx, r in discs:
        y0, y1 = max(0, cy - r), min(h, cy + r + 1)
        x0, x1 = max(0, cx - r), min(w, cx + r + 1)
        if y0 >= y1 or x0 >= x1:
            continue
        yy = torch.arange(y0, y1, device=dem.device).view(-1, 1)
        xx = torch.arange(x0, x1, device=dem.device).view(1, -1)
        m = (yy - cy) ** 2 + (xx - cx) ** 2 <= r * r
        sub = dem[y0:y1, x0:x1]
        sub[m] = value
    return dem

"""The prepare step on the GPU: NoData inference and push-pull void fill.

Host mirror of preprocess_dem_to_cog (io/dem_preprocess.py:462-727) for a
raster already resident in device memory: the same arguments and steps, with
the file read and the COG write left to the caller (they are I/O, not compute).

    out, info = preprocess_dem(raster, src_nodata=None, fill_mode="enclosed")

1. NEAREST sample of the raster at (ch k, cw k) (_coarse_shape, oversample k).
2. Undeclared NoData: the dominant sentinel / extreme value of the sample
   (_detect_sentinel_nodata, :241-275) and the dominant constant border
   (_detect_border_nodata, :191-238), from device mode statistics.
3. NaN-aware coarse average of the sample (_nan_aware_coarse_average).
4. If the coarse grid has voids: push-pull fill of the coarse grid
   (algorithms/_pyramid_fill.py) and, for fill_mode 'enclosed', the voids
   4-connected to the grid edge (_edge_connected_mask) kept as exterior.
5. Every NoData pixel -> NaN, then filled from the bilinearly sampled coarse
   surface (_apply_band_fill, _sample_coarse).

Each step is a kernel of csrc/prepare.hip; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _native as N
from . import ops

FILL_MODES = ("none", "enclosed", "all")
_FILL_CODE = {"none": N.FILL_NONE, "enclosed": N.FILL_ENCLOSED, "all": N.FILL_ALL}

# io/dem_preprocess.py:231-238
NODATA_SENTINELS = (
    0.0, -9999.0, -99999.0, 9999.0,
    -32768.0, 32767.0, -32767.0, 65535.0,
    -2147483648.0, 2147483647.0, -2147483647.0,
    -128.0, 127.0, 255.0,
)


def _s():
    return torch.cuda.current_stream().cuda_stream


def coarse_shape(width: int, height: int, coarse_max: int) -> Tuple[int, int]:
    """_coarse_shape (io/dem_preprocess.py:278-282): (ch, cw)."""
    scale = max(width / coarse_max, height / coarse_max, 1.0)
    return max(1, int(round(height / scale))), max(1, int(round(width / scale)))


def coarse_oversample(ch: int, cw: int, start: int = 4, max_intermediate: int = 4096) -> int:
    """io/dem_preprocess.py:530-534 and _nan_aware_coarse_average :162-164."""
    k = max(2, int(start))
    while k > 2 and max(ch * k, cw * k) > int(max_intermediate):
        k -= 1
    return k


def nearest_sample(raster: torch.Tensor, oh: int, ow: int) -> torch.Tensor:
    out = torch.empty((oh, ow), dtype=torch.float32, device=raster.device)
    wd = ops.window(raster)
    N.check(N.lib().fsg_prepare_sample(C.byref(wd), oh, ow, out.data_ptr(), out.stride(0), _s()))
    return out


def value_mode(values: torch.Tensor, src_nodata: Optional[float]) -> Dict[str, float]:
    """Mode statistics of the finite values that are not the declared NoData."""
    v = values.contiguous().reshape(-1)
    n = int(v.numel())
    if n == 0:
        return {"mode": 0.0, "count": 0, "n_valid": 0, "vmin": 0.0, "vmax": 0.0, "n_distinct": 0}
    nb = int(N.lib().fsg_value_mode_workspace_size(n))
    if nb <= 0:
        raise RuntimeError("fsg_value_mode_workspace_size failed")
    ws = torch.empty(nb, dtype=torch.uint8, device=v.device)
    st = torch.zeros(C.sizeof(N.fsg_mode_stats), dtype=torch.uint8, device=v.device)
    has = src_nodata is not None and math.isfinite(float(src_nodata))
    N.check(N.lib().fsg_value_mode(v.data_ptr(), n, 1 if has else 0,
                                   float(src_nodata) if has else 0.0, st.data_ptr(),
                                   ws.data_ptr(), nb, _s()))
    s = N.fsg_mode_stats.from_buffer_copy(bytes(st.cpu().numpy().tobytes()))
    return {"mode": s.mode, "count": int(s.count), "n_valid": int(s.n_valid), "vmin": s.vmin,
            "vmax": s.vmax, "n_distinct": int(s.n_distinct)}


def count_equal(values: torch.Tensor, value: float) -> int:
    v = values.contiguous().reshape(-1)
    cnt = torch.zeros(1, dtype=torch.int64, device=v.device)
    N.check(N.lib().fsg_count_equal(v.data_ptr(), int(v.numel()), float(value), cnt.data_ptr(),
                                    _s()))
    return int(cnt.item())


def detect_sentinel_nodata(sample: torch.Tensor, src_nodata, *,
                           min_fraction: float = 0.05) -> Optional[float]:
    """_detect_sentinel_nodata (io/dem_preprocess.py:241-275) from device
    mode statistics of the sample's valid finite values."""
    st = value_mode(sample, src_nodata)
    if st["n_valid"] < 64:
        return None
    cand = float(st["mode"])
    if st["count"] / float(st["n_valid"]) < float(min_fraction):
        return None
    known = any(abs(cand - s) <= 1e-6 * max(1.0, abs(s)) for s in NODATA_SENTINELS)
    very_extreme = abs(cand) >= 1e30
    is_range_extreme = cand == float(st["vmin"]) or cand == float(st["vmax"])
    if known or very_extreme or is_range_extreme:
        return cand
    return None


def detect_border_nodata(sample: torch.Tensor, src_nodata, *, min_border_fraction: float = 0.5,
                         min_total_fraction: float = 0.02) -> Optional[float]:
    """_detect_border_nodata (io/dem_preprocess.py:191-238): the dominant
    value of the sample's outer ring, if it also covers min_total_fraction of
    the whole sample."""
    h, w = sample.shape
    if h < 8 or w < 8:
        return None
    ring = torch.cat([sample[0, :], sample[-1, :], sample[1:-1, 0], sample[1:-1, -1]])
    st = value_mode(ring, src_nodata)
    if st["n_valid"] == 0:
        return None
    cand = float(st["mode"])
    if st["count"] / float(st["n_valid"]) < float(min_border_fraction):
        return None
    if count_equal(sample, cand) / float(h * w) < float(min_total_fraction):
        return None
    return cand


def coarse_average(sample: torch.Tensor, ch: int, cw: int, k: int, nodata: List[float]):
    """_nan_aware_coarse_average: (coarse, cvalid uint8, number of voids)."""
    coarse = torch.empty((ch, cw), dtype=torch.float32, device=sample.device)
    cvalid = torch.empty((ch, cw), dtype=torch.uint8, device=sample.device)
    n_void = torch.zeros(1, dtype=torch.int64, device=sample.device)
    arr = (C.c_float * max(1, len(nodata)))(*[float(v) for v in nodata])
    N.check(N.lib().fsg_coarse_average(sample.data_ptr(), sample.stride(0), ch, cw, k, arr,
                                       len(nodata), coarse.data_ptr(), coarse.stride(0),
                                       cvalid.data_ptr(), n_void.data_ptr(), _s()))
    return coarse, cvalid, int(n_void.item())


def pushpull_fill(coarse: torch.Tensor, cvalid: torch.Tensor, n_void: int) -> torch.Tensor:
    """pushpull_fill (algorithms/_pyramid_fill.py:24-87) of the coarse grid."""
    h, w = coarse.shape
    out = torch.empty_like(coarse)
    nb = int(N.lib().fsg_pushpull_workspace_size(h, w))
    ws = torch.empty(max(1, nb), dtype=torch.uint8, device=coarse.device)
    N.check(N.lib().fsg_pushpull_fill(coarse.data_ptr(), coarse.stride(0), cvalid.data_ptr(), h,
                                      w, h * w - int(n_void), out.data_ptr(), out.stride(0),
                                      ws.data_ptr(), nb, _s()))
    return out


def edge_connected(mask: torch.Tensor) -> Tuple[torch.Tensor, int]:
    """_edge_connected_mask (utils/nodata_handler.py:13-25): (uint8 mask of
    the cells 4-connected to the edge, number of sweeps)."""
    m = mask.to(torch.uint8).contiguous()
    h, w = m.shape
    ext = torch.empty_like(m)
    flag = torch.zeros(1, dtype=torch.int32, device=m.device)
    sweeps = C.c_int32(0)
    N.check(N.lib().fsg_edge_connected(m.data_ptr(), h, w, ext.data_ptr(), flag.data_ptr(),
                                       C.byref(sweeps), _s()))
    return ext, int(sweeps.value)


def _isclose(a: float, b: float) -> bool:
    return abs(a - b) <= 1e-8 + 1e-5 * abs(b)


def preprocess_dem(raster: torch.Tensor, *, src_nodata: Optional[float] = None,
                   fill_mode: str = "enclosed", coarse_max: int = 2048,
                   nodata_override: Optional[float] = None, detect_nodata: bool = True,
                   nodata_border_fraction: float = 0.5, nodata_sentinel_fraction: float = 0.05,
                   out: Optional[torch.Tensor] = None):
    """preprocess_dem_to_cog (io/dem_preprocess.py:462-727) on a float32 GPU
    raster: returns (prepared float32 raster, info).  ``out`` may be the
    raster itself (in place)."""
    fill_mode = str(fill_mode).lower()
    if fill_mode not in FILL_MODES:
        raise ValueError(f"Unknown fill_mode={fill_mode!r}. Choose from {FILL_MODES}.")
    ops._check_window(raster, "raster")
    height, width = raster.shape
    ch, cw = coarse_shape(width, height, coarse_max)
    extra: List[float] = []
    if nodata_override is not None and math.isfinite(float(nodata_override)):
        extra.append(float(nodata_override))
    k = coarse_oversample(ch, cw)
    nn = nearest_sample(raster, ch * k, cw * k)
    detected = []
    if detect_nodata:
        for cand, why in ((detect_sentinel_nodata(nn, src_nodata,
                                                  min_fraction=nodata_sentinel_fraction),
                           "sentinel"),
                          (detect_border_nodata(nn, src_nodata,
                                                min_border_fraction=nodata_border_fraction),
                           "border")):
            if cand is None:
                continue
            already = any(_isclose(cand, v) for v in extra)
            declared = src_nodata is not None and _isclose(cand, float(src_nodata))
            if not already and not declared:
                extra.append(float(cand))
                detected.append((why, float(cand)))
    nodata = ([float(src_nodata)] if src_nodata is not None and math.isfinite(float(src_nodata))
              else []) + [v for v in extra if math.isfinite(v)]
    if len(nodata) > N.PREPARE_MAX_NODATA:
        raise ValueError(f"at most {N.PREPARE_MAX_NODATA} NoData values")
    coarse, cvalid, n_void = coarse_average(nn, ch, cw, k, nodata)
    do_fill = fill_mode != "none" and n_void > 0
    surface = exterior = None
    sweeps = 0
    if do_fill:
        surface = pushpull_fill(coarse, cvalid, n_void)
        if fill_mode == "enclosed":
            exterior, sweeps = edge_connected(cvalid == 0)
    if out is None:
        out = torch.empty_like(raster)
    elif out.shape != raster.shape or out.dtype != torch.float32:
        raise ValueError("out must be a float32 raster of the same shape")
    wd = ops.window(raster)
    arr = (C.c_float * max(1, len(nodata)))(*nodata)
    N.check(N.lib().fsg_prepare_fill(
        C.byref(wd), arr, len(nodata), _FILL_CODE[fill_mode], 1 if do_fill else 0,
        surface.data_ptr() if surface is not None else None,
        exterior.data_ptr() if exterior is not None else None, ch, cw, out.data_ptr(),
        out.stride(0), _s()))
    info = {"coarse_shape": (ch, cw), "oversample": k, "extra_nodata": extra,
            "detected": detected, "do_fill": do_fill, "n_void": n_void, "sweeps": sweeps,
            "coarse": coarse, "cvalid": cvalid, "surface": surface, "exterior": exterior}
    return out, info

"""CPU oracle of the ``prepare`` step (NoData inference and push-pull void
fill) -- TEST INFRASTRUCTURE ONLY.

Restates ``preprocess_dem_to_cog`` (io/dem_preprocess.py:462-727) for a raster
already in memory, without the file reads and the COG write: the NEAREST
coarse sample, the sentinel and border NoData rules, the NaN-aware coarse
average, the push-pull fill of the coarse grid (algorithms/_pyramid_fill.py),
the edge-connected exterior (utils/nodata_handler.py:13-25) and the per-pixel
fill of the full raster from the bilinearly sampled coarse surface.  NumPy /
SciPy with the reference's float32 dtypes and operation order.

Pinning (tests/test_oracle_pins.py): the reference's own tests in
tests/test_dem_preprocess.py (void fill preserves valid cells and stays in
range, a flat plateau stays flat, an all-void grid gives zeros, the coarse
average does not blend a 0 exterior into 5000 m terrain, the coarse shape
caps the longest side), brute-force 4-connected flood fill for the exterior,
np.repeat for the integer-ratio NEAREST sample, constructed rasters for both
detection rules.  Parity unpinned: GDAL's NEAREST RasterIO for non-integer
ratios (restated as floor((i + 0.5) * n_src / n_dst); GDAL is absent here).
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np
from scipy import ndimage as ndi

F32 = np.float32

FILL_MODES = ("none", "enclosed", "all")

# io/dem_preprocess.py:231-238
NODATA_SENTINELS = (
    0.0, -9999.0, -99999.0, 9999.0,
    -32768.0, 32767.0, -32767.0, 65535.0,
    -2147483648.0, 2147483647.0, -2147483647.0,
    -128.0, 127.0, 255.0,
)


def coarse_shape(width: int, height: int, coarse_max: int) -> Tuple[int, int]:
    """_coarse_shape (io/dem_preprocess.py:278-282): (ch, cw)."""
    scale = max(width / coarse_max, height / coarse_max, 1.0)
    cw = max(1, int(round(width / scale)))
    ch = max(1, int(round(height / scale)))
    return ch, cw


def coarse_oversample(ch: int, cw: int, start: int = 4, max_intermediate: int = 4096) -> int:
    """The NEAREST sample's oversampling (io/dem_preprocess.py:530-534 and
    _nan_aware_coarse_average :162-164)."""
    k = max(2, int(start))
    while k > 2 and max(ch * k, cw * k) > int(max_intermediate):
        k -= 1
    return k


def nearest_sample(raster: np.ndarray, oh: int, ow: int) -> np.ndarray:
    """src.read(out_shape=(oh, ow), resampling=nearest): source pixel
    floor((i + 0.5) * n_src / n_dst) per axis (GDAL RasterIO NEAREST)."""
    a = np.asarray(raster, dtype=F32)
    h, w = a.shape
    ri = np.minimum(np.floor((np.arange(oh) + 0.5) * (h / oh)).astype(np.int64), h - 1)
    ci = np.minimum(np.floor((np.arange(ow) + 0.5) * (w / ow)).astype(np.int64), w - 1)
    return a[ri[:, None], ci[None, :]]


def declared_mask(values: np.ndarray, src_nodata: Optional[float]) -> np.ndarray:
    """The masked read's mask: cells equal to the declared NoData (NaN when
    the declared NoData is NaN)."""
    if src_nodata is None:
        return np.zeros(values.shape, dtype=bool)
    if not np.isfinite(src_nodata):
        return np.isnan(values)
    return values == F32(src_nodata)


def detect_border_nodata(sample, valid, *, min_border_fraction=0.5,
                         min_total_fraction=0.02) -> Optional[float]:
    """_detect_border_nodata (io/dem_preprocess.py:191-238)."""
    h, w = sample.shape
    if h < 8 or w < 8:
        return None
    ring = np.concatenate([sample[0, :], sample[-1, :], sample[1:-1, 0], sample[1:-1, -1]])
    ring_valid = np.concatenate([valid[0, :], valid[-1, :], valid[1:-1, 0], valid[1:-1, -1]])
    ring = ring[ring_valid & np.isfinite(ring)]
    if ring.size == 0:
        return None
    vals, counts = np.unique(ring, return_counts=True)
    i = int(np.argmax(counts))
    cand = float(vals[i])
    if counts[i] / ring.size < float(min_border_fraction):
        return None
    if float(np.mean(sample == np.float32(cand))) < float(min_total_fraction):
        return None
    return cand


def detect_sentinel_nodata(sample, valid, *, min_fraction=0.05) -> Optional[float]:
    """_detect_sentinel_nodata (io/dem_preprocess.py:241-275)."""
    finite = sample[valid & np.isfinite(sample)]
    if finite.size < 64:
        return None
    vals, counts = np.unique(finite, return_counts=True)
    i = int(np.argmax(counts))
    cand = float(vals[i])
    if counts[i] / float(finite.size) < float(min_fraction):
        return None
    known = any(abs(cand - s) <= 1e-6 * max(1.0, abs(s)) for s in NODATA_SENTINELS)
    very_extreme = abs(cand) >= 1e30
    is_range_extreme = (cand == float(vals[0]) or cand == float(vals[-1]))
    if known or very_extreme or is_range_extreme:
        return cand
    return None


def nan_aware_coarse_average(values, mask, ch, cw, k, src_nodata, extra_nodata):
    """_nan_aware_coarse_average (io/dem_preprocess.py:135-188) on the
    (ch k, cw k) NEAREST sample: (coarse, cmask)."""
    a = np.asarray(values, dtype=F32)
    m = np.asarray(mask, dtype=bool) | ~np.isfinite(a)
    if src_nodata is not None and np.isfinite(src_nodata):
        m = m | (a == np.float32(src_nodata))
    for v in extra_nodata:
        if np.isfinite(v):
            m = m | (a == np.float32(v))
    a = np.where(m, np.nan, a).reshape(ch, k, cw, k)
    finite = np.isfinite(a)
    cnt = finite.sum(axis=(1, 3))
    s = np.nansum(a, axis=(1, 3))
    cmask = cnt == 0
    coarse = np.where(cmask, np.float32(0.0), s / np.maximum(cnt, 1)).astype(np.float32)
    return coarse, cmask


def _zoom(a, zy, zx):
    return ndi.zoom(a, zoom=(zy, zx), order=1, mode="nearest")


def pushpull_fill(coarse, valid) -> np.ndarray:
    """pushpull_fill (algorithms/_pyramid_fill.py:24-87) with scipy's zoom."""
    f32 = np.float32
    out = np.asarray(coarse).astype(f32, copy=True)
    valid = np.asarray(valid, dtype=bool)
    if bool(valid.all()):
        return out
    if not bool(valid.any()):
        return np.zeros_like(out, dtype=f32)
    eps = f32(1e-6)
    w = valid.astype(f32)
    vw = np.where(valid, out, f32(0.0)).astype(f32)
    vws, ws = [vw], [w]
    while max(vws[-1].shape[:2]) > 1 and float(ws[-1].min()) <= 1e-6:
        ch, cw = vws[-1].shape[:2]
        nh, nw = max(1, ch // 2), max(1, cw // 2)
        zy, zx = nh / float(ch), nw / float(cw)
        num = _zoom(vws[-1], zy, zx)
        den = _zoom(ws[-1], zy, zx)
        wv = np.minimum(den, f32(1.0))
        with np.errstate(invalid="ignore", divide="ignore"):
            mean = np.where(den > eps, num / np.maximum(den, eps), f32(0.0))
        vws.append((mean * wv).astype(f32))
        ws.append(wv.astype(f32))
    with np.errstate(invalid="ignore", divide="ignore"):
        filled = np.where(ws[-1] > eps, vws[-1] / np.maximum(ws[-1], eps),
                          f32(0.0)).astype(f32)
    for lvl in range(len(vws) - 2, -1, -1):
        th, tw = vws[lvl].shape[:2]
        fh, fw = filled.shape[:2]
        up = _zoom(filled, th / float(fh), tw / float(fw))
        pad_h, pad_w = max(0, th - up.shape[0]), max(0, tw - up.shape[1])
        if pad_h or pad_w:
            up = np.pad(up, ((0, pad_h), (0, pad_w)), mode="edge")
        up = up[:th, :tw]
        wl = ws[lvl]
        with np.errstate(invalid="ignore", divide="ignore"):
            vl = np.where(wl > eps, vws[lvl] / np.maximum(wl, eps), f32(0.0))
        filled = np.where(wl > eps, vl, up).astype(f32)
    return np.where(valid, out, filled).astype(f32)


def pushpull_levels(h: int, w: int, wmins: List[float]) -> List[Tuple[int, int]]:
    """Level shapes of the push phase given each level's min weight."""
    shapes = [(h, w)]
    i = 0
    while max(shapes[-1]) > 1 and wmins[i] <= 1e-6:
        ch, cw = shapes[-1]
        shapes.append((max(1, ch // 2), max(1, cw // 2)))
        i += 1
    return shapes


def fill_coarse_surface(coarse, valid) -> np.ndarray:
    """_fill_coarse_surface (io/dem_preprocess.py:79-132)."""
    if np.asarray(valid).all():
        return np.asarray(coarse).astype(np.float32, copy=True)
    if not np.asarray(valid).any():
        return np.zeros_like(coarse, dtype=np.float32)
    return pushpull_fill(coarse, valid)


def edge_connected_mask(mask) -> np.ndarray:
    """_edge_connected_mask (utils/nodata_handler.py:13-25): masked cells
    4-connected (binary_propagation's default cross) to the grid edge."""
    mask = np.asarray(mask, dtype=bool)
    if not np.any(mask):
        return np.zeros_like(mask, dtype=bool)
    seeds = np.zeros_like(mask, dtype=bool)
    seeds[0, :] = mask[0, :]
    seeds[-1, :] = mask[-1, :]
    seeds[:, 0] = mask[:, 0]
    seeds[:, -1] = mask[:, -1]
    if not np.any(seeds):
        return np.zeros_like(mask, dtype=bool)
    return ndi.binary_propagation(seeds, mask=mask).astype(bool, copy=False)


def sample_coarse(surface, exterior, height, width, ch, cw):
    """_sample_coarse (io/dem_preprocess.py:729-767) for the whole raster:
    map_coordinates order 1, mode 'nearest', at float32 pixel-centre
    coordinates in the global coarse grid."""
    rr = (np.arange(height, dtype=np.float32) + np.float32(0.5)) * np.float32(ch / height) \
        - np.float32(0.5)
    cc = (np.arange(width, dtype=np.float32) + np.float32(0.5)) * np.float32(cw / width) \
        - np.float32(0.5)
    coords = np.empty((2, height, width), dtype=np.float32)
    coords[0] = rr[:, None]
    coords[1] = cc[None, :]
    fill_vals = ndi.map_coordinates(surface, coords, order=1, mode="nearest", output=np.float32)
    if exterior is None:
        ext = np.zeros((height, width), dtype=bool)
    else:
        ext = ndi.map_coordinates(exterior.astype(np.float32), coords, order=1,
                                  mode="nearest") > 0.5
    return fill_vals, ext


def apply_fill(raster, declared, *, do_fill, surface, exterior, extra_nodata, fill_mode, ch, cw):
    """_apply_band_fill (io/dem_preprocess.py:770-826) over the whole raster
    (the bands sample the global coarse grid, so one band is the same)."""
    arr = np.array(raster, dtype=np.float32, copy=True)
    height, width = arr.shape
    wmask = np.asarray(declared, dtype=bool) | ~np.isfinite(arr)
    for v in extra_nodata:
        wmask = wmask | (arr == np.float32(v))
    if wmask.any():
        arr[wmask] = np.float32(np.nan)
    if do_fill and wmask.any():
        fill_vals, ext = sample_coarse(surface, exterior, height, width, ch, cw)
        fill_here = wmask & ~ext if fill_mode == "enclosed" else wmask
        arr[fill_here] = fill_vals[fill_here]
    if fill_mode == "all":
        if not np.isfinite(arr).all():
            raise RuntimeError("fill_mode='all' left non-finite pixels")
    else:
        arr[~np.isfinite(arr)] = np.float32(np.nan)
    return arr


def prepare(raster, *, src_nodata=None, fill_mode="enclosed", coarse_max=2048,
            nodata_override=None, detect_nodata=True, nodata_border_fraction=0.5,
            nodata_sentinel_fraction=0.05):
    """preprocess_dem_to_cog (io/dem_preprocess.py:462-727) on an in-memory
    raster: returns (filled float32 raster, info)."""
    fill_mode = str(fill_mode).lower()
    if fill_mode not in FILL_MODES:
        raise ValueError(f"Unknown fill_mode={fill_mode!r}")
    a = np.asarray(raster, dtype=np.float32)
    height, width = a.shape
    ch, cw = coarse_shape(width, height, coarse_max)
    extra: List[float] = []
    if nodata_override is not None and np.isfinite(nodata_override):
        extra.append(float(nodata_override))
    k = coarse_oversample(ch, cw)
    nn = nearest_sample(a, ch * k, cw * k)
    nmask = declared_mask(nn, src_nodata)
    detected = []
    if detect_nodata:
        nn_valid = ~nmask & np.isfinite(nn)
        for cand, why in ((detect_sentinel_nodata(nn, nn_valid,
                                                  min_fraction=nodata_sentinel_fraction),
                           "sentinel"),
                          (detect_border_nodata(nn, nn_valid,
                                                min_border_fraction=nodata_border_fraction),
                           "border")):
            if cand is None:
                continue
            already = any(np.isclose(cand, v) for v in extra)
            declared = src_nodata is not None and np.isclose(cand, float(src_nodata))
            if not already and not declared:
                extra.append(float(cand))
                detected.append((why, float(cand)))
    coarse, cmask = nan_aware_coarse_average(nn, nmask, ch, cw, k, src_nodata, extra)
    cvalid = (~cmask) & np.isfinite(coarse)
    has_holes = bool((~cvalid).any())
    do_fill = fill_mode != "none" and has_holes
    surface = exterior = None
    if do_fill:
        surface = fill_coarse_surface(coarse, cvalid)
        if fill_mode == "enclosed":
            exterior = edge_connected_mask(~cvalid)
    out = apply_fill(a, declared_mask(a, src_nodata), do_fill=do_fill, surface=surface,
                     exterior=exterior, extra_nodata=extra, fill_mode=fill_mode, ch=ch, cw=cw)
    info = {"coarse_shape": (ch, cw), "oversample": k, "extra_nodata": extra,
            "detected": detected, "do_fill": do_fill, "coarse": coarse, "cvalid": cvalid,
            "surface": surface, "exterior": exterior}
    return out, info

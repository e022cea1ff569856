"""Float32-faithful CPU restatement of the reference's per-tile compute.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU baseline.  Each function cites the reference file:line it
restates (paths relative to the FujiShaderGPU package of geoign/FujiShaderGPU
1.0.1).  ``cupy`` -> ``numpy`` and ``cupyx.scipy.ndimage`` -> ``scipy.ndimage``.

Type-promotion conventions (CuPy's, which the reference runs under):
* Python scalars are weak: float32_array <op> python_float -> float32, the
  scalar first rounded to float32.
* 0-d CuPy arrays are arrays: float32_array * float64_0d -> float64.  This is
  why hillshade's light-vector products are evaluated in float64
  (_impl_hillshade.py:46-52); the result is cast to float32 at :52.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np
from scipy import ndimage as ndi

F32 = np.float32
GAMMA = 1 / 2.2                      # Constants.DEFAULT_GAMMA (_base.py:13)
DEFAULT_AZIMUTH = 315                # _base.py:14
DEFAULT_ALTITUDE = 45                # _base.py:15


def _err():
    return np.errstate(invalid="ignore", divide="ignore", over="ignore")


# ---------------------------------------------------------------------------
# NaN helpers (_nan_utils.py)
# ---------------------------------------------------------------------------

def handle_nan_with_gaussian(block: np.ndarray, sigma: float, mode: str = "nearest"):
    """_nan_utils.py:18-31."""
    block = np.asarray(block, dtype=F32)
    nan_mask = np.isnan(block)
    if not nan_mask.any():
        return ndi.gaussian_filter(block, sigma=sigma, mode=mode), nan_mask
    filled = np.where(nan_mask, F32(0), block)
    valid = (~nan_mask).astype(F32)
    sv = ndi.gaussian_filter(filled * valid, sigma=sigma, mode=mode)
    sw = ndi.gaussian_filter(valid, sigma=sigma, mode=mode)
    with _err():
        smoothed = np.where(sw > 0, sv / sw, F32(0)).astype(F32)
    return smoothed, nan_mask


def handle_nan_with_uniform(block: np.ndarray, size: int, mode: str = "nearest"):
    """_nan_utils.py:34-47."""
    block = np.asarray(block, dtype=F32)
    nan_mask = np.isnan(block)
    if not nan_mask.any():
        return ndi.uniform_filter(block, size=size, mode=mode), nan_mask
    filled = np.where(nan_mask, F32(0), block)
    valid = (~nan_mask).astype(F32)
    sv = ndi.uniform_filter(filled * valid, size=size, mode=mode)
    sw = ndi.uniform_filter(valid, size=size, mode=mode)
    with _err():
        mean = np.where(sw > 0, sv / sw, F32(0)).astype(F32)
    return mean, nan_mask


def _row_steps(v, pixel_size, signed):
    """A per-row x scale array (no reference counterpart: geographic grids with
    each row's own latitude) with the 1e-9 fallback applied row by row."""
    a = np.asarray(v, dtype=np.float64)
    fb = float(pixel_size if pixel_size else 1.0)
    a = np.where(np.abs(a) < 1e-9, fb, a if signed else np.abs(a))
    return a


def resolve_steps(pixel_size=1.0, pixel_scale_x=None, pixel_scale_y=None):
    """|step| resolution of handle_nan_for_gradient (_nan_utils.py:67-72).
    pixel_scale_x may also be one value per row (returned as an array)."""
    step_y = abs(float(pixel_scale_y if pixel_scale_y is not None else pixel_size))
    if step_y < 1e-9:
        step_y = float(pixel_size if pixel_size else 1.0)
    if pixel_scale_x is not None and np.ndim(pixel_scale_x) == 1:
        return _row_steps(pixel_scale_x, pixel_size, signed=False), step_y
    step_x = abs(float(pixel_scale_x if pixel_scale_x is not None else pixel_size))
    if step_x < 1e-9:
        step_x = float(pixel_size if pixel_size else 1.0)
    return step_x, step_y


def gradient_f32(f: np.ndarray, step_y: float, step_x):
    """np/cp.gradient(f, step_y, step_x, edge_order=2) on float32 with the
    Python spacings as weak scalars (numpy function_base gradient, uniform
    spacing branch).  Returns (d/drow, d/dcol).  A 1-D step_x (one spacing per
    row) takes each row's x derivative with its own spacing -- the same
    formula row by row, so equal spacings give the scalar result bit for bit."""
    f = np.asarray(f, dtype=F32)
    if np.ndim(step_x) == 0:
        return tuple(np.gradient(f, float(step_y), float(step_x), edge_order=2))
    d_row = np.gradient(f, float(step_y), axis=0, edge_order=2)
    d_col = np.empty_like(d_row)
    for i in range(f.shape[0]):
        d_col[i] = np.gradient(f[i], float(step_x[i]), edge_order=2)
    return d_row, d_col


def handle_nan_for_gradient(block, scale=1.0, pixel_size=1.0,
                            pixel_scale_x=None, pixel_scale_y=None):
    """_nan_utils.py:50-74."""
    block = np.asarray(block, dtype=F32)
    nan_mask = np.isnan(block)
    if nan_mask.any():
        if (~nan_mask).any():
            local_fill, _ = handle_nan_with_gaussian(block, sigma=1.0, mode="nearest")
            filled = np.where(nan_mask, local_fill, block)
        else:
            filled = np.zeros_like(block)
    else:
        filled = block
    step_x, step_y = resolve_steps(pixel_size, pixel_scale_x, pixel_scale_y)
    dy, dx = gradient_f32(filled * scale, step_y, step_x)
    return dy, dx, nan_mask


def restore_nan(result: np.ndarray, nan_mask: np.ndarray) -> np.ndarray:
    """_nan_utils.py:701-705."""
    if nan_mask.any():
        result[nan_mask] = np.nan
    return result


# ---------------------------------------------------------------------------
# Hillshade / slope (A2, A3)
# ---------------------------------------------------------------------------

def light_vector(azimuth=DEFAULT_AZIMUTH, altitude=DEFAULT_ALTITUDE):
    """0-d float64 light vector (_impl_hillshade.py:34,47-50)."""
    alt = np.radians(np.float64(altitude))
    az = np.radians(np.float64(float(azimuth)))
    lx = np.sin(az) * np.cos(alt)
    ly = np.cos(az) * np.cos(alt)
    lz = np.sin(alt)
    return np.float64(lx), np.float64(ly), np.float64(lz)


def hillshade(block, *, azimuth=DEFAULT_AZIMUTH, altitude=DEFAULT_ALTITUDE,
              z_factor=1.0, pixel_size=1.0, pixel_scale_x=None, pixel_scale_y=None):
    """compute_hillshade_block (_impl_hillshade.py:20-54)."""
    if z_factor is None:
        z_factor = 1.0
    dy, dx, nan_mask = handle_nan_for_gradient(
        block, scale=z_factor, pixel_size=pixel_size,
        pixel_scale_x=pixel_scale_x, pixel_scale_y=pixel_scale_y)
    sx0 = (pixel_scale_x if (pixel_scale_x is None or np.ndim(pixel_scale_x) == 0)
           else np.asarray(pixel_scale_x).reshape(-1)[0])
    sign_x = 1.0 if (sx0 is None or float(sx0) >= 0.0) else -1.0
    sign_y = 1.0 if (pixel_scale_y is None or float(pixel_scale_y) >= 0.0) else -1.0
    e = dx * F32(sign_x)
    n = dy * F32(sign_y)
    norm = np.sqrt(e * e + n * n + F32(1.0))
    lx, ly, lz = light_vector(azimuth, altitude)
    # float64 from here: 0-d float64 arrays promote (module docstring).
    hs = ((-e).astype(np.float64) * lx + (-n).astype(np.float64) * ly + lz) / norm.astype(np.float64)
    hs = np.clip(hs, 0.0, 1.0).astype(F32)
    return restore_nan(hs, nan_mask)


def slope(block, *, unit="degree", pixel_size=1.0, pixel_scale_x=None, pixel_scale_y=None):
    """compute_slope_block (_impl_slope.py:19-35)."""
    dy, dx, nan_mask = handle_nan_for_gradient(
        block, scale=1, pixel_size=pixel_size,
        pixel_scale_x=pixel_scale_x, pixel_scale_y=pixel_scale_y)
    rad = np.arctan(np.sqrt(dx * dx + dy * dy))
    if unit == "degree":
        # CuPy's degrees kernel multiplies by the double constant 180/pi.
        s = (rad.astype(np.float64) * (180.0 / math.pi)).astype(F32)
    elif unit == "percent":
        s = np.tan(rad) * F32(100)
    else:
        s = rad
    s = restore_nan(s.astype(F32), nan_mask)
    return s.astype(F32)


# ---------------------------------------------------------------------------
# Curvature (A4)
# ---------------------------------------------------------------------------

def curvature_steps(pixel_size=1.0, pixel_scale_x=None, pixel_scale_y=None):
    """Signed steps of compute_curvature_block (_impl_curvature.py:23-28);
    pixel_scale_x may be one value per row."""
    step_y = float(pixel_scale_y if pixel_scale_y is not None else pixel_size)
    if abs(step_y) < 1e-9:
        step_y = float(pixel_size if pixel_size else 1.0)
    if pixel_scale_x is not None and np.ndim(pixel_scale_x) == 1:
        return _row_steps(pixel_scale_x, pixel_size, signed=True), step_y
    step_x = float(pixel_scale_x if pixel_scale_x is not None else pixel_size)
    if abs(step_x) < 1e-9:
        step_x = float(pixel_size if pixel_size else 1.0)
    return step_x, step_y


def curvature(block, *, curvature_type="mean", pixel_size=1.0,
              pixel_scale_x=None, pixel_scale_y=None):
    """compute_curvature_block (_impl_curvature.py:19-57)."""
    block = np.asarray(block, dtype=F32)
    nan_mask = np.isnan(block)
    step_x, step_y = curvature_steps(pixel_size, pixel_scale_x, pixel_scale_y)
    dy, dx, _ = handle_nan_for_gradient(block, pixel_scale_x=step_x, pixel_scale_y=step_y)
    dyy, dyx = gradient_f32(dy, step_y, step_x)
    dxy, dxx = gradient_f32(dx, step_y, step_x)
    one, two = F32(1), F32(2)
    with _err():
        if curvature_type == "mean":
            p, q, r = dx, dy, dxx
            s = (dxy + dyx) / two
            t = dyy
            den = np.power(one + p * p + q * q, F32(1.5))
            num = (one + q * q) * r - two * p * q * s + (one + p * p) * t
            c = -num / (two * den + F32(1e-10))
        elif curvature_type == "gaussian":
            b = one + dx * dx + dy * dy
            c = (dxx * dyy - dxy * dxy) / (b * b)
        elif curvature_type == "planform":
            c = (-(dy * dy * dxx - two * dx * dy * dxy + dx * dx * dyy)
                 / (np.power(dx * dx + dy * dy, F32(1.5)) + F32(1e-10)))
        else:
            c = (-(dx * dx * dxx + two * dx * dy * dxy + dy * dy * dyy)
                 / ((dx * dx + dy * dy) * np.power(one + dx * dx + dy * dy, F32(1.5)) + F32(1e-10)))
        cn = np.tanh(c * F32(100))
        res = (cn + one) / two
        res = np.power(res, F32(GAMMA))
    res = restore_nan(res.astype(F32), nan_mask)
    return res.astype(F32)


# ---------------------------------------------------------------------------
# Down / up sampling (_nan_utils.py:555-698)
# ---------------------------------------------------------------------------

def radius_to_downsample_factor(radius, *, pixel_size=1.0, algorithm_name="default",
                                base_radius=24.0, max_factor=16) -> int:
    """_nan_utils.py:555-601 (block_shape does not enter)."""
    r = max(1.0, float(radius))
    px = max(1e-3, float(pixel_size) if pixel_size else 1.0)
    algo_factor = {
        "topousm_fast": 1.15, "hillshade": 1.0, "slope": 1.0, "specular": 1.4,
        "atmospheric_scattering": 1.05, "curvature": 1.1, "ambient_occlusion": 1.5,
        "openness": 1.4, "multi_light_uncertainty": 1.25,
    }.get(str(algorithm_name), 1.0)
    resolution_factor = max(1.0, 1.0 / px)
    score = (r / max(1.0, base_radius)) * algo_factor * 1.0 * (resolution_factor ** 0.35)
    if score <= 1.0:
        return 1
    factor = 2 ** int(np.floor(np.log2(score)))
    return int(max(1, min(factor, max_factor)))


def downsample_nan_aware(block, factor: int) -> np.ndarray:
    """_nan_utils.py:604-668.  The cell sums are formed exactly (float64) and
    rounded once to float32; CuPy's float32 reduction order is unpinned."""
    block = np.asarray(block, dtype=F32)
    if factor <= 1:
        return block
    f = int(factor)
    h, w = block.shape
    out_h = max(1, (h + f - 1) // f)
    out_w = max(1, (w + f - 1) // f)
    work = block
    pad_h, pad_w = out_h * f - h, out_w * f - w
    if pad_h or pad_w:
        work = np.pad(work, ((0, pad_h), (0, pad_w)), mode="constant", constant_values=np.nan)
    cells = work.reshape(out_h, f, out_w, f)
    finite = np.isfinite(cells)
    cnt = finite.sum(axis=(1, 3)).astype(F32)
    total = np.where(finite, cells, F32(0)).astype(np.float64).sum(axis=(1, 3)).astype(F32)
    with _err():
        coarse = np.where(cnt > 0, total / np.maximum(cnt, F32(1)), F32(np.nan)).astype(F32)
    cnan = np.isnan(coarse)
    if cnan.any():
        cvalid = (~cnan).astype(F32)
        sigma = max(1.0, float(min(coarse.shape[:2])) / 64.0)
        sv = ndi.gaussian_filter(np.where(cnan, F32(0), coarse).astype(F32), sigma=sigma, mode="nearest")
        sw = ndi.gaussian_filter(cvalid, sigma=sigma, mode="nearest")
        enclosed = cnan & (sw > F32(0.5))
        with _err():
            smooth = sv / np.maximum(sw, F32(1e-6))
        coarse = np.where(enclosed, smooth, coarse).astype(F32)
    return coarse.astype(F32)


def upsample_to_shape(block, target_shape) -> np.ndarray:
    """_nan_utils.py:671-698 (zoom order=1, mode='nearest', corner-aligned)."""
    block = np.asarray(block, dtype=F32)
    th, tw = int(target_shape[0]), int(target_shape[1])
    h, w = block.shape
    if h == th and w == tw:
        return block
    sy, sx = th / max(1, h), tw / max(1, w)
    nan_mask = np.isnan(block)
    if not nan_mask.any():
        out = ndi.zoom(block, zoom=(sy, sx), order=1, mode="nearest").astype(F32)
    else:
        valid = (~nan_mask).astype(F32)
        filled = np.where(nan_mask, F32(0), block).astype(F32)
        num = ndi.zoom(filled, zoom=(sy, sx), order=1, mode="nearest")
        den = ndi.zoom(valid, zoom=(sy, sx), order=1, mode="nearest")
        with _err():
            out = np.where(den > F32(1e-3), num / np.maximum(den, F32(1e-6)), F32(np.nan)).astype(F32)
    pad_h, pad_w = max(0, th - out.shape[0]), max(0, tw - out.shape[1])
    if pad_h or pad_w:
        out = np.pad(out, ((0, pad_h), (0, pad_w)), mode="edge")
    return out[:th, :tw]


def bilinear_sample_coarse(coarse, r0, r1, c0, c1, full_h, full_w) -> np.ndarray:
    """_nan_utils.py:255-281 (float32 coordinates, map_coordinates order=1, nearest)."""
    coarse = np.asarray(coarse, dtype=F32)
    ch, cw = coarse.shape
    h, w = int(r1 - r0), int(c1 - c0)
    rr = (np.arange(r0, r1, dtype=F32) + F32(0.5)) * F32(ch / float(full_h)) - F32(0.5)
    cc = (np.arange(c0, c1, dtype=F32) + F32(0.5)) * F32(cw / float(full_w)) - F32(0.5)
    coords = np.empty((2, h, w), dtype=F32)
    coords[0] = rr[:, None]
    coords[1] = cc[None, :]
    return ndi.map_coordinates(coarse, coords, order=1, mode="nearest").astype(F32)


# ---------------------------------------------------------------------------
# TopoUSM Fast (A5-A7)
# ---------------------------------------------------------------------------

def topousm_fast_block(block, *, radii=None, weights=None, pixel_size=1.0) -> np.ndarray:
    """compute_topousm_fast_efficient_block (_impl_topousm_fast.py:49-100)."""
    block = np.asarray(block, dtype=F32)
    if radii is None:
        radii = [4, 16, 64]
    nan_mask = np.isnan(block)
    if weights is None:
        weights = np.array([1.0 / len(radii)] * len(radii), dtype=F32)
    else:
        weights = np.asarray(weights, dtype=F32)
        if len(weights) != len(radii):
            raise ValueError(
                f"Length of weights ({len(weights)}) must match length of radii ({len(radii)})")
    combined = None
    for radius, weight in zip(radii, weights):
        ds = radius_to_downsample_factor(float(radius), pixel_size=pixel_size,
                                         algorithm_name="topousm_fast")
        if ds > 1:
            small = downsample_nan_aware(block, ds)
            r_small = max(1, int(round(float(radius) / ds)))
            if r_small <= 1:
                mean_small, _ = handle_nan_with_gaussian(small, sigma=1.0, mode="nearest")
            else:
                mean_small, _ = handle_nan_with_uniform(small, size=2 * r_small + 1, mode="reflect")
            mean = upsample_to_shape(mean_small, block.shape)
        elif radius <= 1:
            mean, _ = handle_nan_with_gaussian(block, sigma=1.0, mode="nearest")
        else:
            mean, _ = handle_nan_with_uniform(block, size=2 * int(radius) + 1, mode="reflect")
        diff = weight * (block - mean)
        combined = diff if combined is None else combined + diff
    return restore_nan(combined.astype(F32), nan_mask)


def topousm_large_coarse_field(coarse_dem, *, large_radii, large_weights, decimation) -> np.ndarray:
    """compute_topousm_fast_large_coarse_field (_impl_topousm_fast.py:133-155)."""
    coarse_dem = np.asarray(coarse_dem, dtype=F32)
    field = None
    for r, w in zip(large_radii, large_weights):
        rc = max(1, int(round(float(r) / max(decimation, 1.0))))
        if rc <= 1:
            mean_c, _ = handle_nan_with_gaussian(coarse_dem, sigma=1.0, mode="nearest")
        else:
            mean_c, _ = handle_nan_with_uniform(coarse_dem, size=2 * rc + 1, mode="reflect")
        term = F32(w) * mean_c
        field = term if field is None else field + term
    return field.astype(F32)


def _area_weights(n_src: int, n_dst: int):
    """n_dst x n_src sparse float64 matrix: row i holds each source pixel's
    overlap length with [i*n_src/n_dst, (i+1)*n_src/n_dst)."""
    from scipy import sparse
    ratio = n_src / float(n_dst)
    rows, cols, vals = [], [], []
    for i in range(n_dst):
        lo, hi = i * ratio, min((i + 1) * ratio, float(n_src))
        for k in range(int(math.floor(lo)), min(n_src, int(math.ceil(hi)))):
            w = min(hi, k + 1.0) - max(lo, float(k))
            if w > 0:
                rows.append(i)
                cols.append(k)
                vals.append(w)
    return sparse.csr_matrix((vals, (rows, cols)), shape=(n_dst, n_src), dtype=np.float64)


def overview_average(raster: np.ndarray, oh: int, ow: int) -> np.ndarray:
    """The overview read of core/tile_processor.py:575-590 and
    read_overview_coarse_dem (_nan_utils.py:836-871):
    src.read(out_shape=(oh, ow), resampling=Resampling.average, masked=True),
    restated as the area-weighted average of the finite full-resolution
    pixels (each pixel weighs its overlap with the output cell on each axis),
    summed and divided in float64 and rounded once to float32; NaN where a
    cell holds no finite pixel.  GDAL's own arithmetic (and its choice of an
    internal overview level to resample from) is not restated."""
    a = np.asarray(raster, dtype=F32)
    fin = np.isfinite(a)
    v = np.where(fin, a, F32(0)).astype(np.float64)
    wy = _area_weights(a.shape[0], int(oh))
    wx = _area_weights(a.shape[1], int(ow))
    num = np.asarray((wx @ np.asarray(wy @ v).T).T)
    den = np.asarray((wx @ np.asarray(wy @ fin.astype(np.float64)).T).T)
    with _err():
        out = np.where(den > 0, num / np.maximum(den, 1e-300), np.nan)
    return out.astype(F32)


def overview(raster: np.ndarray, sample_max: int = 2048):
    """Overview shape and scale of core/tile_processor.py:575-580
    (scale = max(W, H)/sample_max, at least 1; shape round(H/scale) x
    round(W/scale)) with overview_average.  Returns (coarse, scale)."""
    H, W = raster.shape
    scale = max(W / sample_max, H / sample_max, 1.0)
    sw_, sh_ = max(1, int(round(W / scale))), max(1, int(round(H / scale)))
    if sh_ == H and sw_ == W:
        return np.asarray(raster, dtype=F32).copy(), scale
    return overview_average(raster, sh_, sw_), scale


def overview_block_mean(raster: np.ndarray, sample_max: int = 2048):
    """Overview read of _compute_topousm_fast_overview_coarse_field_tile
    (core/tile_processor.py:575-590) for an integer decimation: GDAL
    Resampling.average over whole cells, NaN excluded.  Returns (coarse, scale)."""
    H, W = raster.shape
    scale = max(W / sample_max, H / sample_max, 1.0)
    sw_, sh_ = max(1, int(round(W / scale))), max(1, int(round(H / scale)))
    f = int(round(scale))
    if abs(scale - f) > 1e-12 or sh_ * f != H or sw_ * f != W:
        raise NotImplementedError("non-integer overview decimation is unpinned")
    if f == 1:
        return raster.astype(F32).copy(), scale
    cells = raster.astype(np.float64).reshape(sh_, f, sw_, f)
    fin = np.isfinite(cells)
    tot = np.where(fin, cells, 0.0).sum(axis=(1, 3))
    cnt = fin.sum(axis=(1, 3))
    with _err():
        out = np.where(cnt > 0, tot / np.maximum(cnt, 1), np.nan).astype(F32)
    return out, scale


def topousm_stat(data) -> Tuple[float]:
    """topousm_fast_stat_func (_normalization.py:22-32), percentile in FP64."""
    data = np.asarray(data, dtype=F32).ravel()
    v = data[~np.isnan(data)]
    if len(v) > 0:
        scale = float(np.percentile(np.abs(v).astype(np.float64), 99.0))
        if scale > 1e-9:
            return (scale,)
        sd = float(np.std(v.astype(np.float64)))
        return (sd if sd > 1e-9 else 1.0,)
    return (1.0,)


def normalize_div(block, scale) -> np.ndarray:
    """apply_global_normalization + topousm_fast_norm_func
    (_global_stats.py:123-153, _normalization.py:35-41)."""
    block = np.asarray(block, dtype=F32)
    nan_mask = np.isnan(block)
    if scale > 0:
        out = block / F32(scale)
    else:
        out = np.zeros_like(block)
    return restore_nan(out.astype(F32), nan_mask).astype(F32)


def topousm_direct(block, *, radii, weights, pixel_size=1.0, coarse_field=None,
                   small_radii=None, small_weights=None, w_large=0.0,
                   field_offset=(0, 0), full_shape=None, stats=None) -> np.ndarray:
    """_direct_topousm_fast (tile/dask_bridge.py:225-271)."""
    block = np.asarray(block, dtype=F32)
    if coarse_field is not None:
        off_r, off_c = field_offset
        full_h, full_w = full_shape if full_shape is not None else block.shape
        large = F32(w_large) * block - bilinear_sample_coarse(
            coarse_field, int(off_r), int(off_r) + block.shape[0],
            int(off_c), int(off_c) + block.shape[1], int(full_h), int(full_w))
        if small_radii:
            t = topousm_fast_block(block, radii=small_radii, weights=small_weights,
                                   pixel_size=pixel_size) + large
        else:
            t = large.astype(F32)
    else:
        t = topousm_fast_block(block, radii=radii, weights=weights, pixel_size=pixel_size)
    if not (isinstance(stats, (tuple, list)) and len(stats) >= 1 and float(stats[0]) > 1e-9):
        stats = topousm_stat(t)
    return normalize_div(t, float(stats[0]))


# ---------------------------------------------------------------------------
# Openness (A8) and ambient occlusion (A9)
# ---------------------------------------------------------------------------

def openness_distances(max_distance: int) -> np.ndarray:
    """_impl_openness.py:67-68."""
    d = np.unique((np.linspace(0.1, 1.0, 10) * max_distance).astype(int))
    return d[d > 0]


def _abs_steps(pixel_size, pixel_scale_x, pixel_scale_y):
    sx = abs(float(pixel_scale_x)) if pixel_scale_x is not None else float(pixel_size)
    sy = abs(float(pixel_scale_y)) if pixel_scale_y is not None else float(pixel_size)
    if sx < 1e-9:
        sx = float(pixel_size) if pixel_size else 1.0
    if sy < 1e-9:
        sy = float(pixel_size) if pixel_size else 1.0
    return sx, sy


def openness_ray_table(num_directions=16, max_distance=50, pixel_size=1.0,
                       pixel_scale_x=None, pixel_scale_y=None):
    """Per (direction, distance) offsets and physical distances
    (_impl_openness.py:58-109).  Returns list of lists of (ox, oy, phys)|None."""
    angles = np.linspace(0, 2 * np.pi, num_directions, endpoint=False)
    directions = np.stack([np.cos(angles), np.sin(angles)], axis=1)
    distances = openness_distances(max_distance)
    sx, sy = _abs_steps(pixel_size, pixel_scale_x, pixel_scale_y)
    table = []
    for d in range(num_directions):
        row = []
        for r in distances:
            ox = int(round(float(r) * float(directions[d][0])))
            oy = int(round(float(r) * float(directions[d][1])))
            if ox == 0 and oy == 0:
                row.append(None)
                continue
            phys = max(float(np.hypot(float(ox) * sx, float(oy) * sy)), 1e-9)
            row.append((ox, oy, phys))
        table.append(row)
    return table, distances


def openness(block, *, openness_type="positive", num_directions=16, max_distance=50,
             pixel_size=1.0, pixel_scale_x=None, pixel_scale_y=None,
             raw_degrees=False) -> np.ndarray:
    """compute_openness_vectorized (_impl_openness.py:31-132).  With
    raw_degrees, the mean angle op = angle_sum / max(dir_count, 1) of
    _impl_openness.py:126 in degrees, before its / (pi/2), clip and gamma."""
    block = np.asarray(block, dtype=F32)
    h, w = block.shape
    nan_mask = np.isnan(block)
    init_val = F32(-np.pi / 2 if openness_type == "positive" else np.pi / 2)
    table, distances = openness_ray_table(num_directions, max_distance, pixel_size,
                                          pixel_scale_x, pixel_scale_y)
    D = int(max(distances)) if distances.size else 0
    if D > 0:
        padded_all = np.pad(np.where(nan_mask, F32(0.0), block), D, mode="edge")
        padded_valid = np.pad(~nan_mask, D, mode="constant", constant_values=False)
    angle_sum = np.zeros((h, w), dtype=F32)
    dir_count = np.zeros((h, w), dtype=F32)
    for d in range(num_directions):
        dir_ext = np.full((h, w), init_val, dtype=F32)
        dir_valid = np.zeros((h, w), dtype=bool)
        for entry in table[d]:
            if entry is None:
                continue
            ox, oy, phys = entry
            shifted = padded_all[D + oy:D + oy + h, D + ox:D + ox + w]
            sv = padded_valid[D + oy:D + oy + h, D + ox:D + ox + w]
            with _err():
                ang = np.arctan((shifted - block) / F32(phys))
            valid = sv & ~nan_mask
            if openness_type == "positive":
                dir_ext = np.where(valid, np.maximum(dir_ext, ang), dir_ext)
            else:
                dir_ext = np.where(valid, np.minimum(dir_ext, ang), dir_ext)
            dir_valid |= valid
        dir_angle = (F32(np.pi / 2) - dir_ext if openness_type == "positive"
                     else F32(np.pi / 2) + dir_ext)
        angle_sum += np.where(dir_valid, dir_angle, F32(0.0))
        dir_count += dir_valid.astype(F32)
    op = angle_sum / np.maximum(dir_count, F32(1.0))
    if raw_degrees:
        deg = (op.astype(np.float64) * (180.0 / np.pi)).astype(F32)
        return restore_nan(deg, nan_mask).astype(F32)
    op = np.clip(op / F32(np.pi / 2), F32(0), F32(1))
    res = np.power(op, F32(GAMMA))
    res = restore_nan(res.astype(F32), nan_mask)
    return res.astype(F32)


def ao_sample_table(num_samples=16, radius=10.0, pixel_size=1.0,
                    pixel_scale_x=None, pixel_scale_y=None):
    """Ring sample table of compute_ambient_occlusion_block
    (_impl_ambient_occlusion.py:54-81).  Returns [(dx, dy, phys, df)|None]."""
    angles = np.linspace(0, 2 * np.pi, num_samples, endpoint=False)
    directions = np.stack([np.cos(angles), np.sin(angles)], axis=1)
    sx, sy = _abs_steps(pixel_size, pixel_scale_x, pixel_scale_y)
    out = []
    for r_factor in (0.25, 0.5, 0.75, 1.0):
        r = radius * r_factor
        dx_all = np.round(r * directions[:, 0]).astype(int)
        dy_all = np.round(r * directions[:, 1]).astype(int)
        for i in range(num_samples):
            dx, dy = int(dx_all[i]), int(dy_all[i])
            if dx == 0 and dy == 0:
                out.append(None)
                continue
            dist = max(float(np.hypot(float(dx) * sx, float(dy) * sy)), 1e-9)
            out.append((dx, dy, dist, 1.0 - (r_factor * 0.3)))
    return out


def ambient_occlusion(block, *, num_samples=16, radius=10.0, intensity=1.0,
                      pixel_size=1.0, pixel_scale_x=None, pixel_scale_y=None) -> np.ndarray:
    """compute_ambient_occlusion_block (_impl_ambient_occlusion.py:32-121)."""
    block = np.asarray(block, dtype=F32)
    h, w = block.shape
    nan_mask = np.isnan(block)
    D = max(1, int(round(float(radius))))
    padded_all = np.pad(block, D, mode="edge")
    total = np.zeros((h, w), dtype=F32)
    count = np.zeros((h, w), dtype=F32)
    max_angle = F32(np.pi / 4)
    for entry in ao_sample_table(num_samples, radius, pixel_size, pixel_scale_x, pixel_scale_y):
        if entry is None:
            continue
        dx, dy, dist, df = entry
        shifted = padded_all[D + dy:D + dy + h, D + dx:D + dx + w]
        with _err():
            ang = np.arctan((shifted - block) / F32(dist))
            occ = np.maximum(F32(0), ang) / max_angle
        occ = np.minimum(occ, F32(1.0))
        valid = ~(np.isnan(shifted) | nan_mask)
        total += np.where(valid, occ * F32(df), F32(0))
        count += np.where(valid, F32(1.0), F32(0))
    count = np.maximum(count, F32(1.0))
    mean = total / count
    ao = F32(1.0) - mean * F32(intensity)
    ao = np.clip(ao, F32(0), F32(1))
    if nan_mask.any():
        ao = ndi.gaussian_filter(np.where(nan_mask, F32(1.0), ao).astype(F32), sigma=1.0, mode="nearest")
    else:
        ao = ndi.gaussian_filter(ao, sigma=1.0, mode="nearest")
    res = np.power(ao, F32(GAMMA))
    res = restore_nan(res.astype(F32), nan_mask)
    return res.astype(F32)


# ---------------------------------------------------------------------------
# Statistics, stretch, encoding, NoData (A7, A10-A13)
# ---------------------------------------------------------------------------

def percentile64(values, q) -> float:
    """cp.percentile(method='linear') evaluated in FP64 on float32 values."""
    v = np.asarray(values, dtype=np.float64).ravel()
    return float(np.percentile(v, q))


def stretch_stat(values) -> Tuple[float, float]:
    """robust_unsigned_stretch_stat_func (_global_stats.py:181-203)."""
    if values is None:
        return (0.0, 0.0)
    v = np.asarray(values, dtype=F32).ravel()
    v = v[np.isfinite(v)]
    if v.size == 0:
        return (0.0, 0.0)
    lo = percentile64(v, 1.0)
    hi = percentile64(v, 99.0)
    scale = hi - lo
    if not (scale > 1e-12):
        return (lo, 0.0)
    return (lo, scale)


def display_stretch(result, stats) -> np.ndarray:
    """_apply_display_stretch_block (tile/dask_bridge.py:173-187)."""
    if not (isinstance(stats, (tuple, list)) and len(stats) >= 2):
        return result
    lo, scale = float(stats[0]), float(stats[1])
    if not (scale > 1e-12):
        return result
    with _err():
        return np.maximum((np.asarray(result, F32) - F32(lo)) / F32(scale), F32(0.0)).astype(F32)


def quantize(arr, a_coef, b_coef, dn_min, dn_max, dtype) -> np.ndarray:
    """quantize_array (io/output_encoding.py:178-190)."""
    a = np.asarray(arr, dtype=F32)
    with _err():
        dn = np.rint(F32(a_coef) * a + F32(b_coef))
    dn = np.clip(dn, dn_min, dn_max)
    dn = np.where(np.isfinite(a), dn, 0.0)
    return dn.astype(np.dtype(dtype))


def nodata_mask(data, nodata) -> Optional[np.ndarray]:
    """_build_nodata_mask (core/tile_processor.py:185-196)."""
    if nodata is None:
        return None
    if isinstance(nodata, float) and math.isnan(nodata):
        return np.isnan(data)
    return np.isclose(data, float(nodata), rtol=0.0, atol=1e-6) | np.isnan(data)


def format_output(result_core, algorithm) -> np.ndarray:
    """_format_algorithm_output (core/tile_processor.py:606-624)."""
    arr = np.asarray(result_core, dtype=F32)
    if algorithm == "hillshade":
        arr = np.clip(arr, 0.0, 1.0)
    return arr


# ---------------------------------------------------------------------------
# Window driver: the reference tile loop for one raster, in memory
# ---------------------------------------------------------------------------

def run_window(algorithm: str, window: np.ndarray, params: dict) -> np.ndarray:
    """Full-window compute of one tile (tile/dask_bridge.py:_process_direct)."""
    p = params
    if algorithm == "hillshade":
        return hillshade(window, azimuth=p.get("azimuth", DEFAULT_AZIMUTH),
                         altitude=p.get("altitude", DEFAULT_ALTITUDE),
                         z_factor=p.get("z_factor", 1.0),
                         pixel_size=p.get("pixel_size", 1.0),
                         pixel_scale_x=p.get("pixel_scale_x"), pixel_scale_y=p.get("pixel_scale_y"))
    if algorithm == "slope":
        return slope(window, unit=p.get("unit", "degree"), pixel_size=p.get("pixel_size", 1.0),
                     pixel_scale_x=p.get("pixel_scale_x"), pixel_scale_y=p.get("pixel_scale_y"))
    if algorithm == "curvature":
        return curvature(window, curvature_type=p.get("curvature_type", "mean"),
                         pixel_size=p.get("pixel_size", 1.0),
                         pixel_scale_x=p.get("pixel_scale_x"), pixel_scale_y=p.get("pixel_scale_y"))
    if algorithm == "openness":
        out = openness(window, openness_type=p.get("openness_type", "positive"),
                       num_directions=p.get("num_directions", 16),
                       max_distance=p.get("max_distance", 50),
                       pixel_size=p.get("pixel_size", 1.0),
                       pixel_scale_x=p.get("pixel_scale_x"), pixel_scale_y=p.get("pixel_scale_y"))
        return display_stretch(out, p.get("global_stats"))
    if algorithm == "ambient_occlusion":
        out = ambient_occlusion(window, num_samples=p.get("num_samples", 16),
                                radius=p.get("radius", 10.0), intensity=p.get("intensity", 1.0),
                                pixel_size=p.get("pixel_size", 1.0),
                                pixel_scale_x=p.get("pixel_scale_x"), pixel_scale_y=p.get("pixel_scale_y"))
        return display_stretch(out, p.get("global_stats"))
    if algorithm == "topousm_fast":
        return topousm_direct(
            window, radii=p.get("radii"), weights=p.get("weights"),
            pixel_size=p.get("pixel_size", 1.0),
            coarse_field=p.get("_topousm_fast_coarse_field"),
            small_radii=p.get("_topousm_fast_small_radii"),
            small_weights=p.get("_topousm_fast_small_weights"),
            w_large=p.get("_topousm_fast_w_large", 0.0),
            field_offset=p.get("_topousm_fast_field_offset", (0, 0)),
            full_shape=p.get("_topousm_fast_full_shape"),
            stats=p.get("global_stats"))
    raise ValueError(f"oracle has no algorithm {algorithm!r}")


# ---------------------------------------------------------------------------
# Spatial mode: multi-radius hillshade / slope / curvature (next row F1)
# ---------------------------------------------------------------------------
# The reference's tile backend sends a spatial-mode tile through Dask as one
# chunk (algorithms/tile/dask_bridge.py:395-405): gpu_da = the tile window.

SPATIAL_AUTO_RADII = (2, 8, 32, 128, 512, 2048)   # common/spatial_mode.py


def clean_normalized_weights(weights):
    """_clean_normalized_weights (_nan_utils.py:148-161)."""
    cleaned = []
    for w in weights:
        try:
            fv = float(w)
        except (TypeError, ValueError):
            return None
        cleaned.append(fv if np.isfinite(fv) and fv > 0 else 0.0)
    s = float(sum(cleaned))
    if s <= 0:
        return None
    return [v / s for v in cleaned]


def resolve_spatial_radii_weights(radii, weights, pixel_size=1.0):
    """_resolve_spatial_radii_weights with _normalize_spatial_radii
    (_nan_utils.py:77-129) for the explicit-radii case, the 2**n profile
    (common/spatial_mode.py:78-84) when the weights do not fit."""
    if radii is None:
        radii = list(SPATIAL_AUTO_RADII)
    out = []
    for r in radii:
        try:
            rv = int(round(float(r)))
        except (TypeError, ValueError):
            continue
        if rv > 0 and rv not in out:
            out.append(rv)
    if not out:
        out = list(SPATIAL_AUTO_RADII)
    if weights is not None and not isinstance(weights, (str, bytes)) and len(weights) == len(out):
        user = clean_normalized_weights(weights)
        if user is not None:
            return out, user
    raw = [2.0 ** (len(out) - 1 - i) for i in range(len(out))]
    tot = sum(raw)
    return out, [v / tot for v in raw]


def dask_reflect_pad(block, depth: int) -> np.ndarray:
    """dask.array.overlap.reflect on every side: [d c b a | a b c d | d c b a]."""
    return np.pad(np.asarray(block, dtype=F32), int(depth), mode="symmetric")


def smooth_for_radius(block, radius, *, pixel_size=1.0, algorithm_name="default"):
    """_smooth_for_radius (_nan_utils.py:527-552)."""
    block = np.asarray(block, dtype=F32)
    r = max(1.0, float(radius))
    if r <= 1.0:
        return block
    factor = radius_to_downsample_factor(r, pixel_size=pixel_size, algorithm_name=algorithm_name)
    if factor <= 1:
        smoothed, _ = handle_nan_with_gaussian(block, sigma=max(0.5, r / 2.0), mode="nearest")
        return smoothed
    reduced = downsample_nan_aware(block, factor)
    small, _ = handle_nan_with_gaussian(reduced, sigma=max(0.5, (r / factor) / 2.0), mode="nearest")
    return upsample_to_shape(small, block.shape)


def dask_edge_depth(algorithm: str, params: dict) -> int:
    """map_overlap depth of the Dask backend's local path (_impl_hillshade.py:133-134,
    _impl_slope.py:71-72, _impl_curvature.py:92-93, _impl_openness.py:204-207,
    _impl_ambient_occlusion.py:191-194)."""
    if algorithm in ("hillshade", "slope"):
        return 1
    if algorithm == "curvature":
        return 2
    if algorithm == "openness":
        return int(params.get("max_distance", 50)) + 1
    if algorithm == "ambient_occlusion":
        return int(float(params.get("radius", 10.0)) + 1)
    raise ValueError(algorithm)


def dask_single_chunk(algorithm: str, raster, params: dict) -> np.ndarray:
    """The Dask backend's local-mode result on a raster held as one chunk:
    map_overlap(block_fn, depth=D, boundary='reflect') pads the chunk with the
    symmetric rule (dask.array.overlap.reflect), runs the block function and
    trims D (core/dask_processor.py:1473-1488)."""
    d = dask_edge_depth(algorithm, params)
    res = run_window(algorithm, dask_reflect_pad(raster, d), params)
    return np.asarray(res, dtype=F32)[d:-d, d:-d]


RAY_SPATIAL = ("openness", "ambient_occlusion")


def spatial_depth(algorithm: str, r: float) -> int:
    """depth_for_scale of the spatial algorithms (_impl_hillshade.py:102,
    _impl_slope.py:64, _impl_curvature.py:84, _impl_openness.py:195,
    _impl_ambient_occlusion.py:180)."""
    if algorithm in RAY_SPATIAL:
        return int(r) + 1
    if algorithm == "curvature":
        return max(3, int(float(r) * 2 + 2))
    return max(2, int(float(r) * 2 + 1))


def ray_spatial_block(algorithm, block, radius, params):
    """compute_openness_spatial_block (_impl_openness.py:135-164) and
    compute_ambient_occlusion_spatial_block (_impl_ambient_occlusion.py:123-157):
    the local ray function of the block, or of its NaN-aware decimated copy
    (radius and pixel scales scaled by the factor) upsampled to the block."""
    block = np.asarray(block, dtype=F32)
    px = params.get("pixel_size", 1.0)
    psx, psy = params.get("pixel_scale_x"), params.get("pixel_scale_y")

    def local(b, r, px_, psx_, psy_):
        if algorithm == "openness":
            return openness(b, openness_type=params.get("openness_type", "positive"),
                            num_directions=params.get("num_directions", 16), max_distance=r,
                            pixel_size=px_, pixel_scale_x=psx_, pixel_scale_y=psy_)
        return ambient_occlusion(b, num_samples=params.get("num_samples", 16), radius=r,
                                 intensity=params.get("intensity", 1.0), pixel_size=px_,
                                 pixel_scale_x=psx_, pixel_scale_y=psy_)

    f = radius_to_downsample_factor(float(radius), pixel_size=px, algorithm_name=algorithm)
    if f <= 1:
        return local(block, radius, px, psx, psy)
    small = downsample_nan_aware(block, f)
    dpsx = float(abs(float(psx)) * f) if psx is not None else None
    dpsy = float(abs(float(psy)) * f) if psy is not None else None
    if algorithm == "openness":
        r_s = max(2, int(round(float(radius) / float(f))))
    else:
        r_s = max(1.0, float(radius) / float(f))
    res = local(small, r_s, float(px) * float(f), dpsx, dpsy)
    return upsample_to_shape(res, block.shape)


def spatial_block(algorithm, block, radius, params):
    """compute_{hillshade,slope}_spatial_block (_impl_hillshade.py:59-70,
    _impl_slope.py:38-45) and curvature's _curv_spatial (_impl_curvature.py:72-78):
    the local block function of the smoothed block; the ray algorithms'
    spatial blocks for openness and ambient occlusion."""
    if algorithm in RAY_SPATIAL:
        return ray_spatial_block(algorithm, block, radius, params)
    sm = smooth_for_radius(block, radius, pixel_size=params.get("pixel_size", 1.0),
                           algorithm_name=algorithm)
    return run_window(algorithm, sm, params)


def coarsen_factor_for_shape(shape, coarse_max=2048) -> int:
    """_nan_utils.py:231-236."""
    longest = max(int(shape[0]), int(shape[1]))
    if longest <= int(coarse_max):
        return 1
    return 1 << int(np.ceil(np.log2(longest / float(coarse_max))))


def spatial_window(algorithm, window, params, *, overview=None, decimation=None,
                   tile_origin=None, full_shape=None) -> np.ndarray:
    """{Hillshade,Slope,Curvature,Openness,AmbientOcclusion}Algorithm.process in
    spatial mode (before the display stretch) on a single-chunk Dask array (the
    tile window): multiscale_response_fields
    (_nan_utils.py:455-524) and the combine of _impl_hillshade.py:110-131 /
    _combine_multiscale_dask (_nan_utils.py:182-213)."""
    window = np.asarray(window, dtype=F32)
    h, w = window.shape
    px = params.get("pixel_size", 1.0)
    agg = str(params.get("agg", "mean") or "mean").lower()
    user_w = params.get("weights")
    radii, resolved_w = resolve_spatial_radii_weights(params.get("radii"), user_w, px)
    if algorithm == "hillshade":
        weights = resolved_w if user_w is None else user_w
    else:
        weights = resolved_w
    thr = int(max(256, min(h, w) // 16))                         # large_radius_threshold
    F = coarsen_factor_for_shape(window.shape)
    coarse_ok = overview is not None and tile_origin is not None
    chunk_cap = max(1, min(h, w) - 1)
    base = {k: v for k, v in params.items() if k not in ("radii", "weights", "mode", "agg")}
    if algorithm == "openness":                       # _impl_openness.py:193
        radii = [int(max(2, round(float(r)))) for r in radii]
    elif algorithm == "ambient_occlusion":            # _impl_ambient_occlusion.py:178
        radii = [max(1, int(round(float(r)))) for r in radii]
    fields = []
    for r in [float(v) for v in radii]:
        large = int(round(r)) > thr
        if large and (F > 1 or coarse_ok):
            if overview is None:
                raise NotImplementedError("window-coarsen branch without an overview")
            fac = float(decimation)
            kw = dict(base)
            kw["pixel_size"] = float(px) * fac
            if kw.get("pixel_scale_x") is not None:
                kw["pixel_scale_x"] = float(kw["pixel_scale_x"]) * fac
            if kw.get("pixel_scale_y") is not None:
                kw["pixel_scale_y"] = float(kw["pixel_scale_y"]) * fac
            resp = spatial_block(algorithm, overview, max(1, int(round(r / fac))), kw).astype(F32)
            r0, c0 = int(tile_origin[0]), int(tile_origin[1])
            fh, fw = (int(full_shape[0]), int(full_shape[1])) if full_shape else (h, w)
            up = bilinear_sample_coarse(resp, r0, r0 + h, c0, c0 + w, fh, fw)
            fields.append(np.where(np.isnan(window), F32(np.nan), up).astype(F32))
        else:
            d = max(1, min(spatial_depth(algorithm, r), chunk_cap))
            res = spatial_block(algorithm, dask_reflect_pad(window, d), r, base)
            fields.append(np.asarray(res, dtype=F32)[d:d + h, d:d + w])
    if agg == "stack":
        return np.stack(fields, axis=0)
    if algorithm != "hillshade" and len(fields) == 1:
        return fields[0]
    if agg == "max":
        return np.max(np.stack(fields), axis=0)
    if agg == "min":
        return np.min(np.stack(fields), axis=0)
    if agg == "sum" and algorithm != "hillshade":
        return np.sum(np.stack(fields), axis=0, dtype=F32)
    if algorithm == "hillshade":
        if (agg == "mean" and isinstance(weights, (list, tuple)) and len(weights) == len(fields)):
            wv = np.asarray(weights, dtype=F32)
            if np.isfinite(wv).all() and wv.sum() > 0:
                wv = wv / wv.sum()
                out = fields[0] * F32(wv[0])
                for i in range(1, len(fields)):
                    out = out + fields[i] * F32(wv[i])
                return out.astype(F32)
    elif weights is not None and len(weights) == len(fields):
        clean = clean_normalized_weights(weights)
        if clean is not None:
            out = fields[0] * F32(clean[0])
            for i in range(1, len(fields)):
                out = out + fields[i] * F32(clean[i])
            return out.astype(F32)
    return (np.sum(np.stack(fields), axis=0, dtype=F32) / F32(len(fields))).astype(F32)

"""Parameter and geometry rules of the tile pipeline (host logic, no GPU).

Each function states the reference rule it reproduces.  These are integer /
list rules; they decide which windows, radii, weights and halos the HIP path
runs with, so they are restated exactly.
"""
from __future__ import annotations

import math
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

# ----------------------------------------------------------------- constants
DEFAULT_GAMMA = 1 / 2.2          # algorithms/_base.py:13
DEFAULT_AZIMUTH = 315            # algorithms/_base.py:14
DEFAULT_ALTITUDE = 45            # algorithms/_base.py:15
MAX_DEPTH = 150                  # algorithms/_base.py:16

# algorithms/common/spatial_mode.py:21-57
AUTO_RADII_SEQUENCE: Tuple[int, ...] = (2, 8, 32, 128, 512, 2048)
AUTO_RADIUS_MAX = 2048
RADII_DRIVEN_ALGOS = frozenset({
    "topousm_fast", "hillshade", "slope", "specular", "atmospheric_scattering",
    "curvature", "ambient_occlusion", "openness", "multi_light_uncertainty",
    "npr_edges", "structure_tensor", "frangi",
})
MULTISCALE_REQUIRED_ALGOS = frozenset({
    "fractal_anomaly", "scale_space_surprise", "visual_saliency", "scale_drift",
    "phase_congruency",
})
LOCAL_RADII = [1]
LOCAL_WEIGHTS = [1.0]

# core/tile_processor.py:89-93
SPATIAL_TILE_ALGORITHMS = frozenset({
    "hillshade", "slope", "specular", "atmospheric_scattering", "curvature",
    "ambient_occlusion", "openness", "multi_light_uncertainty", "npr_edges",
    "structure_tensor", "frangi",
})


# ------------------------------------------------------------ spatial radii

def auto_spatial_radii(short_side_px: Optional[float]) -> List[int]:
    """algorithms/common/spatial_mode.py:60-75."""
    limit = float(AUTO_RADIUS_MAX) if short_side_px is None else min(
        float(AUTO_RADIUS_MAX), float(short_side_px) / 10.0)
    radii = [r for r in AUTO_RADII_SEQUENCE if float(r) <= limit]
    return list(radii) if radii else [AUTO_RADII_SEQUENCE[0]]


def auto_spatial_weights(n: int) -> List[float]:
    """algorithms/common/spatial_mode.py:78-84."""
    if n <= 0:
        return []
    raw = [2.0 ** (n - 1 - i) for i in range(n)]
    total = sum(raw)
    return [w / total for w in raw]


def auto_spatial_profile(short_side_px, radii=None):
    """algorithms/common/spatial_mode.py:87-101."""
    if radii is None:
        radii = auto_spatial_radii(short_side_px)
    else:
        radii = [int(round(float(r))) for r in radii]
    return radii, auto_spatial_weights(len(radii))


def deduplicate_radii_weights(radii: Iterable[int], weights):
    """core/tile_compute.py:9-39."""
    radii_list = list(radii)
    order: List[int] = []
    positions: Dict[int, int] = {}
    for r in radii_list:
        if r not in positions:
            positions[r] = len(order)
            order.append(r)
    if weights is None:
        return order, None
    weights_list = list(weights)
    if len(weights_list) != len(radii_list):
        return order, None
    agg = [0.0] * len(order)
    for r, v in zip(radii_list, weights_list):
        try:
            w = float(v)
        except (TypeError, ValueError):
            w = 0.0
        if not math.isfinite(w) or w <= 0:
            w = 0.0
        agg[positions[r]] += w
    total = sum(agg)
    if total <= 0:
        return order, None
    return order, [w / total for w in agg]


def normalize_topousm_radii_weights(target_distances, weights, pixel_size: float,
                                    manual_radii=None, manual_weights=None):
    """_normalize_topousm_fast_radii_and_weights (core/tile_compute.py:42-90)."""
    if pixel_size <= 0:
        pixel_size = 1.0
    radii_raw = manual_radii if manual_radii is not None else target_distances
    if radii_raw is None:
        return None, None
    resolved: List[int] = []
    for value in radii_raw:
        try:
            numeric = float(value)
        except (TypeError, ValueError):
            continue
        if not math.isfinite(numeric):
            continue
        radius = int(round(numeric / pixel_size)) if manual_radii is None else int(round(numeric))
        if radius < 1:
            radius = 1
        resolved.append(radius)
    if not resolved:
        return None, None
    weights_raw = manual_weights if manual_weights is not None else weights
    return deduplicate_radii_weights(resolved, weights_raw)


def sanitize_spatial_radii_weights_for_tile(algorithm, radii, weights, tile_size):
    """_sanitize_spatial_radii_weights_for_tile (core/tile_processor.py:102-165)."""
    if not isinstance(radii, (list, tuple)) or len(radii) == 0:
        return None, weights, None
    if algorithm not in {"ambient_occlusion", "openness"}:
        out = []
        for v in radii:
            try:
                fv = float(v)
            except (TypeError, ValueError):
                continue
            if fv <= 0:
                continue
            out.append(max(1, int(round(fv))))
        if not out:
            return None, weights, None
        r, w = deduplicate_radii_weights(out, weights)
        return r, w, None
    out_r, out_w = [], []
    for i, rv in enumerate(radii):
        try:
            r = int(round(float(rv)))
        except (TypeError, ValueError):
            continue
        if r <= 0:
            continue
        out_r.append(max(r, 2))
        if isinstance(weights, (list, tuple)) and i < len(weights):
            try:
                out_w.append(float(weights[i]))
            except (TypeError, ValueError):
                out_w.append(0.0)
    if not out_r:
        return None, None, None
    r, w = deduplicate_radii_weights(out_r, out_w if len(out_w) == len(out_r) else None)
    return r, w, None


# --------------------------------------------------------- TopoUSM helpers

def radius_to_downsample_factor(radius, *, pixel_size=1.0, algorithm_name="default",
                                base_radius=24.0, max_factor=16) -> int:
    """_radius_to_downsample_factor (algorithms/_nan_utils.py:555-601)."""
    r = max(1.0, float(radius))
    px = max(1e-3, float(pixel_size) if pixel_size else 1.0)
    algo_factor = {
        "topousm_fast": 1.15, "hillshade": 1.0, "slope": 1.0, "specular": 1.4,
        "atmospheric_scattering": 1.05, "curvature": 1.1, "ambient_occlusion": 1.5,
        "openness": 1.4, "multi_light_uncertainty": 1.25,
    }.get(str(algorithm_name), 1.0)
    score = (r / max(1.0, base_radius)) * algo_factor * 1.0 * (max(1.0, 1.0 / px) ** 0.35)
    if score <= 1.0:
        return 1
    factor = 2 ** int(math.floor(math.log2(score)))
    return int(max(1, min(factor, max_factor)))


def topousm_large_radius_threshold(min_chunk: int) -> int:
    """algorithms/_impl_topousm_fast.py:103-110."""
    return int(max(256, int(min_chunk) // 16))


def split_radii_by_threshold(radii, weights, threshold):
    """algorithms/_impl_topousm_fast.py:113-130."""
    n = len(radii)
    if weights is None or len(weights) != n:
        weights = [1.0 / n] * n
    sr, sw, lr, lw = [], [], [], []
    for r, w in zip(radii, weights):
        if int(r) > int(threshold):
            lr.append(int(r))
            lw.append(float(w))
        else:
            sr.append(int(r))
            sw.append(float(w))
    return sr, sw, lr, lw


def topousm_default_radii(pixel_size: float) -> List[int]:
    """TopoUSMFastAlgorithm._determine_optimal_radii (_impl_topousm_fast.py:335-347)."""
    radii = []
    for dist in (5, 20, 80, 320):
        radius = int(dist / pixel_size)
        radii.append(max(2, min(radius, 256)))
    return sorted(set(radii))


def overview_geometry(width: int, height: int, sample_max: int = 2048):
    """Overview shape/scale of _compute_topousm_fast_overview_coarse_field_tile
    (core/tile_processor.py:575-580): (scale, sample_h, sample_w)."""
    scale = max(width / sample_max, height / sample_max, 1.0)
    return scale, max(1, int(round(height / scale))), max(1, int(round(width / scale)))


# -------------------------------------------------------------- padding

def required_padding(algorithm: str, algo_params: dict, sigma: float, pixel_size: float,
                     target_distances=None, tile_size: int = 1024) -> int:
    """_required_padding_for_algorithm (core/tile_processor.py:207-383), for the
    algorithms on this path (hillshade, slope, curvature, topousm_fast,
    openness, ambient_occlusion)."""
    try:
        base = int(math.ceil(max(float(sigma), 0.0) * 5.0))
    except Exception:
        base = 32
    required = max(32, base)
    mode = str(algo_params.get("mode", "spatial")).lower()
    overview_active = (mode == "spatial" and bool(algo_params.get("radii"))
                       and algorithm != "topousm_fast")
    thr_switch = max(256, int(tile_size) // 16)
    if algorithm == "topousm_fast":
        radii, _ = normalize_topousm_radii_weights(
            target_distances, algo_params.get("weights"), pixel_size,
            manual_radii=algo_params.get("radii"), manual_weights=algo_params.get("weights"))
        if radii:
            required = max(required, int(max(radii) + 16))
    if mode == "spatial" and algorithm in SPATIAL_TILE_ALGORITHMS:
        radii = algo_params.get("radii") or auto_spatial_radii(None)
        try:
            radii_f = [float(r) for r in radii if float(r) > 0]
            if overview_active:
                small = [r for r in radii_f if int(round(r)) <= thr_switch] or [min(radii_f)]
                max_radius = int(round(max(small)))
            else:
                max_radius = max(int(round(r)) for r in radii_f)
        except Exception:
            max_radius = 32
        if algorithm in {"ambient_occlusion", "openness"}:
            required = max(required, int(max_radius + 16))
        else:
            required = max(required, int(max_radius * 2 + 2))
    return max(32, ((required + 31) // 32) * 32)


# ----------------------------------------------------------------- tiling

def tile_grid(width: int, height: int, tile_size: int, padding: int):
    """Tile list of process_dem_tiles (core/tile_processor.py:1591-1627):
    (ty, tx, core_x, core_y, core_w, core_h, win_x, win_y, win_w, win_h)."""
    if tile_size <= 0:
        raise ValueError(f"tile_size must be a positive integer, got {tile_size!r}")
    if padding < 0:
        raise ValueError(f"padding must be >= 0, got {padding!r}")
    nx, ny = math.ceil(width / tile_size), math.ceil(height / tile_size)
    out = []
    for ty in range(ny):
        for tx in range(nx):
            cx, cy = tx * tile_size, ty * tile_size
            cw, ch = min(tile_size, width - cx), min(tile_size, height - cy)
            wx, wy = max(cx - padding, 0), max(cy - padding, 0)
            wxe, wye = min(cx + cw + padding, width), min(cy + ch + padding, height)
            out.append((ty, tx, cx, cy, cw, ch, wx, wy, wxe - wx, wye - wy))
    return out


# ------------------------------------------------------- norm-stat pre-pass

def stratified_windows(width, height, by0, by1, bx0, bx1, *, grid=3, tile=4096):
    """algorithms/_norm_stats.py:64-100 -> [(wy0, wx0, win_w, win_h)]."""
    cell_h = max(1, (int(by1) - int(by0)) // int(grid))
    cell_w = max(1, (int(bx1) - int(bx0)) // int(grid))
    out, seen = [], set()
    for gy in range(int(grid)):
        for gx in range(int(grid)):
            ccy = int(by0) + gy * cell_h + cell_h // 2
            ccx = int(bx0) + gx * cell_w + cell_w // 2
            wy0 = int(min(max(0, ccy - tile // 2), max(0, int(height) - tile)))
            wx0 = int(min(max(0, ccx - tile // 2), max(0, int(width) - tile)))
            tw, th = min(tile, int(width) - wx0), min(tile, int(height) - wy0)
            key = (wy0, wx0, tw, th)
            if key in seen:
                continue
            seen.add(key)
            out.append(key)
    return out


def norm_stat_max_scale(merged: dict) -> float:
    """algorithms/_norm_stats.py:103-121 (keys used on this path)."""
    vals = []
    for key in ("radii", "scales"):
        v = merged.get(key)
        if isinstance(v, (list, tuple)) and v:
            vals.append(max(float(x) for x in v))
    for key in ("kernel_size", "radius"):
        value = merged.get(key)
        if value:
            vals.append(float(value))
    if merged.get("iterations"):
        vals.append(float(merged.get("iterations")))
    if merged.get("max_distance"):
        vals.append(float(merged.get("max_distance")))
    return max(vals) if vals else 16.0


def norm_stat_window_geometry(algorithm: str, merged: dict, max_tile: int = 4096):
    """_norm_stat_halo_pixels + _norm_stat_window_geometry
    (algorithms/_norm_stats.py:124-162) -> (margin, tile)."""
    margin = max(1, int(norm_stat_max_scale(merged) + 16))
    floor_ = min(2048, max(1, int(max_tile)))
    return margin, max(floor_, 4 * margin)


NORM_STAT_ALGOS = frozenset({"topousm_fast", "openness", "ambient_occlusion"})


def coarse_radius(radius: int, ds: int) -> int:
    """Radius on the ds-grid (_impl_topousm_fast.py:72-91): max(1, round(r / ds))."""
    return max(1, int(round(float(radius) / float(ds))))


def _axis_shared_ok(o: int, n: int, full: int, rs: int, rl: int, ds: int, margin: int) -> bool:
    if o % ds or n % ds or n < 1:
        return False
    nc = n // ds
    z = (nc - 1) / (n - 1) if n != 1 else 1.0
    lo = min(max(rs * z, 0.0), nc - 1.0)
    hi = min(max((rs + rl - 1) * z, 0.0), nc - 1.0)
    i_lo = int(math.floor(lo))
    i_hi = min(int(math.floor(hi)) + 1, nc - 1)
    if o == 0 and o + n == full:
        return True                      # the window spans the raster on this axis
    lo_ok = i_lo - margin >= 0 or (o == 0 and margin < nc)
    hi_ok = i_hi + margin <= nc - 1 or (o + n == full and margin < nc)
    return lo_ok and hi_ok


def shared_coarse_ok(win_origin, win_shape, region, ds: int, r_coarse: int, raster_shape) -> bool:
    """Whether a raster-wide coarse mean grid (ds, r_coarse) holds exactly the
    values a window would build for itself over the cells its region samples.

    True when the window origin and shape are multiples of ds (its ds-grid is a
    sub-grid of the raster's, cell for cell) and the box (or the radius-4
    Gaussian for r_coarse <= 1) around every sampled cell stays inside the
    window's grid, except across a window edge that is also a raster edge,
    where both grids reflect alike.  The raster must hold no NaN (no
    enclosed-void fill on either grid); the caller checks that.
    region = (r0, c0, h, w) in window coordinates.
    """
    margin = r_coarse if r_coarse >= 2 else 4
    r0, c0, rh, rw = region
    return (_axis_shared_ok(int(win_origin[0]), int(win_shape[0]), int(raster_shape[0]), r0, rh,
                            ds, margin)
            and _axis_shared_ok(int(win_origin[1]), int(win_shape[1]), int(raster_shape[1]), c0,
                                rw, ds, margin))

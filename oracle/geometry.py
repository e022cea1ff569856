"""Tile and pre-pass window geometry of the reference -- TEST INFRASTRUCTURE ONLY.

The integer rules that decide which pixels each tile window and each
statistics pre-pass window sees, restated from the reference for the checker:

* ``required_padding``  -- _required_padding_for_algorithm
  (core/tile_processor.py:207-383), for hillshade, slope, curvature,
  topousm_fast, openness and ambient_occlusion;
* ``tile_grid``         -- the tile loop of process_dem_tiles
  (core/tile_processor.py:1591-1627);
* ``valid_bbox``        -- the coarse-overview bounding box of
  _compute_norm_stats_tiled (algorithms/_norm_stats.py:254-264);
* ``stratified_windows`` and ``prepass_geometry`` -- algorithms/_norm_stats.py
  :64-162 and the margin trim of :271-276.

Pinned in tests/test_oracle_geometry.py by window lists written out by hand
for the 16384^2 TopoUSM raster (16 tiles, 9 pre-pass windows), the 32768^2
local curvature raster, a ragged 20000 x 13001 raster and the spatial-mode
halos.  Only tests/, the smoke check and bench.py's CPU leg use this module.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

# SPATIAL_TILE_ALGORITHMS (core/tile_processor.py:89-93), the members on the path
_SPATIAL_TILE = ("hillshade", "slope", "curvature", "openness", "ambient_occlusion")
_RAY = ("openness", "ambient_occlusion")
_AUTO_RADII = (2, 8, 32, 128, 512, 2048)       # algorithms/common/spatial_mode.py:21


def _round32(n: int) -> int:
    """Up to the next multiple of 32 (core/tile_processor.py:382)."""
    return -(-int(n) // 32) * 32


def _topousm_radii(radii, pixel_size: float) -> List[int]:
    """The radii _normalize_topousm_fast_radii_and_weights resolves from
    user radii (core/tile_compute.py:42-90): each rounded, at least 1, first
    occurrence kept."""
    out: List[int] = []
    for v in radii or ():
        try:
            x = float(v)
        except (TypeError, ValueError):
            continue
        if not math.isfinite(x):
            continue
        r = max(1, int(round(x)))
        if r not in out:
            out.append(r)
    return out


def required_padding(algorithm: str, params: dict, *, sigma: float = 10.0,
                     pixel_size: float = 1.0, tile_size: int = 1024) -> int:
    """The tile halo (core/tile_processor.py:207-383) for the six algorithms
    on the path: max(32, ceil(5 sigma), the algorithm's own need), rounded up
    to a multiple of 32.  TopoUSM needs its largest radius + 16.  A spatial
    run of the stencils needs 2r + 2, of openness / AO r + 16, where r is the
    largest radius at most max(256, tile // 16) when the overview serves the
    larger ones (spatial mode with radii given), else the largest radius."""
    need = [32, int(math.ceil(max(float(sigma), 0.0) * 5.0))]
    mode = str(params.get("mode", "spatial")).lower()
    radii = params.get("radii")
    if algorithm == "topousm_fast":
        rs = _topousm_radii(radii, pixel_size)
        if rs:
            need.append(max(rs) + 16)
    if mode == "spatial" and algorithm in _SPATIAL_TILE:
        rs = [float(r) for r in (radii or _AUTO_RADII) if float(r) > 0]
        if radii:   # the overview serves the radii above the switch
            switch = max(256, int(tile_size) // 16)
            small = [r for r in rs if int(round(r)) <= switch]
            r_max = int(round(max(small) if small else min(rs)))
        else:
            r_max = max(int(round(r)) for r in rs)
        need.append(r_max + 16 if algorithm in _RAY else 2 * r_max + 2)
    return max(32, _round32(max(need)))


def tile_grid(width: int, height: int, tile: int, padding: int) -> List[Tuple[int, ...]]:
    """(ty, tx, core_x, core_y, core_w, core_h, win_x, win_y, win_w, win_h)
    in row-major tile order (core/tile_processor.py:1591-1627): cores of
    ``tile`` pixels (the last row and column ragged), windows the core grown
    by ``padding`` on every side and clipped to the raster."""
    xs = list(range(0, int(width), int(tile)))
    ys = list(range(0, int(height), int(tile)))
    out = []
    for ty, y0 in enumerate(ys):
        y1 = min(y0 + tile, height)
        wy0, wy1 = max(0, y0 - padding), min(height, y1 + padding)
        for tx, x0 in enumerate(xs):
            x1 = min(x0 + tile, width)
            wx0, wx1 = max(0, x0 - padding), min(width, x1 + padding)
            out.append((ty, tx, x0, y0, x1 - x0, y1 - y0, wx0, wy0, wx1 - wx0, wy1 - wy0))
    return out


def valid_bbox(raster: np.ndarray) -> Optional[Tuple[int, int, int, int]]:
    """Valid-data box (by0, by1, bx0, bx1) from a nearest overview of
    H // cov x W // cov samples, cov = max(1, max(W, H) // 512)
    (algorithms/_norm_stats.py:254-264).  The samples are the pixel centres
    floor((i + 0.5) * H / oh); GDAL's own nearest pick at a non-integer
    ratio is not restated."""
    H, W = raster.shape
    cov = max(1, max(W, H) // 512)
    oh, ow = max(1, H // cov), max(1, W // cov)
    ri = np.minimum(np.floor((np.arange(oh) + 0.5) * (H / oh)).astype(np.int64), H - 1)
    ci = np.minimum(np.floor((np.arange(ow) + 0.5) * (W / ow)).astype(np.int64), W - 1)
    ok = np.isfinite(raster[np.ix_(ri, ci)])
    if not ok.any():
        return None
    rows = np.flatnonzero(ok.any(axis=1))
    cols = np.flatnonzero(ok.any(axis=0))
    return (int(rows[0]) * cov, min(H, (int(rows[-1]) + 1) * cov),
            int(cols[0]) * cov, min(W, (int(cols[-1]) + 1) * cov))


def stratified_windows(width: int, height: int, bbox: Sequence[int], *, grid: int = 3,
                       tile: int = 4096) -> List[Tuple[int, int, int, int]]:
    """(wy0, wx0, win_w, win_h) of a grid x grid layout of tile-sized windows
    centred on the cells of the bounding box, clamped into the raster,
    duplicates dropped, first occurrence kept (algorithms/_norm_stats.py
    :64-100)."""
    by0, by1, bx0, bx1 = (int(v) for v in bbox)
    ch, cw = max(1, (by1 - by0) // grid), max(1, (bx1 - bx0) // grid)
    half = tile // 2

    def origin(c, n):
        return min(max(0, c - half), max(0, n - tile))

    out: List[Tuple[int, int, int, int]] = []
    for gy in range(grid):
        wy0 = origin(by0 + gy * ch + ch // 2, height)
        for gx in range(grid):
            wx0 = origin(bx0 + gx * cw + cw // 2, width)
            key = (wy0, wx0, min(tile, width - wx0), min(tile, height - wy0))
            if key not in out:
                out.append(key)
    return out


def prepass_geometry(algorithm: str, merged: dict) -> Tuple[int, int]:
    """(margin, tile) of the pre-pass windows (algorithms/_norm_stats.py
    :103-162) for topousm_fast, openness and ambient_occlusion: margin =
    int(s + 16) with s the largest of the radii / scales lists, radius,
    iterations and max_distance (16 when none); tile = max(2048, 4 margin)."""
    s = []
    for key in ("radii", "scales"):
        v = merged.get(key)
        if isinstance(v, (list, tuple)) and v:
            s.append(max(float(x) for x in v))
    for key in ("kernel_size", "radius", "iterations", "max_distance"):
        if merged.get(key):
            s.append(float(merged[key]))
    margin = max(1, int((max(s) if s else 16.0) + 16))
    return margin, max(2048, 4 * margin)


def prepass_windows(raster: np.ndarray, algorithm: str, merged: dict,
                    grid: int = 3) -> List[Tuple[int, int, int, int, int]]:
    """The pre-pass windows with their trim: (wy0, wx0, win_w, win_h, m),
    m = min(margin, win_h // 3, win_w // 3) (algorithms/_norm_stats.py
    :252-276).  The valid-fraction test (>= 0.02) is left to the caller."""
    H, W = raster.shape
    box = valid_bbox(raster)
    if box is None:
        return []
    margin, tile = prepass_geometry(algorithm, merged)
    return [(wy, wx, tw, th, int(min(margin, th // 3, tw // 3)))
            for wy, wx, tw, th in stratified_windows(W, H, box, grid=grid,
                                                     tile=min(tile, max(W, H)))]


def plan_tiles(algorithm: str, params: dict, width: int, height: int, tile: int,
               sigma: float = 10.0) -> List[Tuple[int, ...]]:
    """The tile list of a run: tile_grid with required_padding's halo.
    ``params`` are the run's resolved parameters (mode, and the radii after
    TopoUSM's overview split)."""
    pad = required_padding(algorithm, params, sigma=sigma,
                           pixel_size=float(params.get("pixel_size", 1.0) or 1.0),
                           tile_size=tile)
    return tile_grid(width, height, tile, pad)

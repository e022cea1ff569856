"""Metric pixel scales (io/raster_info.py:13-99) without rasterio.

The geotransform is passed as its two pixel-size terms (a = x size, e = y
size, negative on north-up rasters) and the CRS as an ``is_geographic`` flag
(or None when the raster has no CRS).
"""
from __future__ import annotations

import logging
import math
from typing import Optional, Tuple

logger = logging.getLogger(__name__)


def meters_per_degree(lat_deg: float) -> Tuple[float, float]:
    """WGS84 series (io/raster_info.py:13-35): (m/deg lon, m/deg lat)."""
    lat = math.radians(float(lat_deg))
    m_lat = (111132.92 - 559.82 * math.cos(2.0 * lat) + 1.175 * math.cos(4.0 * lat)
             - 0.0023 * math.cos(6.0 * lat))
    m_lon = (111412.84 * math.cos(lat) - 93.5 * math.cos(3.0 * lat)
             + 0.118 * math.cos(5.0 * lat))
    return max(1e-6, float(m_lon)), float(m_lat)


def metric_pixel_scales(*, transform_a: float, transform_e: float,
                        is_geographic: Optional[bool], bounds) -> Tuple[float, float, float, bool, Optional[float]]:
    """metric_pixel_scales_from_metadata (io/raster_info.py:38-99).

    ``is_geographic=None`` means "no CRS"; ``bounds`` = (left, bottom, right, top).
    Returns (scale_x, scale_y, mean_m, is_geographic, lat_center)."""
    sx, sy = float(transform_a), float(transform_e)
    abs_x, abs_y = abs(sx), abs(sy)
    left, bottom, right, top = (float(b) for b in bounds)
    if is_geographic is None:
        plausible_lonlat = (-180.0 <= left <= 180.0 and -180.0 <= right <= 180.0
                            and -90.0 <= bottom <= 90.0 and -90.0 <= top <= 90.0)
        plausible_deg = 0.0 < abs_x <= 1.0 and 0.0 < abs_y <= 1.0
        if plausible_lonlat and plausible_deg:
            raise ValueError(
                "Raster has no CRS and its extent/pixel size look geographic; "
                "cannot safely interpret degree-sized pixels as meters. Assign a CRS first.")
        logger.warning("Raster has no CRS; treating coordinate units as meters because the "
                       "metadata does not look like longitude/latitude.")
    if is_geographic:
        lat_center = 0.5 * (bottom + top)
        m_lon, m_lat = meters_per_degree(lat_center)
        scale_x = math.copysign(abs_x * m_lon, sx if sx != 0 else 1.0)
        scale_y = math.copysign(abs_y * m_lat, sy if sy != 0 else -1.0)
        return float(scale_x), float(scale_y), 0.5 * (abs(scale_x) + abs(scale_y)), True, float(lat_center)
    scale_x = sx if sx != 0 else 1.0
    scale_y = sy if sy != 0 else -1.0
    return float(scale_x), float(scale_y), 0.5 * (abs(scale_x) + abs(scale_y)), False, None


def tile_geographic_scales(transform_a: float, transform_e: float, lat_center_tile: float):
    """Per-tile latitude scaling (core/tile_processor.py:917-929)."""
    m_lon, m_lat = meters_per_degree(lat_center_tile)
    px_x = math.copysign(abs(transform_a) * m_lon, transform_a if transform_a != 0 else 1.0)
    px_y = math.copysign(abs(transform_e) * m_lat, transform_e if transform_e != 0 else -1.0)
    return float(px_x), float(px_y)


def row_geographic_scales_x(transform_a: float, transform_e: float, transform_f: float,
                            row0: int, n_rows: int):
    """Signed x pixel scale (metres) of raster rows row0 .. row0 + n_rows - 1 of a
    geographic grid, each at the latitude of its own row centre
    (f + (row + 0.5) e) with the WGS84 series of meters_per_degree.  No
    reference counterpart: the reference evaluates one latitude per tile
    (core/tile_processor.py:910-932)."""
    import numpy as np
    lat = transform_f + (np.arange(row0, row0 + n_rows, dtype=np.float64) + 0.5) * transform_e
    r = np.radians(lat)
    m_lon = 111412.84 * np.cos(r) - 93.5 * np.cos(3.0 * r) + 0.118 * np.cos(5.0 * r)
    m_lon = np.maximum(m_lon, 1e-6)
    return np.copysign(abs(transform_a) * m_lon, transform_a if transform_a != 0 else 1.0)

"""Cloud-optimized GeoTIFF output and windowed input for device rasters.

The on-disk side of the tile pipeline, without GDAL: ``write_cog`` writes the
output COG that the reference builds with GDAL's COG driver from its tile
files (io/cog_builder.py:658-716) with the options of
core/dask_processor.py:201-228 (ZSTD level 1, 512 tiles, predictor 3 for
float32 and 2 for int16, AVERAGE overviews, 8 levels at most, BigTIFF);
``read_window`` is read_tile_window (core/tile_io.py:49-51) for COGs and
tiled GeoTIFFs (ZSTD, DEFLATE or none), returning a device tensor.

Both run in libfsg (csrc/cog.hip): overviews and predictors on the GPU,
compression and decompression on host threads.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence, Tuple

import torch

from . import _native as N

_DT = {torch.float32: N.OUT_F32, torch.int16: N.OUT_I16, torch.uint8: N.OUT_U8}
_TD = {v: k for k, v in _DT.items()}


def _s():
    return torch.cuda.current_stream().cuda_stream


def write_cog(path: str, raster: torch.Tensor, *, transform: Optional[Sequence[float]] = None,
              epsg: Optional[int] = None, geographic: bool = False,
              nodata: Optional[float] = None, block: int = 512, zstd_level: int = 1,
              overview_count: int = 8, threads: int = 0) -> dict:
    """Write a 2-D float32 / int16 / uint8 GPU raster as a COG.

    transform: the affine (a, b, c, d, e, f) of rasterio (north-up: b = d = 0),
    i.e. x = c + a * col, y = f + e * row.  Returns the writer's report."""
    if raster.dim() != 2 or not raster.is_cuda or raster.dtype not in _DT:
        raise ValueError("raster must be a 2-D float32/int16/uint8 GPU tensor")
    if raster.stride(1) != 1:
        raster = raster.contiguous()
    r = N.fsg_cog_raster()
    r.data = raster.data_ptr()
    r.ld = raster.stride(0)
    r.h, r.w = raster.shape
    r.dtype = _DT[raster.dtype]
    if transform is not None:
        a, b, c, d, e, f = (float(v) for v in transform[:6])
        if b != 0.0 or d != 0.0:
            raise ValueError("only north-up transforms (no rotation) are written")
        r.has_geotransform = 1
        for k, v in enumerate((c, a, 0.0, f, 0.0, e)):
            r.geotransform[k] = v
    r.epsg = int(epsg or 0)
    r.epsg_geographic = 1 if geographic else 0
    if nodata is not None:
        r.has_nodata = 1
        r.nodata = float(nodata)
    o = N.fsg_cog_options(int(block), int(zstd_level), int(overview_count), int(threads))
    rep = N.fsg_cog_report()
    N.check(N.lib().fsg_cog_write(str(path).encode(), C.byref(r), C.byref(o), C.byref(rep), _s()))
    return {"levels": rep.n_levels, "data_start": rep.data_start,
            "compressed_bytes": rep.compressed_bytes, "raw_bytes": rep.raw_bytes}


class CogReader:
    """An open tiled TIFF: ``info(level)`` and ``read(window, level)``."""

    def __init__(self, path: str):
        h = C.c_void_p()
        N.check(N.lib().fsg_cog_open(str(path).encode(), C.byref(h)))
        self._h = h
        self.path = str(path)

    def info(self, level: int = 0) -> dict:
        i = N.fsg_cog_info_t()
        N.check(N.lib().fsg_cog_info(self._h, int(level), C.byref(i)))
        g = list(i.geotransform)
        return {"levels": i.n_levels, "width": i.width, "height": i.height,
                "tile": (i.tile_h, i.tile_w), "dtype": _TD[i.dtype],
                "compression": i.compression, "predictor": i.predictor,
                "transform": ((g[1], g[2], g[0], g[4], g[5], g[3])
                              if i.has_geotransform else None),
                "epsg": i.epsg or None, "geographic": bool(i.epsg_geographic),
                "nodata": i.nodata if i.has_nodata else None}

    def read(self, window: Optional[Tuple[int, int, int, int]] = None, level: int = 0,
             out: Optional[torch.Tensor] = None, device="cuda", threads: int = 0) -> torch.Tensor:
        """window = (col_off, row_off, width, height) as rasterio's Window."""
        inf = self.info(level)
        if window is None:
            window = (0, 0, inf["width"], inf["height"])
        c0, r0, w, h = (int(v) for v in window)
        if out is None:
            out = torch.empty((h, w), dtype=inf["dtype"], device=device)
        elif out.shape != (h, w) or out.dtype != inf["dtype"] or out.stride(1) != 1:
            raise ValueError("out must match the window and the file's sample type")
        N.check(N.lib().fsg_cog_read(self._h, int(level), r0, c0, h, w, out.data_ptr(),
                                     out.stride(0), int(threads), _s()))
        return out

    def close(self):
        if self._h:
            N.lib().fsg_cog_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_window(path: str, window: Tuple[int, int, int, int], level: int = 0) -> torch.Tensor:
    """read_tile_window (core/tile_io.py:49-51) as float32 on the GPU."""
    with CogReader(path) as rd:
        t = rd.read(window, level)
    return t if t.dtype == torch.float32 else t.float()

"""MI355X-native FujiShaderGPU raster hot path.

Per-tile terrain compute (hillshade, slope, curvature, TopoUSM Fast, openness,
ambient occlusion), NoData handling, integer encoding and global statistics as
hand-written HIP kernels for gfx950 (libfsg.so, C ABI in include/fsg.h),
driven from Python with the reference's operator interface:

* ``ALGORITHMS[name].process(dem_gpu, **params)`` -- one tile window
  (algorithms/tile_shared.py, algorithms/tile/dask_bridge.py);
* ``process_dem(raster_gpu, algorithm, ...)`` -- the tile pipeline over a
  device-resident raster (core/tile_processor.py without file I/O).
"""
__version__ = "0.1.0"


def __getattr__(name):
    if name in ("ALGORITHMS", "TileAlgorithm", "load_algorithm", "run_tile_algorithm",
                "apply_nodata_mask", "UnsupportedOnBackend"):
        from . import algorithms
        return getattr(algorithms, name)
    if name in ("process_dem", "plan", "TileResult"):
        from . import scheduler
        return getattr(scheduler, name)
    raise AttributeError(name)

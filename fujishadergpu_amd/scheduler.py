"""Device-resident tile scheduler: process_dem_tiles without the file I/O.

Restates the orchestration of process_dem_tiles / process_single_tile
(core/tile_processor.py:822-1745) for a raster that already lives in HBM:
output encoding, mode and radii resolution, the TopoUSM large-radius split and
overview field, halo sizing, the global-statistics pre-pass, the tile grid,
NoData skipping, per-tile geographic scales, and the per-tile compute.  Each
tile computes only its core region, straight into the output raster, with the
crop, NoData restore, normalisation and quantisation fused into the kernel
epilogue.  GDAL decode/encode, per-tile GeoTIFFs and the COG build are not
part of this path.
"""
from __future__ import annotations

import logging
import math
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N
from . import ops
from .algorithms import (ALGORITHMS, SUPPORTED, CoarsePyramid, OutSpec, TopoUSMFastAlgorithm,
                         UnsupportedOnBackend, load_algorithm)
from .encoding import SUPPORTED_OUTPUT_DTYPES, quantize_params, resolve_output_range
from .params import (LOCAL_RADII, LOCAL_WEIGHTS, MULTISCALE_REQUIRED_ALGOS, NORM_STAT_ALGOS,
                     RADII_DRIVEN_ALGOS, SPATIAL_TILE_ALGORITHMS, auto_spatial_profile,
                     normalize_topousm_radii_weights, overview_geometry, required_padding,
                     sanitize_spatial_radii_weights_for_tile, split_radii_by_threshold,
                     tile_grid, topousm_large_radius_threshold)
from .raster_info import metric_pixel_scales, row_geographic_scales_x, tile_geographic_scales

# algorithms whose gradient takes a per-row x cell width
ROW_SCALED = ("hillshade", "slope", "curvature", "hillshade_slope")
# algorithms whose spatial mode (multi-radius smoothing) runs on the GPU
SPATIAL_GPU = ("hillshade", "slope", "curvature", "openness", "ambient_occlusion")
from .stats import norm_stats_prepass, prepass_pyramid_pairs, valid_bbox_begin

logger = logging.getLogger(__name__)


@dataclass
class TileResult:
    """utils/types.py:7-14 (success / skipped / error per tile)."""
    tile_y: int
    tile_x: int
    success: bool
    filename: Optional[str] = None
    error_message: Optional[str] = None
    skipped_reason: Optional[str] = None


@dataclass
class RunPlan:
    algorithm: str
    params: Dict[str, Any]
    tile_size: int
    padding: int
    tiles: List[tuple]
    output_dtype: str
    qp: Optional[dict]
    stats: Any = None
    extras: Dict[str, Any] = field(default_factory=dict)


def _validate_sizes(tile_size, padding):
    """core/tile_processor.py:1053-1073."""
    if tile_size is not None:
        try:
            parsed = int(tile_size)
        except (TypeError, ValueError) as exc:
            raise ValueError(f"tile_size must be a positive integer, got {tile_size!r}") from exc
        if isinstance(tile_size, bool) or parsed != tile_size or parsed <= 0:
            raise ValueError(f"tile_size must be a positive integer, got {tile_size!r}")
        tile_size = parsed
    if padding is not None:
        try:
            parsed = int(padding)
        except (TypeError, ValueError) as exc:
            raise ValueError(f"padding must be >= 0, got {padding!r}") from exc
        if isinstance(padding, bool) or parsed != padding or parsed < 0:
            raise ValueError(f"padding must be >= 0, got {padding!r}")
        padding = parsed
    return tile_size, padding


def plan(raster: torch.Tensor, algorithm: str, *, tile_size: Optional[int] = 4096,
         padding: Optional[int] = None, sigma: float = 10.0, multiscale_mode: bool = True,
         pixel_size: Optional[float] = None, transform: Optional[Sequence[float]] = None,
         is_geographic: Optional[bool] = False, nodata: Optional[float] = None,
         geographic_rows: bool = False, **algo_params) -> RunPlan:
    """Parameter resolution of process_dem_tiles (core/tile_processor.py:1050-1590).

    ``transform`` = affine (a, b, c, d, e, f); without it the raster is taken as
    projected, north-up, with ``pixel_size`` cells.  ``geographic_rows`` (no
    reference counterpart) gives a geographic raster's gradient algorithms the
    x cell width of every row's own latitude instead of the tile centre's."""
    if algorithm not in ALGORITHMS:
        raise ValueError(f"Algorithm {algorithm} not found or not available on this platform")
    if algorithm not in SUPPORTED:
        ALGORITHMS[algorithm].process(raster)  # raises UnsupportedOnBackend
    tile_size, padding = _validate_sizes(tile_size, padding)
    H, W = int(raster.shape[0]), int(raster.shape[1])
    algo_params = dict(algo_params)
    user_radii = algo_params.get("radii") is not None
    user_weights = algo_params.get("weights") is not None

    output_dtype = str(algo_params.pop("output_dtype", "float32") or "float32").lower()
    output_range = algo_params.pop("output_range", None)
    if output_dtype not in SUPPORTED_OUTPUT_DTYPES:
        raise ValueError(
            f"Unsupported output_dtype={output_dtype!r}. Choose from {SUPPORTED_OUTPUT_DTYPES}.")
    qp = None
    if output_dtype in ("int16", "uint8"):
        vr = resolve_output_range(algorithm, params=algo_params, override=output_range)
        if vr is None:
            logger.warning("No fixed output range for %s (unit=%s); writing float32 instead of %s.",
                           algorithm, algo_params.get("unit", ""), output_dtype)
            output_dtype = "float32"
        else:
            qp = quantize_params(float(vr[0]), float(vr[1]), output_dtype)

    if pixel_size is None:
        pixel_size = abs(float(transform[0])) if transform is not None else 1.0
    if algorithm == "hillshade":
        algo_params.setdefault("z_factor", 1.0)

    mode_now = str(algo_params.get("mode", "spatial")).lower()
    is_local = mode_now == "local" or (algorithm == "topousm_fast" and not multiscale_mode)
    if is_local and algorithm in MULTISCALE_REQUIRED_ALGOS:
        algo_params["mode"] = "spatial"
        mode_now, is_local = "spatial", False
    if is_local:
        algo_params["radii"] = list(LOCAL_RADII)
        algo_params["weights"] = list(LOCAL_WEIGHTS)
    elif algorithm in RADII_DRIVEN_ALGOS and not algo_params.get("radii") and mode_now == "spatial":
        r, w = auto_spatial_profile(min(W, H))
        algo_params["radii"] = r
        if not algo_params.get("weights"):
            algo_params["weights"] = w

    mode_norm = str(algo_params.get("mode", "spatial")).lower()
    if mode_norm == "spatial" and algorithm in SPATIAL_TILE_ALGORITHMS:
        if algo_params.get("radii") is None:
            r, w = auto_spatial_profile(None)
            algo_params["radii"] = r
            if algo_params.get("weights") is None and w:
                algo_params["weights"] = w
        ar, aw, _ = sanitize_spatial_radii_weights_for_tile(
            algorithm, algo_params.get("radii"), algo_params.get("weights"), int(tile_size))
        if ar is not None:
            algo_params["radii"] = ar
            if aw is not None:
                algo_params["weights"] = aw
            elif algo_params.get("weights") is not None:
                algo_params["weights"] = None

    extras: Dict[str, Any] = {}
    if algorithm == "topousm_fast":
        full_r, full_w = normalize_topousm_radii_weights(
            None, None, float(pixel_size), manual_radii=algo_params.get("radii"),
            manual_weights=algo_params.get("weights"))
        if full_r:
            thr = topousm_large_radius_threshold(int(tile_size))
            sr, sw, lr, lw = split_radii_by_threshold(full_r, full_w, thr)
            if lr:
                extras["large"] = (lr, lw)
                algo_params["_topousm_fast_full_radii"] = list(full_r)
                algo_params["_topousm_fast_full_weights"] = list(full_w) if full_w else None
                algo_params["radii"] = sr
                algo_params["weights"] = sw
                algo_params["_topousm_fast_small_radii"] = sr
                algo_params["_topousm_fast_small_weights"] = sw
                algo_params["_topousm_fast_w_large"] = float(sum(lw))
                algo_params["_topousm_fast_full_shape"] = (H, W)

    algo_params.setdefault("is_geographic_dem", bool(is_geographic))
    algo_params["_geographic_rows"] = bool(geographic_rows)
    req = required_padding(algorithm, algo_params, sigma, float(pixel_size), None, int(tile_size))
    padding = req if padding is None else max(padding, req)

    # small DEM -> local (core/tile_processor.py:1442-1452)
    if (str(algo_params.get("mode", "spatial")).lower() == "spatial" and not user_radii
            and not user_weights and min(W, H) <= 1024):
        algo_params["mode"] = "local"

    if transform is not None:
        a, e = float(transform[0]), float(transform[4])
        left, top = float(transform[2]), float(transform[5])
        bounds = (left, top + H * e, left + W * a, top)
        psx, psy, _, is_geo, _ = metric_pixel_scales(transform_a=a, transform_e=e,
                                                     is_geographic=is_geographic, bounds=bounds)
        algo_params["pixel_scale_x"], algo_params["pixel_scale_y"] = psx, psy
        algo_params["is_geographic_dem"] = bool(is_geo)
        extras["transform"] = tuple(float(t) for t in transform)
    else:
        algo_params.setdefault("pixel_scale_x", float(pixel_size))
        algo_params.setdefault("pixel_scale_y", -float(pixel_size))
    algo_params["pixel_size"] = float(pixel_size)
    tiles = tile_grid(W, H, int(tile_size), int(padding))
    return RunPlan(algorithm, algo_params, int(tile_size), int(padding), tiles, output_dtype,
                   qp, extras=extras)


def overview_factor(H: int, W: int) -> Optional[int]:
    """The overview's integer decimation factor, or None when it is fractional."""
    scale, sh, sw = overview_geometry(W, H)
    f = int(round(scale))
    if abs(scale - f) > 1e-12 or sh * f != H or sw * f != W:
        return None
    return f


def overview(raster: torch.Tensor) -> Tuple[torch.Tensor, float]:
    """The raster's overview read (core/tile_processor.py:575-590,
    read_overview_coarse_dem _nan_utils.py:836-871): scale = max(W, H)/2048
    (at least 1), shape round(H/scale) x round(W/scale), GDAL's average
    resampling restated as the area-weighted NaN-aware average.  Returns
    (overview, scale)."""
    H, W = raster.shape
    scale, sh, sw = overview_geometry(W, H)
    f = overview_factor(H, W)
    if f == 1:
        return raster.contiguous().clone(), scale
    if f is not None:
        return ops.block_mean(raster, f), scale
    return ops.overview_average(raster, sh, sw), scale


def overview_field(raster: torch.Tensor, large_radii, large_weights,
                   coarse: Optional[torch.Tensor] = None,
                   nan_free: bool = False) -> Tuple[torch.Tensor, float]:
    """Large-radius coarse field from the overview (core/tile_processor.py:552-602,
    _impl_topousm_fast.py:133-155).  The overview read is GDAL average
    resampling onto round(H/s) x round(W/s) cells, s = max(W, H)/2048: the
    area-weighted NaN-aware average (``overview``), which for an integer
    decimation is the block mean (``coarse``, when the caller already has
    it).  ``nan_free``: the raster holds no NaN (so neither does the
    overview)."""
    H, W = raster.shape
    scale, _, _ = overview_geometry(W, H)
    if coarse is None:
        coarse, scale = overview(raster)
    rc = [max(1, int(round(float(r) / max(scale, 1.0)))) for r in large_radii]
    if nan_free and min(rc) >= 2 and len(rc) <= N.COARSE_JOB_MAX_RADII:
        # no NaN in the overview: the box means of every radius in one batched
        # pair of launches, then the weighted sum in one pass (same values)
        means = ops.topousm_coarse_means_batch([(coarse, 1, rc)])[0]
        return ops.topousm_field_from_means([means[i] for i in range(len(rc))],
                                            large_weights), scale
    return ops.topousm_coarse_field(coarse, rc, large_weights), scale


def _pyramid_pairs(P, with_prepass: bool):
    """(ds, coarse radius) levels of the raster-wide pyramid the tiles (and,
    when the statistics pre-pass runs, its windows) read."""
    pairs = CoarsePyramid.pairs(TopoUSMFastAlgorithm().params_struct(P.params))
    if with_prepass:
        pairs = pairs + prepass_pyramid_pairs(P.params)
    return pairs


def _fused_scan(src: torch.Tensor, P, pairs):
    """ops.raster_scan for the factors this job reads (the integer overview
    and the decimations of the pyramid levels ``pairs``), or None when it has
    none of them."""
    H, W = src.shape
    factors = []
    f_ov = overview_factor(H, W) if "large" in P.extras else None
    if f_ov in (4, 8, 16):
        factors.append(f_ov)
    for ds, _ in pairs:
        if ds in (4, 8, 16) and ds not in factors and len(factors) < 3:
            factors.append(ds)
    if not factors:
        return None
    cnt, grids = ops.raster_scan(src, factors)
    by_f = dict(zip(factors, grids))
    return {"nan": cnt, "overview": by_f.get(f_ov) if f_ov in (4, 8, 16) else None,
            "grids": {f: g for f, g in by_f.items()}}


EDGE_MODES = ("tile", "reflect")


def dask_edge_depth(algorithm: str, params: dict) -> int:
    """The map_overlap depth of the Dask backend's local-mode path
    (_impl_hillshade.py:133-134, _impl_slope.py:71-72, _impl_curvature.py:92-93,
    _impl_openness.py:204-207, _impl_ambient_occlusion.py:191-194)."""
    if algorithm in ("hillshade", "slope", "hillshade_slope"):
        return 1
    if algorithm == "curvature":
        return 2
    if algorithm == "openness":
        return int(params.get("max_distance", 50)) + 1
    if algorithm == "ambient_occlusion":
        return int(float(params.get("radius", 10.0)) + 1)
    raise UnsupportedOnBackend(
        f"{algorithm}: edge_mode='reflect' is not defined here (the Dask backend aligns its "
        "decimation grids to its own chunks, which the tile windows do not reproduce)")


def check_results(results: Sequence[TileResult]) -> None:
    """The end-of-run policy of process_dem_tiles (core/tile_processor.py
    :1689-1695): any failed tile aborts the run with the first three error
    messages; a run in which no tile succeeded (all skipped) raises
    ValueError."""
    errors = [r for r in results if not r.success and not r.skipped_reason]
    if errors:
        detail = "\n".join(f"Tile ({r.tile_y}, {r.tile_x}): {r.error_message}" for r in errors[:3])
        raise RuntimeError(f"Tile processing error:\n{detail}")
    if not any(r.success for r in results):
        raise ValueError("No tiles were processed")


def _run_tile(algo, src: torch.Tensor, win_geom, core_geom, pads, spec, p) -> None:
    """One tile on the current stream.  With edge_mode='reflect' the padded
    copy of an edge window is allocated and filled here, on the same stream as
    the kernel that reads it: the caching allocator then ties its block to that
    stream, so neither the pad nor a later reuse of the block can race the
    reader."""
    wy, wx, wh, ww = win_geom
    cy, cx, ch, cw = core_geom
    top, bottom, left, right = pads
    win = src[wy:wy + wh, wx:wx + ww]
    if top or bottom or left or right:
        win = ops.pad_reflect(win, top, bottom, left, right)
    algo.run_region(win, ops.region(cy - wy + top, cx - wx + left, ch, cw), spec, p)


def process_dem(raster: torch.Tensor, algorithm: str, *, nodata: Optional[float] = None,
                nodata_threshold: float = 1.0, out: Optional[torch.Tensor] = None,
                out_f32: Optional[torch.Tensor] = None, global_stats=None,
                copy_input: bool = True, return_plan: bool = False, edge_mode: str = "tile",
                streams: int = 2, **kwargs):
    """Run one algorithm over a device-resident raster, tile by tile.

    Returns the output raster (float32 with NaN NoData, or int16/uint8 with 0
    NoData), and the per-tile TileResult list.  ``out_f32`` optionally also
    receives the float32 result when the main output is integer-encoded.

    edge_mode: "tile" (the tile backend: windows end at the raster edge and
    each block function applies its own edge rule there) or "reflect" (the
    Dask backend's outer ring: a local-mode window that touches a raster edge
    is extended beyond it by the algorithm's map_overlap depth with the
    symmetric rule of boundary='reflect', core/dask_processor.py:1473-1488).
    Spatial mode already reflects at every window edge (each tile is one
    map_overlap chunk), so there both modes are the same.

    streams: the number of streams the tiles alternate over (the caller's
    current stream plus streams - 1 side streams); 1 runs every tile on the
    caller's stream.  The outputs are the same bits either way."""
    edge_mode = str(edge_mode).lower()
    if edge_mode not in EDGE_MODES:
        raise ValueError(f"edge_mode must be one of {EDGE_MODES}, got {edge_mode!r}")
    P = plan(raster, algorithm, nodata=nodata, **kwargs)
    H, W = raster.shape
    src = raster.clone() if copy_input else raster
    windows = [(t[7], t[6], t[9], t[8]) for t in P.tiles]
    stats = global_stats if global_stats is not None else P.params.get("global_stats")
    # Without a NoData sentinel to convert, one fused read of the raster gives
    # its NaN count, the overview and the coarse-pyramid cell grids; the
    # per-window counts are then only needed when the raster holds NaN.
    pairs = _pyramid_pairs(P, stats is None) if algorithm == "topousm_fast" else []
    scan = _fused_scan(src, P, pairs) if nodata is None else None
    counts = ops.nodata_to_nan(src, nodata, windows) if scan is None else None
    if "large" in P.extras and scan is None:
        # queued before the counts are read back, so the read overlaps it
        lr, lw = P.extras["large"]
        fld, scale = overview_field(src, lr, lw)
        P.params["_topousm_fast_coarse_field"] = fld
        P.extras["overview_scale"] = scale
    bbox_handle = None
    if stats is None and algorithm in NORM_STAT_ALGOS:
        # the pre-pass overview, read back with the counts (one synchronisation)
        bbox_handle = valid_bbox_begin(src)
    if scan is not None:
        if int(scan["nan"].item()) == 0:
            nan_counts = np.zeros(len(windows), dtype=np.int64)
        else:
            nan_counts = ops.nodata_to_nan(src, None, windows).cpu().numpy()
    else:
        nan_counts = counts.cpu().numpy()
    # the tile windows cover the raster: all zero means the raster holds no NaN
    raster_nan_free = bool((nan_counts == 0).all())
    if "large" in P.extras and scan is not None:
        # queued first after the synchronisation: it runs while the host
        # prepares the rest (only the main tiles read it)
        lr, lw = P.extras["large"]
        # the scan's cell means round their sum first; they equal the
        # overview's float64 means wherever every cell is whole (NaN-free)
        fld, scale = overview_field(src, lr, lw,
                                    coarse=scan["overview"] if raster_nan_free else None,
                                    nan_free=raster_nan_free)
        P.params["_topousm_fast_coarse_field"] = fld
        P.extras["overview_scale"] = scale
    spatial_tiles = (algorithm in SPATIAL_GPU and
                     str(P.params.get("mode", "local")).lower() == "spatial")
    if spatial_tiles and P.params.get("radii"):
        # the unified overview coarse source of the spatial algorithms
        # (core/tile_processor.py:1519-1545, read_overview_coarse_dem
        # _nan_utils.py:836-871): GDAL average resampling, restated as the
        # area-weighted NaN-aware average (the block mean at an integer ratio)
        ov, _ = overview(src)
        sh, sw = int(ov.shape[0]), int(ov.shape[1])
        P.params["_overview_coarse_dem"] = ov
        P.params["_overview_decimation"] = 0.5 * (W / float(sw) + H / float(sh))
        P.params["_tile_full_shape"] = (int(H), int(W))
        P.params["_overview_cache"] = {}
    pyramid = None
    if algorithm == "topousm_fast" and raster_nan_free:
        pyramid = CoarsePyramid(src, decimated=scan["grids"] if scan else None)
        pyramid.require(pairs)
    reflect_d = 0
    if edge_mode == "reflect" and not spatial_tiles:
        reflect_d = dask_edge_depth(algorithm, P.params)
        if P.params.get("_geographic_rows"):
            raise ValueError("edge_mode='reflect' takes one pixel scale per tile "
                             "(geographic_rows is not defined for it)")
    if stats is None:
        stats = norm_stats_prepass(src, algorithm, P.params, raster_nan_free=raster_nan_free,
                                   pyramid=pyramid, bbox_handle=bbox_handle)
    if stats is not None:
        P.params["global_stats"] = stats
    P.stats = stats
    dt = {"float32": torch.float32, "int16": torch.int16, "uint8": torch.uint8}[P.output_dtype]
    # agg='stack' in spatial mode: band-first C x H x W, one band per radius
    # (_nan_utils.py:191-194; the tile writer keeps it as C bands,
    # core/tile_processor.py:944-971,606-624)
    shape: Tuple[int, ...] = (H, W)
    if spatial_tiles and str(P.params.get("agg", "mean") or "mean").lower() == "stack":
        from . import spatial
        shape = (spatial.n_fields(algorithm, P.params), H, W)
    if out is None:
        out = torch.empty(shape, dtype=dt, device=raster.device)
    elif tuple(out.shape) != shape:
        raise ValueError(f"out has shape {tuple(out.shape)}, expected {shape}")
    if out_f32 is not None and tuple(out_f32.shape) != shape:
        raise ValueError(f"out_f32 has shape {tuple(out_f32.shape)}, expected {shape}")
    # tiles alternate between two streams (independent windows, one algorithm
    # object, hence workspace, per stream): each tile's last workgroups overlap
    # the next tile's first.  Spatial-mode tiles share the overview responses
    # they build as they go; each entry carries an event the readers wait on.
    if int(streams) < 1:
        raise ValueError(f"streams must be >= 1, got {streams!r}")
    ring = ops.StreamRing(int(streams)) if len(P.tiles) > 1 and int(streams) > 1 else None
    algos = [load_algorithm(algorithm) for _ in range(len(ring) if ring else 1)]
    results: List[TileResult] = []
    transform = P.extras.get("transform")
    for k_tile, (t, nn) in enumerate(zip(P.tiles, nan_counts)):
        algo = algos[k_tile % len(algos)]
        ty, tx, cx, cy, cw, ch, wx, wy, ww, wh = t
        ratio = float(nn) / float(ww * wh) if ww * wh else 1.0
        o_view = out[..., cy:cy + ch, cx:cx + cw]
        f_view = out_f32[..., cy:cy + ch, cx:cx + cw] if out_f32 is not None else None
        if nn > 0 and ratio >= nodata_threshold:
            o_view.fill_(float("nan") if dt == torch.float32 else 0)
            if f_view is not None:
                f_view.fill_(float("nan"))
            results.append(TileResult(ty, tx, False, skipped_reason=(
                f"NoData covers {ratio:.1%} (threshold: {nodata_threshold:.1%})")))
            continue
        p = dict(P.params)
        p["_window_nan_free"] = int(nn) == 0
        if spatial_tiles:
            p["_tile_origin"] = (int(wy), int(wx))
        if algorithm == "topousm_fast":
            if "_topousm_fast_coarse_field" in p:
                p["_topousm_fast_field_offset"] = (int(wy), int(wx))
            if pyramid is not None:
                p["_topousm_fast_pyramid"] = pyramid
                p["_window_origin"] = (int(wy), int(wx))
        if transform is not None and p.get("is_geographic_dem"):
            a, _, _, _, e, f = transform
            top = f + cy * e
            bottom = f + (cy + ch) * e
            p["pixel_scale_x"], p["pixel_scale_y"] = tile_geographic_scales(
                a, e, 0.5 * (top + bottom))
            if p.get("_geographic_rows") and algorithm in ROW_SCALED and not spatial_tiles:
                p["pixel_scale_x_rows"] = row_geographic_scales_x(a, e, f, int(wy), int(wh))
        pads = (0, 0, 0, 0)
        if reflect_d:
            pads = (min(reflect_d, wh) if wy == 0 else 0,
                    min(reflect_d, wh) if wy + wh == H else 0,
                    min(reflect_d, ww) if wx == 0 else 0,
                    min(reflect_d, ww) if wx + ww == W else 0)
        spec = OutSpec(f32=o_view if dt == torch.float32 else f_view,
                       q=o_view if dt != torch.float32 else None, qp=P.qp,
                       clip01=(algorithm == "hillshade"))
        try:
            if ring is not None:
                with torch.cuda.stream(ring.stream(k_tile)):
                    _run_tile(algo, src, (wy, wx, wh, ww), (cy, cx, ch, cw), pads, spec, p)
            else:
                _run_tile(algo, src, (wy, wx, wh, ww), (cy, cx, ch, cw), pads, spec, p)
            results.append(TileResult(ty, tx, True))
        except UnsupportedOnBackend:
            if ring is not None:
                ring.join()
            raise
        except Exception as exc:  # per-tile failure record (core/tile_processor.py:1022-1024)
            logger.error("Tile (%d, %d) processing failed: %s", ty, tx, exc)
            results.append(TileResult(ty, tx, False, error_message=str(exc)))
    if ring is not None:
        ring.join()
    check_results(results)
    if return_plan:
        return out, results, P
    return out, results

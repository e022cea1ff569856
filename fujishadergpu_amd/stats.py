"""Global normalisation statistics on the GPU (the stratified pre-pass).

Restates _compute_norm_stats_tiled / inject_global_stats
(algorithms/_norm_stats.py:176-350) for an in-memory raster already resident
in HBM: the valid-data bounding box, the 3x3 stratified windows, the raw block
function on each window (normalisation off), the margin trim, the pooled
values and the stat function -- all on the device.  The statistic stays on the
device (float64) and feeds the normalisation epilogue of the main pass
directly; nothing is copied back unless the caller asks.
"""
from __future__ import annotations

import logging
from typing import Optional

import numpy as np
import torch

from . import _native as N
from . import ops
from .algorithms import (AmbientOcclusionAlgorithm, CoarsePyramid, OpennessAlgorithm, OutSpec,
                         TopoUSMFastAlgorithm)
from .params import (NORM_STAT_ALGOS, coarse_radius, norm_stat_window_geometry,
                     stratified_windows)

logger = logging.getLogger(__name__)

# Defaults merged before the block-function parameters (ALGORITHMS[..].get_default_params()).
_DEFAULTS = {
    "topousm_fast": TopoUSMFastAlgorithm().get_default_params,
    "openness": OpennessAlgorithm().get_default_params,
    "ambient_occlusion": AmbientOcclusionAlgorithm().get_default_params,
}
# Parameters each raw block function accepts (inspect.signature in the reference).
_BLOCK_KEYS = {
    "topousm_fast": ("radii", "weights", "pixel_size"),
    "openness": ("openness_type", "num_directions", "max_distance", "pixel_size",
                 "pixel_scale_x", "pixel_scale_y"),
    "ambient_occlusion": ("num_samples", "radius", "intensity", "pixel_size",
                          "pixel_scale_x", "pixel_scale_y"),
}


def valid_bbox_begin(raster: torch.Tensor):
    """Start valid_bbox: the <= 512^2 overview samples are gathered on the
    device and copied into pinned host memory on the current stream; the
    returned handle is finished by valid_bbox_end once the stream got there
    (e.g. after any later synchronisation)."""
    H, W = raster.shape
    cov = max(1, max(W, H) // 512)
    oh, ow = max(1, H // cov), max(1, W // cov)
    ri = torch.clamp(((torch.arange(oh, device=raster.device, dtype=torch.float64) + 0.5) * (H / oh))
                     .floor().long(), max=H - 1)
    ci = torch.clamp(((torch.arange(ow, device=raster.device, dtype=torch.float64) + 0.5) * (W / ow))
                     .floor().long(), max=W - 1)
    ov = raster.index_select(0, ri).index_select(1, ci)
    if not raster.is_cuda:
        return ov, None, cov, H, W
    host = torch.empty((oh, ow), dtype=raster.dtype, pin_memory=True)
    host.copy_(ov, non_blocking=True)
    ev = torch.cuda.Event()
    ev.record()
    return host, ev, cov, H, W


def valid_bbox_end(handle):
    host, ev, cov, H, W = handle
    if ev is not None:
        ev.synchronize()
    vm = np.isfinite(host.numpy())
    ys, xs = np.nonzero(vm.any(axis=1))[0], np.nonzero(vm.any(axis=0))[0]
    if ys.size == 0:
        return None
    by0, by1 = int(ys[0]) * cov, min(H, (int(ys[-1]) + 1) * cov)
    bx0, bx1 = int(xs[0]) * cov, min(W, (int(xs[-1]) + 1) * cov)
    return by0, by1, bx0, bx1


def valid_bbox(raster: torch.Tensor):
    """Valid-data bounding box from a ~512 px nearest-neighbour overview
    (_norm_stats.py:254-264).  GDAL's nearest pick for non-integer ratios is
    unpinned; the overview samples pixel centres floor((i + 0.5) * H / oh).
    The bounding box of the overview's finite cells is host logic."""
    return valid_bbox_end(valid_bbox_begin(raster))


def prepass_params(algorithm: str, params: dict):
    """(merged parameters, block-function keywords) of the pre-pass windows:
    the defaults under the caller's parameters, topousm_fast with its full
    radius list (the large radii included)."""
    merged = {**_DEFAULTS[algorithm](), **(params or {})}
    if merged.get("_topousm_fast_full_radii"):
        merged["radii"] = list(merged["_topousm_fast_full_radii"])
        if merged.get("_topousm_fast_full_weights"):
            merged["weights"] = list(merged["_topousm_fast_full_weights"])
    kw = {k: merged[k] for k in _BLOCK_KEYS[algorithm] if merged.get(k) is not None}
    return merged, kw


def prepass_pyramid_pairs(params: dict):
    """(ds, coarse radius) pairs the topousm_fast pre-pass windows read."""
    _, kw = prepass_params("topousm_fast", params)
    return CoarsePyramid.pairs(TopoUSMFastAlgorithm().params_struct(kw))


def _own_coarse_means(raster, base, regions, pyramid):
    """Coarse means of the pre-pass windows that the raster-wide grids do not
    cover, built for all windows in a few batched launches: per region index,
    {radius index: grid}.  ``base`` is the windows' parameter struct."""
    jobs, keys = [], []
    for ri, (wy, wx, th, tw, m, nan_free) in enumerate(regions):
        if not nan_free:
            continue
        prm = type(base).from_buffer_copy(base)
        pyramid.attach(prm, (wy, wx), (th, tw), ops.region(m, m, th - 2 * m, tw - 2 * m))
        by_ds = {}
        for i in range(prm.n_radii):
            ds = int(prm.ds[i])
            if ds > 1 and not prm.mean[i]:
                by_ds.setdefault(ds, []).append(i)
        for ds, idx in by_ds.items():
            rs = [coarse_radius(int(prm.radii[i]), ds) for i in idx]
            if ds in (4, 8, 16) and min(rs) >= 2 and len(rs) <= N.COARSE_JOB_MAX_RADII:
                jobs.append((raster[wy:wy + th, wx:wx + tw], ds, rs))
                keys.append((ri, idx))
    own = {}
    for (ri, idx), g in zip(keys, ops.topousm_coarse_means_batch(jobs)):
        for k, i in enumerate(idx):
            own.setdefault(ri, {})[i] = g[k]
    return own


def norm_stats_prepass(raster: torch.Tensor, algorithm: str, params: dict, *,
                       grid: int = 3, max_tile: int = 4096, min_valid_frac: float = 0.02,
                       return_pool: bool = False, raster_nan_free: bool = False,
                       pyramid=None, bbox_handle=None):
    """Device-side _compute_norm_stats_tiled.  ``raster`` holds NaN at NoData.
    Returns a float64 GPU tensor (stats) or None (no valid window).

    ``raster_nan_free`` (known from the NoData scan) skips the per-window NaN
    count; ``pyramid`` (a CoarsePyramid of this raster, NaN-free rasters only)
    lets topousm_fast windows share raster-wide coarse means where exact;
    ``bbox_handle`` (valid_bbox_begin of this raster, queued earlier) saves
    the overview read its own synchronisation."""
    if algorithm not in NORM_STAT_ALGOS:
        return None
    merged, kw = prepass_params(algorithm, params)
    margin, tile = norm_stat_window_geometry(algorithm, merged, max_tile)
    H, W = raster.shape
    bbox = valid_bbox_end(bbox_handle) if bbox_handle is not None else valid_bbox(raster)
    if bbox is None:
        return None
    wins = stratified_windows(W, H, *bbox, grid=grid, tile=min(tile, max(W, H)))
    # valid fraction of every window in one device pass, one host sync
    if raster_nan_free:
        nan_counts = np.zeros(len(wins), dtype=np.int64)
    else:
        counts = ops.nodata_to_nan(raster, None, [(wy, wx, th, tw) for wy, wx, tw, th in wins])
        nan_counts = counts.cpu().numpy()
    regions = []
    for (wy, wx, tw, th), nn in zip(wins, nan_counts):
        if (th * tw - int(nn)) / float(th * tw) < min_valid_frac:
            continue
        m = int(min(margin, th // 3, tw // 3))
        regions.append((wy, wx, th, tw, m, int(nn) == 0))
    if not regions:
        return None
    total = sum((th - 2 * m) * (tw - 2 * m) for _, _, th, tw, m, _ in regions)
    pool = torch.empty(total, dtype=torch.float32, device=raster.device)
    algo = {"topousm_fast": TopoUSMFastAlgorithm, "openness": OpennessAlgorithm,
            "ambient_occlusion": AmbientOcclusionAlgorithm}[algorithm]()
    bp = dict(kw)
    own = {}
    base = algo.params_struct(bp) if algorithm == "topousm_fast" else None
    if pyramid is not None and algorithm == "topousm_fast":
        pyramid.require(pyramid.pairs(base))
        own = _own_coarse_means(raster, base, regions, pyramid)
    # the windows alternate between two streams (one algorithm object, hence
    # workspace, per stream; disjoint pool slices)
    ring = ops.StreamRing(2) if len(regions) > 1 else None
    algos = [algo] + [type(algo)() for _ in range(len(ring) - 1)] if ring else [algo]
    try:
        off = 0
        for ri, (wy, wx, th, tw, m, nan_free) in enumerate(regions):
            win = raster[wy:wy + th, wx:wx + tw]
            rh, rw = th - 2 * m, tw - 2 * m
            dst = pool[off:off + rh * rw].view(rh, rw)
            reg = ops.region(m, m, rh, rw)
            if algorithm == "topousm_fast":
                wp = {**bp, "_window_nan_free": nan_free, "_topousm_fast_prm": base}
                if pyramid is not None:
                    wp["_topousm_fast_pyramid"] = pyramid
                    wp["_window_origin"] = (wy, wx)
                if ri in own:
                    wp["_topousm_fast_means"] = own[ri]
                if ring is not None:
                    with torch.cuda.stream(ring.stream(ri)):
                        algos[ri % len(algos)].run_region(win, reg, OutSpec(f32=dst), wp,
                                                          normalize=False)
                else:
                    algo.run_region(win, reg, OutSpec(f32=dst), wp, normalize=False)
            elif ring is not None:
                with torch.cuda.stream(ring.stream(ri)):
                    algos[ri % len(algos)].run_region(win, reg, OutSpec(f32=dst),
                                                      {**bp, "mode": "local"}, stretch=False)
            else:
                algo.run_region(win, reg, OutSpec(f32=dst), {**bp, "mode": "local"}, stretch=False)
            off += rh * rw
    finally:
        if ring is not None:
            ring.join()
    if algorithm == "topousm_fast":
        pct, nval, mom = ops.percentile(pool, [99.0], ops.PCT_ABS)
        if not raster_nan_free and int(nval.item()) == 0:
            # every pooled value is NoData: the reference pools nothing and
            # returns None (_norm_stats.py:280-283), so the main pass falls
            # back to per-tile statistics
            return (None, pool) if return_pool else None
        stats = ops.finalize_stats(0, pct, mom, nval)
    else:
        # openness / AO: an empty or degenerate pool gives a zero scale, the
        # no-op stretch the reference's None leaves (_norm_stats.py:286-289,
        # tile/dask_bridge.py:173-187)
        pct, nval, mom = ops.percentile(pool, [1.0, 99.0], ops.PCT_FINITE)
        stats = ops.finalize_stats(1, pct, mom, nval)
    if return_pool:
        return stats, pool
    return stats


class RayPrepass:
    """The openness / ambient_occlusion pre-pass of ``norm_stats_prepass``
    planned once and replayable: the window list, the pool and the percentile
    buffers are fixed at construction (one host synchronisation for the
    valid-data box and the windows' NoData counts); ``run`` enqueues the
    windows' block functions, the pooled p1/p99 and the stretch statistics
    with no allocation and no host synchronisation, so a graph can capture
    it.  The windows, trims, per-window calls and percentile are those of
    ``norm_stats_prepass`` (_compute_norm_stats_tiled,
    algorithms/_norm_stats.py:176-350), so ``run`` gives the same statistics
    bit for bit."""

    def __init__(self, raster: torch.Tensor, algorithm: str, params: dict, *,
                 raster_nan_free: bool = False, grid: int = 3, max_tile: int = 4096,
                 min_valid_frac: float = 0.02):
        if algorithm not in ("openness", "ambient_occlusion"):
            raise ValueError(f"RayPrepass covers openness and ambient_occlusion, not {algorithm!r}")
        self.algorithm = algorithm
        self.raster = raster
        merged, kw = prepass_params(algorithm, params)
        self.kw = {**kw, "mode": "local"}
        margin, tile = norm_stat_window_geometry(algorithm, merged, max_tile)
        H, W = raster.shape
        self.regions = []
        self.pool = None
        self.buffers = None
        bbox = valid_bbox(raster)
        if bbox is None:
            return
        wins = stratified_windows(W, H, *bbox, grid=grid, tile=min(tile, max(W, H)))
        if raster_nan_free:
            nan_counts = np.zeros(len(wins), dtype=np.int64)
        else:
            nan_counts = ops.nodata_to_nan(
                raster, None, [(wy, wx, th, tw) for wy, wx, tw, th in wins]).cpu().numpy()
        for (wy, wx, tw, th), nn in zip(wins, nan_counts):
            if (th * tw - int(nn)) / float(th * tw) < min_valid_frac:
                continue
            m = int(min(margin, th // 3, tw // 3))
            self.regions.append((wy, wx, th, tw, m))
        if not self.regions:
            return
        total = sum((th - 2 * m) * (tw - 2 * m) for _, _, th, tw, m in self.regions)
        self.pool = torch.empty(total, dtype=torch.float32, device=raster.device)
        self.buffers = ops.PercentileBuffers(total, [1.0, 99.0], raster.device)
        self._algo = {"openness": OpennessAlgorithm,
                      "ambient_occlusion": AmbientOcclusionAlgorithm}[algorithm]()

    def run(self) -> Optional[torch.Tensor]:
        """Enqueue the pre-pass on the current stream; the float64 device
        statistics (or None when no window has enough valid data)."""
        if self.pool is None:
            return None
        off = 0
        for wy, wx, th, tw, m in self.regions:
            rh, rw = th - 2 * m, tw - 2 * m
            win = self.raster[wy:wy + th, wx:wx + tw]
            self._algo.run_region(win, ops.region(m, m, rh, rw),
                                  OutSpec(f32=self.pool[off:off + rh * rw].view(rh, rw)),
                                  self.kw, stretch=False)
            off += rh * rw
        return self.buffers.run(self.pool, ops.PCT_FINITE, 1)

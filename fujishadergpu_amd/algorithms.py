"""Drop-in tile algorithms: the reference's per-tile operator interface on HIP.

Mirrors ``TileAlgorithm`` (algorithms/tile_shared.py:19-44) and the direct
per-tile paths of ``DaskSharedTileAdapter`` (algorithms/tile/dask_bridge.py:
72-271, 334-405): ``process(dem_gpu, **params)`` takes the tile window as a
2-D float32 GPU tensor (NaN = NoData) and returns a new float32 GPU tensor of
the same shape.  Parameter names, defaults, None handling and errors follow
the reference.  Every computation is a libfsg kernel; there is no CPU path.

``run_region`` is the scheduler's entry: it writes only a sub-rectangle of
the window (the tile core) straight into caller-owned outputs, fusing the
reference's crop, NoData restore, statistics post-processing and integer
encoding into the kernel epilogue.
"""
from __future__ import annotations

import logging
import math
from abc import ABC, abstractmethod
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N
from . import ops
from .params import (DEFAULT_ALTITUDE, DEFAULT_AZIMUTH, coarse_radius,
                     normalize_topousm_radii_weights, radius_to_downsample_factor,
                     shared_coarse_ok, topousm_default_radii)

logger = logging.getLogger(__name__)


class UnsupportedOnBackend(NotImplementedError):
    """A reference algorithm or mode that this MI355X build does not provide."""


# ------------------------------------------------------------- step helpers

def gradient_steps(pixel_size=1.0, pixel_scale_x=None, pixel_scale_y=None):
    """|step| of handle_nan_for_gradient (algorithms/_nan_utils.py:67-72)."""
    sy = abs(float(pixel_scale_y if pixel_scale_y is not None else pixel_size))
    sx = abs(float(pixel_scale_x if pixel_scale_x is not None else pixel_size))
    if sy < 1e-9:
        sy = float(pixel_size if pixel_size else 1.0)
    if sx < 1e-9:
        sx = float(pixel_size if pixel_size else 1.0)
    return sx, sy


def curvature_steps(pixel_size=1.0, pixel_scale_x=None, pixel_scale_y=None):
    """Signed steps of compute_curvature_block (_impl_curvature.py:23-28)."""
    sy = float(pixel_scale_y if pixel_scale_y is not None else pixel_size)
    sx = float(pixel_scale_x if pixel_scale_x is not None else pixel_size)
    if abs(sy) < 1e-9:
        sy = float(pixel_size if pixel_size else 1.0)
    if abs(sx) < 1e-9:
        sx = float(pixel_size if pixel_size else 1.0)
    return sx, sy


class RowScales:
    """Per-row signed x pixel scales resolved once for the device: the float64
    tensor the kernels read, and the first row's value on the host (its sign
    orients hillshade; every row shares it)."""

    def __init__(self, rows, device, pixel_size: float = 1.0):
        a = np.asarray(rows.cpu() if isinstance(rows, torch.Tensor) else rows,
                       dtype=np.float64).reshape(-1)
        ps = float(pixel_size or 1.0)
        a = np.where(np.abs(a) < 1e-9, ps, a)   # the 1e-9 fallback of _nan_utils.py:69-72
        if a.size == 0 or ((a > 0).any() and (a < 0).any()):
            raise ValueError("pixel_scale_x_rows must be non-empty and keep one sign")
        self.n = int(a.size)
        self.first = float(a[0])
        self.t = torch.from_numpy(np.ascontiguousarray(a)).to(device)


def x_steps_rows(p: dict, h: int, device) -> Optional[RowScales]:
    """The optional ``pixel_scale_x_rows`` parameter -- one signed x pixel
    scale per window row, for geographic grids whose cell width follows the
    latitude of each row -- resolved for the device (a RowScales passes
    through without a host round trip).  No reference counterpart: the
    reference uses one latitude per tile (core/tile_processor.py:910-932)."""
    rows = p.get("pixel_scale_x_rows")
    if rows is None:
        return None
    if not isinstance(rows, RowScales):
        rows = RowScales(rows, device, p.get("pixel_size", 1.0))
    if rows.n != h:
        raise ValueError(f"pixel_scale_x_rows has {rows.n} values for a window of {h} rows")
    return rows


def shade_steps(p: dict, win: torch.Tensor):
    """(|step_x|, |step_y|, sign_x, sign_y, per-row x steps or None) of the
    hillshade/slope gradient (_nan_utils.py:65-73, _impl_hillshade.py:39-42)."""
    psx, psy = p.get("pixel_scale_x"), p.get("pixel_scale_y")
    rows = x_steps_rows(p, win.shape[0], win.device)
    if rows is not None:
        psx = rows.first
    sx, sy = gradient_steps(p.get("pixel_size", 1.0), psx, psy)
    sgx = 1.0 if (psx is None or float(psx) >= 0.0) else -1.0
    sgy = 1.0 if (psy is None or float(psy) >= 0.0) else -1.0
    return sx, sy, sgx, sgy, (rows.t if rows is not None else None)


def ray_steps(pixel_size=1.0, pixel_scale_x=None, pixel_scale_y=None):
    """|step| of the openness/AO physical distance (_impl_openness.py:70-75)."""
    sx = abs(float(pixel_scale_x)) if pixel_scale_x is not None else float(pixel_size)
    sy = abs(float(pixel_scale_y)) if pixel_scale_y is not None else float(pixel_size)
    if sx < 1e-9:
        sx = float(pixel_size) if pixel_size else 1.0
    if sy < 1e-9:
        sy = float(pixel_size) if pixel_size else 1.0
    return sx, sy


def openness_table(num_directions, max_distance, pixel_size=1.0, pixel_scale_x=None,
                   pixel_scale_y=None):
    """Ray table of compute_openness_vectorized (_impl_openness.py:58-109)."""
    angles = np.linspace(0, 2 * np.pi, int(num_directions), endpoint=False)
    dirs = np.stack([np.cos(angles), np.sin(angles)], axis=1)
    dist = np.unique((np.linspace(0.1, 1.0, 10) * max_distance).astype(int))
    dist = dist[dist > 0]
    sx, sy = ray_steps(pixel_size, pixel_scale_x, pixel_scale_y)
    table = []
    for d in range(int(num_directions)):
        row = []
        for r in dist:
            ox = int(round(float(r) * float(dirs[d][0])))
            oy = int(round(float(r) * float(dirs[d][1])))
            if ox == 0 and oy == 0:
                row.append(None)
                continue
            row.append((ox, oy, max(float(np.hypot(float(ox) * sx, float(oy) * sy)), 1e-9)))
        table.append(row)
    return table


def ao_table(num_samples, radius, pixel_size=1.0, pixel_scale_x=None, pixel_scale_y=None):
    """Ring table of compute_ambient_occlusion_block (_impl_ambient_occlusion.py:54-81)."""
    angles = np.linspace(0, 2 * np.pi, int(num_samples), endpoint=False)
    dirs = np.stack([np.cos(angles), np.sin(angles)], axis=1)
    sx, sy = ray_steps(pixel_size, pixel_scale_x, pixel_scale_y)
    out = []
    for rf in (0.25, 0.5, 0.75, 1.0):
        r = radius * rf
        dxa = np.round(r * dirs[:, 0]).astype(int)
        dya = np.round(r * dirs[:, 1]).astype(int)
        for i in range(int(num_samples)):
            dx, dy = int(dxa[i]), int(dya[i])
            if dx == 0 and dy == 0:
                out.append(None)
                continue
            out.append((dx, dy, max(float(np.hypot(float(dx) * sx, float(dy) * sy)), 1e-9),
                        1.0 - (rf * 0.3)))
    return out


def _stats_tuple_ok(stats, n=1, tiny=1e-9) -> bool:
    return isinstance(stats, (tuple, list)) and len(stats) >= n


class OutSpec:
    """Where and how run_region writes: float32 and/or integer outputs.

    The outputs are region-shaped 2-D tensors, or band-first C x h x w stacks
    for agg='stack' (one band per radius, the reference's stack contract,
    _nan_utils.py:191-194, tile/dask_bridge.py:32-36)."""

    def __init__(self, f32: Optional[torch.Tensor] = None, q: Optional[torch.Tensor] = None,
                 qp: Optional[dict] = None, clip01: bool = False):
        self.f32, self.q, self.qp, self.clip01 = f32, q, qp, clip01

    @property
    def n_bands(self) -> Optional[int]:
        """The band count of a C x h x w stack output, None for 2-D outputs."""
        t = self.f32 if self.f32 is not None else self.q
        return int(t.shape[0]) if t is not None and t.dim() == 3 else None

    def band(self, i: int) -> "OutSpec":
        return OutSpec(f32=self.f32[i] if self.f32 is not None else None,
                       q=self.q[i] if self.q is not None else None, qp=self.qp,
                       clip01=self.clip01)

    def descriptors(self, post=N.POST_NONE, host_stats=(0.0, 0.0), dev_stats=None):
        if self.n_bands is not None:
            raise ValueError("a band stack (C x H x W) output is written only by agg='stack' "
                             "in spatial mode")
        out = []
        if self.f32 is not None:
            out.append(ops.output(self.f32, clip01=self.clip01, post=post,
                                  host_stats=host_stats, dev_stats=dev_stats))
        if self.q is not None:
            out.append(ops.output(self.q, qp=self.qp, clip01=self.clip01, post=post,
                                  host_stats=host_stats, dev_stats=dev_stats))
        return out


def _post_from_stats(kind: str, stats):
    """(post_kind, host_stats, dev_stats) for a global_stats value that is a
    tuple (host) or a float64 GPU tensor (device, from the pre-pass)."""
    if stats is None:
        return N.POST_NONE, (0.0, 0.0), None
    if isinstance(stats, torch.Tensor):
        return (N.POST_DIV if kind == "div" else N.POST_STRETCH), (0.0, 0.0), stats
    if kind == "div":
        return N.POST_DIV, (float(stats[0]), 0.0), None
    if not (isinstance(stats, (tuple, list)) and len(stats) >= 2):
        return N.POST_NONE, (0.0, 0.0), None
    return N.POST_STRETCH, (float(stats[0]), float(stats[1])), None


def _run_spatial(name: str, win, reg, out: "OutSpec", p: dict,
                 post=(N.POST_NONE, (0.0, 0.0), None), direct: bool = False) -> None:
    """spatial.run through ``out``: one descriptor set for a 2-D output, one
    per band for a band-first stack."""
    from . import spatial
    if out.n_bands is None:
        spatial.run(name, win, reg, out.descriptors(*post), p, direct=direct)
    else:
        spatial.run(name, win, reg, None, p, direct=direct,
                    bands=[out.band(i).descriptors(*post) for i in range(out.n_bands)])


# ------------------------------------------------------------------- bases

class TileAlgorithm(ABC):
    """algorithms/tile_shared.py:19-44."""

    @abstractmethod
    def get_default_params(self) -> Dict[str, Any]:
        ...

    @abstractmethod
    def process(self, dem_gpu: torch.Tensor, **params) -> torch.Tensor:
        ...


class NativeTileAlgorithm(TileAlgorithm):
    name = ""
    clip01_output = False

    def merged(self, params: dict) -> dict:
        """_merged_params (tile/dask_bridge.py:16-25)."""
        p = dict(self.get_default_params())
        p.update(params)
        return p

    # algorithms whose spatial mode runs here (spatial.py); the others raise
    spatial_ok = False

    def check_mode(self, p: dict) -> None:
        """Spatial-mode requests take the reference's multi-radius path
        (tile/dask_bridge.py:340-346); built for hillshade, slope, curvature,
        openness and ambient occlusion (spatial.py), not for the others."""
        if str(p.get("mode", "local")).lower() == "spatial" and not self.spatial_ok:
            raise UnsupportedOnBackend(
                f"{self.name}: mode='spatial' (multi-radius smoothing) is not built on the "
                "MI355X backend; use mode='local'")

    def is_spatial(self, p: dict) -> bool:
        return self.spatial_ok and str(p.get("mode", "local")).lower() == "spatial"

    def stack_bands(self, p: dict) -> Optional[int]:
        """The band count when ``p`` asks for agg='stack' on a multi-radius
        path (spatial mode, or hillshade's local multi-radius branch):
        band-first C x H x W, one band per radius, even for one radius
        (_nan_utils.py:191-194, tile/dask_bridge.py:32-36).  None when the
        output is one band (the local block functions ignore agg)."""
        if str(p.get("agg", "mean") or "mean").lower() != "stack":
            return None
        from . import spatial
        if self.is_spatial(p):
            return spatial.n_fields(self.name, p)
        return None

    def process(self, dem_gpu: torch.Tensor, **params) -> torch.Tensor:
        ops._check_window(dem_gpu, "dem_gpu")
        h, w = dem_gpu.shape
        p = self.merged(params)
        c = self.stack_bands(p)
        shape = (h, w) if c is None else (c, h, w)
        out = torch.empty(shape, dtype=torch.float32, device=dem_gpu.device)
        self.run_region(dem_gpu, ops.full_region(dem_gpu), OutSpec(f32=out), p)
        return out

    @abstractmethod
    def run_region(self, win: torch.Tensor, reg: N.fsg_region, out: OutSpec, p: dict) -> None:
        ...


class HillshadeAlgorithm(NativeTileAlgorithm):
    """_direct_hillshade (tile/dask_bridge.py:72-111) -> compute_hillshade_block."""
    name = "hillshade"

    def get_default_params(self):
        return {"azimuth": DEFAULT_AZIMUTH, "altitude": DEFAULT_ALTITUDE, "z_factor": 1.0,
                "pixel_size": 1.0, "multiscale": False, "radii": None, "weights": None,
                "agg": "mean", "mode": "local"}

    spatial_ok = True

    @staticmethod
    def _direct_radii(p: dict) -> List[float]:
        radii = p.get("radii", [1])
        if not isinstance(radii, (list, tuple)) or len(radii) == 0:
            radii = [1]
        return [max(1.0, float(r)) for r in radii]

    def stack_bands(self, p: dict) -> Optional[int]:
        c = super().stack_bands(p)
        if c is None and str(p.get("agg", "mean") or "mean").lower() == "stack" \
                and not self.is_spatial(p) and len(self._direct_radii(p)) > 1:
            c = len(self._direct_radii(p))        # _direct_hillshade's multi-radius branch
        return c

    def run_region(self, win, reg, out, p):
        self.check_mode(p)
        if self.is_spatial(p):
            _run_spatial("hillshade", win, reg, out, p)
            return
        radii = self._direct_radii(p)
        if (bool(p.get("multiscale", False)) or len(radii) > 1) and len(radii) > 1:
            # _direct_hillshade's multi-radius branch (tile/dask_bridge.py:95-105)
            _run_spatial("hillshade", win, reg, out, {**p, "radii": radii}, direct=True)
            return
        z = p.get("z_factor", 1.0)
        z = 1.0 if z is None else z
        sx, sy, sgx, sgy, rows = shade_steps(p, win)
        for o in out.descriptors():
            ops.hillshade_slope(win, reg, step_x=sx, step_y=sy, sign_x=sgx, sign_y=sgy,
                                z_factor=float(z), azimuth=float(p.get("azimuth", DEFAULT_AZIMUTH)),
                                altitude=float(p.get("altitude", DEFAULT_ALTITUDE)), hs_out=o,
                                nan_free=bool(p.get("_window_nan_free", False)),
                                step_x_rows=rows)


class SlopeAlgorithm(NativeTileAlgorithm):
    """_direct_slope (tile/dask_bridge.py:114-124) -> compute_slope_block."""
    name = "slope"

    def get_default_params(self):
        return {"unit": "degree", "pixel_size": 1.0, "mode": "local", "radii": None,
                "weights": None}

    spatial_ok = True

    def run_region(self, win, reg, out, p):
        self.check_mode(p)
        if self.is_spatial(p):
            _run_spatial("slope", win, reg, out, p)
            return
        sx, sy, _, _, rows = shade_steps(p, win)
        for o in out.descriptors():
            ops.hillshade_slope(win, reg, step_x=sx, step_y=sy, unit=p.get("unit", "degree"),
                                sl_out=o, nan_free=bool(p.get("_window_nan_free", False)),
                                step_x_rows=rows)


class HillshadeSlopeFused(NativeTileAlgorithm):
    """Hillshade and slope from one read of the window (one launch, two
    outputs); each output equals the separate algorithm's."""
    name = "hillshade_slope"

    def get_default_params(self):
        d = HillshadeAlgorithm().get_default_params()
        d.update({"unit": "degree"})
        return d

    def run_region(self, win, reg, out, p, out_slope: Optional[OutSpec] = None):
        self.check_mode(p)
        z = p.get("z_factor", 1.0)
        z = 1.0 if z is None else z
        sx, sy, sgx, sgy, rows = shade_steps(p, win)
        hs = out.descriptors()
        sl = out_slope.descriptors() if out_slope is not None else []
        n = max(len(hs), len(sl))
        for i in range(n):
            ops.hillshade_slope(win, reg, step_x=sx, step_y=sy, sign_x=sgx, sign_y=sgy,
                                z_factor=float(z), azimuth=float(p.get("azimuth", DEFAULT_AZIMUTH)),
                                altitude=float(p.get("altitude", DEFAULT_ALTITUDE)),
                                unit=p.get("unit", "degree"),
                                hs_out=hs[i] if i < len(hs) else None,
                                sl_out=sl[i] if i < len(sl) else None,
                                nan_free=bool(p.get("_window_nan_free", False)),
                                step_x_rows=rows)


class CurvatureAlgorithm(NativeTileAlgorithm):
    """_direct_curvature (tile/dask_bridge.py:141-151) -> compute_curvature_block."""
    name = "curvature"

    def get_default_params(self):
        return {"curvature_type": "mean", "pixel_size": 1.0, "mode": "local", "radii": None,
                "weights": None}

    spatial_ok = True

    def run_region(self, win, reg, out, p):
        self.check_mode(p)
        if self.is_spatial(p):
            _run_spatial("curvature", win, reg, out, p)
            return
        sx, sy = curvature_steps(p.get("pixel_size", 1.0), p.get("pixel_scale_x"),
                                 p.get("pixel_scale_y"))
        rows = x_steps_rows(p, win.shape[0], win.device)
        if rows is not None:
            sx = rows.first
        for o in out.descriptors():
            ops.curvature(win, reg, step_x=sx, step_y=sy,
                          curvature_type=p.get("curvature_type", "mean"), out=o,
                          nan_free=bool(p.get("_window_nan_free", False)),
                          step_x_rows=rows.t if rows is not None else None)


class OpennessAlgorithm(NativeTileAlgorithm):
    """_direct_openness (tile/dask_bridge.py:208-222): openness + display stretch."""
    name = "openness"

    def get_default_params(self):
        return {"openness_type": "positive", "num_directions": 16, "max_distance": 50,
                "pixel_size": 1.0, "mode": "local", "radii": None, "weights": None}

    spatial_ok = True

    def run_region(self, win, reg, out, p, stretch=True):
        self.check_mode(p)
        post = _post_from_stats("stretch", p.get("global_stats") if stretch else None)
        if self.is_spatial(p):
            _run_spatial("openness", win, reg, out, p, post=post)
            return
        table = openness_table(p.get("num_directions", 16), p.get("max_distance", 50),
                               p.get("pixel_size", 1.0), p.get("pixel_scale_x"),
                               p.get("pixel_scale_y"))
        for o in out.descriptors(*post):
            ops.openness(win, reg, table=table,
                         openness_type=p.get("openness_type", "positive"), out=o)


class AmbientOcclusionAlgorithm(NativeTileAlgorithm):
    """_direct_ambient_occlusion (tile/dask_bridge.py:190-205): AO + display stretch."""
    name = "ambient_occlusion"

    def get_default_params(self):
        return {"num_samples": 16, "radius": 10.0, "intensity": 1.0, "pixel_size": 1.0,
                "mode": "local", "radii": None, "weights": None}

    spatial_ok = True

    def run_region(self, win, reg, out, p, stretch=True):
        self.check_mode(p)
        post = _post_from_stats("stretch", p.get("global_stats") if stretch else None)
        if self.is_spatial(p):
            _run_spatial("ambient_occlusion", win, reg, out, p, post=post)
            return
        radius = p.get("radius", 10.0)
        table = ao_table(p.get("num_samples", 16), radius, p.get("pixel_size", 1.0),
                         p.get("pixel_scale_x"), p.get("pixel_scale_y"))
        for o in out.descriptors(*post):
            ops.ambient_occlusion(win, reg, table=table, intensity=p.get("intensity", 1.0),
                                  pad_d=max(1, int(round(float(radius)))), out=o)


class CoarsePyramid:
    """Raster-wide coarse means for topousm_fast's decimated radii, resident
    in HBM and shared by every window whose own coarse means they equal.

    Each window of compute_topousm_fast_efficient_block
    (_impl_topousm_fast.py:72-91) decimates itself and box-filters the result;
    for a NaN-free raster, windows on the ds lattice whose sampled cells stay
    clear of their edges (shared_coarse_ok) read the same values from one
    raster-wide grid per (ds, coarse radius) instead.
    """

    def __init__(self, raster: torch.Tensor, decimated: Optional[Dict[int, torch.Tensor]] = None):
        """decimated: {ds: the raster's NaN-free ds x ds cell means}, already
        computed (e.g. by ops.raster_scan); those levels skip their own read."""
        self.raster = raster
        self.decimated = dict(decimated or {})
        self.grids: Dict[Tuple[int, int], torch.Tensor] = {}

    @staticmethod
    def pairs(prm) -> List[Tuple[int, int]]:
        return [(int(prm.ds[i]), coarse_radius(int(prm.radii[i]), int(prm.ds[i])))
                for i in range(prm.n_radii) if prm.ds[i] > 1]

    def require(self, pairs) -> None:
        by_ds: Dict[int, List[int]] = {}
        for ds, rs in pairs:
            if (ds, rs) not in self.grids and rs not in by_ds.setdefault(ds, []):
                by_ds[ds].append(rs)
        todo = [(ds, rss) for ds, rss in by_ds.items() if rss]
        # one batched call (two factors 2x / 4x apart share one read of the
        # raster) where every level fits it, else one call per level
        if todo and all(ds in (4, 8, 16) and len(rss) <= N.COARSE_JOB_MAX_RADII
                        and min(rss) >= 2 for ds, rss in todo):
            res = ops.topousm_coarse_means_batch(
                [(self.decimated[ds], 1, rss) if ds in self.decimated else (self.raster, ds, rss)
                 for ds, rss in todo])
        else:
            res = [ops.topousm_coarse_means(self.raster, ds, rss) for ds, rss in todo]
        for (ds, rss), g in zip(todo, res):
            for i, rs in enumerate(rss):
                self.grids[(ds, rs)] = g[i]

    def attach(self, prm, win_origin, win_shape, reg) -> int:
        """Point prm.mean[i] at the shared grid wherever that is exact;
        returns how many radii were attached."""
        H, W = self.raster.shape
        n = 0
        for i in range(prm.n_radii):
            ds = int(prm.ds[i])
            if ds <= 1:
                continue
            rs = coarse_radius(int(prm.radii[i]), ds)
            g = self.grids.get((ds, rs))
            if g is None or not shared_coarse_ok(win_origin, win_shape,
                                                 (reg.r0, reg.c0, reg.h, reg.w), ds, rs, (H, W)):
                continue
            v = g[int(win_origin[0]) // ds:, int(win_origin[1]) // ds:]
            prm.mean[i] = v.data_ptr()
            prm.mean_ld[i] = int(g.stride(0))
            n += 1
        return n


class TopoUSMFastAlgorithm(NativeTileAlgorithm):
    """_direct_topousm_fast (tile/dask_bridge.py:225-271)."""
    name = "topousm_fast"

    def __init__(self):
        self.ws = ops.Workspace()

    def get_default_params(self):
        return {"mode": "radius", "radii": None, "weights": None, "sigmas": None,
                "agg": "mean", "auto_sigma": False}

    def _determine_optimal_radii(self, pixel_size: float) -> List[int]:
        return topousm_default_radii(pixel_size)

    def params_struct(self, p: dict, win_origin=(0, 0)):
        pixel_size = p.get("pixel_size", 1.0)
        radii = p.get("radii", None)
        weights = p.get("weights", None)
        if radii is None:
            radii = self._determine_optimal_radii(pixel_size)
        field = p.get("_topousm_fast_coarse_field", None)
        if field is not None:
            radii = list(p.get("_topousm_fast_small_radii", []) or [])
            weights = p.get("_topousm_fast_small_weights", None)
        if weights is None:
            weights = [np.float32(1.0 / len(radii))] * len(radii) if radii else []
        elif len(weights) != len(radii):
            raise ValueError(
                f"Length of weights ({len(weights)}) must match length of radii ({len(radii)})")
        ds = [radius_to_downsample_factor(float(r), pixel_size=pixel_size,
                                          algorithm_name="topousm_fast") for r in radii]
        if field is not None:
            if not isinstance(field, torch.Tensor):
                raise TypeError("_topousm_fast_coarse_field must be a GPU tensor")
            off = p.get("_topousm_fast_field_offset", (0, 0))
            full = p.get("_topousm_fast_full_shape", None)
            return ops.topousm_params(radii, weights, ds, field=field, full_shape=full,
                                      win_origin=(int(off[0]), int(off[1])),
                                      w_large=float(p.get("_topousm_fast_w_large", 0.0)))
        return ops.topousm_params(radii, weights, ds)

    def run_region(self, win, reg, out, p, normalize=True):
        # a caller running many windows with the same parameters builds the
        # struct once (params_struct of p) and passes it in
        base = p.get("_topousm_fast_prm")
        prm = type(base).from_buffer_copy(base) if base is not None else self.params_struct(p)
        # the scheduler knows each window's NoData count: a NaN-free window
        # may use the plain-mean kernel variant (identical results)
        if p.get("_window_nan_free", False):
            prm.nan_hint = 0
        pyr = p.get("_topousm_fast_pyramid")
        if pyr is not None and p.get("_window_origin") is not None:
            pyr.attach(prm, p["_window_origin"], tuple(win.shape), reg)
        # this window's own coarse means, built beforehand for many windows at
        # once (ops.topousm_coarse_means_batch): radius index -> grid
        for i, g in (p.get("_topousm_fast_means") or {}).items():
            prm.mean[int(i)] = g.data_ptr()
            prm.mean_ld[int(i)] = int(g.stride(0))
        if prm.field and p.get("_topousm_fast_full_shape") is None:
            prm.full_h, prm.full_w = int(win.shape[0]), int(win.shape[1])
        stats = p.get("global_stats", None) if normalize else None
        if normalize and not isinstance(stats, torch.Tensor) and not (
                isinstance(stats, (tuple, list)) and len(stats) >= 1 and float(stats[0]) > 1e-9):
            logger.warning(
                "topousm_fast: global_stats missing on tile direct path; estimating from this "
                "tile only. Run through the normal pipeline so the pre-pass can provide "
                "seam-free global stats.")
            raw = torch.empty((win.shape[0], win.shape[1]), dtype=torch.float32, device=win.device)
            ops.topousm(win, ops.full_region(win), prm, ops.output(raw), self.ws)
            pct, nval, mom = ops.percentile(raw, [99.0], ops.PCT_ABS)
            stats = ops.finalize_stats(0, pct, mom, nval)
            src = raw[reg.r0:reg.r0 + reg.h, reg.c0:reg.c0 + reg.w]
            for o in out.descriptors(N.POST_DIV, (0.0, 0.0), stats):
                ops.encode(src, ops.full_region(src), o)
            return
        post = _post_from_stats("div", stats) if normalize else (N.POST_NONE, (0.0, 0.0), None)
        descs = out.descriptors(*post)
        for i in range(0, len(descs), 2):
            ops.topousm(win, reg, prm, descs[i], self.ws,
                        out2=descs[i + 1] if i + 1 < len(descs) else None)


class UnsupportedAlgorithm(TileAlgorithm):
    """A reference algorithm outside this build's hot path (registry name kept)."""

    def __init__(self, name: str):
        self.name = name

    def get_default_params(self):
        return {}

    def process(self, dem_gpu, **params):
        raise UnsupportedOnBackend(
            f"algorithm {self.name!r} is not available on the MI355X backend "
            "(hot path: hillshade, slope, curvature, topousm_fast, openness, "
            "ambient_occlusion)")


_NATIVE = {
    "topousm_fast": TopoUSMFastAlgorithm,
    "hillshade": HillshadeAlgorithm,
    "slope": SlopeAlgorithm,
    "curvature": CurvatureAlgorithm,
    "ambient_occlusion": AmbientOcclusionAlgorithm,
    "openness": OpennessAlgorithm,
}
# algorithms/dask_registry.py:27-49 (same names, same order)
_REFERENCE_NAMES = [
    "topousm_fast", "hillshade", "slope", "specular", "atmospheric_scattering",
    "multiscale_terrain", "blur", "curvature", "visual_saliency", "npr_edges",
    "ambient_occlusion", "openness", "fractal_anomaly", "scale_space_surprise",
    "multi_light_uncertainty", "structure_tensor", "frangi", "lic", "phase_congruency",
    "tv_decomposition", "scale_drift",
]
ALGORITHMS: Dict[str, TileAlgorithm] = {
    n: (_NATIVE[n]() if n in _NATIVE else UnsupportedAlgorithm(n)) for n in _REFERENCE_NAMES
}
SUPPORTED = tuple(n for n in _REFERENCE_NAMES if n in _NATIVE)


def load_algorithm(name: str) -> TileAlgorithm:
    """_load_algorithm (core/tile_processor.py:807-820)."""
    if name in _NATIVE:
        return _NATIVE[name]()
    raise ValueError(f"Algorithm {name} not found or not available on this platform")


def run_tile_algorithm(algo_instance, algorithm: str, dem_gpu: torch.Tensor, sigma: float,
                       multiscale_mode: bool, pixel_size: float, algo_params: Dict[str, Any]):
    """core/tile_compute.py:93-129."""
    if algorithm == "topousm_fast":
        radii, w = normalize_topousm_radii_weights(None, None, pixel_size,
                                                   manual_radii=algo_params.get("radii"),
                                                   manual_weights=algo_params.get("weights"))
        params = {"multiscale_mode": multiscale_mode, "radii": radii, "weights": w,
                  "pixel_size": pixel_size, "sigma": sigma}
        for key in ("global_stats", "_topousm_fast_coarse_field", "_topousm_fast_small_radii",
                    "_topousm_fast_small_weights", "_topousm_fast_w_large",
                    "_topousm_fast_full_shape", "_topousm_fast_field_offset"):
            if key in algo_params and algo_params[key] is not None:
                params[key] = algo_params[key]
        return algo_instance.process(dem_gpu, **params)
    params = {"sigma": sigma, "pixel_size": pixel_size, **algo_params}
    return algo_instance.process(dem_gpu, **params)


def apply_nodata_mask(result_gpu: torch.Tensor, mask_nodata, nodata):
    """core/tile_compute.py:132-152 (mask may be a bool tensor or NumPy array)."""
    if mask_nodata is None:
        return result_gpu
    mask = torch.as_tensor(np.asarray(mask_nodata) if not isinstance(mask_nodata, torch.Tensor)
                           else mask_nodata, device=result_gpu.device).bool()
    fill = float(nodata if nodata is not None else 0)
    if result_gpu.dim() == 2:
        result_gpu[mask] = fill
    elif result_gpu.dim() == 3:
        if tuple(result_gpu.shape[:2]) == tuple(mask.shape):
            result_gpu[mask, :] = fill
        elif tuple(result_gpu.shape[-2:]) == tuple(mask.shape):
            result_gpu[:, mask] = fill
        else:
            raise ValueError(
                f"Unsupported result/mask shapes for nodata masking: "
                f"result={tuple(result_gpu.shape)}, mask={tuple(mask.shape)}")
    else:
        raise ValueError(f"Unsupported result ndim for nodata masking: {result_gpu.dim()}")
    return result_gpu

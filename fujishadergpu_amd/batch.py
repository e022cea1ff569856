"""Batched fused tile launches (hillshade + slope + curvature + topousm_fast).

The reference walks its tiles one by one and runs one algorithm per call
(``core/tile_processor.py:1635-1683`` -> ``run_tile_algorithm``,
``core/tile_compute.py:93-129``).  A ``FusedBatch`` plans a whole set of
independent tiles -- ragged sizes, projected or geographic -- once, and runs
every tile's four products through one persistent kernel that drains a device
work queue (``fsg_batch_run``), optionally captured once in a hipGraph and
replayed (``fsg_batch_capture`` / ``fsg_batch_replay``).  ``run_unfused`` is
the per-tile, per-algorithm path through the same algorithm classes, and the
two produce identical bits.  ``tune`` picks the persistent launch shape by
measuring it.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from . import _native as N
from . import ops
from .algorithms import (CurvatureAlgorithm, HillshadeAlgorithm, OutSpec, SlopeAlgorithm,
                         TopoUSMFastAlgorithm, _post_from_stats, curvature_steps, gradient_steps,
                         x_steps_rows)
from .params import DEFAULT_ALTITUDE, DEFAULT_AZIMUTH


@dataclass
class TileJob:
    """One independent tile: its window (the whole tile is the region), the
    reference parameter keys, and float32 output rasters of the tile's shape
    (None = product not wanted)."""
    dem: torch.Tensor
    params: dict = field(default_factory=dict)
    hillshade: Optional[torch.Tensor] = None
    slope: Optional[torch.Tensor] = None
    curvature: Optional[torch.Tensor] = None
    topousm: Optional[torch.Tensor] = None
    nan_free: bool = False


def _empty_out() -> N.fsg_output:
    return N.fsg_output()


def _tile_struct(job: TileJob, keep: list) -> N.fsg_batch_tile:
    """The tile's C descriptor; tensors it points at are appended to keep."""
    p = job.params
    t = N.fsg_batch_tile()
    t.win = ops.window(job.dem)
    t.reg = ops.full_region(job.dem)
    psx, psy = p.get("pixel_scale_x"), p.get("pixel_scale_y")
    rs = x_steps_rows(p, job.dem.shape[0], job.dem.device)
    rows = None
    if rs is not None:
        keep.append(rs)
        psx, rows = rs.first, rs.t
    if job.hillshade is not None or job.slope is not None:
        sx, sy = gradient_steps(p.get("pixel_size", 1.0), psx, psy)
        z = p.get("z_factor", 1.0)
        t.shade = ops.shade_params(
            step_x=sx, step_y=sy, sign_x=1.0 if (psx is None or float(psx) >= 0.0) else -1.0,
            sign_y=1.0 if (psy is None or float(psy) >= 0.0) else -1.0,
            z_factor=1.0 if z is None else float(z),
            azimuth=float(p.get("azimuth", DEFAULT_AZIMUTH)),
            altitude=float(p.get("altitude", DEFAULT_ALTITUDE)), unit=p.get("unit", "degree"),
            nan_free=job.nan_free, step_x_rows=rows)
    t.hillshade = ops.output(job.hillshade, clip01=True) if job.hillshade is not None \
        else _empty_out()
    t.slope = ops.output(job.slope) if job.slope is not None else _empty_out()
    if job.curvature is not None:
        cx, cy = curvature_steps(p.get("pixel_size", 1.0), psx, psy)
        t.curv = ops.curvature_params(step_x=cx, step_y=cy,
                                      curvature_type=p.get("curvature_type", "mean"),
                                      nan_free=job.nan_free, step_x_rows=rows)
        t.curvature = ops.output(job.curvature)
    else:
        t.curvature = _empty_out()
    if job.topousm is not None:
        prm = TopoUSMFastAlgorithm().params_struct(p)
        prm.nan_hint = 0 if job.nan_free else 1
        t.usm = prm
        post = _post_from_stats("div", p.get("global_stats"))
        t.topousm = ops.output(job.topousm, post=post[0], host_stats=post[1], dev_stats=post[2])
    else:
        t.topousm = _empty_out()
    t.topousm2 = _empty_out()
    return t


class FusedBatch:
    """A planned batch (``fsg_batch_create``); keeps the tiles' tensors alive."""

    def __init__(self, jobs: Sequence[TileJob]):
        self.jobs = list(jobs)
        self._keep = []
        arr = (N.fsg_batch_tile * len(self.jobs))(*[_tile_struct(j, self._keep)
                                                    for j in self.jobs])
        h = C.c_void_p()
        N.check(N.lib().fsg_batch_create(arr, len(self.jobs), C.byref(h)))
        self._h = h
        self._arr = arr
        self.config = N.fsg_batch_config(3, 0, 0, 8)

    def info(self) -> Dict[str, int]:
        v = (C.c_int64 * 8)()
        N.check(N.lib().fsg_batch_info(self._h, v))
        keys = ("usm_items", "stencil_items", "fused_usm_tiles", "fallback_usm_tiles",
                "stencil_tiles", "lds_bytes", "box_radii", "device_bytes")
        return dict(zip(keys, [int(x) for x in v]))

    def _cfg(self, config):
        if config is None:
            return self.config
        if isinstance(config, N.fsg_batch_config):
            return config
        wg, order = config[0], config[1]
        chunk = config[2] if len(config) > 2 else 8
        return N.fsg_batch_config(int(wg), int(order), 0, int(chunk))

    def run(self, config=None, stream=None) -> None:
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        N.check(N.lib().fsg_batch_run(self._h, C.byref(self._cfg(config)), s))

    def capture(self, stream: torch.cuda.Stream, config=None) -> None:
        """Record one pass into a hipGraph on a non-default stream."""
        N.check(N.lib().fsg_batch_capture(self._h, C.byref(self._cfg(config)),
                                          stream.cuda_stream))

    def replay(self, stream: torch.cuda.Stream) -> None:
        N.check(N.lib().fsg_batch_replay(self._h, stream.cuda_stream))

    def close(self) -> None:
        if self._h:
            N.lib().fsg_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_unfused(jobs: Sequence[TileJob]) -> None:
    """Per tile, per algorithm: the reference's loop shape, through the same
    algorithm classes the tile scheduler uses."""
    hs, sl, cv, usm = (HillshadeAlgorithm(), SlopeAlgorithm(), CurvatureAlgorithm(),
                       TopoUSMFastAlgorithm())
    for j in jobs:
        reg = ops.full_region(j.dem)
        p = {**j.params, "_window_nan_free": j.nan_free}
        # the stencils run in local mode: TopoUSM's radii are not theirs
        ps = {k: v for k, v in p.items() if k not in ("radii", "weights")}
        ps["mode"] = "local"
        if j.hillshade is not None:
            hs.run_region(j.dem, reg, OutSpec(f32=j.hillshade, clip01=True), ps)
        if j.slope is not None:
            sl.run_region(j.dem, reg, OutSpec(f32=j.slope), ps)
        if j.curvature is not None:
            cv.run_region(j.dem, reg, OutSpec(f32=j.curvature), ps)
        if j.topousm is not None:
            usm.run_region(j.dem, reg, OutSpec(f32=j.topousm), p)


def _time(fn, reps: int) -> float:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune(batch: FusedBatch, candidates: Sequence[Tuple[int, int, int]] = (
        (2, 0, 8), (4, 0, 8), (4, 0, 32), (4, 1, 8), (8, 0, 16)),
         reps: int = 3) -> Tuple[Tuple[int, ...], Dict[Tuple[int, ...], float]]:
    """Measure each (workgroups per CU, item order, items per queue claim) on
    the batch itself and keep the fastest as the batch's configuration."""
    times = {}
    for c in candidates:
        batch.run(c)                       # warm
        times[tuple(c)] = _time(lambda: batch.run(c), reps)
    best = min(times, key=times.get)
    batch.config = batch._cfg(best)
    return best, times

"""Batched fused tile launches (hillshade + slope + curvature + topousm_fast,
or hillshade + slope + curvature + openness).

The reference walks its tiles one by one and runs one algorithm per call
(``core/tile_processor.py:1635-1683`` -> ``run_tile_algorithm``,
``core/tile_compute.py:93-129``).  A ``FusedBatch`` plans a whole set of
independent tiles -- ragged sizes, projected or geographic -- once, and runs
every tile's four products as one batched launch over the NaN-free stencil
tiles' interiors (row-walking items) and one persistent kernel that drains a
device work queue over everything else (``fsg_batch_run``), optionally
captured once in a hipGraph and replayed (``fsg_batch_capture`` /
``fsg_batch_replay``).  ``run_unfused`` is
the per-tile, per-algorithm path through the same algorithm classes, and the
two produce identical bits.  ``tune`` picks the persistent launch shape by
measuring it.

``RasterJob`` is the whole-raster form: one device raster, several
local-mode products (hillshade, slope, curvature, openness), its tiles --
ragged edge tiles included -- planned once and run as one NoData pass, the
openness stretch pre-pass and the batch's launches, which a graph captures
and replays.  Its outputs equal ``scheduler.process_dem``'s for each product
bit for bit.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from . import _native as N
from . import ops
from .algorithms import (CurvatureAlgorithm, HillshadeAlgorithm, OpennessAlgorithm, OutSpec,
                         SlopeAlgorithm, TopoUSMFastAlgorithm, _post_from_stats, curvature_steps,
                         gradient_steps, openness_table, x_steps_rows)
from .params import DEFAULT_ALTITUDE, DEFAULT_AZIMUTH


@dataclass
class TileJob:
    """One independent tile: its window, the region computed (``region`` =
    (r0, c0, h, w) in window coordinates; None: the whole window), the
    reference parameter keys, and float32 outputs of the region's shape
    (None = product not wanted).  ``product_params`` optionally gives a
    product ("hillshade", "slope", "curvature", "openness", "topousm") its
    own parameter dict instead of ``params``; ``openness_stats`` is the
    openness display stretch's (p1, p99 - p1): a tuple, or a float64 device
    tensor filled before the batch runs (None: no stretch)."""
    dem: torch.Tensor
    params: dict = field(default_factory=dict)
    hillshade: Optional[torch.Tensor] = None
    slope: Optional[torch.Tensor] = None
    curvature: Optional[torch.Tensor] = None
    topousm: Optional[torch.Tensor] = None
    nan_free: bool = False
    openness: Optional[torch.Tensor] = None
    region: Optional[Tuple[int, int, int, int]] = None
    product_params: Dict[str, dict] = field(default_factory=dict)
    openness_stats: object = None


def _empty_out() -> N.fsg_output:
    return N.fsg_output()


def _tile_struct(job: TileJob, keep: list) -> N.fsg_batch_tile:
    """The tile's C descriptor; tensors and arrays it points at are appended
    to keep."""
    pp = job.product_params
    p = job.params
    t = N.fsg_batch_tile()
    t.win = ops.window(job.dem)
    t.reg = ops.region(*job.region) if job.region is not None else ops.full_region(job.dem)

    resolved = {}

    def scales(q):
        # one resolution per parameter dict: products sharing it share the
        # per-row table (one tile's stencil reads one)
        if id(q) in resolved:
            return resolved[id(q)]
        psx, psy = q.get("pixel_scale_x"), q.get("pixel_scale_y")
        rs = x_steps_rows(q, job.dem.shape[0], job.dem.device)
        rows = None
        if rs is not None:
            keep.append(rs)
            psx, rows = rs.first, rs.t
        resolved[id(q)] = (psx, psy, rows)
        return resolved[id(q)]

    if job.hillshade is not None or job.slope is not None:
        # one stencil pass gives both: the hillshade keys from the hillshade
        # parameters, the slope unit from the slope parameters
        ph, ps = pp.get("hillshade", p), pp.get("slope", p)
        if job.hillshade is not None and job.slope is not None:
            for k, d in (("pixel_size", 1.0), ("pixel_scale_x", None), ("pixel_scale_y", None)):
                a, b = ph.get(k, d), ps.get(k, d)
                if (d if a is None else a) != (d if b is None else b):
                    raise ValueError(f"hillshade and slope of one tile differ in {k!r}")
        p = ph if job.hillshade is not None else ps
        psx, psy, rows = scales(p)
        sx, sy = gradient_steps(p.get("pixel_size", 1.0), psx, psy)
        z = p.get("z_factor", 1.0)
        t.shade = ops.shade_params(
            step_x=sx, step_y=sy, sign_x=1.0 if (psx is None or float(psx) >= 0.0) else -1.0,
            sign_y=1.0 if (psy is None or float(psy) >= 0.0) else -1.0,
            z_factor=1.0 if z is None else float(z),
            azimuth=float(p.get("azimuth", DEFAULT_AZIMUTH)),
            altitude=float(p.get("altitude", DEFAULT_ALTITUDE)),
            unit=ps.get("unit", "degree"), nan_free=job.nan_free, step_x_rows=rows)
    t.hillshade = ops.output(job.hillshade, clip01=True) if job.hillshade is not None \
        else _empty_out()
    t.slope = ops.output(job.slope) if job.slope is not None else _empty_out()
    if job.curvature is not None:
        p = pp.get("curvature", job.params)
        psx, psy, rows = scales(p)
        cx, cy = curvature_steps(p.get("pixel_size", 1.0), psx, psy)
        t.curv = ops.curvature_params(step_x=cx, step_y=cy,
                                      curvature_type=p.get("curvature_type", "mean"),
                                      nan_free=job.nan_free, step_x_rows=rows)
        t.curvature = ops.output(job.curvature)
    else:
        t.curvature = _empty_out()
    if job.topousm is not None:
        p = pp.get("topousm", job.params)
        prm = TopoUSMFastAlgorithm().params_struct(p)
        prm.nan_hint = 0 if job.nan_free else 1
        t.usm = prm
        post = _post_from_stats("div", p.get("global_stats"))
        t.topousm = ops.output(job.topousm, post=post[0], host_stats=post[1], dev_stats=post[2])
    else:
        t.topousm = _empty_out()
    t.topousm2 = _empty_out()
    if job.openness is not None:
        p = pp.get("openness", job.params)
        table = openness_table(p.get("num_directions", 16), p.get("max_distance", 50),
                               p.get("pixel_size", 1.0), p.get("pixel_scale_x"),
                               p.get("pixel_scale_y"))
        prm, arr = ops.openness_params(table, p.get("openness_type", "positive"), False)
        keep.append(arr)
        t.open = prm
        post = _post_from_stats("stretch", job.openness_stats)
        t.openness = ops.output(job.openness, post=post[0], host_stats=post[1],
                                dev_stats=post[2])
    else:
        t.openness = _empty_out()
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
        v = (C.c_int64 * 12)()
        N.check(N.lib().fsg_batch_info(self._h, v))
        keys = ("usm_items", "stencil_items", "fused_usm_tiles", "fallback_usm_tiles",
                "stencil_tiles", "lds_bytes", "box_radii", "device_bytes", "openness_items",
                "openness_tiles", "row_tiles", "row_items")
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


RASTER_JOB_PRODUCTS = ("hillshade", "slope", "curvature", "openness")


class RasterJob:
    """Several local-mode products of one device raster in one persistent
    launch.

    The reference runs ``process_dem_tiles`` once per algorithm: the NoData
    scan, the normalisation pre-pass (``_compute_norm_stats_tiled``,
    algorithms/_norm_stats.py:176-350, openness only), then its tile loop
    with one ``run_tile_algorithm`` call per tile
    (core/tile_processor.py:1591-1683).  Here the job is planned once on the
    host -- the tile grid of ``scheduler.plan`` (ragged last row and column
    included), every product's skip decision from its own windows' NoData
    counts (the reference's threshold rule), the pre-pass windows -- and
    ``run`` enqueues the device work only:

    1. the NoData sentinel -> NaN pass with the per-window counts;
    2. the openness stretch pre-pass (``stats.RayPrepass``);
    3. NaN fills of skipped tiles;
    4. the batch: row items over the interiors of NaN-free stencil tiles,
       then ONE persistent ``k_batch_fused`` launch over the other stencil
       blocks (NaN-aware ones included) and openness blocks, each tile
       reading one window with the widest product halo -- or, with
       ``set_split(True)`` (the tuner's choice when it measures faster), the
       per-window kernels tile by tile, the calls ``process_dem`` makes.

    ``capture`` records ``run`` in a graph and ``replay`` launches it.  The
    raster is converted in place (``process_dem(copy_input=False)``'s
    contract); replays reconvert nothing new, so every replay gives the same
    bits.  Outputs are float32 rasters; each equals
    ``process_dem(raster, product, ...)`` bit for bit.  Projected rasters only
    (one pixel scale per tile)."""

    def __init__(self, raster: torch.Tensor, products: Dict[str, dict], *, tile_size: int = 4096,
                 nodata: Optional[float] = None, nodata_threshold: float = 1.0,
                 pixel_size: float = 1.0, outputs: Optional[Dict[str, torch.Tensor]] = None,
                 config=None):
        from .scheduler import plan
        from .stats import RayPrepass
        if not products:
            raise ValueError("RasterJob needs at least one product")
        for name in products:
            if name not in RASTER_JOB_PRODUCTS:
                raise ValueError(f"RasterJob products are {RASTER_JOB_PRODUCTS}, not {name!r}")
        if raster.dim() != 2 or raster.dtype != torch.float32 or not raster.is_cuda:
            raise ValueError("raster must be a 2-D float32 GPU tensor")
        H, W = int(raster.shape[0]), int(raster.shape[1])
        self.raster = raster
        self.shape = (H, W)
        plans = {n: plan(raster, n, nodata=nodata, tile_size=tile_size, pixel_size=pixel_size,
                         **dict(kw)) for n, kw in products.items()}
        for n, P in plans.items():
            if str(P.params.get("mode", "local")).lower() != "local":
                raise ValueError(f"RasterJob runs local mode only ({n})")
            if P.params.get("is_geographic_dem") or P.params.get("_geographic_rows"):
                raise ValueError("RasterJob takes projected rasters (one pixel scale per tile)")
            if P.output_dtype != "float32":
                raise ValueError("RasterJob writes float32 outputs")
            if n == "hillshade" and len(HillshadeAlgorithm._direct_radii(P.params)) > 1:
                # _direct_hillshade's multi-radius branch (tile/dask_bridge.py:95-105) is
                # a spatial filter bank, not the 3x3 stencil a batch item runs
                raise ValueError("RasterJob fuses the single-radius hillshade: pass radii=[1] "
                                 f"(resolved radii {P.params.get('radii')})")
        self.plans = plans
        widest = max(plans.values(), key=lambda P: P.padding)
        cores = [tuple(t[:6]) for t in widest.tiles]
        for n, P in plans.items():
            if [tuple(t[:6]) for t in P.tiles] != cores:
                raise ValueError("products plan different tile grids")
        self.tiles = list(widest.tiles)
        # NoData: the sentinel pass once, then each product's windows counted
        # (a product skips a tile by its own window's NoData share)
        win_w = [(t[7], t[6], t[9], t[8]) for t in self.tiles]
        counts_w = ops.nodata_to_nan(raster, nodata, win_w).cpu().numpy()
        counts, skip = {}, {}
        for n, P in plans.items():
            wins = [(t[7], t[6], t[9], t[8]) for t in P.tiles]
            counts[n] = counts_w if wins == win_w else \
                ops.nodata_to_nan(raster, None, wins).cpu().numpy()
            skip[n] = [bool(c > 0 and float(c) / float(t[8] * t[9]) >= nodata_threshold)
                       for c, t in zip(counts[n], P.tiles)]
        raster_nan_free = bool((counts_w == 0).all())
        # the replayed NoData pass: sentinel -> NaN and the window counts
        self._nodata = nodata
        self._win_flat = (C.c_int32 * max(1, 4 * len(win_w)))(
            *[int(v) for wd in win_w for v in wd])
        self._n_win = len(win_w)
        self._counts = torch.zeros(max(1, len(win_w)), dtype=torch.int64, device=raster.device)
        # the openness stretch pre-pass (unless the caller fixed the statistics)
        self.prepass = None
        open_stats = None
        if "openness" in plans:
            given = plans["openness"].params.get("global_stats")
            if given is not None:
                open_stats = given
            else:
                self.prepass = RayPrepass(raster, "openness", plans["openness"].params,
                                          raster_nan_free=raster_nan_free)
                if self.prepass.pool is not None:
                    open_stats = self.prepass.buffers.stats
        self.outputs = {}
        for n in plans:
            o = (outputs or {}).get(n)
            if o is None:
                o = torch.empty((H, W), dtype=torch.float32, device=raster.device)
            elif tuple(o.shape) != (H, W) or o.dtype != torch.float32:
                raise ValueError(f"output {n!r} must be a float32 {H}x{W} tensor")
            self.outputs[n] = o
        # per tile, one tile job per distinct product window (each product
        # reads exactly the window its own tile loop gives it: the results
        # near a core's edge depend on where its window ends)
        self._fills = []
        jobs = []
        for k, t in enumerate(self.tiles):
            ty, tx, cx, cy, cw, ch = t[:6]
            groups: Dict[tuple, list] = {}
            for n, P in plans.items():
                v = self.outputs[n][cy:cy + ch, cx:cx + cw]
                if skip[n][k]:
                    self._fills.append(v)
                    continue
                wx, wy, ww, wh = P.tiles[k][6:10]
                groups.setdefault((wx, wy, ww, wh, int(counts[n][k]) == 0), []).append((n, v))
            if ch == 0 or cw == 0:
                continue
            for (wx, wy, ww, wh, nan_free), members in groups.items():
                views = dict(members)
                pp = {n: {**plans[n].params, "mode": "local"} for n in views}
                jobs.append(TileJob(dem=raster[wy:wy + wh, wx:wx + ww], params={},
                                    hillshade=views.get("hillshade"), slope=views.get("slope"),
                                    curvature=views.get("curvature"),
                                    openness=views.get("openness"), nan_free=nan_free,
                                    region=(cy - wy, cx - wx, ch, cw), product_params=pp,
                                    openness_stats=open_stats))
        self.batch = FusedBatch(jobs) if jobs else None
        if self.batch is not None and config is not None:
            self.batch.config = self.batch._cfg(config)
        self._jobs = jobs
        self._calls = None
        self.split = False
        self.graph = None

    def set_split(self, split: bool) -> None:
        """The job's execution shape after the shared NoData pass and
        pre-pass: ONE persistent launch over every tile's items (False), or
        the per-window kernels tile by tile and product by product (True) --
        the calls ``process_dem`` makes, with this job's windows, statistics
        and NaN flags, so the same bits either way.  The tuner measures both."""
        self.split = bool(split)
        self.graph = None

    def _split_calls(self):
        if self._calls is None:
            from .algorithms import ALGORITHMS, OutSpec
            calls = []
            for j in self._jobs:
                reg = ops.region(*j.region)
                for n in RASTER_JOB_PRODUCTS:
                    v = getattr(j, n)
                    if v is None:
                        continue
                    q = dict(j.product_params[n])
                    q["_window_nan_free"] = bool(j.nan_free)
                    if n == "openness":
                        q["global_stats"] = j.openness_stats
                    spec = OutSpec(f32=v, clip01=(n == "hillshade"))
                    calls.append((ALGORITHMS[n], j.dem, reg, spec, q))
            self._calls = calls
        return self._calls

    def run(self) -> None:
        """Enqueue the whole job on the current stream."""
        self._counts.zero_()
        has = self._nodata is not None
        r = self.raster
        N.check(N.lib().fsg_nodata_to_nan(r.data_ptr(), int(r.stride(0)), self.shape[0],
                                          self.shape[1], int(has),
                                          float(self._nodata) if has else 0.0, self._win_flat,
                                          self._n_win, self._counts.data_ptr(), ops._stream()))
        if self.prepass is not None:
            self.prepass.run()
        for v in self._fills:
            v.fill_(float("nan"))
        if self.split:
            for algo, dem, reg, spec, q in self._split_calls():
                algo.run_region(dem, reg, spec, q)
        elif self.batch is not None:
            self.batch.run()

    def capture(self) -> None:
        """Record ``run`` into a graph (after one eager warm-up pass on the
        capture stream)."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.run()
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s, capture_error_mode="relaxed"):
            self.run()
        self.graph = g

    def replay(self) -> None:
        if self.graph is None:
            raise RuntimeError("RasterJob.replay: nothing captured")
        self.graph.replay()

    def info(self) -> Dict[str, int]:
        d = self.batch.info() if self.batch is not None else {}
        d["split"] = int(self.split)
        d["tiles"] = len(self.tiles)
        d["skipped_products"] = len(self._fills)
        d["prepass_windows"] = len(self.prepass.regions) if self.prepass is not None else 0
        return d

    def close(self) -> None:
        self.graph = None
        if self.batch is not None:
            self.batch.close()


def tune_raster_job(raster: torch.Tensor, products: Dict[str, dict], *,
                    tile_sizes: Sequence[int] = (2048, 4096, 8192), reps: int = 1,
                    launch_candidates: Sequence[Tuple[int, int, int]] = (
                        (4, 0, 8), (4, 1, 8), (6, 0, 8), (8, 0, 8), (8, 1, 8), (8, 0, 32)),
                    **kw) -> Tuple[RasterJob, Dict[str, object]]:
    """The measuring tuner for a RasterJob: each tile size is planned and its
    job timed (one warm run, then ``reps``); the fastest tile size's persistent
    launch shape (workgroups per CU, item order, items per queue claim) is
    then measured the same way.  The halo is not searched: it is each
    product's padding rule (a narrower one would change results).  Returns
    (the tuned job, {"tile_ms": {...}, "launch_ms": {...}, "tile_size": ...,
    "config": ...})."""
    tile_ms = {}
    best_job, best_t = None, None
    for ts in tile_sizes:
        job = RasterJob(raster, products, tile_size=int(ts), **kw)
        job.run()
        t = _time(job.run, reps)
        tile_ms[int(ts)] = t * 1e3
        if best_t is None or t < best_t:
            if best_job is not None:
                best_job.close()
            best_job, best_t = job, t
        else:
            job.close()
    launch_ms = {}
    if best_job.batch is not None:
        best_cfg, times = tune(best_job.batch, launch_candidates, reps=max(1, reps))
        launch_ms = {str(k): v * 1e3 for k, v in times.items()}
    else:
        best_cfg = None
    # the execution shape: the persistent launch at its best shape against
    # the per-window kernels tile by tile (the whole job timed both ways)
    best_job.run()
    t_fused = _time(best_job.run, max(1, reps))
    best_job.set_split(True)
    best_job.run()
    t_split = _time(best_job.run, max(1, reps))
    if t_split >= t_fused:
        best_job.set_split(False)
    best_ts = min(tile_ms, key=tile_ms.get)
    return best_job, {"tile_ms": tile_ms, "launch_ms": launch_ms, "tile_size": best_ts,
                      "config": best_cfg,
                      "shape_ms": {"persistent": t_fused * 1e3, "per_tile": t_split * 1e3},
                      "shape": "per_tile" if best_job.split else "persistent"}

#!/usr/bin/env python3
"""Benchmark: TopoUSM Fast over a 16384 x 16384 float32 fBm DEM on one MI355X.

Workload (the headline configuration): seeded spectral fBm (Hurst 0.75,
relief 0-3776 m, 5 m cells), radii 2, 8, 32, 128, 512, 2048 with the 2^n
weight profile normalised to 1, 4096 px tiles with the reference halo, spatial
mode; outputs float32 and uint8.  One step is the whole job with the DEM
resident in HBM: NoData scan, overview field for the large radii, the
stratified global-statistics pre-pass (p99 on the GPU), and the tiled main
pass writing both encodings.

Prints ONE JSON line.  ``roofline`` is measured with HIP events around every
launch of the dominant kernel (k_usm_main) inside the timed steps;
``cpu_baseline`` times the CPU oracle (NumPy/SciPy restatement of the
reference's tile compute) on a bounded set of the same tile windows.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Gpix/s & %HBM peak per algorithm; surrogate vs exact openness at ≤1° MAE"
HBM_PEAK_GBS = 8000.0
N_PX = 16384
RADII = [2, 8, 32, 128, 512, 2048]
WEIGHTS = [2.0 ** (5 - i) / 63.0 for i in range(6)]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=None,
                    help="raster side: C2 (default 16384) or the stencils workload (32768)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-tiles", type=int, default=0, help="tiles in the CPU sample (0 = auto)")
    ap.add_argument("--only-c2", action="store_true",
                    help="skip the per-algorithm workloads (C1 hillshade+slope, C3 curvature and "
                         "openness, C5 fused batch)")
    ap.add_argument("--workload", choices=["c1", "c3", "c4", "c5", "f1", "f2", "f4", "stencils"],
                    default=None,
                    help="run only this per-algorithm workload and print its JSON (a profiling "
                         "aid; the bench line is the default run)")
    ap.add_argument("--streams", type=int, default=2,
                    help="streams the tiles alternate over (1: the caller's stream only)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed steps, copy a 8192x8192 float32 region once with "
                         "k_copy_region (256 MiB read + 256 MiB written): the reference point "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_usm_main_topousm_16384_traffic.md")


def measured_traffic():
    """HBM bytes per k_usm_main launch from the committed counter summary
    (the line 'traffic_bytes_per_launch: N'), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1].strip().split()[0])
    except (OSError, ValueError):
        pass
    return None


def make_dem(n, device):
    from fujishadergpu_amd import synth
    return synth.fbm(n, n, 0.75, 3776.0, seed=2, device=device)


def run_step(dem, out_u8, out_f32, streams=2):
    from fujishadergpu_amd.scheduler import process_dem
    return process_dem(dem, "topousm_fast", tile_size=4096, pixel_size=5.0, mode="spatial",
                       radii=RADII, weights=WEIGHTS, output_dtype="uint8", out=out_u8,
                       out_f32=out_f32, copy_input=False, return_plan=True, streams=streams)


_CPU_JOB = {}


def _cpu_tile(t):
    """One tile window of the oracle restatement, in a worker process (the
    raster and the coarse field memory-mapped from the files the parent
    wrote)."""
    import numpy as np
    from oracle import ref
    job = _CPU_JOB
    if "host" not in job:
        job["host"] = np.load(job["host_path"], mmap_mode="r")
        job["field"] = np.load(job["field_path"])
    host, p = job["host"], job["params"]
    ty, tx, cx, cy, cw, ch, wx, wy, ww, wh = t
    win = np.ascontiguousarray(host[wy:wy + wh, wx:wx + ww])
    out = ref.topousm_direct(win, radii=p["radii"], weights=p["weights"], pixel_size=5.0,
                             coarse_field=job["field"], small_radii=p["small_radii"],
                             small_weights=p["small_weights"], w_large=p["w_large"],
                             field_offset=(wy, wx), full_shape=tuple(host.shape),
                             stats=(job["scale"],))
    return t, out[cy - wy:cy - wy + ch, cx - wx:cx - wx + cw]


def _cpu_init(job):
    _CPU_JOB.clear()
    _CPU_JOB.update(job)


def cpu_baseline(dem, plan, stats_scale, gpu_f32, n_tiles):
    """The oracle's per-tile compute on the host: one worker process per
    available core (os.sched_getaffinity), each taking whole tile windows,
    plus the same restatement on one core for one tile."""
    import multiprocessing as mp
    import tempfile
    import numpy as np
    from oracle import ref

    host = dem.cpu().numpy()
    lr, lw = plan.extras["large"]
    coarse, scale = ref.overview_block_mean(host)
    field = ref.topousm_large_coarse_field(coarse, large_radii=lr, large_weights=lw,
                                           decimation=scale)
    p = plan.params
    tiles = plan.tiles
    interior = [t for t in tiles if 0 < t[0] < max(x[0] for x in tiles)
                and 0 < t[1] < max(x[1] for x in tiles)]
    cores = len(os.sched_getaffinity(0))
    if n_tiles <= 0:
        n_tiles = min(len(tiles), cores)
    sample = (interior + [t for t in tiles if t not in interior])[:n_tiles]
    tmp = tempfile.mkdtemp(prefix="fsg_cpu_")
    host_path, field_path = os.path.join(tmp, "dem.npy"), os.path.join(tmp, "field.npy")
    np.save(host_path, host)
    np.save(field_path, field)
    job = {"host_path": host_path, "field_path": field_path, "scale": stats_scale,
           "params": {"radii": p["radii"], "weights": p["weights"],
                      "small_radii": p["_topousm_fast_small_radii"],
                      "small_weights": p["_topousm_fast_small_weights"],
                      "w_large": p["_topousm_fast_w_large"]}}
    workers = max(1, min(cores, len(sample)))
    try:
        ctx = mp.get_context("spawn")
        with ctx.Pool(workers, initializer=_cpu_init, initargs=(job,)) as pool:
            pool.map(abs, range(workers))            # workers started and imported
            t0 = time.perf_counter()
            res = pool.map(_cpu_tile, sample, chunksize=1)
            wall = time.perf_counter() - t0
        # one core, one tile
        _cpu_init(job)
        t0 = time.perf_counter()
        _cpu_tile(sample[0])
        wall1 = time.perf_counter() - t0
    finally:
        _CPU_JOB.clear()
        for f in (host_path, field_path):
            if os.path.exists(f):
                os.remove(f)
        os.rmdir(tmp)
    px = sum(t[4] * t[5] for t in sample)
    # parity of the GPU float32 output on the sampled cores
    g = gpu_f32.cpu().numpy()
    maxd = 0.0
    for t, core in res:
        ty, tx, cx, cy, cw, ch = t[:6]
        d = np.abs(g[cy:cy + ch, cx:cx + cw].astype(np.float64) - core)
        maxd = max(maxd, float(np.nanmax(d)))
    return {
        "value": px / wall / 1e9, "unit": "Gpix/s", "cores": workers, "kind": "port",
        "host_cores_available": cores, "host_cpu_count": os.cpu_count(),
        "value_1_core": sample[0][4] * sample[0][5] / wall1 / 1e9,
        "sample": f"{len(sample)} tile windows (cores {sample[0][4]}x{sample[0][5]} px, "
                  f"halo {plan.padding}) of the 16384^2 job, oracle topousm_direct with the "
                  f"GPU-computed p99 scale; {workers} worker processes, one tile each at a "
                  f"time; value_1_core: one tile on one core",
        "seconds": wall, "seconds_1_core": wall1, "max_abs_diff_vs_gpu": maxd,
    }


# ---------------------------------------------------------------- per algorithm

def _gpu_time(fn, reps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def bench_c1(dev):
    """Hillshade + slope, local, 2048^2 fBm (H 0.8, 0-3000 m, 10 m, seed 1):
    one fused launch; the oracle's hillshade + slope timed on the host."""
    import numpy as np
    import torch
    from fujishadergpu_amd import ops, synth
    from fujishadergpu_amd.algorithms import HillshadeSlopeFused, OutSpec
    n = 2048
    dem = synth.fbm(n, n, 0.8, 3000.0, seed=1, device=dev)
    hs = torch.empty_like(dem)
    sl = torch.empty_like(dem)
    algo = HillshadeSlopeFused()
    p = {"pixel_size": 10.0, "azimuth": 315.0, "altitude": 45.0, "_window_nan_free": True}
    run = lambda: algo.run_region(dem, ops.full_region(dem), OutSpec(f32=hs, clip01=True), p,  # noqa
                                  out_slope=OutSpec(f32=sl))
    run()
    t = _gpu_time(run, 50)
    # the launch's own duration (HIP events on its stream), without host overhead
    from fujishadergpu_amd import _native as N
    L = N.lib()
    tot, cnt = ctypes.c_double(), ctypes.c_int64()
    L.fsg_timing_read(b"k_hillshade_slope", ctypes.byref(tot), ctypes.byref(cnt))
    L.fsg_timing_enable(1)
    for _ in range(50):
        run()
    torch.cuda.synchronize()
    L.fsg_timing_enable(0)
    N.check(L.fsg_timing_read(b"k_hillshade_slope", ctypes.byref(tot), ctypes.byref(cnt)))
    tk = tot.value / 1e3 / max(1, cnt.value)
    from oracle import ref
    host = dem.cpu().numpy()
    t0 = time.perf_counter()
    ehs = ref.hillshade(host, pixel_size=10.0)
    esl = ref.slope(host, pixel_size=10.0)
    tc = time.perf_counter() - t0
    d = max(float(np.abs(hs.cpu().numpy() - ehs).max()), float(np.abs(sl.cpu().numpy() - esl).max()))
    px = n * n
    return {"workload": "hillshade+slope fused, local, 2048x2048 f32 fBm (H=0.8, 0-3000 m, 10 m, "
                        "seed 1), no NoData", "gpix_s": round(px / t / 1e9, 3),
            "ms": round(t * 1e3, 4), "hbm_frac": round(12 * px / t / (HBM_PEAK_GBS * 1e9), 4),
            "alg_bytes_per_px": 12, "kernel_ms": round(tk * 1e3, 4),
            "kernel_gpix_s": round(px / tk / 1e9, 3),
            "kernel_hbm_frac": round(12 * px / tk / (HBM_PEAK_GBS * 1e9), 4),
            "cpu_baseline": {"value": round(px / tc / 1e9, 4), "unit": "Gpix/s", "cores": 1,
                             "kind": "port", "sample": "whole 2048^2 raster, oracle hillshade + "
                                                       "slope (NumPy, one thread of calls)"},
            "max_abs_diff_vs_oracle": d}


def _kernel_ms(name: bytes, fn, reps: int) -> float:
    """Average duration of the launches timed under ``name`` (HIP events on
    the launching stream) over ``reps`` calls of fn."""
    import torch
    from fujishadergpu_amd import _native as N
    L = N.lib()
    tot, cnt = ctypes.c_double(), ctypes.c_int64()
    L.fsg_timing_read(name, ctypes.byref(tot), ctypes.byref(cnt))
    L.fsg_timing_enable(1)
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    L.fsg_timing_enable(0)
    N.check(L.fsg_timing_read(name, ctypes.byref(tot), ctypes.byref(cnt)))
    return tot.value / max(1, cnt.value)


def bench_stencils(dev, n=32768):
    """Hillshade + slope (one fused launch) and curvature (mean), local mode,
    over a whole 32768^2 f32 fBm (H 0.8, 0-3000 m, 10 m, seed 1) as one
    window, NaN-free and with 1% NoData discs (seed 11): the kernel-only
    duration (HIP events) against the algorithmic bytes (one read of the
    raster, one float32 write per product)."""
    import torch
    from fujishadergpu_amd import ops, synth
    from fujishadergpu_amd.algorithms import CurvatureAlgorithm, HillshadeSlopeFused, OutSpec
    dem = synth.fbm(n, n, 0.8, 3000.0, seed=1, device=dev)
    hs = torch.empty_like(dem)
    sl = torch.empty_like(dem)
    res = {}
    px = n * n
    for label in ("nan_free", "nodata_1pct"):
        if label == "nodata_1pct":
            synth.apply_discs(dem, synth.nodata_discs(n, n, 0.01, seed=11))
        nan_free = label == "nan_free"
        reg = ops.full_region(dem)
        p = {"pixel_size": 10.0, "azimuth": 315.0, "altitude": 45.0, "curvature_type": "mean",
             "_window_nan_free": nan_free}
        shade = HillshadeSlopeFused()
        curv = CurvatureAlgorithm()
        run_hs = lambda: shade.run_region(dem, reg, OutSpec(f32=hs, clip01=True), p,  # noqa
                                          out_slope=OutSpec(f32=sl))
        run_cv = lambda: curv.run_region(dem, reg, OutSpec(f32=hs), p)  # noqa
        for name, fn, key, nbytes in ((b"k_hillshade_slope", run_hs, "hillshade_slope", 12),
                                      (b"k_curvature", run_cv, "curvature", 8)):
            fn()
            ms = _kernel_ms(name, fn, 5)
            res[f"{key}_{label}"] = {
                "kernel_ms": round(ms, 3), "gpix_s": round(px / (ms / 1e3) / 1e9, 2),
                "alg_bytes_per_px": nbytes,
                "hbm_frac": round(nbytes * px / (ms / 1e3) / (HBM_PEAK_GBS * 1e9), 4)}
    res["workload"] = (f"{n}x{n} f32 fBm (H=0.8, 0-3000 m, 10 m, seed 1) as one window, local "
                       "mode; NoData case: 1% discs r 5-50 px (seed 11) as NaN")
    del dem, hs, sl
    torch.cuda.empty_cache()
    return res


def _c3_dem(dev, h, w):
    from fujishadergpu_amd import synth
    dem = synth.fbm(h, w, 0.85, 2500.0, seed=3, device=dev)
    synth.apply_discs(dem, synth.nodata_discs(h, w, 0.01, seed=33), -9999.0)
    return dem


C3_PRODUCTS = {"curvature": {"curvature_type": "mean", "mode": "local"},
               "openness": {"openness_type": "positive", "num_directions": 8,
                            "max_distance": 64, "mode": "local"}}


def bench_c3(dev, n=32768):
    """Curvature (mean) + openness (positive, N=8, D=64), local mode, 10 m
    cells, on a 32768^2 fBm (H 0.85, 0-2500 m, seed 3) with 1% NoData discs
    (seed 33): the whole job -- NoData pass, openness stretch pre-pass, and
    ONE persistent launch over every tile's curvature and openness blocks --
    captured in a graph and replayed.  The tuner measures the tile size (the
    halo is each product's padding rule) and the persistent launch shape.
    The same products through the per-tile scheduler (one launch per tile and
    product) are timed beside it, and a 30001 x 32767 raster gives the ragged
    grid (last tile row 1329 px, last column 4095 px) at the 4096 core."""
    import torch
    from fujishadergpu_amd.batch import RasterJob, tune_raster_job
    from fujishadergpu_amd.scheduler import process_dem
    res = {}
    dem = _c3_dem(dev, n, n)
    print("#   c3 tuning tile size and launch shape", file=sys.stderr, flush=True)
    job, rep = tune_raster_job(dem, C3_PRODUCTS, tile_sizes=(2048, 4096, 8192), reps=1,
                               nodata=-9999.0, pixel_size=10.0)
    # both execution shapes captured and replayed; the job keeps the tuner's
    shape_graph_ms = {}
    chosen = job.split
    for split in (not chosen, chosen):
        job.set_split(split)
        job.capture()
        shape_graph_ms["per_tile" if split else "persistent"] = \
            round(_gpu_time(job.replay, 3) * 1e3, 2)
    t = shape_graph_ms[rep["shape"]] / 1e3
    px = n * n
    res["graph_job"] = {"gpix_s": round(px / t / 1e9, 3), "ms": round(t * 1e3, 2),
                        "hbm_frac": round(12 * px / t / (HBM_PEAK_GBS * 1e9), 4),
                        "alg_bytes_per_px": 12, "tile_size": rep["tile_size"],
                        "shape": rep["shape"], "graph_ms_by_shape": shape_graph_ms,
                        "launch_config": list(rep["config"]) if rep["config"] else None,
                        "tuner_tile_ms": {str(k): round(v, 2) for k, v in rep["tile_ms"].items()},
                        "tuner_launch_ms": {k: round(v, 2) for k, v in rep["launch_ms"].items()},
                        "tuner_shape_ms": {k: round(v, 2) for k, v in rep["shape_ms"].items()},
                        "items": job.info()}
    outs = job.outputs
    job.close()
    del job
    # the per-tile scheduler, product by product, at the same tile size
    ref_ms = {}
    for name, kw in C3_PRODUCTS.items():
        o = torch.empty_like(dem)

        def run(_n=name, _kw=kw, _o=o):
            process_dem(dem, _n, tile_size=rep["tile_size"], pixel_size=10.0, nodata=-9999.0,
                        out=_o, copy_input=False, **_kw)
        run()
        ref_ms[name] = _gpu_time(run, 2) * 1e3
        res.setdefault("same_bits_as_per_tile", True)
        a, b = o, outs[name]
        same = bool(torch.equal(torch.isnan(a), torch.isnan(b)) and
                    torch.equal(torch.nan_to_num(a, 0.0), torch.nan_to_num(b, 0.0)))
        res["same_bits_as_per_tile"] = res["same_bits_as_per_tile"] and same
        del o
    tot = sum(ref_ms.values()) / 1e3
    res["per_tile_launches"] = {"ms": {k: round(v, 2) for k, v in ref_ms.items()},
                                "gpix_s_both": round(px / tot / 1e9, 3)}
    del dem, outs
    torch.cuda.empty_cache()
    # the ragged grid
    hr, wr = 30001, 32767
    dem = _c3_dem(dev, hr, wr)
    job = RasterJob(dem, C3_PRODUCTS, tile_size=4096, nodata=-9999.0, pixel_size=10.0,
                    config=rep["config"])
    job.set_split(rep["shape"] == "per_tile")
    job.capture()
    t = _gpu_time(job.replay, 3)
    res["graph_job_ragged"] = {"raster": [hr, wr], "tile_size": 4096, "shape": rep["shape"],
                               "gpix_s": round(hr * wr / t / 1e9, 3), "ms": round(t * 1e3, 2),
                               "tiles": job.info()["tiles"]}
    job.close()
    del dem, job
    torch.cuda.empty_cache()
    res["workload"] = (f"{n}x{n} f32 fBm (H=0.85, 0-2500 m, 10 m, seed 3), 1% NoData discs "
                       "r 5-50 px (seed 33, sentinel -9999), local mode; curvature (mean) + "
                       "openness (positive, 8 directions, 64 px) as one job")
    return res


def c5_jobs(dev, n_tiles=256):
    """256 independent tiles, side log-uniform in [256, 4096] (seeds 50000+i),
    fBm H ~ U[0.6, 0.9], relief ~ U[500, 3000] m; even tiles projected 10 m,
    odd tiles geographic 1 arc-second centred at a latitude ~ U[30, 60] N, every
    row with the x cell width of its own latitude.  Each tile
    is synthesised at the next power of two and cropped (five FFT sizes in
    all, so the FFT library compiles five kernels, not hundreds)."""
    import math

    import numpy as np
    import torch
    from fujishadergpu_amd import synth
    from fujishadergpu_amd.batch import TileJob
    from fujishadergpu_amd.algorithms import RowScales
    from fujishadergpu_amd.raster_info import meters_per_degree, row_geographic_scales_x
    jobs = []
    for i in range(n_tiles):
        r = np.random.default_rng(50000 + i)
        side = int(round(math.exp(r.uniform(math.log(256), math.log(4096)))))
        hurst, relief = r.uniform(0.6, 0.9), r.uniform(500.0, 3000.0)
        big = 1 << max(0, (side - 1).bit_length())
        dem = synth.fbm(big, big, hurst, 1.0, seed=50000 + i, device=dev)[:side, :side]
        lo, hi = dem.min(), dem.max()
        dem = ((dem - lo) / torch.clamp(hi - lo, min=1e-30) * float(relief)).contiguous()
        if i % 32 == 31:
            print(f"#   c5 tiles synthesised: {i + 1}", file=sys.stderr, flush=True)
        p = {"radii": [2, 8, 32, 128], "weights": [8 / 15, 4 / 15, 2 / 15, 1 / 15],
             "curvature_type": "mean", "azimuth": 315.0, "altitude": 45.0}
        if i % 2 == 0:
            p["pixel_size"] = 10.0
        else:
            # 1 arc-second rows, each with the cell width of its own latitude
            lat_c = float(r.uniform(30.0, 60.0))
            m_lon, m_lat = meters_per_degree(lat_c)
            sx, sy = m_lon / 3600.0, m_lat / 3600.0
            rows = RowScales(row_geographic_scales_x(1.0 / 3600.0, -1.0 / 3600.0,
                                                     lat_c + 0.5 * side / 3600.0, 0, side),
                             dev)
            p.update(pixel_scale_x_rows=rows, pixel_scale_y=-sy, pixel_size=0.5 * (sx + sy))
        e = lambda: torch.empty((side, side), dtype=torch.float32, device=dev)  # noqa: E731
        jobs.append(TileJob(dem=dem, params=p, hillshade=e(), slope=e(), curvature=e(),
                            topousm=e(), nan_free=True))
    return jobs


def bench_c5(dev):
    """Fused batch (the batched wavefront main pass, the row-walking stencil
    interiors and one persistent launch over the rest, as one hipGraph
    replay) vs per-tile, per-algorithm launches over the 256 ragged tiles."""
    import torch
    from fujishadergpu_amd.batch import FusedBatch, run_unfused, tune
    from fujishadergpu_amd.stats import norm_stats_prepass
    jobs = c5_jobs(dev)
    # TopoUSM normalisation: one p99 scale (pre-pass over the first tile) for all
    st = norm_stats_prepass(jobs[0].dem, "topousm_fast",
                            {"radii": [2, 8, 32, 128], "weights": jobs[0].params["weights"],
                             "pixel_size": jobs[0].params["pixel_size"]})
    for j in jobs:
        j.params["global_stats"] = st
    px = sum(j.dem.numel() for j in jobs)
    print("#   c5 unfused per-tile runs", file=sys.stderr, flush=True)
    run_unfused(jobs)
    t_unf = _gpu_time(lambda: run_unfused(jobs), 2)
    sample = [0, 1, 2, 3, len(jobs) // 2, len(jobs) - 1]
    ref_out = {i: [t.clone() for t in (jobs[i].hillshade, jobs[i].slope, jobs[i].curvature,
                                       jobs[i].topousm)] for i in sample}
    print("#   c5 fused batch: plan, tune, capture", file=sys.stderr, flush=True)
    b = FusedBatch(jobs)
    info = b.info()
    best, times = tune(b, reps=2)
    t_eager = times[best]
    s = torch.cuda.Stream()
    b.capture(s)
    for j in (jobs[i] for i in sample):
        for t in (j.hillshade, j.slope, j.curvature, j.topousm):
            t.fill_(-7.0)
    b.replay(s)
    s.synchronize()
    same = all(torch.equal(a.view(torch.int32), c.view(torch.int32))
               for i in sample for a, c in zip(ref_out[i], (jobs[i].hillshade, jobs[i].slope,
                                                            jobs[i].curvature, jobs[i].topousm)))

    def replay():
        b.replay(s)
        s.synchronize()
    t_graph = _gpu_time(replay, 5)
    b.close()
    res = {"workload": "256 tiles, side log-U[256, 4096] (seeds 50000+i), fBm H~U[0.6,0.9], "
                       "half projected 10 m, half geographic 1 arc-second at 30-60 N (per-row "
                       "cos(lat) cell width); hillshade+slope+curvature(mean)+topousm{2,8,32,128}, "
                       "f32",
           "tiles": len(jobs), "mpx": round(px / 1e6, 1),
           "gpix_s": round(px / t_graph / 1e9, 3), "ms_graph": round(t_graph * 1e3, 3),
           "hbm_frac": round(20 * px / t_graph / (HBM_PEAK_GBS * 1e9), 4), "alg_bytes_per_px": 20,
           "ms_fused_eager": round(t_eager * 1e3, 3), "ms_unfused_per_tile": round(t_unf * 1e3, 3),
           "speedup_vs_unfused": round(t_unf / t_graph, 2),
           "tuned_config": {"wg_per_cu": best[0], "order": best[1], "chunk": best[2]},
           "tuner_ms": {"/".join(map(str, k)): round(v * 1e3, 3) for k, v in times.items()},
           "batch": info, "identical_to_unfused": bool(same)}
    del jobs, ref_out
    torch.cuda.empty_cache()
    return res


def bench_f1(dev, n=16384):
    """Spatial mode (the CLI default mode) over a 16384^2 raster with the
    automatic radii of that raster (2, 8, 32, 128, 512; 2**n weights), tiled
    4096 with the reference's padding rule, as whole jobs: hillshade (the
    headline line: Gaussian radii, a decimated radius and an overview radius),
    then slope, curvature, openness and ambient occlusion (the last two with
    their p1/p99 display-stretch pre-pass); the oracle on one bounded window for
    the CPU rate."""
    import torch
    from fujishadergpu_amd.scheduler import process_dem
    dem = make_dem(n, dev)
    kw = dict(mode="spatial", tile_size=4096, pixel_size=5.0)
    per = {}
    P = None
    for algo in ("hillshade", "slope", "curvature", "openness", "ambient_occlusion"):
        out, _, Pa = process_dem(dem, algo, return_plan=True, **kw)
        torch.cuda.synchronize()
        ta = _gpu_time(lambda: process_dem(dem, algo, out=out, **kw), 3)
        per[algo] = {"ms": round(ta * 1e3, 2), "gpix_s": round(n * n / ta / 1e9, 3),
                     "halo": Pa.padding}
        print(f"#   f1 {algo}: {ta * 1e3:.2f} ms", file=sys.stderr, flush=True)
        if algo == "hillshade":
            P, t = Pa, ta
        del out
    # CPU: the oracle's spatial window for one 1024^2 core of the same radii
    from oracle import ref
    pad = P.padding
    host = dem[:1024 + 2 * pad, :1024 + 2 * pad].cpu().numpy()
    prm = {k: v for k, v in P.params.items()
           if not isinstance(v, (torch.Tensor, dict)) and not str(k).startswith("_")}
    small = [r for r in P.params["radii"] if r <= 256]
    prm_s = dict(prm, radii=small, weights=None)
    t0 = time.perf_counter()
    ref.spatial_window("hillshade", host, prm_s)
    tc = time.perf_counter() - t0
    del dem
    torch.cuda.empty_cache()
    return {"workload": f"spatial-mode hillshade, {n}x{n} f32 fBm (H=0.75, 0-3776 m, 5 m, "
                        f"seed 2), automatic radii {P.params['radii']}, tile 4096 + halo "
                        f"{P.padding}, overview decimation "
                        f"{P.params.get('_overview_decimation')}",
            "gpix_s": round(n * n / t / 1e9, 3), "ms": round(t * 1e3, 2),
            "algorithms": per,
            "cpu_baseline": {"value": round(1024 * 1024 / tc / 1e9, 5), "unit": "Gpix/s",
                             "cores": 1, "kind": "port",
                             "sample": f"oracle spatial window, one 1024^2 core (+{pad} halo), "
                                       f"the Gaussian/decimated radii {small}"}}


def bench_f2(dev, n=16384):
    """COG output of a 16384^2 float32 result raster resident in HBM (ZSTD level
    1, 512 tiles, predictor 3, AVERAGE overviews while a level exceeds one
    tile, BigTIFF) to local disk, then 4096^2 window reads back to the GPU;
    the oracle's pure-Python decoder on one level-2 overview for the CPU rate."""
    import os
    import tempfile
    import torch
    from fujishadergpu_amd import cogio
    z = make_dem(n, dev)
    d = tempfile.mkdtemp(prefix="fsg_cog_")
    path = os.path.join(d, "out.tif")
    rep = cogio.write_cog(path, z, transform=(5.0, 0.0, 300000.0, 0.0, -5.0, 4000000.0),
                          epsg=32654, nodata=float("nan"))
    t_w = _gpu_time(lambda: cogio.write_cog(path, z, transform=(5.0, 0.0, 300000.0, 0.0, -5.0,
                                                                4000000.0),
                                            epsg=32654, nodata=float("nan")), 2)
    size = os.path.getsize(path)
    rd = cogio.CogReader(path)
    win = torch.empty((4096, 4096), dtype=torch.float32, device=dev)
    rd.read((0, 0, 4096, 4096), out=win)
    reps = 4
    t_r = _gpu_time(lambda: [rd.read((4096 * (k % 4), 4096 * (k // 4 % 4), 4096, 4096), out=win)
                             for k in range(reps)], 1) / reps
    ok = bool(torch.equal(rd.read((4096, 8192, 4096, 4096)), z[8192:12288, 4096:8192]))
    rd.close()
    from oracle import tiff
    t0 = time.perf_counter()
    P = tiff.parse(path)
    tiff.decode(path, 2, P)
    tc = time.perf_counter() - t0
    ov2 = P["ifds"][2]["tags"][256][0] * P["ifds"][2]["tags"][257][0]
    os.remove(path)
    os.rmdir(d)
    raw = n * n * 4
    del z, win
    torch.cuda.empty_cache()
    return {"workload": f"COG write of a {n}x{n} float32 fBm raster in HBM (ZSTD 1, 512 tiles, "
                        f"predictor 3, {rep['levels'] - 1} AVERAGE overviews, BigTIFF) to local "
                        "disk; 4096^2 window reads back to HBM",
            "write_ms": round(t_w * 1e3, 1), "write_gb_s": round(raw / t_w / 1e9, 3),
            "write_gpix_s": round(n * n / t_w / 1e9, 3),
            "file_mb": round(size / 2 ** 20, 1), "ratio": round(rep["raw_bytes"] / size, 3),
            "read_window_ms": round(t_r * 1e3, 2),
            "read_gpix_s": round(4096 * 4096 / t_r / 1e9, 3), "read_exact": ok,
            "host_threads": os.cpu_count(),
            "cpu_baseline": {"value": round(ov2 / tc / 1e9, 5), "unit": "Gpix/s (decode)",
                             "cores": 1, "kind": "port",
                             "sample": "oracle pure-Python decode of the level-2 overview"}}


def _prepare_raster(n, dev, seed):
    """fBm (H 0.8, 300-2800 m) with an undeclared 0 frame (n/25 wide on the
    top and left edges), a declared -9999 block and 0.5 % NaN discs (r 20-200)."""
    from fujishadergpu_amd import synth
    z = synth.fbm(n, n, 0.8, 2500.0, seed=seed, device=dev) + 300.0
    b = max(8, n // 25)
    z[:b] = 0.0
    z[:, :b] = 0.0
    z[n // 2: n // 2 + n // 16, n // 3: n // 3 + n // 12] = -9999.0
    synth.apply_discs(z, synth.nodata_discs(n, n, 0.005, seed=seed + 1, r_min=max(2, n // 800),
                                            r_max=max(4, n // 80)), float("nan"))
    return z


def bench_f4(dev, n=16384):
    """prepare (NoData inference + push-pull fill, fill_mode 'enclosed') of a
    raster resident in HBM, as a whole job; the oracle's prepare on the same
    construction at 2048^2 (coarse grid 256^2, the same ratio) for the CPU rate."""
    import torch
    from fujishadergpu_amd import prepare as P
    z = _prepare_raster(n, dev, 5)
    out = torch.empty_like(z)
    _, info = P.preprocess_dem(z, src_nodata=-9999.0, out=out)
    torch.cuda.synchronize()
    t = _gpu_time(lambda: P.preprocess_dem(z, src_nodata=-9999.0, out=out), 5)
    nodata_in = torch.isnan(z) | (z == -9999.0) | (z == 0.0)
    filled = int((nodata_in & torch.isfinite(out)).sum())
    from oracle import prepare as prep
    m = 2048
    zs = _prepare_raster(m, dev, 5).cpu().numpy()
    t0 = time.perf_counter()
    prep.prepare(zs, src_nodata=-9999.0, coarse_max=256)
    tc = time.perf_counter() - t0
    del z, out
    torch.cuda.empty_cache()
    return {"workload": f"prepare, {n}x{n} f32 fBm (H 0.8, 300-2800 m, seed 5) with an "
                        f"undeclared 0 frame, a declared -9999 block and 0.5 % NaN discs; "
                        f"fill_mode enclosed, coarse_max 2048",
            "gpix_s": round(n * n / t / 1e9, 3), "ms": round(t * 1e3, 2),
            "coarse_shape": list(info["coarse_shape"]), "oversample": info["oversample"],
            "detected": info["detected"], "n_void_coarse": info["n_void"],
            "edge_sweeps": info["sweeps"], "pixels_filled": filled,
            "cpu_baseline": {"value": round(m * m / tc / 1e9, 5), "unit": "Gpix/s", "cores": 1,
                             "kind": "port",
                             "sample": f"oracle prepare of the same construction at {m}x{m} "
                                       "(coarse_max 256)"}}


def bench_c4(dev, n=65536, weights=None):
    """Learned openness surrogate at radius 1024 (16 directions, 10 m) on an
    n^2 fBm (H 0.7, 0-4000 m, seed 4) against the exact kernel: time of each,
    the 1%-sample judge with exact fallback, full-raster error, and the
    held-out accuracy on 64 seeded tiles (seeds 5000-5063)."""
    import torch
    from fujishadergpu_amd import ops, synth
    from fujishadergpu_amd import surrogate as S
    path = weights or S.WEIGHTS_FILE
    net = S.Net(dev)
    net.load(path)
    print("#   c4 held-out tiles", file=sys.stderr, flush=True)
    te = S.make_tiles(range(5000, 5064), dev)
    held = S.evaluate(net, te)
    del te
    torch.cuda.empty_cache()
    print(f"#   c4 synthesising {n}^2", file=sys.stderr, flush=True)
    dem = synth.fbm(n, n, 0.7, 4000.0, seed=4, device=dev)
    torch.cuda.empty_cache()
    ex = torch.empty_like(dem)
    reg = ops.full_region(dem)
    print("#   c4 exact kernel", file=sys.stderr, flush=True)
    t_exact = _gpu_time(lambda: S.exact_openness(dem, reg, 1024, 16, 10.0, out=ex), 1)
    sur = torch.empty_like(dem)
    print("#   c4 surrogate", file=sys.stderr, flush=True)
    S.surrogate_openness(net, dem, 1024, 16, 10.0, out=sur)
    t_sur = _gpu_time(lambda: S.surrogate_openness(net, dem, 1024, 16, 10.0, out=sur), 2)
    print("#   c4 judge", file=sys.stderr, flush=True)
    # the first call flags and replaces; the timed ones repeat the check
    jd = S.judge(dem, sur, 1024, 16, 10.0)
    t_judge = _gpu_time(lambda: S.judge(dem, sur, 1024, 16, 10.0), 3)
    err = (sur - ex).abs_()
    mae = float(err.mean())
    p99 = float(torch.quantile(err.reshape(-1)[::4099].float(), 0.99))
    px = n * n
    res = {"workload": f"{n}x{n} f32 fBm (H=0.7, 0-4000 m, 10 m, seed 4), positive openness "
                       "r=1024 px, 16 directions, angle in degrees",
           "exact_ms": round(t_exact * 1e3, 1), "exact_gpix_s": round(px / t_exact / 1e9, 3),
           "surrogate_ms": round(t_sur * 1e3, 1),
           "surrogate_gpix_s": round(px / t_sur / 1e9, 3),
           "judge_ms": round(t_judge * 1e3, 1),
           "speedup_surrogate_vs_exact": round(t_exact / t_sur, 2),
           "speedup_with_judge": round(t_exact / (t_sur + t_judge), 2),
           "raster_mae_deg_after_judge": round(mae, 4), "raster_p99_abs_err_deg": round(p99, 4),
           "judge": jd, "held_out": held,
           "network": f"6 conv3x3, 32 ch, GELU (tanh form), bf16 on MFMA (fp32 acc), one "
                      f"fused launch on the {S.GRID}x grid; input: 4-level detrended elevation "
                      "pyramid; outputs a residual over the exact openness of the 16x grid and "
                      "the sensitivity to the pixel's own elevation, composed at full resolution"}
    del dem, ex, sur, err
    torch.cuda.empty_cache()
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path "
              "is design-only in this build", file=sys.stderr)
        sys.exit(2)
    import torch

    from fujishadergpu_amd import _native as N

    dev = torch.device("cuda:0")
    if args.workload:
        fn = {"c1": bench_c1, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5,
              "stencils": bench_stencils,
              "f1": bench_f1, "f2": bench_f2, "f4": bench_f4}[args.workload]
        # --size resizes the stencils raster (its default is 32768^2); the
        # other workloads keep their fixed sizes
        kw = {"n": args.size} if args.workload == "stencils" and args.size else {}
        print(json.dumps({args.workload: fn(dev, **kw)}), flush=True)
        return
    n = args.size or N_PX
    dem = make_dem(n, dev)
    out_u8 = torch.empty((n, n), dtype=torch.uint8, device=dev)
    out_f32 = torch.empty((n, n), dtype=torch.float32, device=dev)
    L = N.lib()
    for _ in range(args.warmup):
        run_step(dem, out_u8, out_f32, args.streams)
    torch.cuda.synchronize()
    L.fsg_timing_read(b"k_usm_main", ctypes.byref(ctypes.c_double()), ctypes.byref(ctypes.c_int64()))
    L.fsg_timing_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, results, plan = run_step(dem, out_u8, out_f32, args.streams)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    L.fsg_timing_enable(0)
    tot_ms, cnt, union_ms = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
    N.check(L.fsg_timing_read_union(b"k_usm_main", ctypes.byref(tot_ms), ctypes.byref(cnt),
                                    ctypes.byref(union_ms)))
    ms_step = wall * 1e3 / args.steps
    gpix = n * n / 1e9 / (ms_step / 1e3)

    # algorithmic bytes of k_usm_main per step: main pass = core px x (4 B read
    # + 4 B f32 + 1 B u8 written); pre-pass = pooled px x (4 B read + 4 B write)
    from fujishadergpu_amd.params import norm_stat_window_geometry, stratified_windows
    main_px = sum(t[4] * t[5] for t in plan.tiles)
    margin, tile = norm_stat_window_geometry("topousm_fast", {"radii": RADII})
    wins = stratified_windows(n, n, 0, n, 0, n, grid=3, tile=min(tile, n))
    pre_px = sum((th - 2 * min(margin, th // 3, tw // 3)) * (tw - 2 * min(margin, th // 3, tw // 3))
                 for _, _, tw, th in wins)
    bytes_step = main_px * 9 + pre_px * 8
    launches_step = cnt.value / args.steps
    # the tiles (and the pre-pass windows) alternate between two streams, so
    # consecutive launches overlap and each launch's own start-to-end duration
    # counts the shared time twice; the union of the launch intervals is the
    # time the GPU spent in the kernel, and its share per launch the launch
    # duration the roofline uses
    busy_ms = union_ms.value / args.steps
    avg_ms = busy_ms / max(1.0, launches_step)
    achieved = bytes_step / (busy_ms / 1e3) / 1e9
    overlapped_avg_ms = tot_ms.value / max(1, cnt.value)
    stats_scale = float(plan.stats[0].item())
    if args.calib_copy:
        from fujishadergpu_amd import ops
        m = min(8192, n)
        dst = torch.empty((m, m), dtype=torch.float32, device=dev)
        ops.copy_region(dem, ops.region(0, 0, m, m), dst)
        torch.cuda.synchronize()
        del dst
    traffic = measured_traffic()

    line = {
        "metric": METRIC, "value": round(gpix, 3), "unit": "Gpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"topousm_fast spatial, {n}x{n} f32 fBm DEM (H=0.75, 0-3776 m, 5 m "
                        f"cells, seed 2), radii {RADII}, 2^n weights, tile 4096 + halo "
                        f"{plan.padding}, outputs float32+uint8, stratified p99 pre-pass on GPU",
            "tiles": len(plan.tiles), "overview_decimation": plan.extras.get("overview_scale"),
            "p99_scale": stats_scale,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(traffic) if traffic is not None else None,
            "kernel": "k_usm_main", "launches_per_step": launches_step,
            "avg_launch_ms": round(avg_ms, 4), "alg_bytes_per_step": bytes_step,
            "alg_bytes_per_launch": round(bytes_step / max(1.0, launches_step)),
            "kernel_busy_ms_per_step": round(busy_ms, 4),
            "duration_basis": "union of the launch intervals per step / launches",
            "overlapped_launch_avg_ms": round(overlapped_avg_ms, 4),
            "kernel_share_of_step": round(busy_ms / ms_step, 3),
        },
        "main_pass_tiles": len(plan.tiles),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(dem, plan, stats_scale, out_f32, args.cpu_tiles)
    if not args.only_c2:
        del dem, out_u8, out_f32
        torch.cuda.empty_cache()
        line["per_algorithm"] = {"topousm_fast_c2": {
            "gpix_s": line["value"], "hbm_frac_whole_job": round(
                9 * n * n / (ms_step / 1e3) / (HBM_PEAK_GBS * 1e9), 4),
            "kernel_hbm_frac": line["roofline"]["frac"]}}
        for name, fn in (("hillshade_slope_c1", bench_c1), ("stencils_32768", bench_stencils),
                         ("batch_fused_c5", bench_c5),
                         ("curvature_openness_c3", bench_c3),
                         ("openness_surrogate_c4", bench_c4),
                         ("spatial_hillshade_f1", bench_f1),
                         ("cog_io_f2", bench_f2),
                         ("prepare_f4", bench_f4)):
            print(f"# {name} ...", file=sys.stderr, flush=True)
            line["per_algorithm"][name] = fn(dev)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

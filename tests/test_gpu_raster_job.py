"""Whole-raster batch (RasterJob): curvature, openness, hillshade and slope
tiles of one raster -- ragged edge tiles and NoData discs included -- in one
persistent launch, equal bit for bit to the per-product tile scheduler, and
through graph replays."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

PRODUCTS = {
    "curvature": {"curvature_type": "mean"},
    "openness": {"openness_type": "positive", "num_directions": 8, "max_distance": 64},
    "hillshade": {"radii": [1]},
    "slope": {"unit": "percent"},
}


def _dem(cuda, h, w, seed=3):
    from fujishadergpu_amd import synth
    dem = synth.fbm(h, w, 0.85, 2500.0, seed=seed, device="cpu")
    synth.apply_discs(dem, synth.nodata_discs(h, w, 0.01, seed=seed + 30), -9999.0)
    return dem.to(cuda)


def _same(a, b, what=""):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), (what, int((na != nb).sum()), np.argwhere(na != nb)[:5])
    m = ~na
    d = a[m] != b[m]
    assert not d.any(), (what, int(d.sum()), float(np.abs(a[m] - b[m]).max()))


def _reference(dem_host, cuda, products, **kw):
    from fujishadergpu_amd.scheduler import process_dem
    out = {}
    for name, p in products.items():
        r = dem_host.to(cuda)
        out[name] = process_dem(r, name, copy_input=False, **kw, **p)[0]
    return out


def test_raster_job_equals_the_tile_scheduler_ragged_with_nodata(cuda):
    from fujishadergpu_amd.batch import RasterJob
    dem = _dem(cuda, 2600, 3100)
    host = dem.cpu()
    kw = dict(tile_size=1024, nodata=-9999.0, pixel_size=10.0)
    ref = _reference(host, cuda, PRODUCTS, **kw)
    job = RasterJob(dem, PRODUCTS, **kw)
    info = job.info()
    print("raster job", info, flush=True)
    # 3 x 4 tiles, the last row 552 and the last column 28 px high / wide
    assert info["tiles"] == 12 and info["openness_tiles"] == 12
    assert info["openness_items"] > 0 and info["stencil_tiles"] >= 12
    job.run()
    torch.cuda.synchronize()
    for name in PRODUCTS:
        _same(job.outputs[name], ref[name], name)
    job.close()


def test_raster_job_graph_replays_equal_the_eager_run(cuda):
    from fujishadergpu_amd.batch import RasterJob
    dem = _dem(cuda, 1500, 1300, seed=5)
    job = RasterJob(dem, {k: PRODUCTS[k] for k in ("curvature", "openness")}, tile_size=512,
                    nodata=-9999.0, pixel_size=10.0)
    job.run()
    torch.cuda.synchronize()
    eager = {k: v.clone() for k, v in job.outputs.items()}
    job.capture()
    for rep in range(3):
        for v in job.outputs.values():
            v.fill_(-5.0)
        x = torch.randn(1 << 20, device=cuda).sum()   # unrelated eager work between replays
        torch.cuda.synchronize()
        job.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(x)
        for k, v in job.outputs.items():
            _same(v, eager[k], k)
    job.close()


def test_raster_job_skips_tiles_by_each_products_threshold(cuda):
    from fujishadergpu_amd.batch import RasterJob
    host = _dem(cuda, 1100, 900, seed=7).cpu()
    host[0:600, 0:520] = -9999.0       # the first tile all NoData, its neighbours partly
    kw = dict(tile_size=512, nodata=-9999.0, pixel_size=10.0, nodata_threshold=0.5)
    ref = _reference(host, cuda, {k: PRODUCTS[k] for k in ("curvature", "openness")}, **kw)
    job = RasterJob(host.to(cuda), {k: PRODUCTS[k] for k in ("curvature", "openness")}, **kw)
    assert job.info()["skipped_products"] >= 2
    job.run()
    torch.cuda.synchronize()
    for k in ("curvature", "openness"):
        _same(job.outputs[k], ref[k], k)
    job.close()


def test_raster_job_tuner_keeps_the_bits(cuda):
    from fujishadergpu_amd.batch import tune_raster_job
    dem = _dem(cuda, 1200, 1000, seed=9)
    host = dem.cpu()
    prods = {k: PRODUCTS[k] for k in ("curvature", "openness")}
    job, rep = tune_raster_job(dem, prods, tile_sizes=(256, 512), nodata=-9999.0,
                               pixel_size=10.0)
    print("tuner", rep, flush=True)
    assert rep["tile_size"] in (256, 512) and rep["launch_ms"]
    ref = _reference(host, cuda, prods, tile_size=rep["tile_size"], nodata=-9999.0,
                     pixel_size=10.0)
    job.run()
    torch.cuda.synchronize()
    for k in prods:
        _same(job.outputs[k], ref[k], k)
    job.close()


def test_raster_job_per_tile_shape_equals_the_persistent_launch(cuda):
    """set_split(True): the per-window kernels tile by tile -- eager and
    through graph replays -- give the persistent launch's bits."""
    from fujishadergpu_amd.batch import RasterJob
    dem = _dem(cuda, 2100, 2700, seed=8)
    job = RasterJob(dem, PRODUCTS, tile_size=1024, nodata=-9999.0, pixel_size=10.0)
    job.run()
    torch.cuda.synchronize()
    fused = {k: v.clone() for k, v in job.outputs.items()}
    job.set_split(True)
    assert job.info()["split"] == 1
    for v in job.outputs.values():
        v.fill_(-5.0)
    job.run()
    torch.cuda.synchronize()
    for name in PRODUCTS:
        _same(job.outputs[name], fused[name], f"{name} eager")
    job.capture()
    for rep in range(2):
        for v in job.outputs.values():
            v.fill_(-5.0)
        torch.cuda.synchronize()
        job.replay()
        torch.cuda.synchronize()
        for name in PRODUCTS:
            _same(job.outputs[name], fused[name], f"{name} replay {rep}")
    job.close()

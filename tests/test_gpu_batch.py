"""Batched fused launch: identical bits to the per-tile, per-algorithm path,
through eager runs and hipGraph replays."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _jobs(cuda, specs):
    from fujishadergpu_amd import synth
    from fujishadergpu_amd.batch import TileJob
    jobs = []
    for i, (h, w, kind) in enumerate(specs):
        dem = synth.fbm(h, w, 0.6 + 0.05 * (i % 6), 500.0 + 300.0 * i, seed=700 + i, device="cpu")
        nan_free = True
        if kind == "nan":
            dem[h // 3:h // 3 + 5, w // 2:w // 2 + 7] = float("nan")
            nan_free = False
        dem = dem.to(cuda)
        p = {"pixel_size": 10.0, "azimuth": 315.0, "altitude": 45.0, "curvature_type": "mean",
             "radii": [2, 8, 32, 128], "weights": [8 / 15, 4 / 15, 2 / 15, 1 / 15],
             "global_stats": (15.0 + i,)}
        if kind == "geo":
            p.update(pixel_scale_x=24.3, pixel_scale_y=-30.8)
        if kind == "georows":   # every row's x cell width at its own latitude
            from fujishadergpu_amd.raster_info import row_geographic_scales_x
            p.update(pixel_scale_x_rows=row_geographic_scales_x(1 / 3600, -1 / 3600, 47.3, 0, h),
                     pixel_scale_y=-30.8)
        if kind == "unsorted":
            p.update(radii=[32, 2, 128, 8], weights=[0.2, 0.4, 0.1, 0.3])
        e = lambda: torch.full((h, w), -7.0, dtype=torch.float32, device=cuda)   # noqa: E731
        jobs.append(TileJob(dem=dem, params=p, hillshade=e() if kind != "cv_only" else None,
                            slope=e() if kind != "cv_only" else None, curvature=e(),
                            topousm=e() if kind != "no_usm" else None, nan_free=nan_free))
    return jobs


SPECS = [(256, 256, "proj"), (300, 517, "geo"), (97, 1030, "proj"), (700, 333, "nan"),
         (64, 64, "proj"), (410, 410, "unsorted"), (129, 257, "no_usm"), (512, 96, "cv_only"),
         (1000, 640, "geo"), (333, 300, "georows")]


def _outs(jobs):
    return [[t.clone() for t in (j.hillshade, j.slope, j.curvature, j.topousm) if t is not None]
            for j in jobs]


def _bits_equal(a, b):
    return torch.equal(a.view(torch.int32), b.view(torch.int32))


def _reset(jobs):
    for j in jobs:
        for t in (j.hillshade, j.slope, j.curvature, j.topousm):
            if t is not None:
                t.fill_(-7.0)


def test_fused_batch_equals_per_tile_launches(cuda):
    from fujishadergpu_amd.batch import FusedBatch, run_unfused
    jobs = _jobs(cuda, SPECS)
    run_unfused(jobs)
    torch.cuda.synchronize()
    ref = _outs(jobs)
    for o in ref:
        for t in o:
            assert not bool((t == -7.0).any())          # every pixel written
    _reset(jobs)
    b = FusedBatch(jobs)
    info = b.info()
    print("batch info", info, flush=True)
    assert info["fallback_usm_tiles"] == 2              # unsorted radii + the NaN tile
    assert info["fused_usm_tiles"] == 7 and info["box_radii"] == 3
    # row-walking interiors, row strides 1030 and 410 floats among them
    assert info["row_tiles"] >= 2 and info["row_items"] > 0
    for cfg in [(3, 0, 8), (1, 1, 1), (8, 0, 32)]:
        b.run(cfg)
        torch.cuda.synchronize()
        got = _outs(jobs)
        for i, (g, r) in enumerate(zip(got, ref)):
            for k, (x, y) in enumerate(zip(g, r)):
                assert _bits_equal(x, y), f"cfg {cfg} tile {i} product {k}"
        _reset(jobs)
    b.close()


def test_fused_batch_graph_replays(cuda):
    from fujishadergpu_amd.batch import FusedBatch, run_unfused, tune
    jobs = _jobs(cuda, SPECS[:5])
    run_unfused(jobs)
    torch.cuda.synchronize()
    ref = _outs(jobs)
    b = FusedBatch(jobs)
    best, times = tune(b, candidates=((2, 0, 4), (3, 1, 16)), reps=2)
    assert best in times
    s = torch.cuda.Stream()
    b.capture(s)
    for rep in range(3):
        _reset(jobs)
        torch.cuda.synchronize()
        b.replay(s)
        s.synchronize()
        got = _outs(jobs)
        for i, (g, r) in enumerate(zip(got, ref)):
            for k, (x, y) in enumerate(zip(g, r)):
                assert _bits_equal(x, y), f"replay {rep} tile {i} product {k}"
        # eager work between replays
        run_unfused(jobs[:1])
        torch.cuda.synchronize()
    b.close()


def test_fused_batch_matches_the_oracle_on_one_tile(cuda):
    from oracle import ref
    from fujishadergpu_amd.batch import FusedBatch
    jobs = _jobs(cuda, [(300, 517, "geo")])
    FusedBatch(jobs).run()
    torch.cuda.synchronize()
    j = jobs[0]
    dem = j.dem.cpu().numpy()
    p = j.params
    kw = dict(pixel_size=10.0, pixel_scale_x=p["pixel_scale_x"], pixel_scale_y=p["pixel_scale_y"])
    hs = ref.hillshade(dem, **kw)
    assert np.abs(j.hillshade.cpu().numpy() - hs).max() <= 1e-6
    sl = ref.slope(dem, **kw)
    assert np.abs(j.slope.cpu().numpy() - sl).max() <= 3e-5
    cv = ref.curvature(dem, curvature_type="mean", **kw)
    assert np.abs(j.curvature.cpu().numpy() - cv).max() <= 4e-4
    usm = ref.topousm_direct(dem, radii=p["radii"], weights=p["weights"], pixel_size=10.0,
                             stats=p["global_stats"])
    assert np.abs(j.topousm.cpu().numpy() - usm).max() <= 2e-5


def test_fused_batch_workgroup_items_equal_per_tile_launches(cuda):
    """With the workgroup kernels forced (variant 1), the batch's TopoUSM tiles
    are items of the persistent kernel instead of the batched wavefront
    launch: the same bits as the per-tile launches."""
    from fujishadergpu_amd import ops
    from fujishadergpu_amd.batch import FusedBatch, run_unfused
    jobs = _jobs(cuda, SPECS)
    run_unfused(jobs)
    torch.cuda.synchronize()
    ref = _outs(jobs)
    _reset(jobs)
    ops.set_topousm_variant(1)
    try:
        b = FusedBatch(jobs)
        info = b.info()
        assert info["fused_usm_tiles"] == 7 and info["usm_items"] > 0
        b.run((3, 0, 8))
        torch.cuda.synchronize()
        got = _outs(jobs)
        b.close()
    finally:
        ops.set_topousm_variant(0)
    for i, (g, r) in enumerate(zip(got, ref)):
        for k, (x, y) in enumerate(zip(g, r)):
            assert _bits_equal(x, y), f"tile {i} product {k}"

"""Host-side logic of the tile pipeline (no GPU): the exactness rule for
sharing raster-wide coarse means between topousm_fast windows."""
import pytest

from fujishadergpu_amd.params import coarse_radius, shared_coarse_ok


def test_coarse_radius_rounds_like_the_reference():
    # max(1, round(r / ds)) with Python's round (half to even)
    assert coarse_radius(128, 4) == 32
    assert coarse_radius(2048, 16) == 128
    assert coarse_radius(300, 8) == 38      # 37.5 -> 38
    assert coarse_radius(20, 8) == 2        # 2.5 -> 2
    assert coarse_radius(3, 16) == 1


def test_stratified_prepass_window_of_the_16384_job_shares():
    # 8256^2 window at (4064, 4064), trimmed by 2064: the radius-2048 box
    # (ds 16, 128 coarse cells) reaches coarse rows 0 .. 515 exactly
    reg = (2064, 2064, 4128, 4128)
    assert shared_coarse_ok((4064, 4064), (8256, 8256), reg, 16, 128, (16384, 16384))
    assert shared_coarse_ok((4064, 4064), (8256, 8256), reg, 4, 32, (16384, 16384))
    # one coarse cell more of radius would reflect at the window edge
    assert not shared_coarse_ok((4064, 4064), (8256, 8256), reg, 16, 129, (16384, 16384))


def test_main_tiles_of_the_16384_job_share():
    # interior tile window (4096 + 2 * 160) and the raster-corner tile
    assert shared_coarse_ok((3936, 7968), (4416, 4416), (160, 160, 4096, 4096), 4, 32,
                            (16384, 16384))
    assert shared_coarse_ok((0, 0), (4256, 4256), (0, 0, 4096, 4096), 4, 32, (16384, 16384))


def test_misaligned_window_does_not_share():
    reg = (160, 160, 4096, 4096)
    assert not shared_coarse_ok((3937, 7968), (4416, 4416), reg, 4, 32, (16384, 16384))
    assert not shared_coarse_ok((3936, 7968), (4417, 4416), reg, 4, 32, (16384, 16384))


def test_region_near_an_inner_window_edge_does_not_share():
    # region starting at the window's first row, window edge inside the raster
    assert not shared_coarse_ok((4000, 4000), (1024, 1024), (0, 0, 512, 512), 4, 8,
                                (8192, 8192))
    # the same region on the raster's own edge reflects alike on both grids
    assert shared_coarse_ok((0, 4000), (1024, 1024), (0, 200, 512, 512), 4, 8, (8192, 8192))


def test_window_spanning_the_raster_always_shares():
    assert shared_coarse_ok((0, 0), (96, 80), (0, 0, 96, 80), 8, 64, (96, 80))


def test_gaussian_coarse_radius_uses_its_truncation_margin():
    # r <= 1 is the sigma=1 Gaussian truncated at 4 cells
    # ds-grid 32 x 32, zoom 31/511: region row 80 -> cell 4, row 431 -> cells 26..27
    assert shared_coarse_ok((64, 64), (512, 512), (80, 80, 352, 352), 16, 1, (1024, 1024))
    # row 64 -> cell 3: the truncated Gaussian would reflect at the window edge
    assert not shared_coarse_ok((64, 64), (512, 512), (64, 64, 384, 384), 16, 1, (1024, 1024))
    # a box of radius 2 needs less margin than the Gaussian
    assert shared_coarse_ok((64, 64), (512, 512), (64, 64, 384, 384), 16, 2, (1024, 1024))


def test_judge_sampling_hash_and_segments():
    from fujishadergpu_amd import ops
    # published splitmix64 outputs for seed 0 (first two draws)
    assert ops.splitmix64(0) == 0xE220A8397B1DCDAF
    assert ops.splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    plan = ops.judge_segments(1000, 700, 256, 256, 0.01, 7, 64)
    assert len(plan) == 4 * 3
    for (y0, x0, th, tw), segs in plan:
        assert len(segs) == max(1, -(-int(round(0.01 * th * tw * 1e6)) // (64 * 10 ** 6)))
        for r, c in segs:
            assert y0 <= r < y0 + th and x0 <= c and c + 64 <= x0 + tw
    # the draw depends on the seed and on the tile
    assert plan[0][1] != ops.judge_segments(1000, 700, 256, 256, 0.01, 8, 64)[0][1]
    assert plan[0][1][:3] != plan[1][1][:3]


def test_row_scales_resolution():
    import numpy as np
    import pytest
    from fujishadergpu_amd.algorithms import RowScales, x_steps_rows
    from fujishadergpu_amd.raster_info import meters_per_degree, row_geographic_scales_x
    rs = RowScales([0.0, 21.5, 1e-12, 22.0], "cpu", pixel_size=7.0)
    assert rs.t.dtype.is_floating_point and rs.t.tolist() == [7.0, 21.5, 7.0, 22.0]
    assert rs.first == 7.0
    with pytest.raises(ValueError):
        RowScales([1.0, -1.0], "cpu")
    with pytest.raises(ValueError):
        x_steps_rows({"pixel_scale_x_rows": [1.0, 2.0]}, 3, "cpu")
    assert x_steps_rows({"pixel_scale_x_rows": rs}, 4, "cpu") is rs
    # every row at its own latitude: row r of a 1 arc-second grid from 45 N
    v = row_geographic_scales_x(1 / 3600, -1 / 3600, 45.0, 10, 3)
    for k in range(3):
        assert v[k] == pytest.approx(meters_per_degree(45.0 - (10 + k + 0.5) / 3600)[0] / 3600,
                                     rel=1e-12)
    assert np.all(np.diff(row_geographic_scales_x(1 / 3600, -1 / 3600, 45.0, 0, 100)) > 0)


# --------------------------------------------- prepare: detection decisions

def _np_stats(vals, src_nodata):
    import numpy as np
    v = np.asarray(vals, dtype=np.float32).reshape(-1)
    v = v[np.isfinite(v)]
    if src_nodata is not None and np.isfinite(src_nodata):
        v = v[v != np.float32(src_nodata)]
    if v.size == 0:
        return {"mode": 0.0, "count": 0, "n_valid": 0, "vmin": 0.0, "vmax": 0.0,
                "n_distinct": 0}
    u, c = np.unique(v, return_counts=True)
    i = int(np.argmax(c))
    return {"mode": float(u[i]), "count": int(c[i]), "n_valid": int(v.size),
            "vmin": float(u[0]), "vmax": float(u[-1]), "n_distinct": int(u.size)}


@pytest.mark.parametrize("seed", range(6))
def test_prepare_detection_decisions_follow_the_oracle(monkeypatch, seed):
    # the host decisions on the device statistics (mode, counts, range) give
    # the oracle's answers; the statistics themselves are GPU-tested
    import numpy as np
    import torch
    from fujishadergpu_amd import prepare as P
    from oracle import prepare as prep
    monkeypatch.setattr(P, "value_mode", lambda v, nd: _np_stats(v.numpy(), nd))
    monkeypatch.setattr(P, "count_equal",
                        lambda v, x: int((v.numpy() == np.float32(x)).sum()))
    rng = np.random.default_rng(seed)
    z = np.round(rng.random((48, 56)) * 300 + 50).astype(np.float32)
    kind = seed % 3
    if kind == 0:
        z[:3] = z[-3:] = z[:, :3] = z[:, -3:] = -32768.0
    elif kind == 1:
        z[rng.random(z.shape) < 0.12] = z.max()
    else:
        z[rng.random(z.shape) < 0.04] = 0.0
    decl = None if seed < 3 else -32768.0
    valid = np.isfinite(z) & ~(prep.declared_mask(z, decl))
    t = torch.from_numpy(z)
    assert P.detect_sentinel_nodata(t, decl) == prep.detect_sentinel_nodata(z, valid)
    assert P.detect_border_nodata(t, decl) == prep.detect_border_nodata(z, valid)


def test_prepare_coarse_geometry_matches_the_oracle():
    from fujishadergpu_amd import prepare as P
    from oracle import prepare as prep
    for w, h, cm in ((240000, 220000, 2048), (16384, 16384, 2048), (700, 600, 160),
                     (5000, 300, 2048), (9000, 9001, 1000)):
        ch, cw = P.coarse_shape(w, h, cm)
        assert (ch, cw) == prep.coarse_shape(w, h, cm)
        assert P.coarse_oversample(ch, cw) == prep.coarse_oversample(ch, cw)


@pytest.mark.parametrize("seed", range(12))
def test_valid_bbox_matches_a_brute_force_overview(seed):
    # _norm_stats.py:254-264: the valid bounding box from a ~512 px nearest
    # overview, recomputed here with numpy loops over the overview samples
    import numpy as np
    import torch
    from fujishadergpu_amd.stats import valid_bbox
    rng = np.random.default_rng(seed)
    H, W = (int(v) for v in rng.integers(40, 2600, 2))
    a = np.full((H, W), np.nan, np.float32)
    if seed % 5:
        y0, x0 = int(rng.integers(0, H)), int(rng.integers(0, W))
        a[y0:int(rng.integers(y0, H)) + 1, x0:int(rng.integers(x0, W)) + 1] = 1.0
    cov = max(1, max(W, H) // 512)
    oh, ow = max(1, H // cov), max(1, W // cov)
    ri = [min(int(np.floor((i + 0.5) * (H / oh))), H - 1) for i in range(oh)]
    ci = [min(int(np.floor((j + 0.5) * (W / ow))), W - 1) for j in range(ow)]
    ok = np.isfinite(a[np.ix_(ri, ci)])
    want = None
    if ok.any():
        ys, xs = np.nonzero(ok.any(axis=1))[0], np.nonzero(ok.any(axis=0))[0]
        want = (ys[0] * cov, min(H, (ys[-1] + 1) * cov), xs[0] * cov, min(W, (xs[-1] + 1) * cov))
    assert valid_bbox(torch.from_numpy(a)) == want


def test_stack_band_counts():
    # agg='stack' is band-first with one band per radius (_nan_utils.py:191-194),
    # only on the multi-radius paths; the local block functions ignore agg
    from fujishadergpu_amd.algorithms import ALGORITHMS
    sp = dict(mode="spatial", agg="stack")
    assert ALGORITHMS["hillshade"].stack_bands({**sp, "radii": [2, 8, 32]}) == 3
    assert ALGORITHMS["slope"].stack_bands({**sp, "radii": [40]}) == 1
    assert ALGORITHMS["curvature"].stack_bands({**sp, "radii": [2, 2, 9]}) == 2   # deduplicated
    assert ALGORITHMS["openness"].stack_bands({**sp, "radii": [1, 2]}) == 2       # max(2, r) kept
    assert ALGORITHMS["ambient_occlusion"].stack_bands({**sp, "radii": None}) == 6
    assert ALGORITHMS["hillshade"].stack_bands({**sp, "radii": [2, 8], "agg": "mean"}) is None
    assert ALGORITHMS["hillshade"].stack_bands(dict(mode="local", agg="stack")) is None
    assert ALGORITHMS["hillshade"].stack_bands(dict(mode="local", agg="stack",
                                                    radii=[2, 9])) == 2
    assert ALGORITHMS["slope"].stack_bands(dict(mode="local", agg="stack", radii=[2, 9])) is None
    assert ALGORITHMS["topousm_fast"].stack_bands({**sp, "radii": [2, 8]}) is None


def test_band_stack_output_spec():
    import torch
    from fujishadergpu_amd.algorithms import OutSpec
    t = torch.zeros((3, 4, 5))
    s = OutSpec(f32=t)
    assert s.n_bands == 3 and s.band(1).f32.shape == (4, 5) and s.band(1).n_bands is None
    with pytest.raises(ValueError):
        s.descriptors()


def test_tile_failures_abort_with_the_first_three_messages():
    # core/tile_processor.py:1689-1695: any failed tile aborts the run; the
    # message lists the first three failures in tile order
    from fujishadergpu_amd.scheduler import TileResult, check_results
    ok = [TileResult(0, 0, True), TileResult(0, 1, True)]
    check_results(ok + [TileResult(0, 2, False, skipped_reason="NoData covers 100.0%")])
    bad = [TileResult(1, i, False, error_message=f"boom {i}") for i in range(5)]
    with pytest.raises(RuntimeError) as e:
        check_results(ok + bad)
    msg = str(e.value)
    assert msg.startswith("Tile processing error:\n")
    assert msg.splitlines()[1:] == ["Tile (1, 0): boom 0", "Tile (1, 1): boom 1",
                                    "Tile (1, 2): boom 2"]
    with pytest.raises(ValueError, match="No tiles were processed"):
        check_results([TileResult(0, 0, False, skipped_reason="NoData covers 100.0%")])


def test_raster_job_rejects_bad_products_and_rasters_before_any_launch():
    import torch
    from fujishadergpu_amd.batch import RasterJob
    with pytest.raises(ValueError, match="products"):
        RasterJob(torch.zeros(8, 8), {"topousm_fast": {}})
    with pytest.raises(ValueError, match="at least one"):
        RasterJob(torch.zeros(8, 8), {})
    with pytest.raises(ValueError, match="GPU tensor"):
        RasterJob(torch.zeros(8, 8), {"curvature": {}})


def test_raster_job_refuses_the_multi_radius_hillshade():
    import torch
    from fujishadergpu_amd.batch import HillshadeAlgorithm
    from fujishadergpu_amd.scheduler import plan
    # the resolved radii of a default local hillshade take the multi-radius branch
    P = plan(torch.zeros(2600, 3100), "hillshade", tile_size=1024, pixel_size=10.0)
    assert len(HillshadeAlgorithm._direct_radii(P.params)) > 1
    P1 = plan(torch.zeros(2600, 3100), "hillshade", tile_size=1024, pixel_size=10.0, radii=[1])
    assert HillshadeAlgorithm._direct_radii(P1.params) == [1.0]

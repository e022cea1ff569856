"""Pins for the oracle's tile and pre-pass geometry (oracle/geometry.py), and
the product's geometry (fujishadergpu_amd/params.py, stats.py) against it.
CPU only.

The expected lists are written out by hand from the reference's rules
(core/tile_processor.py:207-383,1591-1627; algorithms/_norm_stats.py:64-162,
254-276), for the workloads the benchmarks run: the 16384^2 TopoUSM raster
(tile 4096), the 32768^2 local curvature / openness raster, a ragged raster
and the spatial-mode halos.
"""
import itertools

import numpy as np
import pytest

from oracle import geometry as geo


# ------------------------------------------------------------- tile halo (A14)

@pytest.mark.parametrize("algorithm,params,tile,want", [
    # TopoUSM, 16384^2 at 5 m: the radii after the overview split are
    # [2, 8, 32, 128] -> 128 + 16 = 144 -> 160
    ("topousm_fast", dict(mode="spatial", radii=[2, 8, 32, 128]), 4096, 160),
    # full radii when the overview is not used: 2048 + 16 = 2064 -> 2080
    ("topousm_fast", dict(mode="spatial", radii=[2, 8, 32, 128, 512, 2048]), 4096, 2080),
    # local mode: max(32, ceil(5 * 10) = 50) -> 64
    ("curvature", dict(mode="local"), 4096, 64),
    ("openness", dict(mode="local", max_distance=64), 4096, 64),
    ("hillshade", dict(mode="local"), 2048, 64),
    # spatial stencils, radii given: the overview serves r > max(256, 4096 // 16);
    # 2 * 128 + 2 = 258 -> 288
    ("hillshade", dict(mode="spatial", radii=[2, 8, 32, 128, 512]), 4096, 288),
    ("slope", dict(mode="spatial", radii=[4, 40, 300]), 1024, 96),        # 2*40+2 = 82
    ("curvature", dict(mode="spatial", radii=[300, 600]), 1024, 608),     # min radius: 602
    # spatial rays: r + 16
    ("openness", dict(mode="spatial", radii=[6, 48, 300]), 1024, 64),    # 48 + 16
    ("ambient_occlusion", dict(mode="spatial", radii=[2, 8, 32, 128, 512]), 4096, 160),
    # spatial without radii: the automatic list, all of it (2 * 2048 + 2)
    ("hillshade", dict(mode="spatial"), 4096, 4128),
    # a large tile raises the switch: max(256, 8192 // 16 = 512) keeps r = 512
    ("hillshade", dict(mode="spatial", radii=[2, 8, 32, 128, 512]), 8192, 1056),
])
def test_required_padding_table(algorithm, params, tile, want):
    assert geo.required_padding(algorithm, params, sigma=10.0, tile_size=tile) == want


def test_required_padding_sigma_floor():
    # a wide sigma dominates a small halo: ceil(5 * 13.1) = 66 -> 96
    assert geo.required_padding("hillshade", dict(mode="local"), sigma=13.1) == 96
    assert geo.required_padding("hillshade", dict(mode="local"), sigma=0.0) == 32


# ------------------------------------------------------------ tile grid (A14)

def test_tile_grid_topousm_16384():
    tiles = geo.tile_grid(16384, 16384, 4096, 160)
    assert len(tiles) == 16
    origins = [(t[0], t[1], t[2], t[3]) for t in tiles]
    assert origins == [(ty, tx, 4096 * tx, 4096 * ty) for ty in range(4) for tx in range(4)]
    assert all(t[4] == t[5] == 4096 for t in tiles)
    by_pos = {(t[0], t[1]): t[6:] for t in tiles}
    assert by_pos[(0, 0)] == (0, 0, 4256, 4256)               # corner: 4096 + 160
    assert by_pos[(1, 1)] == (3936, 3936, 4416, 4416)         # interior: 4096 + 2 * 160
    assert by_pos[(0, 2)] == (8032, 0, 4416, 4256)            # top edge
    assert by_pos[(3, 3)] == (12128, 12128, 4256, 4256)
    assert sum(t[8] * t[9] for t in tiles) == 4 * 4256 ** 2 + 8 * 4256 * 4416 + 4 * 4416 ** 2


def test_tile_grid_curvature_32768():
    tiles = geo.tile_grid(32768, 32768, 4096, 64)
    assert len(tiles) == 64
    sizes = {(t[8], t[9]) for t in tiles}
    assert sizes == {(4160, 4160), (4160, 4224), (4224, 4160), (4224, 4224)}
    assert tiles[9][6:] == (4032, 4032, 4224, 4224)


def test_tile_grid_ragged():
    # 20000 x 13001 at 4096: 5 x 4 tiles, last column 3616 wide, last row 713 high
    tiles = geo.tile_grid(20000, 13001, 4096, 64)
    assert len(tiles) == 20
    assert [t[4] for t in tiles[:5]] == [4096] * 4 + [3616]
    assert [t[5] for t in tiles[::5]] == [4096] * 3 + [713]
    last = tiles[-1]
    assert last[:6] == (3, 4, 16384, 12288, 3616, 713)
    assert last[6:] == (16320, 12224, 3680, 777)
    assert tiles[4][6:] == (16320, 0, 3680, 4160)
    # the cores partition each axis: consecutive, no gap, no overlap
    xs = sorted({(t[2], t[4]) for t in tiles})
    ys = sorted({(t[3], t[5]) for t in tiles})
    for spans, n in ((xs, 20000), (ys, 13001)):
        assert spans[0][0] == 0 and sum(w for _, w in spans) == n
        assert all(a[0] + a[1] == b[0] for a, b in zip(spans, spans[1:]))


def test_tile_grid_one_tile_and_tiny():
    assert geo.tile_grid(300, 200, 4096, 64) == [(0, 0, 0, 0, 300, 200, 0, 0, 300, 200)]
    assert geo.tile_grid(5, 3, 2, 1) == [
        (0, 0, 0, 0, 2, 2, 0, 0, 3, 3), (0, 1, 2, 0, 2, 2, 1, 0, 4, 3),
        (0, 2, 4, 0, 1, 2, 3, 0, 2, 3), (1, 0, 0, 2, 2, 1, 0, 1, 3, 2),
        (1, 1, 2, 2, 2, 1, 1, 1, 4, 2), (1, 2, 4, 2, 1, 1, 3, 1, 2, 2)]


# ------------------------------------------------ pre-pass windows (A11)

def test_prepass_geometry_topousm_full_radii():
    # the pre-pass uses the full radii: margin 2048 + 16, tile 4 * 2064
    assert geo.prepass_geometry("topousm_fast", {"radii": [2, 8, 32, 128, 512, 2048]}) == \
        (2064, 8256)
    assert geo.prepass_geometry("openness", {"radii": [1], "max_distance": 64}) == (80, 2048)
    assert geo.prepass_geometry("ambient_occlusion", {"radius": 10.0}) == (26, 2048)
    assert geo.prepass_geometry("openness", {}) == (32, 2048)


def test_stratified_windows_topousm_16384():
    # cells 16384 // 3 = 5461; centres 2730, 8191, 13652; half tile 4128;
    # origins clamp into [0, 16384 - 8256 = 8128]
    win = geo.stratified_windows(16384, 16384, (0, 16384, 0, 16384), tile=8256)
    o = (0, 4063, 8128)
    assert win == [(y, x, 8256, 8256) for y in o for x in o]
    # with the trim, m = min(2064, 8256 // 3 = 2752) = 2064: 4128^2 interiors,
    # 9 x 4128^2 = 153.4 M pooled values
    r = np.zeros((16384, 16384), np.float32)
    wins = geo.prepass_windows(r, "topousm_fast", {"radii": [2, 8, 32, 128, 512, 2048]})
    assert wins == [(y, x, 8256, 8256, 2064) for y in o for x in o]
    assert sum((w[2] - 2 * w[4]) * (w[3] - 2 * w[4]) for w in wins) == 153363456


def test_stratified_windows_small_raster_deduplicates():
    # a 3000 x 2000 raster, tile 2048: every grid row clamps to the same
    # origin on y (2000 < 2048 -> 0); x cells 1000, centres 500, 1500, 2500
    # -> origins 0, 476, min(1476, 952) = 952; the other two rows repeat them
    win = geo.stratified_windows(3000, 2000, (0, 2000, 0, 3000), tile=2048)
    assert win == [(0, 0, 2048, 2000), (0, 476, 2048, 2000), (0, 952, 2048, 2000)]


def test_stratified_windows_follow_the_valid_box():
    # valid data only in rows 4096..8191, cols 1024..3071 of a 8192^2 raster
    win = geo.stratified_windows(8192, 8192, (4096, 8192, 1024, 3072), tile=2048)
    # cells 1365 x 682; y centres 4778, 6143, 7508 -> 3754, 5119, 6144 (clamped)
    # x centres 1365, 2047, 2729 -> 341, 1023, 1705
    assert [w[:2] for w in win] == [(y, x) for y in (3754, 5119, 6144)
                                    for x in (341, 1023, 1705)]


def test_valid_bbox_nearest_overview():
    r = np.full((1024, 2048), np.nan, np.float32)
    r[300:500, 1000:1100] = 1.0
    # cov = 4: samples at 4 i + 2; rows 300..499 -> i 75..124, cols 1000..1099 -> 250..274
    assert geo.valid_bbox(r) == (300, 500, 1000, 1100)
    r2 = np.full((100, 100), np.nan, np.float32)
    assert geo.valid_bbox(r2) is None
    r2[7, 9] = 0.0
    assert geo.valid_bbox(r2) == (7, 8, 9, 10)


# ------------------------------------------------ product == oracle geometry

def test_product_tile_grid_equals_the_oracle():
    from fujishadergpu_amd.params import tile_grid
    for (w, h, t, p) in itertools.product((1, 7, 300, 4097, 20000), (1, 13, 2049, 13001),
                                          (64, 1000, 4096), (0, 32, 160)):
        assert tile_grid(w, h, t, p) == geo.tile_grid(w, h, t, p), (w, h, t, p)


@pytest.mark.parametrize("algorithm", ["hillshade", "slope", "curvature", "topousm_fast",
                                       "openness", "ambient_occlusion"])
def test_product_padding_equals_the_oracle(algorithm):
    from fujishadergpu_amd.params import required_padding
    cases = [dict(mode="local"), dict(mode="spatial"),
             dict(mode="spatial", radii=[2, 8, 32, 128, 512]),
             dict(mode="spatial", radii=[4, 40, 300]), dict(mode="spatial", radii=[300, 600]),
             dict(mode="spatial", radii=[1.4, 2.6, 9.5]),
             dict(mode="spatial", radii=[2, 8, 32, 128, 512, 2048])]
    for prm, tile, sigma in itertools.product(cases, (512, 1024, 4096, 8192), (0.0, 10.0, 13.1)):
        got = required_padding(algorithm, prm, sigma, 1.0, None, tile)
        assert got == geo.required_padding(algorithm, prm, sigma=sigma, tile_size=tile), \
            (prm, tile, sigma)


def test_product_stratified_windows_equal_the_oracle():
    from fujishadergpu_amd.params import norm_stat_window_geometry, stratified_windows
    rng = np.random.default_rng(0)
    for _ in range(300):
        W, H = (int(v) for v in rng.integers(1, 40000, 2))
        by0 = int(rng.integers(0, H)); by1 = int(rng.integers(by0 + 1, H + 1))
        bx0 = int(rng.integers(0, W)); bx1 = int(rng.integers(bx0 + 1, W + 1))
        tile = int(rng.choice([2048, 4096, 8256]))
        tile = min(tile, max(W, H))
        assert stratified_windows(W, H, by0, by1, bx0, bx1, grid=3, tile=tile) == \
            geo.stratified_windows(W, H, (by0, by1, bx0, bx1), tile=tile)
    for alg, merged in (("topousm_fast", {"radii": [2, 8, 32, 128, 512, 2048]}),
                        ("openness", {"radii": [1], "max_distance": 64}),
                        ("ambient_occlusion", {"radius": 10.0}), ("openness", {})):
        assert norm_stat_window_geometry(alg, merged) == geo.prepass_geometry(alg, merged)


def test_product_valid_bbox_equals_the_oracle():
    import torch
    from fujishadergpu_amd.stats import valid_bbox
    rng = np.random.default_rng(3)
    for H, W in ((1024, 2048), (3001, 777), (100, 100), (5000, 5000)):
        r = np.full((H, W), np.nan, np.float32)
        y0, x0 = int(rng.integers(0, H // 2)), int(rng.integers(0, W // 2))
        r[y0:y0 + H // 3 + 1, x0:x0 + W // 4 + 1] = 5.0
        assert valid_bbox(torch.from_numpy(r)) == geo.valid_bbox(r)

"""GPU parity of the device-resident tile pipeline against the oracle tile loop.

The oracle side runs the reference's per-tile semantics on the CPU (full
window compute, core crop, NoData restore, encoding) over the same tile grid;
the GPU side computes only each core, straight into the output raster.
"""
import numpy as np
import pytest
import torch

from oracle import geometry as geo
from oracle import ref

pytestmark = pytest.mark.gpu


def _raster(h, w, seed, nodata_discs=0, relief=3000.0, hurst=0.8):
    from fujishadergpu_amd import synth
    d = synth.fbm(h, w, hurst, relief, seed=seed, device="cpu")
    if nodata_discs:
        discs = synth.nodata_discs(h, w, 0.01, seed=seed + 1)[:nodata_discs]
        synth.apply_discs(d, discs, -9999.0)
    return d


def _oracle_tiles(host, P, algorithm, nodata, stats=None, field=None):
    """Reference tile loop (core/tile_processor.py:822-1024) with the oracle."""
    H, W = host.shape
    src = host.copy()
    if nodata is not None:
        src[ref.nodata_mask(src, nodata)] = np.nan
    out = np.full((H, W), np.nan, np.float32)
    tiles = geo.plan_tiles(algorithm, P.params, W, H, P.tile_size)
    assert tiles == P.tiles
    for ty, tx, cx, cy, cw, ch, wx, wy, ww, wh in tiles:
        win = src[wy:wy + wh, wx:wx + ww]
        p = {k: v for k, v in P.params.items() if not isinstance(v, torch.Tensor)}
        if stats is not None:
            p["global_stats"] = stats
        if field is not None:
            p["_topousm_fast_coarse_field"] = field
            p["_topousm_fast_field_offset"] = (wy, wx)
        res = ref.run_window(algorithm, win, p)
        core = res[cy - wy:cy - wy + ch, cx - wx:cx - wx + cw]
        out[cy:cy + ch, cx:cx + cw] = ref.format_output(core, algorithm)
    return out


@pytest.mark.parametrize("algorithm,kw", [
    ("hillshade", dict(mode="local")),
    ("slope", dict(mode="local")),
    ("curvature", dict(mode="local", curvature_type="mean")),
    ("openness", dict(mode="local", num_directions=8, max_distance=32)),
    ("ambient_occlusion", dict(mode="local")),
])
@pytest.mark.parametrize("nodata_discs", [0, 12])
def test_scheduler_matches_oracle_tile_loop(cuda, algorithm, kw, nodata_discs):
    from fujishadergpu_amd.scheduler import process_dem
    host_t = _raster(700, 900, 21, nodata_discs=nodata_discs)
    g = host_t.to(cuda)
    nodata = -9999.0 if nodata_discs else None
    out, results, P = process_dem(g, algorithm, tile_size=256, pixel_size=10.0, nodata=nodata,
                                  return_plan=True, **kw)
    assert all(r.success for r in results)
    stats = None
    if P.stats is not None:
        stats = tuple(float(x) for x in P.stats.cpu().numpy())
    exp = _oracle_tiles(host_t.numpy(), P, algorithm, nodata, stats=stats)
    got = out.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    m = ~np.isnan(exp)
    tol = {"hillshade": 1e-6, "slope": 3e-5, "curvature": 4e-4, "openness": 1e-4,
           "ambient_occlusion": 1e-4}[algorithm]
    assert np.abs(got[m] - exp[m]).max() <= tol


def test_scheduler_topousm_overview_prepass_and_uint8(cuda):
    from fujishadergpu_amd.scheduler import process_dem
    H, W = 1100, 1300
    host_t = _raster(H, W, 23, relief=3776.0, hurst=0.75)
    g = host_t.to(cuda)
    radii, weights = [2, 8, 32, 300], [0.4, 0.3, 0.2, 0.1]
    f32 = torch.empty((H, W), dtype=torch.float32, device=cuda)
    out, results, P = process_dem(g, "topousm_fast", tile_size=512, pixel_size=5.0,
                                  mode="spatial", radii=radii, weights=weights,
                                  output_dtype="uint8", out_f32=f32, return_plan=True)
    assert P.extras["large"][0] == [300]
    assert P.extras["large"][1] == pytest.approx([0.1])
    assert P.padding == 64  # max(32 + 16, 50) -> 64
    host = host_t.numpy()
    # overview field (scale 1: no decimation below 2048 px)
    coarse, scale = ref.overview_block_mean(host)
    field = ref.topousm_large_coarse_field(coarse, large_radii=[300], large_weights=[0.1],
                                           decimation=scale)
    # oracle pre-pass over the same stratified windows, full radii
    pool = []
    for wy, wx, tw, th, m in geo.prepass_windows(host, "topousm_fast", {"radii": radii}):
        raw = ref.topousm_fast_block(host[wy:wy + th, wx:wx + tw], radii=radii, weights=weights,
                                     pixel_size=5.0)
        raw = raw[m:-m, m:-m] if m > 0 else raw
        pool.append(raw[~np.isnan(raw)])
    stats = ref.topousm_stat(np.concatenate(pool))
    assert float(P.stats[0]) == pytest.approx(stats[0], rel=1e-6)
    exp = _oracle_tiles(host, P, "topousm_fast", None, stats=(float(P.stats[0]),), field=field)
    got = f32.cpu().numpy()
    assert np.abs(got - exp).max() <= 2e-5
    qp = P.qp
    dn = ref.quantize(got, qp["a_coef"], qp["b_coef"], qp["dn_min"], qp["dn_max"], "uint8")
    assert np.array_equal(out.cpu().numpy(), dn)


def test_scheduler_topousm_fractional_overview_default_radii(cuda):
    # 6000 x 5200 at 10 m, spatial mode, the automatic radii: [2, 8, 32, 128,
    # 512] (short side / 10 = 520); 512 > max(256, 2048 // 16) goes through the
    # overview at decimation 6000 / 2048 = 2.93 onto 1775 x 2048 cells, which
    # is not an integer block mean
    from fujishadergpu_amd.scheduler import process_dem
    H, W = 5200, 6000
    host_t = _raster(H, W, 71, relief=2500.0, hurst=0.8)
    stats = (37.5,)
    out, results, P = process_dem(host_t.to(cuda), "topousm_fast", tile_size=2048,
                                  pixel_size=10.0, mode="spatial", global_stats=stats,
                                  return_plan=True)
    assert all(r.success for r in results)
    assert P.params["radii"] == [2, 8, 32, 128] and P.extras["large"][0] == [512]
    assert P.extras["overview_scale"] == pytest.approx(6000 / 2048)
    host = host_t.numpy()
    coarse, scale = ref.overview(host)
    assert coarse.shape == (1775, 2048) and scale == pytest.approx(6000 / 2048)
    lr, lw = P.extras["large"]
    field = ref.topousm_large_coarse_field(coarse, large_radii=lr, large_weights=lw,
                                           decimation=scale)
    tiles = geo.plan_tiles("topousm_fast", P.params, W, H, P.tile_size)
    assert tiles == P.tiles and P.padding == 160
    got = out.cpu().numpy()
    p = {k: v for k, v in P.params.items() if not isinstance(v, torch.Tensor)}
    p["global_stats"] = stats
    p["_topousm_fast_coarse_field"] = field
    for ty, tx, cx, cy, cw, ch, wx, wy, ww, wh in (tiles[0], tiles[4], tiles[-1]):
        p["_topousm_fast_field_offset"] = (wy, wx)
        res = ref.run_window("topousm_fast", host[wy:wy + wh, wx:wx + ww], p)
        exp = ref.format_output(res[cy - wy:cy - wy + ch, cx - wx:cx - wx + cw], "topousm_fast")
        d = np.abs(got[cy:cy + ch, cx:cx + cw].astype(np.float64) - exp)
        print("tile", (ty, tx), "max |d|", float(d.max()), flush=True)
        assert d.max() <= 2e-5


def test_scheduler_skips_all_nodata_tiles(cuda):
    """A window (core + halo) that is entirely NoData is skipped
    (core/tile_processor.py:864-874); NoData there reads as DN 0."""
    from fujishadergpu_amd.scheduler import process_dem
    host_t = _raster(300, 600, 29)
    host_t[:, :214] = -9999.0   # covers the windows of tile columns 0 (cols 0..213)
    out, results, P = process_dem(host_t.to(cuda), "hillshade", tile_size=150, pixel_size=10.0,
                                  nodata=-9999.0, mode="local", output_dtype="uint8",
                                  return_plan=True)
    assert P.padding == 64
    skipped = sorted((r.tile_y, r.tile_x) for r in results if r.skipped_reason)
    assert skipped == [(0, 0), (1, 0)]
    assert int(out[:, :150].max()) == 0
    assert int(out[:, 214:].min()) >= 1


@pytest.mark.parametrize("algorithm", ["slope", "hillshade", "curvature"])
@pytest.mark.parametrize("rows", [False, True])
def test_geographic_raster_tile_and_row_latitudes(cuda, algorithm, rows):
    # a 1 arc-second geographic raster at 44-45 N: the reference's per-tile
    # latitude (core/tile_processor.py:910-932), or each row's own latitude
    from fujishadergpu_amd.raster_info import row_geographic_scales_x, tile_geographic_scales
    from fujishadergpu_amd.scheduler import process_dem
    host_t = _raster(600, 700, 33, relief=900.0)
    g = host_t.to(cuda)
    a, e, f = 1 / 3600, -1 / 3600, 45.0
    tr = (a, 0.0, 139.0, 0.0, e, f)
    out, _, P = process_dem(g, algorithm, tile_size=256, transform=tr, is_geographic=True,
                            geographic_rows=rows, mode="local", return_plan=True)
    host = host_t.numpy()
    want = np.full(host.shape, np.nan, np.float32)
    tiles = geo.plan_tiles(algorithm, P.params, host.shape[1], host.shape[0], P.tile_size)
    assert tiles == P.tiles
    for ty, tx, cx, cy, cw, ch, wx, wy, ww, wh in tiles:
        p = {k: v for k, v in P.params.items() if not isinstance(v, torch.Tensor)}
        p["pixel_scale_x"], p["pixel_scale_y"] = tile_geographic_scales(
            a, e, 0.5 * ((f + cy * e) + (f + (cy + ch) * e)))
        if rows:
            p["pixel_scale_x"] = row_geographic_scales_x(a, e, f, wy, wh)
        res = ref.run_window(algorithm, host[wy:wy + wh, wx:wx + ww], p)
        want[cy:cy + ch, cx:cx + cw] = ref.format_output(
            res[cy - wy:cy - wy + ch, cx - wx:cx - wx + cw], algorithm)
    got = out.cpu().numpy()
    if out.dtype != torch.float32:
        assert np.abs(got.astype(np.int64) - want.astype(np.int64)).max() <= 1
    else:
        tol = {"slope": 3e-4, "hillshade": 1e-6, "curvature": 4e-4}[algorithm]
        assert np.nanmax(np.abs(got - want)) <= tol


@pytest.mark.parametrize("algorithm,kw,tol", [
    ("hillshade", dict(mode="local"), 2e-6),
    ("slope", dict(mode="local"), 1e-4),
    ("curvature", dict(mode="local", curvature_type="mean"), 4e-4),
    ("openness", dict(mode="local", num_directions=8, max_distance=20), 1e-5),
    ("ambient_occlusion", dict(mode="local", radius=6.0, num_samples=8), 1e-5),
])
def test_reflect_edge_mode_matches_the_dask_single_chunk(cuda, algorithm, kw, tol):
    # edge_mode='reflect': windows touching the raster edge are extended by the
    # Dask map_overlap depth with the symmetric rule; over a multi-tile raster
    # the result equals the Dask backend's single-chunk result everywhere
    from fujishadergpu_amd.scheduler import process_dem
    host_t = _raster(300, 340, 23)
    stats = (0.0, 1.0)                       # identity stretch for openness / AO
    out, _, P = process_dem(host_t.to(cuda), algorithm, tile_size=128, pixel_size=10.0,
                            edge_mode="reflect", global_stats=stats, return_plan=True, **kw)
    p = {k: v for k, v in P.params.items() if not isinstance(v, torch.Tensor)}
    p["global_stats"] = stats
    exp = ref.format_output(ref.dask_single_chunk(algorithm, host_t.numpy(), p), algorithm)
    got = out.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    d = np.abs(got.astype(np.float64) - exp)
    print(algorithm, "max |d|", float(np.nanmax(d)), "ring max |d|",
          float(max(np.nanmax(d[:3]), np.nanmax(d[-3:]), np.nanmax(d[:, :3]),
                    np.nanmax(d[:, -3:]))), flush=True)
    assert np.nanmax(d) <= tol
    # and the tile backend's own edge rule differs on the outer ring
    tile, _ = process_dem(host_t.to(cuda), algorithm, tile_size=128, pixel_size=10.0,
                          global_stats=stats, **kw)
    assert not torch.equal(tile[:, 0], out[:, 0])


@pytest.mark.parametrize("algorithm,kw", [
    ("curvature", dict(mode="local", curvature_type="mean")),
    ("openness", dict(mode="local", num_directions=8, max_distance=40)),
])
def test_reflect_edge_mode_two_streams_equal_one(cuda, algorithm, kw):
    # 3 x 3 tiles of 1024: every edge tile's padded copy is made on the stream
    # that reads it, so the two-stream run equals the one-stream run bit for bit
    from fujishadergpu_amd.scheduler import process_dem
    host_t = _raster(3072, 3072, 29)
    g = host_t.to(cuda)
    stats = (0.0, 1.0)
    one, _ = process_dem(g, algorithm, tile_size=1024, pixel_size=10.0, edge_mode="reflect",
                         global_stats=stats, streams=1, **kw)
    for _ in range(2):
        two, _ = process_dem(g, algorithm, tile_size=1024, pixel_size=10.0,
                             edge_mode="reflect", global_stats=stats, streams=2, **kw)
        assert torch.equal(torch.isnan(one), torch.isnan(two))
        assert torch.equal(torch.nan_to_num(one), torch.nan_to_num(two))


def test_reflect_edge_mode_rejects_topousm(cuda):
    from fujishadergpu_amd.algorithms import UnsupportedOnBackend
    from fujishadergpu_amd.scheduler import process_dem
    host_t = _raster(256, 256, 3)
    with pytest.raises(UnsupportedOnBackend):
        process_dem(host_t.to(cuda), "topousm_fast", tile_size=128, pixel_size=10.0,
                    edge_mode="reflect", radii=[4, 16], weights=[0.5, 0.5])


@pytest.mark.parametrize("nan_discs", [0, 30])
def test_fused_scan_path_equals_the_count_path(cuda, nan_discs):
    """Without a NoData sentinel to convert, the scheduler reads the raster once
    (NaN count, factor-4 overview, pyramid cells) and builds the overview
    field after the synchronisation; with a sentinel it converts and counts per
    window first and builds the field the NaN-aware way.  An 8192^2 raster run
    both ways (the sentinel absent from the data) gives the same statistics and
    outputs bit for bit: NaN-free (pyramid, batched pre-pass means) and with
    NaN (per-window counts, no pyramid)."""
    from fujishadergpu_amd.scheduler import process_dem
    H = W = 8192
    host_t = _raster(H, W, 53, relief=3776.0, hurst=0.75)
    if nan_discs:
        from fujishadergpu_amd import synth
        synth.apply_discs(host_t, synth.nodata_discs(H, W, 0.01, seed=54)[:nan_discs],
                          float("nan"))
    g = host_t.to(cuda)
    radii = [2, 8, 32, 128, 512, 2048]
    weights = [2.0 ** (5 - i) / 63.0 for i in range(6)]
    kw = dict(tile_size=2048, pixel_size=5.0, mode="spatial", radii=radii, weights=weights,
              output_dtype="float32", copy_input=True, return_plan=True)
    a, ra, Pa = process_dem(g, "topousm_fast", nodata=None, **kw)
    b, rb, Pb = process_dem(g, "topousm_fast", nodata=-1.0e30, **kw)
    assert Pa.extras["overview_scale"] == Pb.extras["overview_scale"] == 4.0
    assert torch.equal(Pa.stats, Pb.stats)
    na, nb = torch.isnan(a), torch.isnan(b)
    assert torch.equal(na, nb)
    assert bool((na.any()) == bool(nan_discs))
    assert torch.equal(a[~na], b[~nb])


def test_a_failing_tile_aborts_the_run(cuda, monkeypatch):
    # a tile whose compute raises becomes an error record; the run then
    # aborts with the first three messages (core/tile_processor.py:1022-1024,
    # 1689-1692), after every other tile ran
    from fujishadergpu_amd import algorithms
    from fujishadergpu_amd.scheduler import process_dem
    host_t = _raster(600, 700, 37)
    calls = []
    real = algorithms.SlopeAlgorithm.run_region

    def flaky(self, win, reg, out, p):
        calls.append(1)
        if len(calls) in (2, 5, 6, 7):
            raise RuntimeError(f"planted failure {len(calls)}")
        return real(self, win, reg, out, p)

    monkeypatch.setattr(algorithms.SlopeAlgorithm, "run_region", flaky)
    with pytest.raises(RuntimeError) as e:
        process_dem(host_t.to(cuda), "slope", tile_size=256, pixel_size=10.0, mode="local")
    assert len(calls) == 9
    lines = str(e.value).splitlines()
    assert lines[0] == "Tile processing error:"
    assert lines[1:] == ["Tile (0, 1): planted failure 2", "Tile (1, 1): planted failure 5",
                         "Tile (1, 2): planted failure 6"]


def test_prepass_with_no_valid_pooled_value_leaves_per_tile_stats(cuda):
    # valid data only in a 40 px frame: the one stratified window (the whole
    # raster, margin 40 + 16) holds 10 % valid pixels, but its trimmed
    # interior is all NoData, so the
    # reference pools nothing, returns no global stats (_norm_stats.py:280-283)
    # and every tile estimates its own p99 (tile/dask_bridge.py:261-270)
    from fujishadergpu_amd.scheduler import process_dem
    H = W = 1500
    host_t = _raster(H, W, 81, relief=1500.0)
    host_t[40:-40, 40:-40] = -9999.0
    radii, weights = [2, 8, 40], [0.5, 0.3, 0.2]
    out, results, P = process_dem(host_t.to(cuda), "topousm_fast", tile_size=512,
                                  pixel_size=5.0, mode="spatial", radii=radii, weights=weights,
                                  nodata=-9999.0, return_plan=True)
    assert P.stats is None and "global_stats" not in P.params
    skipped = [r for r in results if r.skipped_reason]
    assert [(r.tile_y, r.tile_x) for r in skipped] == [(1, 1)]
    exp = _oracle_tiles(host_t.numpy(), P, "topousm_fast", -9999.0, stats=None)
    got = out.cpu().numpy()
    m = ~np.isnan(exp)
    assert np.array_equal(np.isnan(got), ~m)
    assert np.abs(got[m] - exp[m]).max() <= 2e-5

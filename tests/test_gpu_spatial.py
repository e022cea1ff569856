"""Spatial mode (multi-radius hillshade, slope, curvature) against the oracle's
restatement of the reference's single-chunk Dask path and overview path."""
import numpy as np
import pytest
import torch

from oracle import geometry as geo
from oracle import ref

pytestmark = pytest.mark.gpu


def _dem(h, w, seed, nan_discs=0, relief=1800.0, hurst=0.8):
    from fujishadergpu_amd import synth
    d = synth.fbm(h, w, hurst, relief, seed=seed, device="cpu")
    if nan_discs:
        discs = synth.nodata_discs(h, w, 0.01, seed=seed + 1)[:nan_discs]
        synth.apply_discs(d, discs, float("nan"))
    return d


def _cmp(name, got, exp, atol, rtol=0.0):
    got = got.detach().cpu().numpy() if isinstance(got, torch.Tensor) else got
    nm_g, nm_e = np.isnan(got), np.isnan(exp)
    assert np.array_equal(nm_g, nm_e), f"{name}: NaN mask differs ({(nm_g != nm_e).sum()} px)"
    d = np.abs(got.astype(np.float64) - exp.astype(np.float64))[~nm_e]
    lim = atol + rtol * np.abs(exp[~nm_e])
    bad = d > lim
    assert not bad.any(), f"{name}: {bad.sum()} px over tolerance, max |d| = {d.max():.3g}"
    return float(d.max()) if d.size else 0.0


TOL = {"hillshade": (2e-5, 0.0), "slope": (2e-4, 2e-5), "curvature": (4e-4, 0.0)}


@pytest.mark.parametrize("algorithm", ["hillshade", "slope", "curvature"])
@pytest.mark.parametrize("radii,weights,nan_discs", [
    ([3, 12, 30], [0.5, 0.3, 0.2], 0),          # Gaussian sigma r / 2 on the padded window
    ([2, 64], None, 4),                        # r = 64: decimated by 2 (2**n weights), NoData
    ([40], None, 0),                           # one radius: the field itself
])
def test_spatial_window_matches_the_oracle(cuda, algorithm, radii, weights, nan_discs):
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(300, 340, 41, nan_discs=nan_discs)
    kw = dict(mode="spatial", radii=radii, weights=weights, pixel_size=10.0, pixel_scale_x=10.0,
              pixel_scale_y=-10.0)
    if algorithm == "curvature":
        kw["curvature_type"] = "mean"
    got = ALGORITHMS[algorithm].process(dem.to(cuda), **kw)
    exp = ref.spatial_window(algorithm, dem.numpy(), kw)
    mx = _cmp(f"{algorithm}/spatial", got, exp, *TOL[algorithm])
    print(algorithm, radii, "max |d|", mx, flush=True)


@pytest.mark.parametrize("agg", ["max", "min", "mean"])
def test_hillshade_local_multiscale_matches_the_direct_path(cuda, agg):
    # _direct_hillshade: the spatial block on the window itself, _combine_direct
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(200, 230, 43)
    kw = dict(mode="local", multiscale=True, radii=[2, 9], weights=None, agg=agg,
              pixel_size=10.0)
    got = ALGORITHMS["hillshade"].process(dem.to(cuda), **kw).cpu().numpy()
    a = dem.numpy()
    fields = [ref.spatial_block("hillshade", a, r, {"pixel_size": 10.0}) for r in (2.0, 9.0)]
    if agg == "max":
        exp = np.maximum(fields[0], fields[1])
    elif agg == "min":
        exp = np.minimum(fields[0], fields[1])
    else:
        inv = np.float32(1.0 / 2.0)
        exp = (np.zeros_like(fields[0]) + fields[0] * inv) + fields[1] * inv
    _cmp("hillshade/direct", got, exp.astype(np.float32), 2e-5)


@pytest.mark.parametrize("algorithm", ["hillshade", "slope"])
def test_scheduler_spatial_with_the_overview(cuda, algorithm):
    # 2048^2 raster, tiles of 1024: r = 300 exceeds the large-radius threshold
    # (256) and comes from the overview response sampled at each tile window
    from fujishadergpu_amd.scheduler import process_dem
    host_t = _dem(2048, 2048, 47, relief=2500.0)
    kw = dict(mode="spatial", radii=[4, 40, 300], weights=[0.5, 0.3, 0.2])
    out, _, P = process_dem(host_t.to(cuda), algorithm, tile_size=1024, pixel_size=10.0,
                            return_plan=True, **kw)
    host = host_t.numpy()
    ov = host.copy()          # overview decimation 1 at 2048 px (sample_max 2048)
    want = np.full(host.shape, np.nan, np.float32)
    tiles = geo.plan_tiles(algorithm, P.params, host.shape[1], host.shape[0], P.tile_size)
    assert tiles == P.tiles
    for ty, tx, cx, cy, cw, ch, wx, wy, ww, wh in tiles:
        p = {k: v for k, v in P.params.items()
             if not isinstance(v, (torch.Tensor, dict)) and not str(k).startswith("_overview")}
        res = ref.spatial_window(algorithm, host[wy:wy + wh, wx:wx + ww], p, overview=ov,
                                 decimation=1.0, tile_origin=(wy, wx), full_shape=host.shape)
        want[cy:cy + ch, cx:cx + cw] = ref.format_output(
            res[cy - wy:cy - wy + ch, cx - wx:cx - wx + cw], algorithm)
    got = out.cpu().numpy() if out.dtype == torch.float32 else None
    assert got is not None
    _cmp(f"{algorithm}/scheduler", got, want, *TOL[algorithm])


@pytest.mark.parametrize("algorithm", ["hillshade", "curvature"])
def test_scheduler_spatial_with_a_fractional_overview(cuda, algorithm):
    # 2600 x 2200: the overview is read at decimation 2600 / 2048 = 1.27 onto
    # 1733 x 2048 cells (area-weighted average, not a block mean); r = 300 is
    # served from it at every tile
    from fujishadergpu_amd.scheduler import process_dem
    H, W = 2200, 2600
    host_t = _dem(H, W, 83, relief=2500.0)
    kw = dict(mode="spatial", radii=[4, 40, 300], weights=[0.5, 0.3, 0.2])
    out, _, P = process_dem(host_t.to(cuda), algorithm, tile_size=1024, pixel_size=10.0,
                            return_plan=True, **kw)
    host = host_t.numpy()
    ov, _ = ref.overview(host)
    assert ov.shape == (1733, 2048)
    assert tuple(P.params["_overview_coarse_dem"].shape) == ov.shape
    dec = 0.5 * (W / 2048.0 + H / 1733.0)
    assert P.params["_overview_decimation"] == pytest.approx(dec)
    tiles = geo.plan_tiles(algorithm, P.params, W, H, P.tile_size)
    assert tiles == P.tiles
    got = out.cpu().numpy()
    p = {k: v for k, v in P.params.items()
         if not isinstance(v, (torch.Tensor, dict)) and not str(k).startswith("_overview")}
    for ty, tx, cx, cy, cw, ch, wx, wy, ww, wh in (tiles[0], tiles[4], tiles[-1]):
        res = ref.spatial_window(algorithm, host[wy:wy + wh, wx:wx + ww], p, overview=ov,
                                 decimation=dec, tile_origin=(wy, wx), full_shape=host.shape)
        exp = ref.format_output(res[cy - wy:cy - wy + ch, cx - wx:cx - wx + cw], algorithm)
        _cmp(f"{algorithm}/tile{ty}{tx}", got[cy:cy + ch, cx:cx + cw], exp, *TOL[algorithm])


@pytest.mark.parametrize("sigma", [1.0, 4.5, 7.3, 16.0, 40.0])
@pytest.mark.parametrize("nan_discs", [0, 1])
def test_gaussian_nan_matches_the_oracle(cuda, sigma, nan_discs):
    # LDS-tiled two-pass Gaussian: r = 4, 18 (r mod 16 = 2), 29, 64, 160 taps
    # per side, tiles with and without NaN in their span, symmetric padding
    from fujishadergpu_amd import spatial
    dem = _dem(333, 290, 53, nan_discs=nan_discs, relief=2500.0)
    if nan_discs:
        dem[:6, :9] = float("nan")        # one corner tile holds NaN, the others none
    pad = 37
    got = spatial.gaussian_nan(dem.to(cuda), pad, sigma).cpu().numpy()
    exp, _ = ref.handle_nan_with_gaussian(ref.dask_reflect_pad(dem.numpy(), pad), sigma,
                                          mode="nearest")
    assert got.shape == exp.shape
    nm = np.isnan(exp)
    assert not np.isnan(got).any() and not nm.any()
    d = np.abs(got.astype(np.float64) - exp.astype(np.float64))
    lim = 2.4e-7 * np.abs(exp.astype(np.float64)) + 1e-30
    print(f"sigma {sigma} nan {nan_discs}: max |d| {d.max():.3g}, "
          f"identical {np.mean(got == exp):.6f}", flush=True)
    assert (d <= lim).all(), f"{(d > lim).sum()} px beyond 2 ulp"


RAY_TOL = 2e-5


@pytest.mark.parametrize("algorithm", ["openness", "ambient_occlusion"])
@pytest.mark.parametrize("radii,weights,nan_discs", [
    ([3, 12], [0.6, 0.4], 0),                   # direct: the local ray kernel on the pad
    ([5, 64], None, 3),                         # r = 64 decimated (2 / 4), NoData
    ([40], None, 0),                            # one radius: the field itself
])
def test_ray_spatial_window_matches_the_oracle(cuda, algorithm, radii, weights, nan_discs):
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(300, 330, 61, nan_discs=nan_discs)
    kw = dict(mode="spatial", radii=radii, weights=weights, pixel_size=10.0,
              pixel_scale_x=10.0, pixel_scale_y=-10.0)
    if algorithm == "openness":
        kw["num_directions"] = 8
    else:
        kw["num_samples"] = 8
    got = ALGORITHMS[algorithm].process(dem.to(cuda), **kw)
    exp = ref.spatial_window(algorithm, dem.numpy(), kw)
    mx = _cmp(f"{algorithm}/spatial", got, exp, RAY_TOL)
    print(algorithm, radii, "max |d|", mx, flush=True)


@pytest.mark.parametrize("algorithm", ["openness", "ambient_occlusion"])
def test_ray_scheduler_spatial_with_the_overview_and_stretch(cuda, algorithm):
    # 2048^2 raster, tiles of 1024: r = 300 comes from the overview response;
    # the display stretch of the given [p1, p99] is the combine's epilogue
    from fujishadergpu_amd.scheduler import process_dem
    host_t = _dem(2048, 2048, 67, relief=2500.0)
    stats = (0.62, 0.97)
    kw = dict(mode="spatial", radii=[6, 48, 300], weights=[0.5, 0.3, 0.2], global_stats=stats)
    if algorithm == "openness":
        kw["num_directions"] = 8
    else:
        kw["num_samples"] = 8
    out, _, P = process_dem(host_t.to(cuda), algorithm, tile_size=1024, pixel_size=10.0,
                            return_plan=True, **kw)
    host = host_t.numpy()
    want = np.full(host.shape, np.nan, np.float32)
    tiles = geo.plan_tiles(algorithm, P.params, host.shape[1], host.shape[0], P.tile_size)
    assert tiles == P.tiles
    for ty, tx, cx, cy, cw, ch, wx, wy, ww, wh in tiles:
        p = {k: v for k, v in P.params.items()
             if not isinstance(v, (torch.Tensor, dict)) and not str(k).startswith("_overview")}
        res = ref.spatial_window(algorithm, host[wy:wy + wh, wx:wx + ww], p, overview=host.copy(),
                                 decimation=1.0, tile_origin=(wy, wx), full_shape=host.shape)
        res = ref.display_stretch(res, stats)
        want[cy:cy + ch, cx:cx + cw] = ref.format_output(
            res[cy - wy:cy - wy + ch, cx - wx:cx - wx + cw], algorithm)
    assert out.dtype == torch.float32
    _cmp(f"{algorithm}/scheduler", out, want, RAY_TOL)


@pytest.mark.parametrize("algorithm,radii", [("hillshade", [3, 12, 30]), ("slope", [40]),
                                             ("curvature", [2, 64]),
                                             ("openness", [1, 2, 24]),
                                             ("ambient_occlusion", [4, 16])])
def test_stack_is_band_first_one_band_per_radius(cuda, algorithm, radii):
    # agg='stack': C x H x W, band i the field of radius i unchanged, even for
    # one radius (_nan_utils.py:191-194); openness keeps the duplicate radius
    # max(2, r) makes of [1, 2]; openness / AO bands each take the stretch
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(260, 300, 61, nan_discs=3)
    kw = dict(mode="spatial", radii=radii, agg="stack", pixel_size=10.0)
    if algorithm == "openness":
        kw.update(num_directions=8, global_stats=(0.6, 0.35))
    if algorithm == "ambient_occlusion":
        kw.update(num_samples=8, global_stats=(0.5, 0.4))
    got = ALGORITHMS[algorithm].process(dem.to(cuda), **kw)
    assert tuple(got.shape) == (len(radii), 260, 300)
    exp = ref.spatial_window(algorithm, dem.numpy(), kw)
    if algorithm in ("openness", "ambient_occlusion"):
        exp = ref.display_stretch(exp, kw["global_stats"])
    assert exp.shape == tuple(got.shape)
    atol, rtol = TOL.get(algorithm, (RAY_TOL, 0.0))
    for i in range(len(radii)):
        _cmp(f"{algorithm}/band{i}", got[i], exp[i], atol, rtol)
    if algorithm == "openness":
        assert torch.equal(torch.nan_to_num(got[0]), torch.nan_to_num(got[1]))


def test_stack_of_the_direct_hillshade_path(cuda):
    # local mode, several radii: _direct_hillshade's _combine_direct stack
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(200, 230, 43)
    got = ALGORITHMS["hillshade"].process(dem.to(cuda), mode="local", radii=[2, 9], agg="stack",
                                          pixel_size=10.0)
    assert tuple(got.shape) == (2, 200, 230)
    for i, r in enumerate((2.0, 9.0)):
        exp = ref.spatial_block("hillshade", dem.numpy(), r, {"pixel_size": 10.0})
        _cmp(f"direct/band{i}", got[i], exp, 2e-5)
    one = ALGORITHMS["hillshade"].process(dem.to(cuda), mode="local", agg="stack",
                                          pixel_size=10.0)
    assert one.dim() == 2          # the local block function ignores agg


def test_scheduler_stack_with_the_overview_and_uint8(cuda):
    # tiles write every band; r = 300 comes from the overview; DN per band
    from fujishadergpu_amd.scheduler import process_dem
    host_t = _dem(2048, 2048, 47, relief=2500.0)
    kw = dict(mode="spatial", radii=[4, 40, 300], agg="stack")
    f32 = torch.empty((3, 2048, 2048), dtype=torch.float32, device=cuda)
    out, _, P = process_dem(host_t.to(cuda), "hillshade", tile_size=1024, pixel_size=10.0,
                            output_dtype="uint8", out_f32=f32, return_plan=True, **kw)
    assert tuple(out.shape) == (3, 2048, 2048) and out.dtype == torch.uint8
    host = host_t.numpy()
    tiles = geo.plan_tiles("hillshade", P.params, 2048, 2048, P.tile_size)
    assert tiles == P.tiles
    want = np.full((3, 2048, 2048), np.nan, np.float32)
    p = {k: v for k, v in P.params.items()
         if not isinstance(v, (torch.Tensor, dict)) and not str(k).startswith("_overview")}
    for ty, tx, cx, cy, cw, ch, wx, wy, ww, wh in tiles:
        res = ref.spatial_window("hillshade", host[wy:wy + wh, wx:wx + ww], p,
                                 overview=host.copy(), decimation=1.0, tile_origin=(wy, wx),
                                 full_shape=host.shape)
        want[:, cy:cy + ch, cx:cx + cw] = ref.format_output(
            res[:, cy - wy:cy - wy + ch, cx - wx:cx - wx + cw], "hillshade")
    for i in range(3):
        _cmp(f"stack/band{i}", f32[i], want[i], *TOL["hillshade"])
    qp = P.qp
    dn = ref.quantize(f32.cpu().numpy(), qp["a_coef"], qp["b_coef"], qp["dn_min"], qp["dn_max"],
                      "uint8")
    assert np.array_equal(out.cpu().numpy(), dn)

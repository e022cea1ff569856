"""GPU parity: each HIP kernel (through the C ABI) against the CPU oracle.

Tolerances (float32 outputs):
* hillshade: bit-exact except where a NaN fill enters (|d| <= 1e-6);
* slope (degrees): |d| <= 2e-5 + 1e-6*|v|  (atanf implementations differ by ulps);
* curvature display value: |d| <= 4e-4, and |d| <= 2e-6 in the tanh domain
  (one tanh ulp at saturation becomes 3.8e-4 after the 1/2.2 gamma);
* topousm (normalised units): |d| <= 2e-5 (float32 box means of ~3000 m
  elevations: one f32 ulp of a mean is ~2.4e-4 m, ~1.2e-5 after the p99
  divide; observed 0 except for 3.8e-6 and 7.6e-6 in two cases);
* openness / AO display values: |d| <= 1e-5.
NaN masks must match exactly; integer encodings must equal the oracle's
quantisation of the GPU float values bit for bit.
"""
import numpy as np
import pytest
import torch

from oracle import ref

pytestmark = pytest.mark.gpu


def _dem(h, w, seed, relief=3000.0, hurst=0.8, nan_discs=0):
    from fujishadergpu_amd import synth
    d = synth.fbm(h, w, hurst, relief, seed=seed, device="cpu")
    if nan_discs:
        rng = np.random.default_rng(seed + 99)
        for _ in range(nan_discs):
            cy, cx, r = int(rng.integers(0, h)), int(rng.integers(0, w)), int(rng.integers(2, 9))
            yy, xx = np.ogrid[:h, :w]
            m = (yy - cy) ** 2 + (xx - cx) ** 2 <= r * r
            d[torch.from_numpy(m)] = float("nan")
    return d


def _cmp(name, got, exp, atol, rtol=0.0):
    got = got.detach().cpu().numpy() if isinstance(got, torch.Tensor) else got
    nm_g, nm_e = np.isnan(got), np.isnan(exp)
    assert np.array_equal(nm_g, nm_e), f"{name}: NaN mask differs ({(nm_g != nm_e).sum()} px)"
    d = np.abs(got - exp)[~nm_e]
    lim = atol + rtol * np.abs(exp[~nm_e])
    bad = d > lim
    assert not bad.any(), (f"{name}: {bad.sum()} px over tolerance, max |d| = {d.max():.3g}")
    return float(d.max()) if d.size else 0.0, float((d == 0).mean()) if d.size else 1.0


@pytest.mark.parametrize("nan_discs", [0, 6])
@pytest.mark.parametrize("psx,psy", [(10.0, -10.0), (None, None), (3.0, 7.5)])
def test_hillshade_slope_window(cuda, nan_discs, psx, psy):
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(131, 157, 3, nan_discs=nan_discs)
    g = dem.to(cuda)
    kw = dict(pixel_size=10.0, pixel_scale_x=psx, pixel_scale_y=psy)
    hs = ALGORITHMS["hillshade"].process(g, **kw)
    sl = ALGORITHMS["slope"].process(g, **kw)
    a = dem.numpy()
    mx, exact = _cmp("hillshade", hs, ref.hillshade(a, **kw), 1e-6)
    if nan_discs == 0:
        assert exact == 1.0, f"hillshade not bit-exact ({exact:.4f}, max {mx})"
    _cmp("slope", sl, ref.slope(a, **kw), 2e-5, 1e-6)


@pytest.mark.parametrize("nan_discs", [0, 5])
@pytest.mark.parametrize("lat", [31.0, 58.5])
def test_stencils_with_per_row_x_spacing(cuda, nan_discs, lat):
    # a geographic 1 arc-second window: every row's x cell width at its own latitude
    from fujishadergpu_amd.algorithms import ALGORITHMS
    from fujishadergpu_amd.raster_info import meters_per_degree, row_geographic_scales_x
    dem = _dem(150, 170, 31, relief=1800.0, nan_discs=nan_discs)
    g = dem.to(cuda)
    rows = row_geographic_scales_x(1 / 3600, -1 / 3600, lat + 0.02, 0, 150)
    psy = -meters_per_degree(lat)[1] / 3600
    kw = dict(pixel_size=25.0, pixel_scale_y=psy)
    a = dem.numpy()
    hs = ALGORITHMS["hillshade"].process(g, pixel_scale_x_rows=rows, z_factor=1.5, **kw)
    _cmp("hillshade/rows", hs, ref.hillshade(a, pixel_scale_x=rows, z_factor=1.5, **kw), 1e-6)
    sl = ALGORITHMS["slope"].process(g, pixel_scale_x_rows=rows, **kw)
    _cmp("slope/rows", sl, ref.slope(a, pixel_scale_x=rows, **kw), 2e-5, 1e-6)
    for ctype in ("mean", "gaussian"):
        cv = ALGORITHMS["curvature"].process(g, pixel_scale_x_rows=rows, curvature_type=ctype,
                                             **kw).cpu().numpy()
        exp = ref.curvature(a, pixel_scale_x=rows, curvature_type=ctype, **kw)
        _cmp(f"curvature/rows/{ctype}", cv, exp, 4e-4)
        tg = 2.0 * np.power(cv.astype(np.float64), 2.2) - 1.0
        te = 2.0 * np.power(exp.astype(np.float64), 2.2) - 1.0
        _cmp(f"curvature/rows/{ctype}/tanh", tg, te, 2e-6)
    # equal rows give the scalar path's bits
    same = np.full(150, float(rows[75]))
    h1 = ALGORITHMS["hillshade"].process(g, pixel_scale_x_rows=same, **kw)
    h2 = ALGORITHMS["hillshade"].process(g, pixel_scale_x=float(rows[75]), **kw)
    assert torch.equal(h1.view(torch.int32), h2.view(torch.int32))    # NaN bits included


@pytest.mark.parametrize("z,az,alt,unit", [(2.5, 45.0, 30.0, "percent"), (1.0, 200.0, 60.0, "radian")])
def test_hillshade_slope_params(cuda, z, az, alt, unit):
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(64, 70, 5)
    g = dem.to(cuda)
    hs = ALGORITHMS["hillshade"].process(g, z_factor=z, azimuth=az, altitude=alt, pixel_size=5.0)
    sl = ALGORITHMS["slope"].process(g, unit=unit, pixel_size=5.0)
    a = dem.numpy()
    _cmp("hillshade", hs, ref.hillshade(a, z_factor=z, azimuth=az, altitude=alt, pixel_size=5.0), 1e-6)
    exp = ref.slope(a, unit=unit, pixel_size=5.0)
    if unit == "percent":
        # tan(atan(m)) * 100 has condition number ~m on steep cells: compare the
        # angle atan(v / 100), where one ulp of atan/tan is 1e-7 rad.
        got = sl.cpu().numpy()
        _cmp("slope%", np.arctan(got.astype(np.float64) / 100.0),
             np.arctan(exp.astype(np.float64) / 100.0), 1e-6)
    else:
        _cmp("slope", sl, exp, 2e-5, 2e-6)


@pytest.mark.parametrize("ctype", ["mean", "gaussian", "planform", "profile"])
@pytest.mark.parametrize("nan_discs", [0, 5])
def test_curvature_window(cuda, ctype, nan_discs):
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(97, 121, 7, relief=2500.0, hurst=0.85, nan_discs=nan_discs)
    g = dem.to(cuda)
    kw = dict(curvature_type=ctype, pixel_size=10.0, pixel_scale_x=10.0, pixel_scale_y=-10.0)
    out = ALGORITHMS["curvature"].process(g, **kw).cpu().numpy()
    exp = ref.curvature(dem.numpy(), **kw)
    _cmp(f"curvature/{ctype}", out, exp, 4e-4)
    # The display map v = ((t + 1) / 2) ** (1 / 2.2) turns one ulp of tanh at
    # saturation (t = -1 vs -1 + 6e-8) into 3.8e-4 in v; in the tanh domain
    # t = 2 v**2.2 - 1 the two implementations agree to 2e-6.
    tg = 2.0 * np.power(out.astype(np.float64), 2.2) - 1.0
    te = 2.0 * np.power(exp.astype(np.float64), 2.2) - 1.0
    _cmp(f"curvature/{ctype}/tanh", tg, te, 2e-6)
    d = np.abs(out - exp)[~np.isnan(exp)]
    assert (d <= 1e-6).mean() >= 0.9, f"only {(d <= 1e-6).mean():.3f} within 1e-6"


@pytest.mark.parametrize("nan_discs", [0, 6])
@pytest.mark.parametrize("off,ld", [(0, 1104), (1, 1104), (2, 1104), (3, 1104), (0, 1107)])
def test_stencils_wide_windows(cuda, nan_discs, off, ld):
    """Windows wide enough for the four-columns-per-lane interior kernels (a
    wave spans 256 columns).  The window view starts `off` floats past a
    16-byte boundary of a row stride that is a multiple of 4, so the aligned
    wave starts, the one-column side strips, the 8-byte output stores
    (off == 0) and the per-column stores (off != 0) all run; a row stride
    that is not a multiple of 4 (ld 1107) takes the one-column kernel; a
    region inside the window (output origin away from the window's) runs
    them again.  The split size is lowered so that these interiors take the
    row kernels (by default they are small enough for the general kernel)."""
    from fujishadergpu_amd import ops
    prev = ops.set_stencil_split_min(0)
    try:
        _stencils_wide_windows(cuda, nan_discs, off, ld)
    finally:
        ops.set_stencil_split_min(prev)


def _stencils_wide_windows(cuda, nan_discs, off, ld):
    from fujishadergpu_amd import ops
    from fujishadergpu_amd.algorithms import ALGORITHMS, CurvatureAlgorithm, \
        HillshadeSlopeFused, OutSpec
    dem = _dem(600, ld, 21 + off, nan_discs=3 * nan_discs)
    g = dem.to(cuda)[:, off:]
    a = dem[:, off:].numpy()
    kw = dict(pixel_size=10.0)
    hs = ALGORITHMS["hillshade"].process(g, **kw)
    sl = ALGORITHMS["slope"].process(g, **kw)
    ref_hs, ref_sl = ref.hillshade(a, **kw), ref.slope(a, **kw)
    mx, exact = _cmp("hillshade", hs, ref_hs, 1e-6)
    if not nan_discs:
        assert exact == 1.0, f"hillshade not bit-exact ({exact:.4f}, max {mx})"
    _cmp("slope", sl, ref_sl, 2e-5, 1e-6)
    cv = ALGORITHMS["curvature"].process(g, curvature_type="mean", **kw).cpu().numpy()
    ref_cv = ref.curvature(a, curvature_type="mean", **kw)
    _cmp("curvature", cv, ref_cv, 4e-4)
    d = np.abs(cv - ref_cv)[~np.isnan(ref_cv)]
    assert (d <= 1e-6).mean() >= 0.9, f"only {(d <= 1e-6).mean():.3f} within 1e-6"
    # a region of the window: rows 5..504, columns 37 .. 37+900
    r0, c0, rh, rw = 5, 37, 500, 900
    reg = ops.region(r0, c0, rh, rw)
    p = {"pixel_size": 10.0, "curvature_type": "mean", "_window_nan_free": not nan_discs}
    o_hs = torch.empty((rh, rw), dtype=torch.float32, device=cuda)
    o_sl = torch.empty_like(o_hs)
    o_cv = torch.empty_like(o_hs)
    HillshadeSlopeFused().run_region(g, reg, OutSpec(f32=o_hs, clip01=True), p,
                                     out_slope=OutSpec(f32=o_sl))
    CurvatureAlgorithm().run_region(g, reg, OutSpec(f32=o_cv), p)
    sub = (slice(r0, r0 + rh), slice(c0, c0 + rw))
    mx, exact = _cmp("hillshade/region", o_hs, ref_hs[sub], 1e-6)
    if not nan_discs:
        assert exact == 1.0, f"hillshade/region not bit-exact ({exact:.4f}, max {mx})"
    _cmp("slope/region", o_sl, ref_sl[sub], 2e-5, 1e-6)
    _cmp("curvature/region", o_cv, ref_cv[sub], 4e-4)


@pytest.mark.parametrize("radii,weights,px", [
    ([2, 8, 32], [0.5, 0.3, 0.2], 5.0),
    ([2, 8, 32, 128], [8 / 15, 4 / 15, 2 / 15, 1 / 15], 5.0),
    ([1], [1.0], 10.0),
    ([4, 64], None, 10.0),
    # unsorted: a box radius after a decimated one (general kernel)
    ([32, 2, 128, 8], [0.2, 0.4, 0.1, 0.3], 5.0),
    # decimated radii only (no full-resolution box term)
    ([128, 300], [0.6, 0.4], 5.0),
])
@pytest.mark.parametrize("nan_discs", [0, 8])
def test_topousm_window(cuda, radii, weights, px, nan_discs):
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(389, 301, 11, relief=3776.0, hurst=0.75, nan_discs=nan_discs)
    g = dem.to(cuda)
    out = ALGORITHMS["topousm_fast"].process(g, radii=radii, weights=weights, pixel_size=px,
                                             global_stats=(17.0,))
    exp = ref.topousm_direct(dem.numpy(), radii=radii, weights=weights, pixel_size=px,
                             stats=(17.0,))
    mx_, _ = _cmp("topousm", out, exp, 2e-5)
    print("maxdiff topousm", mx_, flush=True)


def test_topousm_field_and_two_outputs(cuda):
    from fujishadergpu_amd import ops
    from fujishadergpu_amd.algorithms import OutSpec, TopoUSMFastAlgorithm
    from fujishadergpu_amd.encoding import quantize_params
    H = W = 512
    dem = _dem(H, W, 13, relief=3776.0, hurst=0.75)
    g = dem.to(cuda)
    coarse_np, scale = ref.overview_block_mean(dem.numpy(), sample_max=128)
    field_np = ref.topousm_large_coarse_field(coarse_np, large_radii=[300, 512],
                                              large_weights=[0.05, 0.02], decimation=scale)
    coarse = ops.block_mean(g, int(scale))
    field = ops.topousm_coarse_field(coarse, [max(1, int(round(r / scale))) for r in (300, 512)],
                                     [0.05, 0.02])
    _cmp("overview", coarse, coarse_np, 0.0)
    mx_, _ = _cmp("field", field, field_np, 2e-5)
    print("maxdiff field", mx_, flush=True)
    # one tile window (rows 100..400, cols 60..460) with the field
    wy, wx, wh, ww = 100, 60, 300, 400
    win = g[wy:wy + wh, wx:wx + ww]
    p = dict(radii=[2, 8, 32], weights=[0.4, 0.3, 0.2], pixel_size=5.0,
             _topousm_fast_coarse_field=field, _topousm_fast_small_radii=[2, 8, 32],
             _topousm_fast_small_weights=[0.4, 0.3, 0.2], _topousm_fast_w_large=0.07,
             _topousm_fast_full_shape=(H, W), _topousm_fast_field_offset=(wy, wx),
             global_stats=(9.0,))
    algo = TopoUSMFastAlgorithm()
    qp = quantize_params(-1.176, 1.176, "uint8")
    f32 = torch.empty((wh, ww), dtype=torch.float32, device=cuda)
    u8 = torch.empty((wh, ww), dtype=torch.uint8, device=cuda)
    algo.run_region(win, ops.full_region(win), OutSpec(f32=f32, q=u8, qp=qp), p)
    exp = ref.topousm_direct(dem.numpy()[wy:wy + wh, wx:wx + ww], radii=[2, 8, 32],
                             weights=[0.4, 0.3, 0.2], pixel_size=5.0, coarse_field=field_np,
                             small_radii=[2, 8, 32], small_weights=[0.4, 0.3, 0.2],
                             w_large=0.07, field_offset=(wy, wx), full_shape=(H, W),
                             stats=(9.0,))
    mx_, _ = _cmp("topousm+field", f32, exp, 2e-5)
    print("maxdiff topousm+field", mx_, flush=True)
    dn = ref.quantize(f32.cpu().numpy(), qp["a_coef"], qp["b_coef"], qp["dn_min"], qp["dn_max"],
                      "uint8")
    assert np.array_equal(u8.cpu().numpy(), dn)


@pytest.mark.parametrize("h,w,oh,ow", [(700, 1000, 287, 410), (2600, 2200, 2048, 1733),
                                       (999, 1001, 998, 1000), (640, 480, 80, 60)])
@pytest.mark.parametrize("nan_discs", [0, 20])
def test_overview_average_matches_the_oracle(cuda, h, w, oh, ow, nan_discs):
    # the area-weighted overview read at fractional (and one integer) ratios
    from fujishadergpu_amd import ops
    dem = _dem(h, w, 41 + h, nan_discs=nan_discs)
    got = ops.overview_average(dem.to(cuda), oh, ow)
    exp = ref.overview_average(dem.numpy(), oh, ow)
    # float64 sums in another order, rounded once: at most one float32 ulp
    mx, _ = _cmp("overview", got, exp, 0.0, rtol=1.2e-7)
    print("overview", (h, w, oh, ow), "max |d|", mx, flush=True)
    if h % oh == 0 and w % ow == 0 and not nan_discs:
        # an integer ratio over whole cells is the block mean, bit for bit
        bm = ops.block_mean(dem.to(cuda), h // oh)
        assert torch.equal(got, bm)


def test_topousm_shared_coarse_means(cuda):
    """A window reading raster-wide coarse means (CoarsePyramid) equals the
    same window building its own, and the oracle."""
    from fujishadergpu_amd import ops
    from fujishadergpu_amd.algorithms import CoarsePyramid, OutSpec, TopoUSMFastAlgorithm
    H = W = 2048
    dem = _dem(H, W, 31, relief=3776.0, hurst=0.75)
    g = dem.to(cuda)
    radii, weights = [2, 8, 32, 128, 200], [0.4, 0.25, 0.15, 0.12, 0.08]
    p = dict(radii=radii, weights=weights, pixel_size=5.0, global_stats=(20.0,),
             _window_nan_free=True)
    algo = TopoUSMFastAlgorithm()
    pyr = CoarsePyramid(g)
    pyr.require(pyr.pairs(algo.params_struct(p)))
    assert sorted(pyr.grids) == [(4, 32), (8, 25)]
    wy, wx, wh, ww = 256, 512, 1536, 1536
    win = g[wy:wy + wh, wx:wx + ww]
    reg = ops.region(256, 256, 1024, 1024)
    assert pyr.attach(algo.params_struct(p), (wy, wx), (wh, ww), reg) == 2
    own = torch.empty((1024, 1024), dtype=torch.float32, device=cuda)
    shared = torch.empty_like(own)
    algo.run_region(win, reg, OutSpec(f32=own), p)
    algo.run_region(win, reg, OutSpec(f32=shared),
                    {**p, "_topousm_fast_pyramid": pyr, "_window_origin": (wy, wx)})
    d = float((own - shared).abs().max())
    print(f"shared vs own coarse means: max |d| = {d:.3g}", flush=True)
    assert d <= 1e-6
    exp = ref.topousm_direct(dem.numpy()[wy:wy + wh, wx:wx + ww], radii=radii, weights=weights,
                             pixel_size=5.0, stats=(20.0,))[256:1280, 256:1280]
    mx_, _ = _cmp("topousm/shared", shared, exp, 2e-5)
    print("maxdiff topousm/shared", mx_, flush=True)


def test_topousm_coarse_means_batch_equals_per_window(cuda):
    """The batched coarse means (one decimation launch with two factors sharing
    a read, one launch per box pass) equal the per-window entry point bit for
    bit: aligned and unaligned window views, ragged edges, single and paired
    factors, several radii per factor; and a pre-pass window reading them
    equals the same window building its own."""
    from fujishadergpu_amd import ops
    from fujishadergpu_amd.algorithms import OutSpec, TopoUSMFastAlgorithm
    dem = _dem(1500, 1700, 37, relief=3776.0, hurst=0.75)
    g = dem.to(cuda)
    w1 = g[3:1403, 5:1617]          # unaligned origin, 1400 x 1612
    w2 = g[64:1364, 128:1152]       # float4-aligned origin, 1300 x 1024
    w3 = g[0:1037, 0:1501]
    jobs = [(w1, 16, [32, 128]), (w1, 4, [32]), (w2, 4, [2, 9]), (w2, 8, [5]),
            (w3, 8, [3, 17]), (w3, 16, [2])]
    got = ops.topousm_coarse_means_batch(jobs)
    for (w, ds, rs), t in zip(jobs, got):
        want = ops.topousm_coarse_means(w, ds, rs)
        assert t.shape == want.shape
        assert torch.equal(t, want), (ds, rs)
    # a window of 1400 x 1612 reading them as its own means
    algo = TopoUSMFastAlgorithm()
    p = dict(radii=[2, 8, 32, 128, 512, 2048], weights=[0.3, 0.2, 0.15, 0.15, 0.1, 0.1],
             pixel_size=5.0, _window_nan_free=True)
    prm = algo.params_struct(p)
    assert [int(prm.ds[i]) for i in range(prm.n_radii)] == [1, 1, 1, 4, 16, 16]
    own = ops.topousm_coarse_means_batch([(w1, 16, [32, 128]), (w1, 4, [32])])
    reg = ops.region(300, 300, 800, 1012)
    a = torch.empty((800, 1012), dtype=torch.float32, device=cuda)
    b = torch.empty_like(a)
    algo.run_region(w1, reg, OutSpec(f32=a), p, normalize=False)
    algo.run_region(w1, reg, OutSpec(f32=b),
                    {**p, "_topousm_fast_means": {3: own[1][0], 4: own[0][0], 5: own[0][1]}},
                    normalize=False)
    assert torch.equal(a, b)


@pytest.mark.parametrize("shape,view", [((1500, 1700), False), ((1037, 2050), True),
                                        ((64, 100), False), ((140001, 37), False)])
def test_raster_scan_equals_block_means_and_nan_count(cuda, shape, view):
    """One read gives the NaN count and the block means at 4, 8 and 16, each
    equal to fsg_block_mean's bit for bit (ragged edges; an unaligned view
    takes the scalar loads); pre-decimated grids (ds 1 jobs) give the same
    coarse means as decimating the raster."""
    from fujishadergpu_amd import ops
    H, W = shape
    dem = _dem(H + 3, W + 5, 41, relief=3776.0, hurst=0.75, nan_discs=5)
    g = dem.to(cuda)
    src = g[1:1 + H, 3:3 + W] if view else g[:H, :W].contiguous()
    n_nan = int(torch.isnan(src).sum())
    assert n_nan > 0
    for factors in ([4, 8, 16], [16], [8, 4]):
        cnt, grids = ops.raster_scan(src, factors)
        assert int(cnt.item()) == n_nan
        for f, t in zip(factors, grids):
            want = ops.block_mean(src, f)
            assert torch.equal(torch.isnan(t), torch.isnan(want)), f
            m = ~torch.isnan(want)
            assert torch.equal(t[m], want[m]), f
    clean = torch.nan_to_num(src, nan=1000.0)
    if min(H, W) >= 1000:
        _, (g4, g16) = ops.raster_scan(clean, [4, 16])
        a = ops.topousm_coarse_means_batch([(g4, 1, [32]), (g16, 1, [32, 9])])
        b = ops.topousm_coarse_means_batch([(clean, 4, [32]), (clean, 16, [32, 9])])
        for x, y in zip(a, b):
            assert torch.equal(x, y)


def test_field_from_batched_means_equals_coarse_field(cuda):
    """The NaN-free overview field (batched box means, one weighted-sum pass)
    equals fsg_topousm_coarse_field bit for bit."""
    from fujishadergpu_amd import ops
    ov = _dem(700, 900, 43, relief=3776.0, hurst=0.75).to(cuda)
    rc, wts = [64, 256, 9], [0.0317, 0.0159, 0.21]
    want = ops.topousm_coarse_field(ov, rc, wts)
    means = ops.topousm_coarse_means_batch([(ov, 1, rc)])[0]
    got = ops.topousm_field_from_means([means[i] for i in range(3)], wts)
    assert torch.equal(got, want)


def test_topousm_coarse_means_batch_rejects_what_it_does_not_batch(cuda):
    from fujishadergpu_amd import _native as N, ops
    g = _dem(300, 300, 5).to(cuda)
    with pytest.raises(N.FsgError):
        ops.topousm_coarse_means_batch([(g, 2, [8])])      # ds 2: per-window entry point
    with pytest.raises(N.FsgError):
        ops.topousm_coarse_means_batch([(g, 4, [1])])      # Gaussian branch
    with pytest.raises(ValueError):
        ops.topousm_coarse_means_batch([(g, 4, [2, 3, 4, 5, 6])])


@pytest.mark.parametrize("otype", ["positive", "negative"])
@pytest.mark.parametrize("nan_discs", [0, 6])
def test_openness_window(cuda, otype, nan_discs):
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(150, 170, 17, relief=2500.0, hurst=0.85, nan_discs=nan_discs)
    g = dem.to(cuda)
    kw = dict(openness_type=otype, num_directions=8, max_distance=64, pixel_size=10.0,
              pixel_scale_x=10.0, pixel_scale_y=-10.0)
    out = ALGORITHMS["openness"].process(g, **kw)
    _cmp(f"openness/{otype}", out, ref.openness(dem.numpy(), **kw), 1e-5)
    st = (0.62, 0.21)
    out2 = ALGORITHMS["openness"].process(g, global_stats=st, **kw)
    _cmp("openness+stretch", out2, ref.display_stretch(ref.openness(dem.numpy(), **kw), st), 1e-4)


@pytest.mark.parametrize("f", [2, 4, 8, 16, 64, 3])
def test_block_mean_factors(cuda, f):
    from fujishadergpu_amd import ops
    dem = _dem(150, 171, 29, nan_discs=4).numpy()
    got = ops.block_mean(torch.from_numpy(dem).to(cuda), f).cpu().numpy()
    oh, ow = -(-150 // f), -(-171 // f)
    want = np.full((oh, ow), np.nan, np.float32)
    for i in range(oh):
        for j in range(ow):
            blk = dem[i * f:(i + 1) * f, j * f:(j + 1) * f].astype(np.float64)
            v = blk[np.isfinite(blk)]
            if v.size:
                want[i, j] = np.float32(np.float32(v.sum()) / np.float32(v.size))
    assert np.array_equal(np.isnan(got), np.isnan(want))
    m = ~np.isnan(want)
    assert np.allclose(got[m], want[m], rtol=1e-7, atol=0)


@pytest.mark.parametrize("nan_discs", [0, 6])
def test_openness_raw_degrees_points_and_judge(cuda, nan_discs):
    from fujishadergpu_amd import ops
    from fujishadergpu_amd.algorithms import openness_table
    dem = _dem(150, 170, 23, relief=2500.0, hurst=0.8, nan_discs=nan_discs)
    g = dem.to(cuda)
    kw = dict(openness_type="positive", num_directions=8, max_distance=48, pixel_size=10.0)
    want = ref.openness(dem.numpy(), raw_degrees=True, **kw)
    tab = openness_table(8, 48, 10.0, None, None)
    out = torch.empty((150, 170), dtype=torch.float32, device=cuda)
    ops.openness(g, ops.full_region(g), table=tab, openness_type="positive",
                 out=ops.output(out), raw_degrees=True)
    got = out.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(want))
    m = ~np.isnan(want)
    assert np.abs(got[m] - want[m]).max() < 2e-4        # degrees
    # points: the same function at listed pixels, bit for bit
    rng = np.random.default_rng(1)
    pts = np.stack([rng.integers(0, 150, 500), rng.integers(0, 170, 500)], 1).astype(np.int32)
    pv = ops.openness_points(g, torch.from_numpy(pts).to(cuda), table=tab,
                             openness_type="positive").cpu().numpy()
    assert np.array_equal(pv, got[pts[:, 0], pts[:, 1]], equal_nan=True)
    # judge: per-tile mean |exact - approx| over hashed row segments
    approx = want + rng.normal(0, 2.0, want.shape).astype(np.float32)
    approx[5, :] = np.nan
    L = 40
    mae = ops.openness_judge(g, torch.from_numpy(approx).to(cuda), tile_h=64, tile_w=96,
                             sample_frac=0.05, seed=1234, seg_len=L, table=tab).cpu().numpy()
    plan = ops.judge_segments(150, 170, 64, 96, 0.05, 1234, L)
    assert len(plan) == mae.size == 3 * 2
    for t_i, ((y0, x0, th, tw), segs) in enumerate(plan):
        e = []
        for (r_, c_) in segs:
            assert y0 <= r_ < y0 + th and x0 <= c_ and c_ + L <= x0 + tw
            d = np.abs(want[r_, c_:c_ + L].astype(np.float64) - approx[r_, c_:c_ + L])
            e += list(d[~np.isnan(d)])
        assert mae[t_i] == pytest.approx(np.mean(e), rel=1e-4, abs=3e-4), t_i


@pytest.mark.parametrize("nan_discs", [0, 6])
def test_ambient_occlusion_window(cuda, nan_discs):
    from fujishadergpu_amd.algorithms import ALGORITHMS
    dem = _dem(120, 140, 19, nan_discs=nan_discs)
    g = dem.to(cuda)
    kw = dict(num_samples=16, radius=10.0, intensity=1.0, pixel_size=10.0)
    out = ALGORITHMS["ambient_occlusion"].process(g, **kw)
    _cmp("ao", out, ref.ambient_occlusion(dem.numpy(), **kw), 1e-5)


@pytest.mark.parametrize("n", [1, 2, 1000, 123457])
def test_percentile_matches_numpy(cuda, n):
    from fujishadergpu_amd import ops
    rng = np.random.default_rng(n)
    v = (rng.standard_normal(n) * 7.0).astype(np.float32)
    if n > 1:
        v[rng.integers(0, n, size=max(1, n // 50))] = np.nan
    t = torch.from_numpy(v).to(cuda)
    pct, nval, mom = ops.percentile(t, [99.0], ops.PCT_ABS)
    fin = v[~np.isnan(v)]
    assert int(nval.item()) == fin.size
    exp = np.percentile(np.abs(fin).astype(np.float64), 99.0)
    assert abs(pct.item() - exp) <= 1e-12 * max(1.0, abs(exp))
    pct2, _, _ = ops.percentile(t, [1.0, 99.0], ops.PCT_FINITE)
    exp2 = np.percentile(fin.astype(np.float64), [1.0, 99.0])
    assert np.allclose(pct2.cpu().numpy(), exp2, rtol=1e-12, atol=1e-12)
    assert abs(mom[0].item() - fin.astype(np.float64).mean()) < 1e-9 * max(1, abs(fin).max())


def test_percentile_ties_and_reference_vectors(cuda):
    from fujishadergpu_amd import ops
    # tests/test_topousm_fast_normalization.py:27-36: 800 x 0.01 + 200 x 0.5 -> p99 = 0.5
    data = np.concatenate([np.full(800, 0.01, np.float32), np.full(200, 0.5, np.float32)])
    pct, _, mom = ops.percentile(torch.from_numpy(data).to(cuda), [99.0], ops.PCT_ABS)
    stats = ops.finalize_stats(0, pct, mom, torch.tensor([1000], device=cuda))
    assert stats[0].item() == pytest.approx(0.5, rel=1e-2)
    # all-equal values: p99 is the value
    pct, _, _ = ops.percentile(torch.full((4096,), 3.25, device=cuda), [99.0], ops.PCT_ABS)
    assert pct.item() == 3.25


def test_percentile_moments_stable_and_deterministic(cuda):
    """The moments behind the std fallback: data far from zero (where a
    one-pass sum of squares cancels) against numpy's two-pass float64 mean and
    std, the same bits on a second run, and the fallback itself: p99 of |v| is
    0 when 99.5 % of the pool is 0, so the scale is the pool's std."""
    from fujishadergpu_amd import ops
    rng = np.random.default_rng(5)
    v = (1.0e4 + rng.standard_normal(3_000_001) * 0.01).astype(np.float32)
    t = torch.from_numpy(v).to(cuda)
    _, _, m1 = ops.percentile(t, [99.0], ops.PCT_ABS)
    _, _, m2 = ops.percentile(t, [99.0], ops.PCT_ABS)
    a, b = m1.cpu().numpy(), m2.cpu().numpy()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (a, b)
    d = v.astype(np.float64)
    assert abs(a[0] - d.mean()) <= 1e-12 * d.mean()
    assert abs(a[1] - d.std()) <= 1e-9 * d.std(), (a[1], d.std())
    # the fallback: p99 = 0, scale = std
    w = np.zeros(200_000, np.float32)
    w[rng.choice(w.size, size=1000, replace=False)] = 1.0
    tw = torch.from_numpy(w).to(cuda)
    pct, nval, mom = ops.percentile(tw, [99.0], ops.PCT_ABS)
    assert pct.item() == 0.0
    stats = ops.finalize_stats(0, pct, mom, nval)
    exp = float(w.astype(np.float64).std())
    assert abs(stats[0].item() - exp) <= 1e-12 * exp, (stats[0].item(), exp)


def test_encode_quantize_bit_exact(cuda):
    from fujishadergpu_amd import ops
    from fujishadergpu_amd.encoding import quantize_params
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.uniform(-1.5, 1.5, 200000).astype(np.float32),
                        np.array([np.inf, -np.inf, np.nan, 1.0, 0.0, -0.0], np.float32)])
    # exact .5 ties after a*v+b
    v = np.concatenate([v, (np.arange(-300, 300, dtype=np.float32) + 0.5) / np.float32(127.0)])
    t = torch.from_numpy(v).to(cuda)
    for lo, hi, dt in ((-1.176, 1.176, "uint8"), (-1.176, 1.176, "int16"), (0.0, 1.0, "uint8"),
                       (0.0, 90.0, "int16")):
        qp = quantize_params(lo, hi, dt)
        got = ops.quantize(t, qp, dt).cpu().numpy()
        exp = ref.quantize(v, qp["a_coef"], qp["b_coef"], qp["dn_min"], qp["dn_max"], dt)
        assert np.array_equal(got, exp), (lo, hi, dt)
    # tests/test_audit_p1_regressions.py:42-49
    qp = {"a_coef": 100.0, "b_coef": 0.0, "dn_min": 0, "dn_max": 255}
    got = ops.quantize(torch.tensor([np.inf, -np.inf, np.nan, 1.0], device=cuda), qp, "uint8")
    assert got.cpu().tolist() == [0, 0, 0, 100]


def test_nodata_to_nan_and_counts(cuda):
    from fujishadergpu_amd import ops
    data = torch.tensor([[1.0, -9999.0, 5.0], [float("nan"), 2.0, -9998.9]], device=cuda)
    counts = ops.nodata_to_nan(data, -9999.0, [(0, 0, 2, 3), (0, 1, 2, 2)])
    mask = torch.isnan(data).cpu().numpy()
    assert mask.tolist() == [[False, True, False], [True, False, False]]
    assert counts.cpu().tolist() == [2, 1]


def test_row_mapped_launches_cover_rasters_taller_than_the_grid_limit(cuda):
    # 140001 rows (more than the 65535 blocks gridDim.y allows): NoData
    # conversion with a sentinel, region copy, encode and the block mean
    # stride over the rows beyond their capped grid
    from fujishadergpu_amd import _native as N
    from fujishadergpu_amd import ops
    from fujishadergpu_amd.encoding import quantize_params
    H, W = 140001, 37
    host = torch.arange(H * W, dtype=torch.float32).reshape(H, W) % 1013.0
    host[::97, 5] = -9999.0
    g = host.to(cuda)
    counts = ops.nodata_to_nan(g, -9999.0, [(0, 0, H, W), (70000, 0, 70001, W)])
    want = host.clone()
    want[want == -9999.0] = float("nan")
    assert torch.equal(torch.isnan(g).cpu(), torch.isnan(want))
    assert counts.cpu().tolist() == [int(torch.isnan(want).sum()),
                                     int(torch.isnan(want[70000:]).sum())]
    reg = ops.region(3, 2, H - 5, 30)
    dst = torch.empty((H - 5, 30), dtype=torch.float32, device=cuda)
    ops.copy_region(g, reg, dst)
    assert torch.equal(torch.nan_to_num(dst.cpu(), nan=-1.0),
                       torch.nan_to_num(want[3:H - 2, 2:32], nan=-1.0))
    qp = quantize_params(0.0, 1012.0, "uint8")
    u8 = torch.empty((H - 5, 30), dtype=torch.uint8, device=cuda)
    ops.encode(g, reg, ops.output(u8, qp=qp))
    exp = ref.quantize(want[3:H - 2, 2:32].numpy(), qp["a_coef"], qp["b_coef"], qp["dn_min"],
                       qp["dn_max"], "uint8")
    assert np.array_equal(u8.cpu().numpy(), exp)
    bm = ops.block_mean(g, 2)                        # 70001 output rows
    cells = want[:140000, :36].reshape(70000, 2, 18, 2).double()
    assert bm.shape == (70001, 19)
    m = torch.isnan(cells).logical_not()
    tot = torch.where(m, cells, torch.zeros_like(cells)).sum(dim=(1, 3))
    cnt = m.sum(dim=(1, 3))
    exp_bm = torch.where(cnt > 0, tot.float() / cnt.float(), torch.tensor(float("nan")))
    got_bm = bm[:70000, :18].cpu()
    assert torch.equal(torch.isnan(got_bm), torch.isnan(exp_bm))
    assert torch.equal(torch.nan_to_num(got_bm), torch.nan_to_num(exp_bm))


@pytest.mark.parametrize("radii,weights,px,shape,reg,field", [
    # 16384^2 tile geometry in small: 3 box radii, ds-4 radius, overview field
    ([2, 8, 32, 128], [8 / 15, 4 / 15, 2 / 15, 1 / 15], 5.0, (700, 1100), (160, 160, 380, 780),
     True),
    # pre-pass geometry: three decimated radii, no field
    ([2, 8, 32, 128, 512, 2048], [32 / 63, 16 / 63, 8 / 63, 4 / 63, 2 / 63, 1 / 63], 5.0,
     (1200, 1300), (200, 230, 800, 840), False),
    # one and two box radii, odd radius (lanes straddle the output edge),
    # region at the window edge (reflected columns and rows), ragged strips
    ([3], [1.0], 10.0, (300, 517), (0, 0, 300, 517), False),
    ([5, 21], [0.7, 0.3], 10.0, (411, 389), (0, 13, 400, 376), True),
    ([2, 8, 30], [0.5, 0.3, 0.2], 10.0, (257, 1000), (1, 3, 255, 990), True),
])
def test_wave_kernel_equals_the_workgroup_kernel(cuda, radii, weights, px, shape, reg, field):
    """The single-wavefront main pass and the workgroup kernel give the same
    bits (float64 prefix differences equal the running sums exactly), and
    both match the oracle."""
    from fujishadergpu_amd import ops
    from fujishadergpu_amd.algorithms import OutSpec, TopoUSMFastAlgorithm
    from fujishadergpu_amd.encoding import quantize_params
    H, W = shape
    dem = _dem(H, W, 91 + H, relief=3776.0, hurst=0.75)
    g = dem.to(cuda)
    p = dict(radii=radii, weights=weights, pixel_size=px, global_stats=(19.5,),
             _window_nan_free=True)
    if field:
        ov = _dem(64, 96, 7, relief=3776.0, hurst=0.75)
        fld = ops.topousm_coarse_field(ov.to(cuda), [9], [0.03])
        p.update(_topousm_fast_coarse_field=fld, _topousm_fast_small_radii=radii,
                 _topousm_fast_small_weights=weights, _topousm_fast_w_large=0.03,
                 _topousm_fast_full_shape=(H * 3, W * 2), _topousm_fast_field_offset=(H, W // 2))
    r0, c0, rh, rw = reg
    qp = quantize_params(-1.176, 1.176, "uint8")
    outs = []
    try:
        for variant in (1, 0):
            ops.set_topousm_variant(variant)
            f32 = torch.full((rh, rw), -7.0, dtype=torch.float32, device=cuda)
            u8 = torch.zeros((rh, rw), dtype=torch.uint8, device=cuda)
            TopoUSMFastAlgorithm().run_region(g, ops.region(r0, c0, rh, rw),
                                              OutSpec(f32=f32, q=u8, qp=qp), p)
            torch.cuda.synchronize()
            outs.append((f32, u8))
    finally:
        ops.set_topousm_variant(0)
    (a32, a8), (b32, b8) = outs
    assert torch.equal(a32, b32), float((a32 - b32).abs().max())
    assert torch.equal(a8, b8)
    kw = dict(radii=radii, weights=weights, pixel_size=px, stats=(19.5,))
    if field:
        kw.update(coarse_field=fld.cpu().numpy(), small_radii=radii, small_weights=weights,
                  w_large=0.03, field_offset=(H, W // 2), full_shape=(H * 3, W * 2))
    exp = ref.topousm_direct(dem.numpy(), **kw)[r0:r0 + rh, c0:c0 + rw]
    mx, _ = _cmp("topousm/wave", b32, exp, 2e-5)
    print("radii", radii, "max |d| vs oracle", mx, flush=True)

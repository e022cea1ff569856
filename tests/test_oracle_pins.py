"""Pins for the CPU oracle (oracle/ref.py), CPU only.

Each oracle function is checked against something other than itself:
closed forms, brute-force FP64 loops written independently of SciPy, and the
values the reference's own tests hold (cited by file).  Every check is chosen
so that a plausible slip -- a sign, a transposed axis, a dropped term, a wrong
boundary mode or rounding rule -- fails it.
"""
import math

import numpy as np
import pytest

from oracle import ref

F32 = np.float32
G = 1.0 / 2.2
ZERO_DISP = 0.5 ** G          # display value of zero curvature / half openness


# --------------------------------------------------------------- gradient (A1)

def test_gradient_exact_on_planes_with_anisotropic_steps():
    y, x = np.mgrid[0:12, 0:15].astype(F32)
    z = (F32(3) * x - F32(2) * y).astype(F32)
    drow, dcol = ref.gradient_f32(z, 2.0, 4.0)        # step_y = 2, step_x = 4
    assert np.all(dcol == F32(0.75)) and np.all(drow == F32(-1.0))


def test_gradient_edge_order_two_is_exact_on_quadratics():
    y, x = np.mgrid[0:9, 0:11].astype(F32)
    drow, dcol = ref.gradient_f32(x * x, 1.0, 1.0)
    assert np.array_equal(dcol, 2 * x) and np.all(drow == 0)   # edges included
    drow, dcol = ref.gradient_f32(y * y, 1.0, 1.0)
    assert np.array_equal(drow, 2 * y) and np.all(dcol == 0)


# ------------------------------------------------------ hillshade, slope (A2-3)

def _plane(a, b, h=16, w=16):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    return (a * x + b * y).astype(F32)


def test_flat_hillshade_is_sin_altitude():
    out = ref.hillshade(np.zeros((8, 8), F32), altitude=45.0)
    assert np.allclose(out, math.sin(math.radians(45.0)), atol=1e-7, rtol=0)
    out = ref.hillshade(np.zeros((8, 8), F32), altitude=30.0)
    assert np.allclose(out, 0.5, atol=1e-7, rtol=0)


@pytest.mark.parametrize("sy", [None, -1.0])
def test_hillshade_on_a_plane_matches_the_lambert_closed_form(sy):
    a, b = 0.5, 0.25
    out = ref.hillshade(_plane(a, b), azimuth=315.0, altitude=45.0, pixel_size=1.0,
                        pixel_scale_y=sy)
    az, alt = math.radians(315.0), math.radians(45.0)
    lx, ly, lz = math.sin(az) * math.cos(alt), math.cos(az) * math.cos(alt), math.sin(alt)
    n = b * (1.0 if sy is None else -1.0)        # north-up rasters flip the row axis
    exp = (-a * lx - n * ly + lz) / math.sqrt(1.0 + a * a + b * b)
    assert np.allclose(out, exp, atol=1e-6, rtol=0)


def test_slope_on_a_plane_in_every_unit():
    z = _plane(0.5, 0.25)
    g = math.hypot(0.5 / 2.0, 0.25 / 2.0)        # pixel size 2
    assert np.allclose(ref.slope(z, unit="degree", pixel_size=2.0), math.degrees(math.atan(g)),
                       atol=2e-5, rtol=0)
    assert np.allclose(ref.slope(z, unit="radians", pixel_size=2.0), math.atan(g),
                       atol=1e-6, rtol=0)
    assert np.allclose(ref.slope(z, unit="percent", pixel_size=2.0), 100.0 * g, atol=1e-4, rtol=0)


def test_nan_fill_does_not_spread_nan():
    z = _plane(0.5, 0.25)
    z[5, 7] = np.nan
    for f in (ref.hillshade, ref.slope, ref.curvature):
        out = f(z)
        assert np.array_equal(np.isnan(out), np.isnan(z))


# ----------------------------------------------------------------- curvature (A4)

def _grid(h=64, w=64):
    return np.mgrid[0:h, 0:w].astype(F32)


def test_planform_zero_profile_nonzero_on_a_parabolic_cylinder():
    # reference tests/test_curvature_analytic.py:18-41
    _, x = _grid()
    z = (F32(0.01) * (x - F32(32.0)) ** 2).astype(F32)
    plan = ref.curvature(z, curvature_type="planform")[20:44, 20:44]
    prof = ref.curvature(z, curvature_type="profile")[20:44, 20:44]
    assert np.abs(plan - ZERO_DISP).max() < 0.02
    assert np.abs(prof - ZERO_DISP).max() > 0.05


# ---------------------------------------------- per-row x spacing (geographic rows)

def _rows_dem(h=40, w=50, seed=3):
    rng = np.random.default_rng(seed)
    return np.cumsum(np.cumsum(rng.normal(size=(h, w)), 0), 1).astype(F32) * F32(0.7)


def test_equal_row_spacings_reproduce_the_scalar_stencils_bit_for_bit():
    z = _rows_dem()
    rows = np.full(z.shape[0], 23.5)
    for fn, kw in ((ref.hillshade, {}), (ref.slope, {}), (ref.curvature, {}),
                   (ref.curvature, {"curvature_type": "gaussian"})):
        a = fn(z, pixel_scale_x=rows, pixel_scale_y=-30.0, **kw)
        b = fn(z, pixel_scale_x=23.5, pixel_scale_y=-30.0, **kw)
        assert np.array_equal(a, b, equal_nan=True), fn.__name__


def test_hillshade_and_slope_rows_use_their_own_spacing():
    # each pixel's hillshade/slope depends only on its own row's x spacing:
    # row i equals the scalar-spacing result with step rows[i]
    z = _rows_dem()
    rows = 20.0 + 5.0 * np.cos(np.linspace(0.5, 1.0, z.shape[0]))
    hs = ref.hillshade(z, pixel_scale_x=rows, pixel_scale_y=-31.0)
    sl = ref.slope(z, pixel_scale_x=rows, pixel_scale_y=-31.0)
    for i in (0, 1, 17, z.shape[0] - 1):
        assert np.array_equal(hs[i], ref.hillshade(z, pixel_scale_x=rows[i],
                                                   pixel_scale_y=-31.0)[i])
        assert np.array_equal(sl[i], ref.slope(z, pixel_scale_x=rows[i], pixel_scale_y=-31.0)[i])


def test_slope_of_a_column_ramp_is_atan_of_the_row_gradient():
    # z = a j: every row's metric gradient is a / sx_i exactly
    h, w, a = 12, 20, 3.0
    z = np.tile((a * np.arange(w)).astype(F32), (h, 1))
    rows = np.linspace(15.0, 25.0, h)
    sl = ref.slope(z, pixel_scale_x=rows, pixel_scale_y=-30.0)
    want = np.degrees(np.arctan(a / rows))[:, None] * np.ones((1, w))
    assert np.allclose(sl, want, rtol=1e-5, atol=0)   # float32 gradient and atan


def test_mean_curvature_of_a_column_profile_is_row_local():
    # z = g(col): dy = 0 on the interior rows (the edge_order=2 edge formula
    # leaves a float32 residue on the first and last), so the mean curvature
    # of interior row i is the scalar result with x step rows[i]: the mixed
    # term is multiplied by dy = 0
    h, w = 10, 30
    g = (0.02 * (np.arange(w) - 13.0) ** 3).astype(F32)
    z = np.tile(g, (h, 1))
    rows = np.linspace(18.0, 26.0, h)
    cv = ref.curvature(z, pixel_scale_x=rows, pixel_scale_y=-30.0)
    for i in range(1, h - 1):
        assert np.array_equal(cv[i], ref.curvature(z, pixel_scale_x=rows[i],
                                                   pixel_scale_y=-30.0)[i])


def test_dome_has_plan_and_profile_curvature():
    # reference tests/test_curvature_analytic.py:44-54
    y, x = _grid()
    dome = (F32(-0.01) * ((x - F32(32)) ** 2 + (y - F32(32)) ** 2)).astype(F32)
    for ct in ("planform", "profile"):
        assert np.abs(ref.curvature(dome, curvature_type=ct)[20:44, 20:44] - ZERO_DISP).max() > 0.05


def _disp(k):
    return ((math.tanh(100.0 * k) + 1.0) / 2.0) ** G


def test_mean_curvature_at_a_bowl_apex():
    # z = c (x^2 + y^2): p = q = s = 0, r = t = 2c -> H = -(r + t) / 2 = -2c
    c = 0.001
    y, x = _grid(33, 33)
    z = (F32(c) * ((x - 16) ** 2 + (y - 16) ** 2)).astype(F32)
    out = ref.curvature(z, curvature_type="mean")
    assert out[16, 16] == pytest.approx(_disp(-2 * c), abs=1e-6)


def test_signed_row_step_flips_the_row_second_derivatives():
    # compute_curvature_block takes its first derivatives with |step| but the
    # second ones with the signed steps (_impl_curvature.py:23-33).  On a
    # north-up raster (pixel_scale_y < 0) the row second derivatives dyy and
    # dxy change sign, dxx and dyx do not.
    c = 0.001
    y, x = _grid(33, 33)
    bowl = (F32(c) * ((x - 16) ** 2 + (y - 16) ** 2)).astype(F32)
    # bowl apex: r = 2c, t = -2c -> H = -(r + t) / 2 = 0 (vs -2c with a positive step)
    neg = ref.curvature(bowl, curvature_type="mean", pixel_scale_x=1.0, pixel_scale_y=-1.0)
    pos = ref.curvature(bowl, curvature_type="mean", pixel_scale_x=1.0, pixel_scale_y=1.0)
    assert neg[16, 16] == pytest.approx(ZERO_DISP, abs=1e-6)
    assert pos[16, 16] == pytest.approx(_disp(-2 * c), abs=1e-6)
    # saddle z = k (x - 16)(y - 16): p = k (y - 16), q = k (x - 16), r = t = 0,
    # dyx = k and dxy = -k with the signed step, so s = 0 and H = 0 everywhere;
    # with a positive step s = k and H = p q s / (1 + p^2 + q^2)^1.5
    k = 0.0625          # a power of two: the surface and its differences are exact
    sad = (F32(k) * (x - 16) * (y - 16)).astype(F32)
    neg = ref.curvature(sad, curvature_type="mean", pixel_scale_x=1.0, pixel_scale_y=-1.0)
    pos = ref.curvature(sad, curvature_type="mean", pixel_scale_x=1.0, pixel_scale_y=1.0)
    assert np.abs(neg - ZERO_DISP).max() < 1e-6
    for (i, j) in ((10, 20), (22, 5), (25, 27)):
        p, q = k * (i - 16), k * (j - 16)
        hk = p * q * k / (1 + p * p + q * q) ** 1.5
        assert pos[i, j] == pytest.approx(_disp(hk), abs=2e-6)
        assert abs(pos[i, j] - ZERO_DISP) > 1e-3


def test_gaussian_curvature_at_an_elliptic_apex():
    # z = c x^2 + d y^2 at the apex: K = r t - s^2 = 4 c d
    c, d = 0.05, 0.03
    y, x = _grid(33, 33)
    z = (F32(c) * (x - 16) ** 2 + F32(d) * (y - 16) ** 2).astype(F32)
    out = ref.curvature(z, curvature_type="gaussian")
    assert out[16, 16] == pytest.approx(_disp(4 * c * d), abs=1e-6)


# -------------------------------------------------- NaN-aware filters (A1, A5)

def _gauss_w(sigma=1.0, truncate=4.0):
    r = int(truncate * sigma + 0.5)
    k = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * k * k / (sigma * sigma))
    return k.astype(int), w / w.sum()


def _idx(i, n, mode):
    if mode == "nearest":
        return min(max(i, 0), n - 1)
    # scipy 'reflect': d c b a | a b c d | d c b a
    period = 2 * n
    i %= period
    return i if i < n else period - 1 - i


def _brute_sep(x, taps, weights, mode):
    """Separable 2D correlation by explicit loops in float64."""
    h, w = x.shape
    out = np.zeros((h, w))
    for i in range(h):
        for j in range(w):
            s = 0.0
            for ki, wi in zip(taps, weights):
                for kj, wj in zip(taps, weights):
                    s += wi * wj * x[_idx(i + ki, h, mode), _idx(j + kj, w, mode)]
            out[i, j] = s
    return out


def _nan_block(seed, h=9, w=11):
    rng = np.random.default_rng(seed)
    b = (rng.random((h, w)) * 100).astype(F32)
    b[2, 3] = b[6, 0] = b[8, 10] = np.nan
    return b


def test_nan_gaussian_fill_is_the_normalised_convolution():
    b = _nan_block(1)
    valid = ~np.isnan(b)
    taps, wts = _gauss_w(1.0)
    sv = _brute_sep(np.where(valid, b, 0).astype(np.float64), taps, wts, "nearest")
    sw = _brute_sep(valid.astype(np.float64), taps, wts, "nearest")
    got, mask = ref.handle_nan_with_gaussian(b, 1.0, mode="nearest")
    assert np.array_equal(mask, ~valid)
    assert np.allclose(got, sv / sw, atol=2e-4, rtol=0)


@pytest.mark.parametrize("mode", ["reflect", "nearest"])
def test_nan_box_mean_matches_brute_force(mode):
    b = _nan_block(2)
    valid = ~np.isnan(b)
    r = 2
    taps, wts = np.arange(-r, r + 1), np.full(2 * r + 1, 1.0 / (2 * r + 1))
    sv = _brute_sep(np.where(valid, b, 0).astype(np.float64), taps, wts, mode)
    sw = _brute_sep(valid.astype(np.float64), taps, wts, mode)
    got, _ = ref.handle_nan_with_uniform(b, 2 * r + 1, mode=mode)
    assert np.allclose(got, sv / sw, atol=2e-4, rtol=0)


# ------------------------------------------------------------ TopoUSM (A5-A7)

def test_topousm_on_a_parabola_is_minus_weighted_r_r_plus_1_over_3():
    # box mean of x^2 over [x-r, x+r] is x^2 + r(r+1)/3
    _, x = np.mgrid[0:24, 0:41].astype(F32)
    z = ((x - F32(20)) ** 2).astype(F32)
    out = ref.topousm_fast_block(z, radii=[2, 5], weights=[0.6, 0.4], pixel_size=100.0)
    exp = -(0.6 * 2 * 3 / 3 + 0.4 * 5 * 6 / 3)
    assert np.allclose(out[:, 6:35], exp, atol=1e-3, rtol=0)


def _large_only(block, field, w_large, off, full):
    """The large-radius part of _direct_topousm_fast on its own
    (tile/dask_bridge.py:237-255): no small radii, scale 1."""
    return ref.topousm_direct(block, radii=[], weights=[], coarse_field=field, small_radii=None,
                              w_large=w_large, field_offset=off, full_shape=full, stats=(1.0,))


def test_topousm_large_part_of_a_constant_is_zero():
    # constant c: every box mean of the overview is c, the field is (sum w) c
    # and W_large * c - bilinear(field) vanishes up to the float32 rounding of
    # the two weighted sums
    c = 1234.75
    r = np.full((256, 320), c, F32)
    coarse, s = ref.overview(r, sample_max=32)          # s = 10: 26 x 32 cells
    lw = [0.125, 0.0625]
    field = ref.topousm_large_coarse_field(coarse, large_radii=[40, 120], large_weights=lw,
                                           decimation=s)
    assert np.allclose(field, sum(lw) * c, rtol=2e-7)
    large = _large_only(r[30:200, 50:300], field, sum(lw), (30, 50), r.shape)
    assert np.abs(large).max() <= 4 * np.spacing(F32(c))


def test_topousm_large_part_of_a_ramp_vanishes_inside():
    # z = a x + b y on an integer-decimated raster (s = 8): the 8 x 8 cell
    # means are z at the cell centres, a box keeps a linear function away from
    # the coarse edges, and the pixel-centre bilinear map u = (x + 0.5) s_c / W
    # - 0.5 returns z itself at every pixel -- so W_large z - bilinear(field)
    # is 0 in the interior.  A field offset at the core origin instead of the
    # window origin, or a corner-aligned map, leaves a ramp residue.
    a, b = 0.75, -0.5
    yy, xx = np.mgrid[0:256, 0:256].astype(np.float64)
    z = (a * xx + b * yy).astype(F32)
    coarse, s = ref.overview(z, sample_max=32)
    assert s == 8.0 and coarse.shape == (32, 32)
    field = ref.topousm_large_coarse_field(coarse, large_radii=[40], large_weights=[0.5],
                                           decimation=s)          # r_c = 5: box of 11
    off = (40, 36)
    large = _large_only(z[40:220, 36:230], field, 0.5, off, z.shape)
    # interior: coarse samples at least 6 cells from the coarse edge -> pixels 52..203
    inner = large[52 - 40:204 - 40, 52 - 36:204 - 36]
    assert np.abs(inner).max() < 2e-4
    wrong = _large_only(z[40:220, 36:230], field, 0.5, (60, 56), z.shape)
    # the shifted sample reads 20 rows and columns further on: 0.5 (20 a + 20 b)
    assert np.abs(wrong[52 - 40:150 - 40, 52 - 36:150 - 36] + 2.5).max() < 2e-4


@pytest.mark.parametrize("r,s,rc", [(4, 8.0, 1), (12, 8.0, 2), (20, 8.0, 2), (28, 8.0, 4),
                                    (44, 8.0, 6), (512, 8.0, 64), (2048, 8.0, 256),
                                    (300, 2.9296875, 102), (7, 5.0, 1)])
def test_topousm_coarse_radius_rounding(r, s, rc):
    # r_c = max(1, round(r / s)) with Python's round (ties to even), and the
    # sigma = 1 Gaussian when r_c <= 1 (_impl_topousm_fast.py:133-155)
    assert max(1, int(round(r / s))) == rc
    rng = np.random.default_rng(r)
    g = (rng.random((10, 13)) * 50).astype(F32)
    f = ref.topousm_large_coarse_field(g, large_radii=[r], large_weights=[1.0], decimation=s)
    if rc <= 1:
        taps, wts, mode = *_gauss_w(1.0), "nearest"
    else:
        taps, wts, mode = np.arange(-rc, rc + 1), np.full(2 * rc + 1, 1.0 / (2 * rc + 1)), \
            "reflect"
    x = g.astype(np.float64)
    for axis in (0, 1):            # separable: one explicit 1-D pass per axis
        n = x.shape[axis]
        y = np.zeros_like(x)
        for i in range(n):
            for k, w in zip(taps, wts):
                src = _idx(i + int(k), n, mode)
                if axis == 0:
                    y[i] += w * x[src]
                else:
                    y[:, i] += w * x[:, src]
        x = y
    exp = x
    assert np.allclose(f, exp, atol=3e-5, rtol=0)


def test_topousm_is_zero_on_planes_and_constants():
    z = _plane(3.0, -2.0, 32, 40)
    out = ref.topousm_fast_block(z, radii=[2, 4], weights=[0.5, 0.5], pixel_size=100.0)
    assert np.abs(out[5:-5, 5:-5]).max() < 1e-3
    c = np.full((20, 30), 123.25, F32)
    assert np.all(ref.topousm_fast_block(c, radii=[2, 9, 40], weights=[0.2, 0.3, 0.5],
                                         pixel_size=5.0) == 0)


def test_downsample_factor_table():
    cases = {(2, 5.0): 1, (32, 5.0): 1, (128, 5.0): 4, (200, 5.0): 8, (512, 5.0): 16,
             (2048, 5.0): 16, (48, 0.5): 2, (48, 1.0): 2, (24, 1.0): 1}
    from fujishadergpu_amd.params import radius_to_downsample_factor as product
    for (r, px), f in cases.items():
        assert ref.radius_to_downsample_factor(r, pixel_size=px, algorithm_name="topousm_fast") == f
        assert product(r, pixel_size=px, algorithm_name="topousm_fast") == f


def test_decimation_is_the_nan_aware_cell_mean():
    rng = np.random.default_rng(3)
    b = (rng.random((10, 13)) * 50).astype(F32)
    b[0, 0] = b[5, 7] = np.nan
    got = ref.downsample_nan_aware(b, 4)
    assert got.shape == (3, 4)
    for i in range(3):
        for j in range(4):
            cell = b[4 * i:4 * i + 4, 4 * j:4 * j + 4].astype(np.float64)
            assert got[i, j] == pytest.approx(np.nanmean(cell), rel=1e-6)


def test_enclosed_empty_cell_is_filled_from_its_neighbours():
    b = np.arange(12 * 12, dtype=F32).reshape(12, 12)
    b[4:8, 4:8] = np.nan                            # one all-NaN cell at factor 4
    got = ref.downsample_nan_aware(b, 4)
    coarse = np.array([[np.nanmean(b[4 * i:4 * i + 4, 4 * j:4 * j + 4].astype(np.float64))
                        if not np.isnan(b[4 * i:4 * i + 4, 4 * j:4 * j + 4]).all() else np.nan
                        for j in range(3)] for i in range(3)])
    valid = ~np.isnan(coarse)
    taps, wts = _gauss_w(1.0)                       # sigma = max(1, 3 / 64)
    sv = _brute_sep(np.where(valid, coarse, 0), taps, wts, "nearest")
    sw = _brute_sep(valid.astype(np.float64), taps, wts, "nearest")
    assert sw[1, 1] > 0.5
    assert got[1, 1] == pytest.approx(sv[1, 1] / sw[1, 1], rel=1e-5)


def test_zoom_up_is_corner_aligned_bilinear():
    rng = np.random.default_rng(4)
    c = (rng.random((4, 5)) * 10).astype(F32)
    th, tw = 10, 13
    got = ref.upsample_to_shape(c, (th, tw))
    for i in range(th):
        for j in range(tw):
            y, x = i * 3 / (th - 1), j * 4 / (tw - 1)
            y0, x0 = min(int(y), 2), min(int(x), 3)
            ty, tx = y - y0, x - x0
            v = ((1 - ty) * ((1 - tx) * c[y0, x0] + tx * c[y0, x0 + 1])
                 + ty * ((1 - tx) * c[y0 + 1, x0] + tx * c[y0 + 1, x0 + 1]))
            assert got[i, j] == pytest.approx(v, abs=2e-6)


def test_overview_field_sample_is_pixel_centre_bilinear_with_clamp():
    rng = np.random.default_rng(5)
    c = (rng.random((6, 8)) * 10).astype(F32)
    H, W = 24, 32                                   # coarse = 1/4 of the raster
    got = ref.bilinear_sample_coarse(c, 3, 11, 0, 32, H, W)
    for i, r in enumerate(range(3, 11)):
        for j in range(32):
            y = min(max((r + 0.5) * 6 / H - 0.5, 0.0), 5.0)
            x = min(max((j + 0.5) * 8 / W - 0.5, 0.0), 7.0)
            y0, x0 = min(int(y), 4), min(int(x), 6)
            ty, tx = y - y0, x - x0
            v = ((1 - ty) * ((1 - tx) * c[y0, x0] + tx * c[y0, x0 + 1])
                 + ty * ((1 - tx) * c[y0 + 1, x0] + tx * c[y0 + 1, x0 + 1]))
            assert got[i, j] == pytest.approx(v, abs=2e-6)


def test_overview_integer_decimation_is_a_block_mean():
    rng = np.random.default_rng(6)
    r = (rng.random((12, 12)) * 7).astype(F32)
    r[1, 1] = np.nan
    got, scale = ref.overview_block_mean(r, sample_max=4)
    assert scale == 3.0 and got.shape == (4, 4)
    f = 3
    for i in range(got.shape[0]):
        for j in range(got.shape[1]):
            assert got[i, j] == pytest.approx(np.nanmean(
                r[f * i:f * i + f, f * j:f * j + f].astype(np.float64)), rel=1e-6)


def _supersampled_block_mean(a, qy, qx, py, px):
    """Nearest-upsample by (qy, qx), then NaN-aware py x px block means: the
    area average at ratios py/qy and px/qx, built without overlap weights."""
    up = np.repeat(np.repeat(a.astype(np.float64), qy, axis=0), qx, axis=1)
    oh, ow = up.shape[0] // py, up.shape[1] // px
    cells = up[:oh * py, :ow * px].reshape(oh, py, ow, px)
    with np.errstate(invalid="ignore"):
        return np.nanmean(cells, axis=(1, 3))


@pytest.mark.parametrize("h,w,oh,ow,qy,qx", [
    (10, 14, 4, 4, 2, 2),      # ratios 5/2 and 7/2
    (15, 8, 6, 3, 2, 3),       # 5/2 and 8/3 (different ratio per axis)
    (12, 20, 5, 8, 5, 2),      # 12/5 and 5/2
])
@pytest.mark.parametrize("with_nan", [False, True])
def test_overview_average_equals_a_supersampled_block_mean(h, w, oh, ow, qy, qx, with_nan):
    rng = np.random.default_rng(h * 100 + w)
    a = (rng.random((h, w)) * 900.0 + 100.0).astype(F32)
    if with_nan:
        a[rng.random((h, w)) < 0.2] = np.nan
        a[: h // 3, : w // 3] = np.nan          # a whole cell may be empty
    py, px = h * qy // oh, w * qx // ow
    assert py * oh == h * qy and px * ow == w * qx
    exp = _supersampled_block_mean(a, qy, qx, py, px)
    got = ref.overview_average(a, oh, ow)
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    np.testing.assert_allclose(got[~np.isnan(got)], exp[~np.isnan(exp)], rtol=2e-7)


def test_overview_average_constant_mass_and_integer_ratio():
    rng = np.random.default_rng(5)
    c = np.full((37, 53), 1234.5, dtype=F32)
    c[rng.random(c.shape) < 0.3] = np.nan
    got = ref.overview_average(c, 9, 17)
    assert np.all(got == F32(1234.5))
    # mass: every source pixel is covered exactly once in total
    a = (rng.random((30, 44)) * 100).astype(F32)
    got = ref.overview_average(a, 7, 13)
    area = (30 / 7) * (44 / 13)
    assert float(got.astype(np.float64).sum() * area) == pytest.approx(
        float(a.astype(np.float64).sum()), rel=1e-6)
    # an integer ratio is the float64 block mean of the finite pixels, rounded
    # once, bit for bit (3 x 3 cells, so most counts are not powers of two)
    b = (rng.random((24, 36)) * 3000).astype(F32)
    b[rng.random(b.shape) < 0.25] = np.nan
    cells = b.astype(np.float64).reshape(8, 3, 12, 3)
    with np.errstate(invalid="ignore"):
        exp = np.nanmean(cells, axis=(1, 3)).astype(F32)
    assert np.array_equal(ref.overview_average(b, 8, 12), exp, equal_nan=True)
    got, scale = ref.overview_block_mean(b, sample_max=12)
    assert scale == 3.0 and np.array_equal(got, exp, equal_nan=True)


def test_overview_geometry_at_a_fractional_decimation():
    # core/tile_processor.py:575-580 for a 10000 x 7000 (W x H) raster
    a = np.zeros((70, 100), dtype=F32)
    got, scale = ref.overview(a, sample_max=20)
    assert scale == 5.0 and got.shape == (14, 20)
    b = np.zeros((7000 // 50, 10000 // 50), dtype=F32)
    got, scale = ref.overview(b, sample_max=48)
    assert scale == pytest.approx(200 / 48) and got.shape == (34, 48)   # round(140/4.1667)=34


def test_p99_stat_reference_vectors():
    # reference tests/test_topousm_fast_normalization.py:6-33
    rng = np.random.default_rng(0)
    base = rng.normal(0.0, 0.01, size=2048).astype(F32)
    base[:3] = [10.0, -8.0, 12.0]
    s = ref.topousm_stat(base)[0]
    assert 0.0 < s < 0.2
    data = np.concatenate([np.full(800, 0.01, F32), np.full(200, 0.5, F32)])
    assert ref.topousm_stat(data)[0] == pytest.approx(0.5, rel=1e-2)
    # linear in-between order statistic (method='linear')
    assert ref.topousm_stat(np.arange(101, dtype=F32))[0] == pytest.approx(99.0)
    assert ref.topousm_stat(np.arange(11, dtype=F32))[0] == pytest.approx(9.9)


def test_normalisation_is_a_plain_divide():
    # reference tests/test_topousm_fast_normalization.py:36-44
    out = ref.normalize_div(np.array([-4.0, -1.0, 0.0, 1.0, 4.0], F32), 2.0)
    assert out.tolist() == [-2.0, -0.5, 0.0, 0.5, 2.0]


# --------------------------------------------------------- openness, AO (A8-A10)

def _radial(n=101):
    yy, xx = np.mgrid[0:n, 0:n]
    return np.sqrt((xx - n // 2) ** 2 + (yy - n // 2) ** 2)


def test_openness_reference_orderings_and_flat():
    # reference tests/test_openness_yokoyama.py:21-47
    r = _radial()
    kw = dict(num_directions=16, max_distance=40)
    peak, pit = (50 - r).astype(F32), (r - 50).astype(F32)
    assert (ref.openness(peak, openness_type="positive", **kw)[50, 50]
            > ref.openness(pit, openness_type="positive", **kw)[50, 50])
    assert (ref.openness(pit, openness_type="negative", **kw)[50, 50]
            > ref.openness(peak, openness_type="negative", **kw)[50, 50])
    flat = np.zeros((101, 101), F32)
    for t in ("positive", "negative"):
        assert ref.openness(flat, openness_type=t, **kw)[50, 50] == pytest.approx(1.0, abs=1e-3)


@pytest.mark.parametrize("k", [1.0, 2.0, 0.5])
def test_openness_at_a_cone_apex(k):
    # every sample of a cone of slope k sits at elevation angle atan(k)
    r = _radial(81)
    pit = (F32(k) * r).astype(F32)
    kw = dict(num_directions=8, max_distance=30)
    ang = (math.pi / 2 - math.atan(k)) / (math.pi / 2)
    assert ref.openness(pit, openness_type="positive", **kw)[40, 40] == pytest.approx(
        ang ** G, abs=2e-6)
    peak = (-F32(k) * r).astype(F32)
    assert ref.openness(peak, openness_type="negative", **kw)[40, 40] == pytest.approx(
        ang ** G, abs=2e-6)


@pytest.mark.parametrize("k", [1.0, 0.25])
def test_openness_raw_degrees_at_a_cone_apex_and_against_the_display_value(k):
    r = _radial(81)
    pit = (F32(k) * r).astype(F32)
    kw = dict(num_directions=8, max_distance=30, openness_type="positive")
    raw = ref.openness(pit, raw_degrees=True, **kw)
    assert raw[40, 40] == pytest.approx(90.0 - math.degrees(math.atan(k)), abs=2e-4)
    # the display value is clip(raw / 90, 0, 1) ** (1 / 2.2) at every pixel
    disp = ref.openness(pit, **kw)
    want = np.clip(raw.astype(np.float64) / 90.0, 0, 1) ** G
    assert np.abs(disp - want).max() < 2e-6


def test_openness_uses_physical_distance():
    # pixel size 2 halves every elevation angle's tangent on the same cone
    r = _radial(81)
    pit = (F32(2.0) * r).astype(F32)
    out = ref.openness(pit, openness_type="positive", num_directions=8, max_distance=30,
                       pixel_size=2.0)
    ang = (math.pi / 2 - math.atan(1.0)) / (math.pi / 2)
    assert out[40, 40] == pytest.approx(ang ** G, abs=2e-6)


def test_ambient_occlusion_flat_and_plane():
    flat = np.zeros((40, 40), F32)
    assert np.allclose(ref.ambient_occlusion(flat, num_samples=8, radius=6.0), 1.0, atol=1e-7)
    # a plane z = k x: every pixel sees the same ring, so the interior is the
    # FP64 ring average below (smoothing a constant field leaves it constant)
    k, n_s, rad = 0.4, 8, 6.0
    y, x = np.mgrid[0:48, 0:48]
    z = (k * x).astype(F32)
    out = ref.ambient_occlusion(z, num_samples=n_s, radius=rad)
    tot = cnt = 0.0
    for rf in (0.25, 0.5, 0.75, 1.0):
        for i in range(n_s):
            a = 2 * math.pi * i / n_s
            dx, dy = round(rad * rf * math.cos(a)), round(rad * rf * math.sin(a))
            if dx == 0 and dy == 0:
                continue
            occ = min(max(0.0, math.atan(k * dx / math.hypot(dx, dy))) / (math.pi / 4), 1.0)
            tot += occ * (1.0 - 0.3 * rf)
            cnt += 1
    exp = max(0.0, min(1.0, 1.0 - tot / cnt)) ** G
    assert np.allclose(out[12:36, 12:36], exp, atol=2e-6)


def test_display_stretch_percentiles():
    lo, sc = ref.stretch_stat(np.arange(101, dtype=F32))
    assert (lo, sc) == (pytest.approx(1.0), pytest.approx(98.0))
    out = ref.display_stretch(np.array([0.0, 1.0, 50.0, 99.0, 200.0], F32), (lo, sc))
    assert np.allclose(out, [0.0, 0.0, 49 / 98, 1.0, 199 / 98])
    assert ref.stretch_stat(np.full(10, 3.0, F32)) == (3.0, 0.0)


# ---------------------------------------------------- encoding, NoData (A12-13)

def test_quantize_reference_vector_and_rounding_rule():
    # reference tests/test_audit_p1_regressions.py:44-50
    dn = ref.quantize(np.array([np.inf, -np.inf, np.nan, 1.0], F32), 100.0, 0.0, 0, 255, "uint8")
    assert dn.tolist() == [0, 0, 0, 100]
    # rint is round-half-to-even, then clip to the code range
    dn = ref.quantize(np.array([0.5, 1.5, 2.5, -0.5, 300.0], F32), 1.0, 0.0, 0, 255, "uint8")
    assert dn.tolist() == [0, 2, 2, 0, 255]
    dn = ref.quantize(np.array([-1.0, 0.0, 1.0], F32), 32767.0, 0.0, -32767, 32767, "int16")
    assert dn.tolist() == [-32767, 0, 32767]


def test_product_quantizer_matches_the_oracle():
    from fujishadergpu_amd.encoding import quantize_params, resolve_output_range
    qp = quantize_params(-1.176, 1.176, "uint8")
    x = np.linspace(-1.5, 1.5, 4001, dtype=F32)
    exp = ref.quantize(x, qp["a_coef"], qp["b_coef"], qp["dn_min"], qp["dn_max"], "uint8")
    assert exp.min() == qp["dn_min"] and exp.max() == qp["dn_max"]
    assert np.all(np.diff(exp.astype(int)) >= 0)
    with pytest.raises(ValueError, match="high > low"):
        resolve_output_range("hillshade", override=(1.0, 0.0))   # test_audit_p5:103-107


def test_nodata_mask_reference_vectors():
    # reference tests/test_audit_p1_regressions.py:20-38
    d = np.array([[1.0, -9999.0], [np.nan, 2.0]], F32)
    assert ref.nodata_mask(d, -9999.0).tolist() == [[False, True], [True, False]]
    d = np.array([-9999.0, -9998.9, 0.0], F32)
    assert ref.nodata_mask(d, -9999.0).tolist() == [True, False, False]


def test_hillshade_output_keeps_nan_through_the_clip():
    # reference tests/test_output_nodata_policy.py:20-41
    out = ref.format_output(np.array([[np.nan, 0.5], [1.2, -0.1]], F32), "hillshade")
    assert np.isnan(out[0, 0]) and out[1, 0] == 1.0 and out[1, 1] == 0.0 and out[0, 1] == 0.5


# ------------------------------------------------- host parameter logic (A15-16)

@pytest.mark.parametrize("short_side,expected", [
    (30000, [2, 8, 32, 128, 512, 2048]), (20480, [2, 8, 32, 128, 512, 2048]),
    (10000, [2, 8, 32, 128, 512]), (5120, [2, 8, 32, 128, 512]), (5119, [2, 8, 32, 128]),
    (300, [2, 8]), (15, [2]),
])
def test_auto_radii_reference_table(short_side, expected):
    # reference tests/test_spatial_auto_radii.py:30-43
    from fujishadergpu_amd.params import auto_spatial_radii
    assert auto_spatial_radii(short_side) == expected


def test_auto_weights_and_dedup_reference_vectors():
    from fujishadergpu_amd.params import (auto_spatial_profile, auto_spatial_weights,
                                          deduplicate_radii_weights)
    for n in (1, 2, 3, 4, 6):
        w = auto_spatial_weights(n)
        assert sum(w) == pytest.approx(1.0)
        assert all(w[i] == pytest.approx(2 * w[i + 1]) for i in range(n - 1))
    r, w = auto_spatial_profile(10000, radii=[4, 16, 64, 256])
    assert r == [4, 16, 64, 256] and w == pytest.approx([8 / 15, 4 / 15, 2 / 15, 1 / 15])
    r, w = deduplicate_radii_weights([4, 4, 8], [0.2, 0.3, 0.5])   # test_audit_p5:20-25
    assert r == [4, 8] and w == pytest.approx([0.5, 0.5])


def test_meters_per_degree_at_35_north():
    # reference tests/test_audit_p1_regressions.py:99-107
    from fujishadergpu_amd.raster_info import meters_per_degree
    m_lon, m_lat = meters_per_degree(35.0)
    assert m_lat == pytest.approx(110940.5, abs=15) and m_lon == pytest.approx(91288.0, abs=15)


# ------------------------------------------------ surrogate network (A18)

def test_cnn_oracle_backward_matches_finite_differences():
    from oracle import cnn
    rng = np.random.default_rng(0)
    params = cnn.init_params(rng, [3, 5, 4, 2])
    x = rng.normal(size=(2, 6, 7, 3))
    base = rng.normal(size=(2, 6, 7))
    label = rng.normal(size=(2, 6, 7))
    label[0, 0, 0] = np.nan                      # unlabelled cell
    label_s = rng.normal(size=(2, 6, 7))
    label_s[1, 2, 3] = np.nan
    weight_s = rng.uniform(0.1, 2.0, size=(2, 6, 7))
    out, cache = cnn.forward(params, x)
    grads = cnn.backward(params, cache, out, base, label, scale=0.3, label_s=label_s,
                         weight_s=weight_s)

    def f(ps):
        return cnn.loss(cnn.forward(ps, x)[0], base, label, scale=0.3, label_s=label_s,
                        weight_s=weight_s)
    eps = 1e-6
    for li in range(len(params)):
        for which in (0, 1):
            arr = params[li][which]
            flat = arr.reshape(-1)
            for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
                old = flat[idx]
                flat[idx] = old + eps
                up = f(params)
                flat[idx] = old - eps
                dn = f(params)
                flat[idx] = old
                num = (up - dn) / (2 * eps)
                ana = grads[li][which].reshape(-1)[idx]
                assert ana == pytest.approx(num, rel=1e-5, abs=1e-7), (li, which, idx)


def test_cnn_oracle_gelu_is_the_tanh_form():
    from oracle import cnn
    x = np.linspace(-8, 8, 4001)
    # the published tanh form, and its distance from the erf form
    tanh_form = 0.5 * x * (1 + np.tanh(math.sqrt(2 / math.pi) * (x + 0.044715 * x ** 3)))
    assert np.abs(cnn.gelu(x) - tanh_form).max() < 1e-12
    assert np.abs(cnn.gelu(x) - cnn.gelu_erf(x)).max() < 5e-4
    assert cnn.gelu(np.array([0.0]))[0] == 0.0
    h = 1e-6
    num = (cnn.gelu(x + h) - cnn.gelu(x - h)) / (2 * h)
    assert np.abs(cnn.gelu_grad(x) - num).max() < 1e-7


def test_cnn_oracle_bilinear_reproduces_affine_fields():
    from oracle import cnn
    hs, ws, f = 7, 9, 4
    gy, gx = np.meshgrid(np.arange(hs, dtype=float), np.arange(ws, dtype=float), indexing="ij")
    src = 3.0 + 0.7 * gy - 1.3 * gx
    up = cnn.bilinear(src, f, hs * f, ws * f)
    ys = (np.arange(hs * f) + 0.5) / f - 0.5
    xs = (np.arange(ws * f) + 0.5) / f - 0.5
    want = 3.0 + 0.7 * np.clip(ys, 0, hs - 1)[:, None] - 1.3 * np.clip(xs, 0, ws - 1)[None, :]
    assert np.abs(up - want).max() < 1e-12
    # pixel centres of the grid cells' own positions are reproduced exactly at f = 1
    assert np.array_equal(cnn.bilinear(src, 1, hs, ws), src)


def test_cnn_oracle_detail_label_recovers_a_known_slope():
    from oracle import cnn
    rng = np.random.default_rng(5)
    hs, ws, f, s0, unit = 6, 8, 4, 0.037, 20.0
    lab = rng.normal(size=(hs, ws))
    zg = rng.normal(size=(hs, ws)) * 10
    dz = rng.normal(size=(hs * f, ws * f))
    z = cnn.bilinear(zg, f, hs * f, ws * f) + dz
    full = cnn.bilinear(lab, f, hs * f, ws * f) + s0 * dz
    ey0, ex0 = f, 2 * f                          # exact known on a sub-rectangle
    exact = full[ey0:ey0 + 3 * f, ex0:ex0 + 4 * f]
    ls, wt = cnn.detail_label(exact, ey0, ex0, lab, zg, f, z, unit)
    inside = np.zeros((hs, ws), bool)
    inside[1:4, 2:6] = True
    assert np.allclose(ls[inside], unit * s0, rtol=1e-10)
    assert np.isnan(ls[~inside]).all() and (wt[~inside] == 0).all()
    cells = dz.reshape(hs, f, ws, f).transpose(0, 2, 1, 3)
    assert np.allclose(wt[inside], (cells ** 2).sum(axis=(2, 3))[inside] / (f * f * unit ** 2))
    # compose with the recovered slope reproduces the field exactly
    out = np.stack([lab, np.where(inside, ls, 0.0)], axis=-1)
    comp = cnn.compose(out, np.zeros((hs, ws)), zg, f, z=z, s_scale=1.0 / unit)
    assert np.allclose(comp[ey0:ey0 + 3 * f, ex0:ex0 + 4 * f], exact, atol=1e-12)


def test_cnn_oracle_conv_is_a_same_correlation():
    from oracle import cnn
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 5, 6, 2))
    w = rng.normal(size=(3, 2, 3, 3))
    b = rng.normal(size=3)
    out = cnn.conv3x3(x, w, b)
    # brute force at every pixel, zero outside the image
    for y in range(5):
        for xx in range(6):
            for co in range(3):
                s = b[co]
                for dy in range(3):
                    for dx in range(3):
                        yy, xc = y + dy - 1, xx + dx - 1
                        if 0 <= yy < 5 and 0 <= xc < 6:
                            s += float(np.dot(x[0, yy, xc], w[co, :, dy, dx]))
                assert out[0, y, xx, co] == pytest.approx(s, abs=1e-12)
    assert cnn.gelu_erf(np.array([3.0]))[0] == pytest.approx(
        3.0 * 0.5 * (1 + math.erf(3 / math.sqrt(2))))


# ------------------------------------------------ spatial mode (multi-radius)

def test_dask_reflect_is_the_symmetric_pad():
    # dask.array.overlap.reflect: [d c b a | a b c d | d c b a]
    a = np.array([[1, 2, 3, 4]], dtype=F32)
    p = ref.dask_reflect_pad(np.repeat(a, 4, axis=0), 3)
    assert p[3].tolist() == [3, 2, 1, 1, 2, 3, 4, 4, 3, 2]


@pytest.mark.parametrize("radius,algo", [(6.0, "hillshade"), (30.0, "slope"), (64.0, "hillshade"),
                                         (150.0, "curvature")])
def test_smooth_for_radius_keeps_a_constant(radius, algo):
    z = np.full((200, 230), 123.25, F32)
    s = ref.smooth_for_radius(z, radius, pixel_size=10.0, algorithm_name=algo)
    assert s.shape == z.shape
    assert np.abs(s - 123.25).max() < 1e-4


@pytest.mark.parametrize("radius", [6.0, 20.0, 40.0])
def test_smooth_for_radius_sigma_is_half_the_radius(radius):
    # _smooth_for_radius (_nan_utils.py:527-552) smooths with sigma = r / 2
    # (truncate 4, no decimation below r = 48 at 10 m for hillshade).  Away from
    # the edges the Gaussian of u^2 is u^2 plus the second moment of the
    # normalised truncated kernel, sum_k g_k (k / 8)^2 for u = (x - 115) / 8;
    # sigma = r / 3 or r would move it by a factor 4/9 or 4
    sigma = radius / 2.0
    R = int(4.0 * sigma + 0.5)
    k = np.arange(-R, R + 1, dtype=np.float64)
    g = np.exp(-0.5 * (k / sigma) ** 2)
    m2 = float((g * (k / 8.0) ** 2).sum() / g.sum())
    x = np.arange(230, dtype=np.float64)
    z = np.tile((((x - 115.0) / 8.0) ** 2).astype(F32), (60, 1))
    s = ref.smooth_for_radius(z, radius, pixel_size=10.0, algorithm_name="hillshade")
    inner = s[:, R + 1:230 - R - 1] - z[:, R + 1:230 - R - 1]
    assert np.abs(inner - m2).max() < 2e-4
    assert abs(m2 - (sigma / 8.0) ** 2) < 0.01 * (sigma / 8.0) ** 2


def test_spatial_hillshade_of_a_plane_interior_is_the_closed_form():
    # a Gaussian keeps a plane away from the padded edges; radii below the
    # decimation threshold (r / 24 < 2 for hillshade at 10 m)
    h, w, a, b = 260, 300, 0.12, -0.07
    yy, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    z = (a * 10.0 * xx + b * 10.0 * yy).astype(F32)
    kw = dict(pixel_size=10.0, pixel_scale_x=10.0, pixel_scale_y=10.0, radii=[3, 12, 30],
              weights=[0.5, 0.3, 0.2])
    out = ref.spatial_window("hillshade", z, kw)
    want = ref.hillshade(z, pixel_size=10.0, pixel_scale_x=10.0, pixel_scale_y=10.0)
    c = out[100:160, 100:200]
    assert np.abs(c - want[100:160, 100:200]).max() < 2e-5


def test_spatial_single_radius_is_the_spatial_block():
    rng = np.random.default_rng(2)
    z = np.cumsum(rng.normal(size=(120, 140)), axis=1).astype(F32)
    kw = dict(pixel_size=10.0, radii=[5], weights=None)
    out = ref.spatial_window("slope", z, kw)
    d = ref.spatial_depth("slope", 5)
    blk = ref.spatial_block("slope", ref.dask_reflect_pad(z, d), 5, {"pixel_size": 10.0})
    assert np.array_equal(out, blk[d:-d, d:-d])


# ------------------------------------- spatial mode of openness and AO (ray blocks)

@pytest.mark.parametrize("algorithm,radius,factor", [("openness", 64, 2),
                                                     ("ambient_occlusion", 64, 4)])
def test_ray_spatial_block_of_a_flat_raster_is_one(algorithm, radius, factor):
    # the decimated branch (factor from the reference's table) keeps a flat
    # raster flat: openness 90 deg -> 1, AO no occlusion -> 1
    assert ref.radius_to_downsample_factor(radius, pixel_size=10.0,
                                           algorithm_name=algorithm) == factor
    z = np.full((150, 170), 812.5, F32)
    out = ref.ray_spatial_block(algorithm, z, radius, {"pixel_size": 10.0})
    assert out.shape == z.shape
    assert np.abs(out - 1.0).max() < 1e-6


def test_ray_spatial_ao_of_a_plane_matches_the_closed_form():
    # AO of the plane z = a * x (metres) in the decimated branch (f = 4): the
    # ring offsets are those of radius 64 / 4 = 16, and dz / dist is the same
    # on the coarse grid only when both the pixel size and the scales are
    # multiplied by f.  Closed form per sample, written out here.
    a, px, r, f = 0.35, 10.0, 64.0, 4
    h, w = 256, 256
    xx = np.arange(w, dtype=np.float64)[None, :].repeat(h, 0)
    z = (a * px * xx).astype(F32)
    p = {"pixel_size": px, "pixel_scale_x": px, "pixel_scale_y": -px, "num_samples": 16,
         "intensity": 1.0}
    out = ref.ray_spatial_block("ambient_occlusion", z, r, p)
    rs = r / f
    ang = np.linspace(0, 2 * np.pi, 16, endpoint=False)
    tot = n = 0.0
    for rf in (0.25, 0.5, 0.75, 1.0):
        for t in ang:
            dx, dy = int(np.round(rs * rf * np.cos(t))), int(np.round(rs * rf * np.sin(t)))
            if dx == 0 and dy == 0:
                continue
            dist = np.hypot(dx * px * f, dy * px * f)
            occ = min(max(np.arctan(a * dx * px * f / dist), 0.0) / (np.pi / 4), 1.0)
            tot += occ * (1.0 - rf * 0.3)
            n += 1
    want = (1.0 - tot / n) ** (1 / 2.2)
    c = out[80:176, 80:176]                 # rays and the 1-sigma blur stay inside
    assert np.abs(c - want).max() < 2e-5


def test_spatial_openness_single_radius_is_the_ray_block():
    rng = np.random.default_rng(5)
    z = np.cumsum(rng.normal(size=(140, 150)), axis=0).astype(F32) * F32(3.0)
    kw = dict(pixel_size=10.0, radii=[12], weights=None, num_directions=8)
    out = ref.spatial_window("openness", z, kw)
    d = ref.spatial_depth("openness", 12)
    assert d == 13
    blk = ref.ray_spatial_block("openness", ref.dask_reflect_pad(z, d), 12,
                                {"pixel_size": 10.0, "num_directions": 8})
    assert np.array_equal(out, blk[d:-d, d:-d])


# ------------------------------------------------------ prepare (NoData + fill)

from oracle import prepare as prep  # noqa: E402


def test_prepare_fill_coarse_surface_fills_voids_and_preserves_valid():
    # reference tests/test_dem_preprocess.py:37-49
    coarse = np.arange(64, dtype=F32).reshape(8, 8)
    valid = np.ones((8, 8), dtype=bool)
    valid[3:5, 3:5] = False
    out = prep.fill_coarse_surface(coarse, valid)
    assert np.isfinite(out).all()
    assert np.allclose(out[valid], coarse[valid])
    assert float(out[~valid].min()) >= float(coarse[valid].min()) - 1e-3
    assert float(out[~valid].max()) <= float(coarse[valid].max()) + 1e-3


def test_prepare_fill_coarse_surface_flat_region_stays_flat():
    # reference tests/test_dem_preprocess.py:52-65
    coarse = np.full((128, 128), 100.0, dtype=F32)
    valid = np.ones((128, 128), dtype=bool)
    valid[40:90, 40:90] = False
    out = prep.fill_coarse_surface(coarse, valid)
    assert np.isfinite(out).all()
    assert np.allclose(out[~valid], 100.0, atol=1e-2)
    assert float(out[~valid].std()) < 1e-2


def test_prepare_fill_coarse_surface_all_invalid_returns_zeros():
    # reference tests/test_dem_preprocess.py:68-75
    out = prep.fill_coarse_surface(np.full((4, 4), np.nan, F32), np.zeros((4, 4), bool))
    assert np.isfinite(out).all() and np.all(out == 0.0)


def test_prepare_nan_aware_coarse_average_no_boundary_blend():
    # reference tests/test_dem_preprocess.py:78-103 (the raster as the NEAREST
    # read of itself: 256^2 at coarse_max 64 is a 4x oversample of 64^2)
    H = W = 256
    data = np.zeros((H, W), F32)
    data[:, W // 2:] = 5000.0
    ch, cw = prep.coarse_shape(W, H, 64)
    k = prep.coarse_oversample(ch, cw)
    assert (ch, cw, k) == (64, 64, 4)
    nn = prep.nearest_sample(data, ch * k, cw * k)
    assert np.array_equal(nn, data)
    coarse, cmask = prep.nan_aware_coarse_average(nn, np.zeros_like(nn, bool), ch, cw, k,
                                                  None, [0.0])
    valid = coarse[~cmask]
    assert valid.size > 0 and valid.min() >= 4999.0 and valid.max() <= 5001.0
    assert cmask.any()
    assert not ((coarse > 1.0) & (coarse < 4000.0)).any()


def test_prepare_coarse_shape_caps_longest_side():
    # reference tests/test_dem_preprocess.py:106-110
    ch, cw = prep.coarse_shape(240000, 220000, 2048)
    assert max(ch, cw) <= 2048
    assert abs((cw / ch) - (240000 / 220000)) < 0.05


def test_prepare_nearest_sample_integer_ratios():
    rng = np.random.default_rng(3)
    a = rng.normal(size=(12, 20)).astype(F32)
    up = prep.nearest_sample(a, 36, 60)                       # x3 up: each pixel repeated
    assert np.array_equal(up, np.repeat(np.repeat(a, 3, 0), 3, 1))
    dn = prep.nearest_sample(a, 6, 5)                         # /2 rows, /4 cols: centre cells
    assert np.array_equal(dn, a[1::2][:, 2::4])


def _bfs_exterior(mask):
    h, w = mask.shape
    ext = np.zeros_like(mask)
    stack = [(i, j) for i in range(h) for j in range(w)
             if mask[i, j] and (i in (0, h - 1) or j in (0, w - 1))]
    for i, j in stack:
        ext[i, j] = True
    while stack:
        i, j = stack.pop()
        for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            u, v = i + di, j + dj
            if 0 <= u < h and 0 <= v < w and mask[u, v] and not ext[u, v]:
                ext[u, v] = True
                stack.append((u, v))
    return ext


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_prepare_edge_connected_is_a_4_connected_flood_fill(seed):
    rng = np.random.default_rng(seed)
    m = rng.random((37, 41)) < 0.45
    assert np.array_equal(prep.edge_connected_mask(m), _bfs_exterior(m))


def test_prepare_sentinel_rule():
    rng = np.random.default_rng(4)
    z = (rng.random((64, 64)) * 800 + 100).astype(F32)
    v = np.ones_like(z, bool)
    s = z.copy()
    s[rng.random(z.shape) < 0.1] = -9999.0                    # a known sentinel, 10 %
    assert prep.detect_sentinel_nodata(s, v) == -9999.0
    s = z.copy()
    s[rng.random(z.shape) < 0.1] = 512.25                     # interior, not a sentinel
    assert prep.detect_sentinel_nodata(s, v) is None
    s = z.copy()
    s[rng.random(z.shape) < 0.1] = 4321.5                     # the range maximum
    assert prep.detect_sentinel_nodata(s, v) == 4321.5
    s = z.copy()
    s[rng.random(z.shape) < 0.02] = -9999.0                   # below min_fraction
    assert prep.detect_sentinel_nodata(s, v) is None


def test_prepare_border_rule():
    rng = np.random.default_rng(5)
    z = (rng.random((64, 64)) * 800 + 100).astype(F32)
    v = np.ones_like(z, bool)
    f = z.copy()
    f[:6, :] = f[-6:, :] = f[:, :6] = f[:, -6:] = 0.0           # a constant frame
    assert prep.detect_border_nodata(f, v) == 0.0
    f = z.copy()
    f[0, :] = 0.0                                             # one edge: under half the ring
    assert prep.detect_border_nodata(f, v) is None
    f = z.copy()
    f[0, :] = f[-1, :] = f[:, 0] = f[:, -1] = 7.0            # whole ring, 6 % of the grid
    assert prep.detect_border_nodata(f, v) == 7.0
    assert prep.detect_border_nodata(f, v, min_total_fraction=0.07) is None


def test_prepare_end_to_end_keeps_the_exterior_and_fills_the_holes():
    # plane with an undeclared 0 frame (exterior, detected: 0 is a known
    # sentinel covering 23 % of the grid) and an interior -9999 hole given as
    # the override
    H, W = 300, 340
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    plane = (1000.0 + 0.5 * xx + 0.25 * yy).astype(F32)
    z = plane.copy()
    z[:20, :] = z[-20:, :] = z[:, :20] = z[:, -20:] = 0.0
    z[140:160, 150:175] = -9999.0
    out, info = prep.prepare(z, coarse_max=128, nodata_override=-9999.0)
    assert info["extra_nodata"] == [-9999.0, 0.0]
    assert info["detected"] == [("sentinel", 0.0)]
    assert info["do_fill"] and info["exterior"] is not None and info["exterior"].any()
    assert np.isnan(out[:10, :]).all() and np.isnan(out[:, -10:]).all()
    inner = out[140:160, 150:175]
    assert np.isfinite(inner).all()
    assert np.abs(inner - plane[140:160, 150:175]).max() < 12.0   # a membrane over the hole
    keep = (z != 0.0) & (z != -9999.0)
    assert np.array_equal(out[keep], z[keep])
    out_all, _ = prep.prepare(z, coarse_max=128, nodata_override=-9999.0, fill_mode="all")
    assert np.isfinite(out_all).all()
    out_none, _ = prep.prepare(z, coarse_max=128, nodata_override=-9999.0, fill_mode="none")
    assert np.isnan(out_none[140:160, 150:175]).all() and np.isnan(out_none[:20]).all()


# ------------------------------------------- Dask edge semantics (reflect)

def test_dask_edge_hillshade_of_a_plane_halves_the_edge_gradient():
    # boundary='reflect' mirrors [a | a b ..]: the central difference at the
    # first column is (z1 - z0) / (2 h), half the plane's slope, where the
    # tile backend's one-sided edge_order=2 stencil gives the full slope
    h, w, a, px = 40, 50, 0.3, 10.0
    z = np.tile((a * px * np.arange(w)).astype(F32), (h, 1))
    kw = dict(pixel_size=px, pixel_scale_x=px, pixel_scale_y=px)
    dk = ref.dask_single_chunk("hillshade", z, kw)
    tile = ref.run_window("hillshade", z, kw)
    half = np.tile((0.5 * a * px * np.arange(w)).astype(F32), (h, 1))
    want_edge = ref.run_window("hillshade", half, kw)[5, 10]    # interior value of slope a / 2
    assert abs(float(dk[5, 0]) - float(want_edge)) < 1e-6
    assert abs(float(dk[5, -1]) - float(want_edge)) < 1e-6
    assert abs(float(tile[5, 0]) - float(tile[5, 10])) < 1e-6   # one-sided: full slope
    assert np.array_equal(dk[2:-2, 2:-2], tile[2:-2, 2:-2])        # the interior is the same


def test_dask_edge_slope_closed_form_at_the_corner():
    h, w, a, b, px = 30, 30, 0.2, 0.1, 5.0
    yy, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    z = (a * px * xx + b * px * yy).astype(F32)
    dk = ref.dask_single_chunk("slope", z, dict(pixel_size=px))
    # at (0, 0) both axes see half the step: atan(hypot(a, b) / 2) in degrees
    assert abs(float(dk[0, 0]) - np.degrees(np.arctan(np.hypot(a, b) / 2))) < 2e-5
    assert abs(float(dk[15, 15]) - np.degrees(np.arctan(np.hypot(a, b)))) < 2e-5


def test_dask_edge_depths():
    assert ref.dask_edge_depth("curvature", {}) == 2
    assert ref.dask_edge_depth("openness", {"max_distance": 20}) == 21
    assert ref.dask_edge_depth("ambient_occlusion", {"radius": 7.5}) == 8


# ------------------------------------------------------- COG reader oracle

def test_tiff_oracle_decodes_libtiff_zstd_predictor3(tmp_path):
    # Pillow's libtiff writes the file; the oracle's pure-Python reader
    # (predictor 3 restated from Adobe TN3) must return the exact input
    from PIL import Image
    from oracle import tiff
    rng = np.random.default_rng(8)
    a = (rng.normal(size=(37, 53)) * 1000).astype(F32)
    a[3, 4] = np.nan
    p = tmp_path / "f.tif"
    Image.fromarray(a).save(p, compression="zstd", tiffinfo={317: 3})
    P = tiff.parse(str(p))
    assert P["ifds"][0]["tags"][317] == [3] and P["ifds"][0]["tags"][259] == [50000]
    got = tiff.decode(str(p), 0, P)
    assert np.array_equal(got, a, equal_nan=True)


def test_tiff_oracle_decodes_libtiff_deflate_predictor2(tmp_path):
    from PIL import Image
    from oracle import tiff
    a = (np.arange(40 * 61) % 251).astype(np.uint8).reshape(40, 61)
    p = tmp_path / "u.tif"
    Image.fromarray(a).save(p, compression="tiff_adobe_deflate", tiffinfo={317: 2})
    P = tiff.parse(str(p))
    assert P["ifds"][0]["tags"][317] == [2]
    assert np.array_equal(tiff.decode(str(p), 0, P), a)


def test_tiff_oracle_overview_average_closed_forms():
    from oracle import tiff
    a = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9]], F32)
    o = tiff.overview_average(a)
    assert o.tolist() == [[3.0, 4.5], [7.5, 9.0]]            # clipped blocks at the edge
    b = a.copy()
    b[0, 0] = np.nan
    assert tiff.overview_average(b)[0, 0] == np.float32((2 + 4 + 5) / 3)
    c = np.array([[-9999, 10], [-9999, -9999]], np.int16)
    assert tiff.overview_average(c, nodata=-9999).tolist() == [[10]]
    assert tiff.overview_average(np.array([[1, 2]], np.uint8)).tolist() == [[2]]   # 1.5 -> 2
    assert tiff.overview_levels(16384, 16384) == [(16384 >> k,) * 2 for k in range(6)]

"""The prepare step (NoData inference, push-pull void fill) on the GPU against
the oracle's restatement of preprocess_dem_to_cog (oracle/prepare.py)."""
import numpy as np
import pytest
import torch

from oracle import prepare as prep

pytestmark = pytest.mark.gpu
F32 = np.float32


def _fbm(h, w, seed, relief=1500.0):
    from fujishadergpu_amd import synth
    return synth.fbm(h, w, 0.8, relief, seed=seed, device="cpu").numpy()


@pytest.mark.parametrize("shape,out", [((100, 120), (400, 480)), ((300, 257), (150, 64)),
                                       ((777, 501), (1024, 640)), ((64, 64), (64, 64))])
def test_nearest_sample_matches_the_oracle(cuda, shape, out):
    from fujishadergpu_amd import prepare as P
    a = np.random.default_rng(1).normal(size=shape).astype(F32)
    got = P.nearest_sample(torch.from_numpy(a).to(cuda), *out).cpu().numpy()
    assert np.array_equal(got, prep.nearest_sample(a, *out))


def _np_mode(vals, src_nodata):
    v = vals[np.isfinite(vals)]
    if src_nodata is not None and np.isfinite(src_nodata):
        v = v[v != F32(src_nodata)]
    u, c = np.unique(v, return_counts=True)
    i = int(np.argmax(c))
    return float(u[i]), int(c[i]), int(v.size), float(u[0]), float(u[-1]), int(u.size)


@pytest.mark.parametrize("case", ["planted", "ties", "zeros", "declared", "nan"])
def test_value_mode_matches_numpy_unique(cuda, case):
    from fujishadergpu_amd import prepare as P
    rng = np.random.default_rng(7)
    v = np.round(rng.normal(size=200_003) * 50).astype(F32)   # many repeats
    decl = None
    if case == "planted":
        v[rng.random(v.size) < 0.07] = -9999.0
    elif case == "ties":
        v = np.repeat(np.arange(-50, 50, dtype=F32), 7)        # all counts equal -> smallest
    elif case == "zeros":
        v[::3] = 0.0
        v[1::3] = -0.0
    elif case == "declared":
        v[rng.random(v.size) < 0.2] = 3.0
        decl = 3.0
    else:
        v[rng.random(v.size) < 0.3] = np.nan
        v[:5] = np.inf
    st = P.value_mode(torch.from_numpy(v).to(cuda), decl)
    mode, cnt, nv, vmin, vmax, nd = _np_mode(v, decl)
    assert (st["mode"], st["count"], st["n_valid"], st["vmin"], st["vmax"], st["n_distinct"]) \
        == (mode, cnt, nv, vmin, vmax, nd)


@pytest.mark.parametrize("k", [2, 3, 4])
def test_coarse_average_matches_the_oracle_bit_for_bit(cuda, k):
    from fujishadergpu_amd import prepare as P
    ch, cw = 61, 83
    rng = np.random.default_rng(k)
    s = (_fbm(ch * k, cw * k, 11) + 800.0).astype(F32)
    s[rng.random(s.shape) < 0.15] = np.nan
    s[:, : 3 * k] = 0.0                        # an undeclared 0 band
    s[20:30, 40:50] = -9999.0                  # the declared NoData
    nodata = [-9999.0, 0.0]
    coarse, cvalid, n_void = P.coarse_average(torch.from_numpy(s).to(cuda), ch, cw, k, nodata)
    exp, cmask = prep.nan_aware_coarse_average(s, s == F32(-9999.0), ch, cw, k, -9999.0, [0.0])
    assert np.array_equal(coarse.cpu().numpy(), exp)
    assert np.array_equal(cvalid.cpu().numpy() == 0, cmask | ~np.isfinite(exp))
    assert n_void == int(cmask.sum())


def _pp_case(name):
    rng = np.random.default_rng(5)
    if name == "ramp":
        c = np.arange(64, dtype=F32).reshape(8, 8)
        v = np.ones((8, 8), bool)
        v[3:5, 3:5] = False
    elif name == "flat":
        c = np.full((128, 128), 100.0, F32)
        v = np.ones((128, 128), bool)
        v[40:90, 40:90] = False
    elif name == "terrain":
        c = (_fbm(301, 517, 13) + 500.0).astype(F32)
        v = rng.random(c.shape) > 0.35
        v[100:220, 200:400] = False
    else:                                      # a lone valid cell in a wide grid
        c = np.full((33, 700), 42.5, F32)
        v = np.zeros(c.shape, bool)
        v[16, 350] = True
    return c, v


@pytest.mark.parametrize("name", ["ramp", "flat", "terrain", "single"])
def test_pushpull_fill_matches_the_oracle(cuda, name):
    from fujishadergpu_amd import prepare as P
    c, v = _pp_case(name)
    got = P.pushpull_fill(torch.from_numpy(c).to(cuda),
                          torch.from_numpy(v.astype(np.uint8)).to(cuda),
                          int((~v).sum())).cpu().numpy()
    exp = prep.fill_coarse_surface(c, v)
    assert np.array_equal(got[v], c[v])
    d = np.abs(got.astype(np.float64) - exp)
    scale = max(1.0, float(np.abs(c[v]).max()))
    print(name, "max |d|", d.max(), "identical", np.mean(got == exp), flush=True)
    # same zoom coordinates, weights and float64 accumulation order as scipy:
    # bit-identical (the bound is what a reordered accumulation could cost)
    assert d.max() <= 4e-7 * scale
    assert np.array_equal(got, exp)


def test_pushpull_all_void_gives_zeros(cuda):
    from fujishadergpu_amd import prepare as P
    c = torch.full((4, 4), float("nan"), device=cuda)
    got = P.pushpull_fill(c, torch.zeros((4, 4), dtype=torch.uint8, device=cuda), 16)
    assert torch.equal(got, torch.zeros_like(got))


@pytest.mark.parametrize("case", ["random", "spiral", "frame"])
def test_edge_connected_matches_the_oracle(cuda, case):
    from fujishadergpu_amd import prepare as P
    rng = np.random.default_rng(9)
    if case == "random":
        m = rng.random((301, 417)) < 0.55
    elif case == "spiral":                     # a one-cell serpentine from the top edge
        h, w = 97, 130
        m = np.zeros((h, w), bool)
        m[0, :] = True
        for r in range(2, h - 1, 2):
            m[r, 1:w - 1] = True
        for i, r in enumerate(range(1, h - 2, 2)):
            m[r, 1 if i % 2 == 0 else w - 2] = True
    else:
        m = np.zeros((500, 640), bool)
        m[:25] = m[-25:] = True
        m[:, :25] = m[:, -25:] = True
        m[200:260, 300:380] = True             # an enclosed void
    ext, sweeps = P.edge_connected(torch.from_numpy(m).to(cuda))
    exp = prep.edge_connected_mask(m)
    print(case, "sweeps", sweeps, flush=True)
    assert np.array_equal(ext.cpu().numpy().astype(bool), exp)


def _raster(h, w, seed):
    z = (_fbm(h, w, seed) + 900.0).astype(F32)
    rng = np.random.default_rng(seed)
    z[: h // 12] = 0.0                         # undeclared 0 exterior (top band)
    z[:, -w // 15:] = 0.0                      # and the right edge
    z[h // 2: h // 2 + 40, w // 3: w // 3 + 55] = -9999.0   # declared NoData hole
    z[rng.random(z.shape) < 0.002] = np.nan    # scattered NaN
    z[h // 3: h // 3 + 30, w // 2: w // 2 + 30] = np.nan     # an enclosed NaN hole
    return z


@pytest.mark.parametrize("fill_mode", ["enclosed", "all", "none"])
def test_preprocess_matches_the_oracle(cuda, fill_mode):
    from fujishadergpu_amd import prepare as P
    z = _raster(600, 720, 21)
    got, gi = P.preprocess_dem(torch.from_numpy(z).to(cuda), src_nodata=-9999.0,
                               fill_mode=fill_mode, coarse_max=160)
    exp, ei = prep.prepare(z, src_nodata=-9999.0, fill_mode=fill_mode, coarse_max=160)
    assert gi["extra_nodata"] == ei["extra_nodata"] and gi["detected"] == ei["detected"]
    assert gi["do_fill"] == ei["do_fill"]
    got = got.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    keep = np.isfinite(z) & (z != 0.0) & (z != -9999.0)
    assert np.array_equal(got[keep], z[keep])
    d = np.abs(got[~keep & np.isfinite(exp)].astype(np.float64) - exp[~keep & np.isfinite(exp)])
    print(fill_mode, "filled px", d.size, "max |d|", d.max() if d.size else 0.0, flush=True)
    assert d.size == 0 or d.max() <= 1e-3
    assert np.array_equal(got, exp, equal_nan=True)


def test_preprocess_full_size_properties(cuda):
    # 8192^2 (coarse 2048^2, oversample 2): valid pixels unchanged, the
    # enclosed holes filled, the exterior frame NaN
    from fujishadergpu_amd import prepare as P
    from fujishadergpu_amd import synth
    n = 8192
    z = synth.fbm(n, n, 0.8, 2000.0, seed=31, device=cuda) + 500.0
    z[:300] = 0.0
    z[:, :300] = 0.0
    z[4000:4400, 5000:5600] = float("nan")
    valid = torch.isfinite(z) & (z != 0.0)
    out, info = P.preprocess_dem(z, fill_mode="enclosed")
    assert info["coarse_shape"] == (2048, 2048) and info["oversample"] == 2
    assert 0.0 in info["extra_nodata"]
    assert torch.equal(out[valid], z[valid])
    assert torch.isfinite(out[4000:4400, 5000:5600]).all()
    assert torch.isnan(out[:280, :]).all() and torch.isnan(out[:, :280]).all()

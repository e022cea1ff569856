"""COG output and windowed input (csrc/cog.hip) against the oracle reader,
the oracle's AVERAGE pyramid and Pillow's libtiff as an independent decoder."""
import numpy as np
import pytest
import torch

from oracle import tiff

pytestmark = pytest.mark.gpu
F32 = np.float32


def _fbm(h, w, seed):
    from fujishadergpu_amd import synth
    return synth.fbm(h, w, 0.8, 2000.0, seed=seed, device="cpu").numpy() + F32(300.0)


@pytest.mark.parametrize("kind", ["f32", "i16", "u8"])
def test_cog_write_layout_pixels_and_overviews(cuda, tmp_path, kind):
    from fujishadergpu_amd import cogio
    h, w = 1300, 1733
    a = _fbm(h, w, 3)
    rng = np.random.default_rng(4)
    if kind == "f32":
        a[rng.random(a.shape) < 0.05] = np.nan
        nodata = float("nan")
    elif kind == "i16":
        a = np.round(a).astype(np.int16)
        a[rng.random(a.shape) < 0.05] = -9999
        nodata = -9999.0
    else:
        a = (a / 10).astype(np.uint8)
        nodata = None
    p = str(tmp_path / f"{kind}.tif")
    tr = (10.0, 0.0, 500000.0, 0.0, -10.0, 4200000.0)
    rep = cogio.write_cog(p, torch.from_numpy(a).to(cuda), transform=tr, epsg=32654,
                          nodata=nodata, block=256, overview_count=8)
    shapes = tiff.overview_levels(h, w, 256, 8)
    assert rep["levels"] == len(shapes) == 4
    P = tiff.parse(p)
    assert P["big"] and len(P["ifds"]) == len(shapes)
    t0 = P["ifds"][0]["tags"]
    pred = {"f32": 3, "i16": 2, "u8": 1}[kind]
    assert t0[259] == [50000] and t0[317] == [pred] and t0[322] == [256] and t0[323] == [256]
    assert t0[33550][:2] == [10.0, 10.0] and t0[33922][3:5] == [500000.0, 4200000.0]
    assert 3072 in t0[34735] and 32654 in t0[34735]
    assert (42113 in t0) == (nodata is not None)
    # COG layout: every IFD ahead of the tile data, the data from the smallest
    # overview to full resolution
    first_data = min(min(d["tags"][324]) for d in P["ifds"])
    assert max(d["offset"] for d in P["ifds"]) < first_data == rep["data_start"]
    starts = [min(d["tags"][324]) for d in P["ifds"]]
    assert starts == sorted(starts, reverse=True)
    for d in P["ifds"][1:]:
        assert d["tags"][254] == [1]
    # pixels: level 0 exact, overviews = the AVERAGE cascade
    want = a
    for lv, shp in enumerate(shapes):
        got = tiff.decode(p, lv, P)
        assert got.shape == shp
        if lv:
            want = tiff.overview_average(want, nodata=nodata if kind != "f32" else None)
        assert np.array_equal(got, want, equal_nan=True), f"level {lv}"


def test_cog_readable_by_libtiff(cuda, tmp_path):
    # an independent decoder: Pillow's libtiff opens the BigTIFF COG, the main
    # image and the overview pages
    from PIL import Image
    from fujishadergpu_amd import cogio
    a = _fbm(700, 900, 5)
    a[10:20, 30:40] = np.nan
    p = str(tmp_path / "c.tif")
    cogio.write_cog(p, torch.from_numpy(a).to(cuda), nodata=float("nan"), block=256)
    im = Image.open(p)
    assert np.array_equal(np.array(im), a, equal_nan=True)
    im.seek(1)
    assert np.array_equal(np.array(im), tiff.overview_average(a), equal_nan=True)


@pytest.mark.parametrize("kind", ["f32", "i16", "u8"])
def test_cog_window_reads(cuda, tmp_path, kind):
    from fujishadergpu_amd import cogio
    a = _fbm(1100, 1500, 7)
    if kind == "i16":
        a = np.round(a).astype(np.int16)
    elif kind == "u8":
        a = (a / 10).astype(np.uint8)
    p = str(tmp_path / "w.tif")
    cogio.write_cog(p, torch.from_numpy(a).to(cuda), block=256, transform=(2, 0, 100, 0, -2, 900),
                    epsg=6677)
    ov = tiff.overview_average(a)
    rng = np.random.default_rng(1)
    with cogio.CogReader(p) as rd:
        inf = rd.info(0)
        assert (inf["width"], inf["height"], inf["tile"]) == (1500, 1100, (256, 256))
        assert inf["transform"] == (2.0, 0.0, 100.0, 0.0, -2.0, 900.0) and inf["epsg"] == 6677
        for _ in range(6):
            c0, r0 = int(rng.integers(0, 1400)), int(rng.integers(0, 1000))
            w, h = int(rng.integers(1, 1500 - c0 + 1)), int(rng.integers(1, 1100 - r0 + 1))
            got = rd.read((c0, r0, w, h)).cpu().numpy()
            assert np.array_equal(got, a[r0:r0 + h, c0:c0 + w])
        full1 = rd.read(None, level=1).cpu().numpy()
        assert np.array_equal(full1, ov)
    assert torch.equal(cogio.read_window(p, (5, 7, 300, 200)).cpu(),
                       torch.from_numpy(a[7:207, 5:305].astype(F32)))


@pytest.mark.parametrize("comp,pred,kind", [("tiff_adobe_deflate", 3, "f32"),
                                             ("zstd", 3, "f32"), ("zstd", 2, "u8"),
                                             ("raw", 1, "u8")])
def test_cog_reader_reads_libtiff_files(cuda, tmp_path, comp, pred, kind):
    # files written by libtiff through Pillow (classic TIFF, strips) read by the
    # GPU reader: an independent encoder for its decompression and predictors
    from PIL import Image
    from fujishadergpu_amd import cogio
    a = _fbm(300, 410, 9)
    if kind == "u8":
        a = (a / 10).astype(np.uint8)
    p = str(tmp_path / "t.tif")
    Image.fromarray(a).save(p, compression=comp, tiffinfo={317: pred} if pred > 1 else {})
    with cogio.CogReader(p) as rd:
        assert rd.info(0)["predictor"] == pred
        got = rd.read().cpu().numpy()
        win = rd.read((17, 33, 200, 150)).cpu().numpy()
    assert np.array_equal(got, a)
    assert np.array_equal(win, a[33:183, 17:217])

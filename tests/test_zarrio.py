"""Zarr v2 I/O (core/dask_io.py:15-97): the layout xarray's to_zarr(zarr_format=2,
consolidated=True) writes, read back with the reference's variable choice,
band squeeze and float32 cast.  No zarr / xarray in this image: the format
pins are the Zarr v2 spec's (metadata keys, chunk keys, edge chunks at the full
chunk shape, fill value encoding) and the codecs' own stream formats."""
import gzip
import json
import os
import zlib

import numpy as np
import pytest
import torch

from fujishadergpu_amd import zarrio


def test_zarr_path_detection():
    # tests/test_zarr_io.py:9-11
    assert zarrio.is_zarr_path("a.zarr")
    assert zarrio.is_zarr_path("A.ZARR")
    assert not zarrio.is_zarr_path("a.tif")


def test_select_data_var_order():
    assert zarrio.select_data_var(["x", "band_data", "dem"]) == "dem"
    assert zarrio.select_data_var(["foo", "data"]) == "data"
    assert zarrio.select_data_var(["foo", "bar"]) == "foo"
    with pytest.raises(ValueError, match="No data variables"):
        zarrio.select_data_var([])


def test_roundtrip_keeps_the_variable_name_and_values(tmp_path):
    # tests/test_zarr_io.py:14-41: the 2 x 2 round trip, name 'dem' kept
    src = str(tmp_path / "input.zarr")
    vals = torch.tensor([[1.0, 2.0], [3.0, 4.0]])
    zarrio.write_zarr(vals, src, name="dem", chunks=(2, 2))
    got, name, info = zarrio.read_zarr(src, device="cpu")
    assert name == "dem" and info["dims"] == ["y", "x"]
    assert got.dtype == torch.float32 and got.tolist() == vals.tolist()
    out = str(tmp_path / "output.zarr")
    zarrio.write_zarr(got, out, name=name)
    back, name2, _ = zarrio.read_zarr(out, device="cpu")
    assert name2 == "dem" and back.tolist() == vals.tolist()


@pytest.mark.parametrize("comp", [None, "zlib", "gzip", "zstd"])
def test_ragged_chunks_and_nan_round_trip_bit_exact(tmp_path, comp):
    rng = np.random.default_rng(3)
    a = rng.normal(1000.0, 300.0, size=(1000, 777)).astype(np.float32)
    a[5:9, 700:777] = np.nan
    p = str(tmp_path / "r.zarr")
    zarrio.write_zarr(torch.from_numpy(a), p, chunks=(256, 200), compressor=comp)
    got, name, _ = zarrio.read_zarr(p, device="cpu")
    assert name == "result"                     # unnamed data -> 'result'
    g = got.numpy()
    assert np.array_equal(np.isnan(g), np.isnan(a))
    assert np.array_equal(g[~np.isnan(a)].view(np.uint32), a[~np.isnan(a)].view(np.uint32))
    # 4 x 4 chunk grid, keys "i.j"
    keys = sorted(f for f in os.listdir(os.path.join(p, "result")) if not f.startswith("."))
    assert keys == sorted(f"{i}.{j}" for i in range(4) for j in range(4))


def test_layout_and_codec_streams(tmp_path):
    a = np.arange(30, dtype=np.float32).reshape(5, 6)
    p = str(tmp_path / "l.zarr")
    zarrio.write_zarr(torch.from_numpy(a), p, name="dem", chunks=(4, 4), compressor="zlib",
                      level=3)
    za = json.load(open(os.path.join(p, "dem", ".zarray")))
    assert za == {"chunks": [4, 4], "compressor": {"id": "zlib", "level": 3}, "dtype": "<f4",
                  "fill_value": "NaN", "filters": None, "order": "C", "shape": [5, 6],
                  "zarr_format": 2}
    assert json.load(open(os.path.join(p, "dem", ".zattrs"))) == {"_ARRAY_DIMENSIONS": ["y", "x"]}
    assert json.load(open(os.path.join(p, ".zgroup"))) == {"zarr_format": 2}
    zm = json.load(open(os.path.join(p, ".zmetadata")))
    assert zm["zarr_consolidated_format"] == 1 and zm["metadata"]["dem/.zarray"] == za
    # chunk (1, 1) = rows 4..7, cols 4..7: the 1 x 2 corner of the data, the
    # rest the NaN fill, as a plain zlib stream of little-endian float32
    raw = zlib.decompress(open(os.path.join(p, "dem", "1.1"), "rb").read())
    blk = np.frombuffer(raw, dtype="<f4").reshape(4, 4)
    assert blk[0, :2].tolist() == [28.0, 29.0] and np.isnan(blk[1:]).all() and \
        np.isnan(blk[0, 2:]).all()


def _hand_store(path, arrays):
    """A store laid out as xarray writes one: {name: (array, dims, attrs, comp)}."""
    os.makedirs(path)
    meta = {".zgroup": {"zarr_format": 2}, ".zattrs": {}}
    for name, (arr, dims, attrs, comp) in arrays.items():
        os.makedirs(os.path.join(path, name))
        za = {"chunks": list(arr.shape) or [], "compressor": comp, "dtype": arr.dtype.str,
              "fill_value": None, "filters": None, "order": "C", "shape": list(arr.shape),
              "zarr_format": 2}
        meta[f"{name}/.zarray"] = za
        meta[f"{name}/.zattrs"] = {"_ARRAY_DIMENSIONS": list(dims), **attrs}
        key = ".".join(["0"] * arr.ndim) if arr.ndim else "0"
        raw = arr.tobytes()
        if comp and comp["id"] == "gzip":
            raw = gzip.compress(raw)
        with open(os.path.join(path, name, key), "wb") as f:
            f.write(raw)
    for rel, doc in meta.items():
        json.dump(doc, open(os.path.join(path, rel), "w"))
    json.dump({"metadata": meta, "zarr_consolidated_format": 1},
              open(os.path.join(path, ".zmetadata"), "w"))


def test_reads_an_xarray_style_store_with_coordinates_and_a_band(tmp_path):
    p = str(tmp_path / "x.zarr")
    band = np.arange(12, dtype=">i2").reshape(1, 3, 4)    # big-endian int16, one band
    _hand_store(p, {
        "x": (np.arange(4, dtype="<f8"), ["x"], {}, None),
        "y": (np.arange(3, dtype="<f8"), ["y"], {}, None),
        "spatial_ref": (np.array(0, dtype="<i8"), [], {}, None),
        "band_data": (band, ["band", "y", "x"], {"coordinates": "spatial_ref"},
                      {"id": "gzip", "level": 1}),
    })
    got, name, info = zarrio.read_zarr(p, device="cpu")
    assert name == "band_data" and info["dims"] == ["y", "x"]
    assert got.dtype == torch.float32 and got.tolist() == band[0].astype(np.float32).tolist()


def test_rejects_what_it_cannot_read(tmp_path):
    p = str(tmp_path / "b.zarr")
    _hand_store(p, {"dem": (np.zeros((2, 2), dtype="<f4"), ["y", "x"], {},
                            {"id": "blosc", "cname": "lz4"})})
    with pytest.raises(NotImplementedError, match="blosc"):
        zarrio.read_zarr(p, device="cpu")
    q = str(tmp_path / "c.zarr")
    _hand_store(q, {"dem": (np.zeros((2, 2, 2), dtype="<f4"), ["t", "y", "x"], {}, None)})
    with pytest.raises(ValueError, match="exactly 2D"):
        zarrio.read_zarr(q, device="cpu")
    with pytest.raises(ValueError, match="compressor"):
        zarrio.write_zarr(torch.zeros(2, 2), str(tmp_path / "d.zarr"), compressor="blosc")

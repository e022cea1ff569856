"""Zarr v2 raster I/O for device-resident DEMs.

The reference's Dask pipeline reads and writes Zarr through xarray
(core/dask_io.py:15-97):

- ``is_zarr_path``: a path ending in ``.zarr``;
- ``select_data_var``: the first of elevation / dem / band_data / data, else
  the first variable;
- ``load_input_dataarray``: a single-band ``band`` dimension squeezed away,
  exactly 2-D required, cast to float32;
- ``write_zarr_output``: ``to_dataset(name=data.name or 'result')`` stored
  with ``zarr_format=2, consolidated=True, mode='w'``.

Neither xarray nor zarr is part of this stack, so this module writes and
reads the Zarr v2 layout itself, as those calls lay it out:

- a group: ``.zgroup``, ``.zattrs`` and the consolidated ``.zmetadata``;
- one array directory per variable: ``.zarray`` (shape, chunks, dtype,
  compressor, fill_value, order "C", filters null) and ``.zattrs`` with
  xarray's ``_ARRAY_DIMENSIONS``;
- chunk files ``"i.j"``, edge chunks stored at the full chunk shape.

Chunks are stored uncompressed or with the ``zlib`` / ``gzip`` / ``zstd``
codecs of numcodecs' registry (a zlib stream, a gzip member, a zstd frame).
Blosc, the zarr-python default, needs a library this image does not have:
reading a Blosc store raises.  The raster leaves and enters HBM in one copy;
the chunks are packed and (de)compressed on host threads (zlib and zstd
release the GIL).
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes as C
import gzip
import json
import math
import os
import shutil
import zlib
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

DATA_VAR_CANDIDATES = ("elevation", "dem", "band_data", "data")
COMPRESSORS = (None, "zlib", "gzip", "zstd")


def is_zarr_path(path) -> bool:
    """core/dask_io.py:15-16."""
    return str(path).lower().endswith(".zarr")


def select_data_var(names: Sequence[str]) -> str:
    """core/dask_io.py:19-25: the first of elevation / dem / band_data / data
    present, else the first variable."""
    names = list(names)
    for candidate in DATA_VAR_CANDIDATES:
        if candidate in names:
            return candidate
    if not names:
        raise ValueError("No data variables found in Zarr dataset.")
    return names[0]


# ------------------------------------------------------------------ codecs

_ZSTD = None


def _zstd():
    """libzstd (the system library the COG writer links) through ctypes."""
    global _ZSTD
    if _ZSTD is None:
        lib = C.CDLL("libzstd.so.1")
        lib.ZSTD_compressBound.restype = C.c_size_t
        lib.ZSTD_compressBound.argtypes = [C.c_size_t]
        lib.ZSTD_compress.restype = C.c_size_t
        lib.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
        lib.ZSTD_decompress.restype = C.c_size_t
        lib.ZSTD_decompress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
        lib.ZSTD_isError.restype = C.c_uint
        lib.ZSTD_isError.argtypes = [C.c_size_t]
        _ZSTD = lib
    return _ZSTD


def _encode(buf: bytes, comp: Optional[dict]) -> bytes:
    if comp is None:
        return buf
    cid = comp.get("id")
    if cid == "zlib":
        return zlib.compress(buf, int(comp.get("level", 1)))
    if cid == "gzip":
        return gzip.compress(buf, int(comp.get("level", 1)), mtime=0)
    if cid == "zstd":
        lib = _zstd()
        cap = lib.ZSTD_compressBound(len(buf))
        out = C.create_string_buffer(cap)
        n = lib.ZSTD_compress(out, cap, buf, len(buf), int(comp.get("level", 1)))
        if lib.ZSTD_isError(n):
            raise RuntimeError("zstd compression failed")
        return out.raw[:n]
    raise NotImplementedError(f"Zarr compressor {cid!r} is not supported")


def _decode(data: bytes, comp: Optional[dict], nbytes: int) -> bytes:
    if comp is None:
        return data
    cid = comp.get("id")
    if cid == "zlib":
        return zlib.decompress(data)
    if cid == "gzip":
        return gzip.decompress(data)
    if cid == "zstd":
        lib = _zstd()
        out = C.create_string_buffer(nbytes)
        n = lib.ZSTD_decompress(out, nbytes, data, len(data))
        if lib.ZSTD_isError(n) or n != nbytes:
            raise ValueError("corrupt zstd chunk")
        return out.raw
    raise NotImplementedError(
        f"Zarr compressor {cid!r} is not supported (zlib, gzip, zstd or none)")


# ------------------------------------------------------------------ metadata

def _fill_json(v):
    """Zarr v2 JSON encoding of a fill value (NaN / Infinity as strings)."""
    if v is None:
        return None
    f = float(v)
    if math.isnan(f):
        return "NaN"
    if math.isinf(f):
        return "Infinity" if f > 0 else "-Infinity"
    return f if not float(f).is_integer() else (int(f) if abs(f) < 2 ** 53 else f)


def _fill_value(v, dtype: np.dtype):
    if v is None:
        return None
    if isinstance(v, str):
        return {"NaN": np.nan, "Infinity": np.inf, "-Infinity": -np.inf}[v]
    return dtype.type(v)


def _read_json(path: str) -> dict:
    with open(path, "r", encoding="utf-8") as f:
        return json.load(f)


def _metadata(path: str) -> Dict[str, dict]:
    """Every metadata document of the store, keyed by relative path: the
    consolidated .zmetadata when present, else the files themselves."""
    zm = os.path.join(path, ".zmetadata")
    if os.path.exists(zm):
        return dict(_read_json(zm)["metadata"])
    meta = {}
    for root, _dirs, files in os.walk(path):
        for fn in files:
            if fn in (".zarray", ".zattrs", ".zgroup"):
                rel = os.path.relpath(os.path.join(root, fn), path).replace(os.sep, "/")
                meta[rel] = _read_json(os.path.join(root, fn))
    return meta


# ------------------------------------------------------------------ read

def read_zarr(path: str, device="cuda", variable: Optional[str] = None
              ) -> Tuple[torch.Tensor, str, dict]:
    """load_input_dataarray for a Zarr v2 store (core/dask_io.py:28-57): the
    data variable chosen as select_data_var does (or ``variable``), a
    single-band ``band`` dimension squeezed, exactly 2-D, cast to float32.
    Returns (raster on ``device``, variable name, {"dims", "attrs"})."""
    if not os.path.isdir(path):
        raise FileNotFoundError(path)
    meta = _metadata(path)
    arrays = sorted(k[:-len("/.zarray")] for k in meta if k.endswith("/.zarray"))
    root_array = ".zarray" in meta
    if root_array:
        name, prefix = "", ""
        za = meta[".zarray"]
        attrs = meta.get(".zattrs", {})
    else:
        # xarray's data variables: every array except the coordinates (a
        # dimension's own array, and the arrays another one names in its
        # "coordinates" attribute, e.g. spatial_ref)
        named = set()
        for a in arrays:
            named.update(str(meta.get(f"{a}/.zattrs", {}).get("coordinates", "")).split())
        data_vars = [a for a in arrays if a not in named and
                     meta.get(f"{a}/.zattrs", {}).get("_ARRAY_DIMENSIONS", [a]) != [a]]
        name = variable if variable is not None else select_data_var(data_vars)
        if f"{name}/.zarray" not in meta:
            raise ValueError(f"Zarr store has no array {name!r}")
        prefix = name + "/"
        za = meta[f"{name}/.zarray"]
        attrs = meta.get(f"{name}/.zattrs", {})
    if int(za.get("zarr_format", 2)) != 2:
        raise ValueError("only Zarr format 2 stores are read")
    if za.get("filters"):
        raise NotImplementedError("Zarr filters are not supported")
    if za.get("order", "C") != "C":
        raise NotImplementedError("only C-order Zarr arrays are read")
    dtype = np.dtype(za["dtype"])
    shape = [int(s) for s in za["shape"]]
    chunks = [int(c) for c in za["chunks"]]
    dims = list(attrs.get("_ARRAY_DIMENSIONS", [f"dim_{i}" for i in range(len(shape))]))
    if "band" in dims and shape[dims.index("band")] == 1 and len(shape) == 3:
        pass  # squeezed below, after the read
    elif len(shape) != 2:
        raise ValueError(f"DEM input must be exactly 2D after selecting a data variable; "
                         f"got dims={tuple(dims)}, shape={tuple(shape)}")
    sep = za.get("dimension_separator", ".")
    comp = za.get("compressor")
    fill = _fill_value(za.get("fill_value"), dtype)
    out = np.empty(shape, dtype=dtype)
    grid = [int(math.ceil(s / c)) if s else 0 for s, c in zip(shape, chunks)]
    nbytes = int(np.prod(chunks)) * dtype.itemsize
    base = os.path.join(path, prefix) if prefix else path

    def one(idx):
        key = sep.join(str(i) for i in idx)
        fn = os.path.join(base, key)
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, shape))
        if not os.path.exists(fn):
            out[sl] = fill if fill is not None else 0
            return
        with open(fn, "rb") as f:
            raw = _decode(f.read(), comp, nbytes)
        if len(raw) != nbytes:
            raise ValueError(f"chunk {key} holds {len(raw)} bytes, expected {nbytes}")
        block = np.frombuffer(raw, dtype=dtype).reshape(chunks)
        out[sl] = block[tuple(slice(0, s.stop - s.start) for s in sl)]

    idxs = list(np.ndindex(*grid)) if all(grid) else []
    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(one, idxs))
    if len(shape) == 3:
        b = dims.index("band")
        out = np.take(out, 0, axis=b)
        dims = [d for d in dims if d != "band"]
    host = torch.from_numpy(np.ascontiguousarray(out.astype(np.float32, copy=False)))
    return host.to(device), (name or "result"), {"dims": dims, "attrs": attrs}


# ------------------------------------------------------------------ write

def write_zarr(raster: torch.Tensor, path: str, *, name: Optional[str] = None,
               chunks: Tuple[int, int] = (512, 512), compressor: Optional[str] = "zlib",
               level: int = 1, dims: Tuple[str, str] = ("y", "x"),
               attrs: Optional[dict] = None) -> None:
    """write_zarr_output (core/dask_io.py:74-97): the raster as the variable
    ``name or 'result'`` of a consolidated Zarr v2 group, written with
    mode 'w' (an existing store at ``path`` is replaced).  float32 (NaN fill)
    or any numeric dtype the tensor has."""
    if compressor not in COMPRESSORS:
        raise ValueError(f"compressor must be one of {COMPRESSORS}, got {compressor!r}")
    if raster.dim() != 2:
        raise ValueError("write_zarr takes a 2-D raster")
    cy, cx = int(chunks[0]), int(chunks[1])
    if cy <= 0 or cx <= 0:
        raise ValueError("chunks must be positive")
    var = name or "result"
    if "/" in var or var.startswith("."):
        raise ValueError(f"bad variable name {var!r}")
    host = raster.detach().to("cpu").numpy()
    dtype = host.dtype.newbyteorder("<") if host.dtype.byteorder == ">" else host.dtype
    host = host.astype(dtype, copy=False)
    H, W = host.shape
    fill = np.nan if np.issubdtype(dtype, np.floating) else 0
    comp = None if compressor is None else {"id": compressor, "level": int(level)}
    zarray = {"chunks": [cy, cx], "compressor": comp, "dtype": dtype.str,
              "fill_value": _fill_json(fill), "filters": None, "order": "C",
              "shape": [H, W], "zarr_format": 2}
    zattrs_var = {"_ARRAY_DIMENSIONS": list(dims), **(attrs or {})}
    if os.path.isdir(path):
        shutil.rmtree(path)
    elif os.path.exists(path):
        os.remove(path)
    os.makedirs(os.path.join(path, var))
    docs = {".zgroup": {"zarr_format": 2}, ".zattrs": {},
            f"{var}/.zarray": zarray, f"{var}/.zattrs": zattrs_var}
    for rel, doc in docs.items():
        with open(os.path.join(path, rel), "w", encoding="utf-8") as f:
            json.dump(doc, f, indent=4, sort_keys=True, allow_nan=False)
    with open(os.path.join(path, ".zmetadata"), "w", encoding="utf-8") as f:
        json.dump({"metadata": docs, "zarr_consolidated_format": 1}, f, indent=4,
                  sort_keys=True, allow_nan=False)

    def one(idx):
        i, j = idx
        blk = host[i * cy:(i + 1) * cy, j * cx:(j + 1) * cx]
        if blk.shape != (cy, cx):   # edge chunk: stored at the full chunk shape
            full = np.full((cy, cx), fill, dtype=dtype)
            full[:blk.shape[0], :blk.shape[1]] = blk
            blk = full
        data = _encode(np.ascontiguousarray(blk).tobytes(), comp)
        with open(os.path.join(path, var, f"{i}.{j}"), "wb") as f:
            f.write(data)

    idxs = [(i, j) for i in range(int(math.ceil(H / cy))) for j in range(int(math.ceil(W / cx)))]
    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(one, idxs))

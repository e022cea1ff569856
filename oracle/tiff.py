"""CPU oracle of the COG output -- TEST INFRASTRUCTURE ONLY.

A pure-Python reader of little-endian TIFF / BigTIFF restated from the TIFF
6.0 and BigTIFF specifications: IFD chains, tiled or stripped single-band
images, ZSTD (libzstd through ctypes) / DEFLATE (zlib) / no compression,
predictor 2 (horizontal differencing) and predictor 3 (Adobe Photoshop TIFF
Technical Note 3: bytes regrouped by significance, most significant first,
then byte-wise differencing), GeoTIFF georeferencing tags.  Plus the AVERAGE
overview pyramid the reference asks GDAL for (OVERVIEW_RESAMPLING=AVERAGE,
core/dask_processor.py:201-228), restated as the float64 mean of the valid
pixels of each 2 x 2 block, level by level.

Pinning (tests/test_oracle_pins.py): the reader decodes files that Pillow's
libtiff writes (ZSTD + predictor 3 float32, DEFLATE + predictor 2 uint8) to
the exact input.  Parity unpinned: GDAL's own AVERAGE overview arithmetic
(GDAL is absent here); the restatement is the mean of valid pixels.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import struct
import zlib
from typing import Dict, List, Optional

import numpy as np

_ZSTD = None


def _zstd():
    global _ZSTD
    if _ZSTD is None:
        name = ctypes.util.find_library("zstd") or "libzstd.so.1"
        lib = ctypes.CDLL(name)
        lib.ZSTD_decompress.restype = ctypes.c_size_t
        lib.ZSTD_decompress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                        ctypes.c_size_t]
        lib.ZSTD_isError.restype = ctypes.c_uint
        lib.ZSTD_isError.argtypes = [ctypes.c_size_t]
        _ZSTD = lib
    return _ZSTD


def zstd_decompress(data: bytes, size: int) -> bytes:
    lib = _zstd()
    out = ctypes.create_string_buffer(size)
    src = ctypes.create_string_buffer(data, len(data))
    r = lib.ZSTD_decompress(out, size, src, len(data))
    if lib.ZSTD_isError(r) or r != size:
        raise ValueError("ZSTD decompression failed")
    return out.raw


_TYPES = {1: ("B", 1), 2: ("s", 1), 3: ("H", 2), 4: ("I", 4), 5: ("II", 8), 6: ("b", 1),
          7: ("B", 1), 8: ("h", 2), 9: ("i", 4), 11: ("f", 4), 12: ("d", 8), 16: ("Q", 8),
          17: ("q", 8)}


def parse(path: str) -> Dict:
    """{'big': bool, 'ifds': [{'offset': int, 'tags': {tag: values}}]}."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:2] != b"II":
        raise ValueError("not little-endian")
    magic = struct.unpack_from("<H", buf, 2)[0]
    big = magic == 43
    if magic not in (42, 43):
        raise ValueError("not a TIFF")
    off = struct.unpack_from("<Q", buf, 8)[0] if big else struct.unpack_from("<I", buf, 4)[0]
    ifds = []
    while off:
        n = struct.unpack_from("<Q" if big else "<H", buf, off)[0]
        p = off + (8 if big else 2)
        tags = {}
        for k in range(n):
            e = p + k * (20 if big else 12)
            tag, ty = struct.unpack_from("<HH", buf, e)
            cnt = struct.unpack_from("<Q" if big else "<I", buf, e + 4)[0]
            fmt, sz = _TYPES[ty]
            nbytes = cnt * sz
            vo = e + (12 if big else 8)
            if nbytes > (8 if big else 4):
                vo = struct.unpack_from("<Q" if big else "<I", buf, vo)[0]
            raw = buf[vo:vo + nbytes]
            if ty == 2:
                tags[tag] = raw.split(b"\0")[0].decode()
            elif ty == 5:
                v = struct.unpack("<" + "I" * (2 * cnt), raw)
                tags[tag] = [v[2 * i] / v[2 * i + 1] for i in range(cnt)]
            else:
                tags[tag] = list(struct.unpack("<" + fmt * cnt, raw))
        ifds.append({"offset": off, "entries": n, "tags": tags})
        nx = p + n * (20 if big else 12)
        off = struct.unpack_from("<Q" if big else "<I", buf, nx)[0]
    return {"big": big, "ifds": ifds, "size": len(buf)}


def undo_predictor3(row: np.ndarray, width: int, bps: int = 4) -> np.ndarray:
    """Floating-point predictor of one row (width samples of bps bytes)."""
    b = np.cumsum(row.astype(np.uint32)) & 0xFF          # undo byte differencing
    planes = b.astype(np.uint8).reshape(bps, width)      # most significant plane first
    out = np.empty((width, bps), np.uint8)
    for k in range(bps):
        out[:, bps - 1 - k] = planes[k]                  # little-endian sample bytes
    return out.reshape(-1)


def _dtype(bps: int, fmt: int):
    return {(32, 3): np.float32, (16, 2): np.int16, (8, 1): np.uint8, (16, 1): np.uint16,
            (32, 2): np.int32}[(bps, fmt)]


def decode(path: str, level: int = 0, parsed: Optional[Dict] = None) -> np.ndarray:
    """The full image of IFD `level`."""
    P = parsed or parse(path)
    t = P["ifds"][level]["tags"]
    w, h = t[256][0], t[257][0]
    bps = t[258][0]
    fmt = t.get(339, [1])[0]
    comp = t.get(259, [1])[0]
    pred = t.get(317, [1])[0]
    dt = np.dtype(_dtype(bps, fmt))
    if 322 in t:
        tw, th = t[322][0], t[323][0]
        offs, cnts = t[324], t[325]
    else:                                                # strips: tiles of the full width
        tw, th = w, t.get(278, [h])[0]
        offs, cnts = t[273], t[279]
    ntx = (w + tw - 1) // tw
    img = np.zeros((((h + th - 1) // th) * th, ntx * tw), dt)
    with open(path, "rb") as f:
        data = f.read()
    for i, (o, c) in enumerate(zip(offs, cnts)):
        raw = data[o:o + c]
        size = tw * th * dt.itemsize
        if comp == 50000:
            raw = zstd_decompress(raw, size)
        elif comp in (8, 32946):
            raw = zlib.decompress(raw)
        elif comp != 1:
            raise ValueError(f"compression {comp}")
        rows_here = len(raw) // (tw * dt.itemsize)
        a = np.frombuffer(raw, np.uint8)[:rows_here * tw * dt.itemsize].reshape(rows_here, -1)
        if pred == 3:
            a = np.stack([undo_predictor3(r, tw, dt.itemsize) for r in a])
        vals = a.copy().view(dt).reshape(rows_here, tw)
        if pred == 2:
            vals = np.cumsum(vals.astype(np.int64), axis=1).astype(dt)
        ty, tx = divmod(i, ntx)
        img[ty * th:ty * th + rows_here, tx * tw:(tx + 1) * tw] = vals
    return img[:h, :w]


def overview_average(a: np.ndarray, nodata: Optional[float] = None) -> np.ndarray:
    """One AVERAGE level: the float64 mean of the valid pixels of each 2 x 2
    block (clipped at the far edges); integer types round to nearest; a block
    with none takes the NoData value (NaN for float without one, else 0)."""
    h, w = a.shape
    hd, wd = (h + 1) // 2, (w + 1) // 2
    pad = np.zeros((2 * hd, 2 * wd), np.float64)
    ok = np.zeros((2 * hd, 2 * wd), bool)
    pad[:h, :w] = a
    v = np.ones((h, w), bool)
    if a.dtype.kind == "f":
        v &= ~np.isnan(a)
    if nodata is not None and not np.isnan(nodata):
        v &= a != a.dtype.type(nodata)
    ok[:h, :w] = v
    pad = np.where(ok, pad, 0.0)
    s = pad.reshape(hd, 2, wd, 2)
    o = ok.reshape(hd, 2, wd, 2)
    # float64 sum in the kernel's order: (0,0), (0,1), (1,0), (1,1)
    tot = ((s[:, 0, :, 0] + s[:, 0, :, 1]) + s[:, 1, :, 0]) + s[:, 1, :, 1]
    n = o.sum(axis=(1, 3))
    with np.errstate(invalid="ignore", divide="ignore"):
        m = tot / np.maximum(n, 1)
    if a.dtype.kind == "f":
        fill = np.float64(nodata) if nodata is not None else np.nan
        return np.where(n > 0, m, fill).astype(a.dtype)
    fill = float(nodata) if nodata is not None else 0.0
    return np.where(n > 0, np.floor(m + 0.5), fill).astype(a.dtype)


def overview_levels(h: int, w: int, block: int = 512, count: int = 8) -> List:
    shapes = [(h, w)]
    while len(shapes) - 1 < count and max(shapes[-1]) > block:
        hh, ww = shapes[-1]
        shapes.append(((hh + 1) // 2, (ww + 1) // 2))
    return shapes

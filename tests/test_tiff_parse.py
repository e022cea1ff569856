"""The COG reader's TIFF / BigTIFF IFD parser against malformed files: built
with the host address and undefined-behaviour sanitizers (g++, no GPU code)
and run over truncations, oversized counts, offsets past the end, huge or zero
sizes, IFD loops and random byte flips of two valid files; the same files go
through the library's fsg_cog_open.  Every file must give OK or a clean error,
and no sanitizer report."""
import ctypes
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fujishadergpu_amd", "csrc")


def classic_tiff(w=16, h=16, tw=16, th=16, next_ifd=0):
    """A valid little-endian tiled float32 TIFF, uncompressed."""
    ntx, nty = (w + tw - 1) // tw, (h + th - 1) // th
    nt = ntx * nty
    tile = np.arange(tw * th, dtype="<f4").tobytes()
    data_off = 8
    data = tile * nt
    offs = [data_off + k * len(tile) for k in range(nt)]
    extra_off = data_off + len(data)
    extra = b""
    entries = []

    def ent(tag, typ, vals, fmt):
        nonlocal extra
        raw = b"".join(struct.pack("<" + fmt, v) for v in vals)
        if len(raw) <= 4:
            entries.append(struct.pack("<HHI", tag, typ, len(vals)) + raw.ljust(4, b"\0"))
        else:
            off = extra_off + len(extra)
            extra += raw
            entries.append(struct.pack("<HHII", tag, typ, len(vals), off))

    ent(256, 4, [w], "I")
    ent(257, 4, [h], "I")
    ent(258, 3, [32], "H")
    ent(259, 3, [1], "H")
    ent(277, 3, [1], "H")
    ent(322, 4, [tw], "I")
    ent(323, 4, [th], "I")
    ent(324, 4, offs, "I")
    ent(325, 4, [len(tile)] * nt, "I")
    ent(339, 3, [3], "H")
    ifd_off = extra_off + len(extra)
    ifd = struct.pack("<H", len(entries)) + b"".join(entries) + struct.pack("<I", next_ifd)
    return b"II" + struct.pack("<HI", 42, ifd_off) + data + extra + ifd, ifd_off


def big_tiff(w=20, h=12, tw=16, th=16):
    """A valid BigTIFF, tiled uint8, uncompressed, 2 x 1 tiles."""
    ntx, nty = (w + tw - 1) // tw, (h + th - 1) // th
    nt = ntx * nty
    tile = bytes(range(256)) * (tw * th // 256)
    data_off = 16
    data = tile * nt
    offs = [data_off + k * len(tile) for k in range(nt)]
    entries = [(256, 16, [w]), (257, 16, [h]), (258, 3, [8]), (259, 3, [1]), (277, 3, [1]),
               (322, 16, [tw]), (323, 16, [th]), (324, 16, offs), (325, 16, [len(tile)] * nt),
               (339, 3, [1])]
    extra_off = data_off + len(data)
    extra = b""
    raw_entries = []
    for tag, typ, vals in entries:
        fmt = "Q" if typ == 16 else "H"
        raw = b"".join(struct.pack("<" + fmt, v) for v in vals)
        if len(raw) <= 8:
            raw_entries.append(struct.pack("<HHQ", tag, typ, len(vals)) + raw.ljust(8, b"\0"))
        else:
            raw_entries.append(struct.pack("<HHQQ", tag, typ, len(vals),
                                           extra_off + len(extra)))
            extra += raw
    ifd_off = extra_off + len(extra)
    ifd = struct.pack("<Q", len(raw_entries)) + b"".join(raw_entries) + struct.pack("<Q", 0)
    return b"II" + struct.pack("<HHHQ", 43, 8, 0, ifd_off) + data + extra + ifd


def _patch_entry(buf, ifd_off, tag, *, count=None, value=None, typ=None):
    """Classic TIFF: rewrite the entry of ``tag``."""
    b = bytearray(buf)
    n = struct.unpack_from("<H", b, ifd_off)[0]
    for k in range(n):
        at = ifd_off + 2 + 12 * k
        t, ty, c, v = struct.unpack_from("<HHII", b, at)
        if t == tag:
            struct.pack_into("<HHII", b, at, t, ty if typ is None else typ,
                             c if count is None else count, v if value is None else value)
            return bytes(b)
    raise KeyError(tag)


def malformed_files(tmp):
    good, ifd = classic_tiff()
    big = big_tiff()
    files = {"good": good, "big": big, "multi": classic_tiff(w=40, h=33)[0]}
    for cut in list(range(0, 40, 3)) + [len(good) - k for k in (1, 2, 5, 13, 30)]:
        files[f"trunc{cut}"] = good[:cut]
    for cut in (20, 60, len(big) - 3, len(big) - 40):
        files[f"bigtrunc{cut}"] = big[:cut]
    bad = bytearray(good)
    struct.pack_into("<H", bad, ifd, 0xFFFF)
    files["entries65535"] = bytes(bad)
    files["width0"] = _patch_entry(good, ifd, 256, value=0)
    files["width2g"] = _patch_entry(good, ifd, 256, value=0xFFFFFFFF)
    files["height2g"] = _patch_entry(good, ifd, 257, value=0x80000000)
    files["tilew1"] = _patch_entry(good, ifd, 322, value=1)      # tables no longer match
    files["tilewhuge"] = _patch_entry(good, ifd, 322, value=1 << 29)
    files["offsets_count_huge"] = _patch_entry(good, ifd, 324, count=1 << 28, value=8)
    files["offsets_count_max"] = _patch_entry(good, ifd, 324, count=0xFFFFFFFF, value=8)
    files["tile_past_eof"] = _patch_entry(good, ifd, 324, value=len(good) + 4096)
    files["tile_bytes_huge"] = _patch_entry(good, ifd, 325, value=0xFFFFFFF0)
    files["bad_type"] = _patch_entry(good, ifd, 258, typ=99)
    files["value_past_eof"] = _patch_entry(good, ifd, 258, count=3, value=len(good) + 100)
    files["compression_jpeg"] = _patch_entry(good, ifd, 259, value=7)
    loop = bytearray(good)
    struct.pack_into("<I", loop, len(good) - 4, ifd)           # next IFD = itself
    files["ifd_loop"] = bytes(loop)
    files["ifd_past_eof"] = good[:4] + struct.pack("<I", len(good) + 10) + good[8:]
    rng = np.random.default_rng(11)
    for k in range(150):
        src = good if k % 2 == 0 else big
        b = bytearray(src)
        for _ in range(int(rng.integers(1, 6))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        files[f"flip{k}"] = bytes(b)
    paths = {}
    for name, data in files.items():
        p = os.path.join(tmp, name + ".tif")
        with open(p, "wb") as f:
            f.write(data)
        paths[name] = p
    return paths


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    return malformed_files(str(tmp_path_factory.mktemp("tiff")))


def test_ifd_parser_under_host_sanitizers(files, tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "tiff_ifd_check")
    src = os.path.join(ROOT, "tests", "cpp", "tiff_ifd_check.cpp")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-I", CSRC, src,
           "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    names = list(files)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=87")
    r = subprocess.run([exe] + [files[n] for n in names], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    out = dict(zip(names, r.stdout.strip().split("\n")))
    assert out["good"] == "OK 1 16 16 16 16"
    assert out["big"] == "OK 1 20 12 16 16"
    assert out["multi"] == "OK 1 40 33 16 16"
    for n in ("entries65535", "width0", "width2g", "height2g", "tilew1", "tilewhuge",
              "offsets_count_huge", "offsets_count_max", "tile_past_eof", "tile_bytes_huge",
              "bad_type", "value_past_eof", "compression_jpeg", "ifd_loop", "ifd_past_eof"):
        assert out[n].startswith("ERR"), (n, out[n])
    for n in names:
        if n.startswith("trunc") or n.startswith("bigtrunc"):
            assert out[n].startswith("ERR"), (n, out[n])
        assert out[n].startswith("OK") or out[n].startswith("ERR"), (n, out[n])


def test_library_open_refuses_malformed_files_cleanly(files):
    from fujishadergpu_amd import _native as N
    L = N.lib()
    for name, path in files.items():
        h = ctypes.c_void_p()
        rc = L.fsg_cog_open(path.encode(), ctypes.byref(h))
        if name in ("good", "big", "multi"):
            assert rc == 0, (name, L.fsg_last_error())
        if rc == 0:
            L.fsg_cog_close(h)
        else:
            assert not h.value
            assert b"fsg_cog_open" in L.fsg_last_error()

"""Device-level wrappers over libfsg.so for torch CUDA (HIP) tensors.

torch supplies device memory and the current stream; every computation is a
libfsg kernel.  Windows are 2-D float32 tensors whose rows may be strided
(a view into a larger raster), i.e. ``stride(1) == 1``.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N

_DTYPE_CODE = {"float32": N.OUT_F32, "int16": N.OUT_I16, "uint8": N.OUT_U8}
_TORCH_DTYPE = {"float32": torch.float32, "int16": torch.int16, "uint8": torch.uint8}


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _check_window(t: torch.Tensor, name="window") -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on the GPU, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on the GPU (got device {t.device})")
    if t.dtype != torch.float32 or t.dim() != 2:
        raise ValueError(f"{name} must be a 2-D float32 tensor, got {t.dtype} {tuple(t.shape)}")
    if t.shape[0] > 0 and t.shape[1] > 1 and t.stride(1) != 1:
        raise ValueError(f"{name} rows must be contiguous (stride(1) == 1)")


def window(t: torch.Tensor) -> N.fsg_window:
    _check_window(t)
    ld = t.stride(0) if t.shape[0] > 1 else max(1, t.shape[1])
    return N.fsg_window(t.data_ptr(), int(ld), int(t.shape[0]), int(t.shape[1]))


def region(r0: int, c0: int, h: int, w: int) -> N.fsg_region:
    return N.fsg_region(int(r0), int(c0), int(h), int(w))


def full_region(t: torch.Tensor) -> N.fsg_region:
    return region(0, 0, t.shape[0], t.shape[1])


def output(t: torch.Tensor, *, qp: Optional[dict] = None, clip01: bool = False,
           post: int = N.POST_NONE, host_stats: Sequence[float] = (0.0, 0.0),
           dev_stats: Optional[torch.Tensor] = None) -> N.fsg_output:
    """Output descriptor for a (region-shaped) tensor of float32/int16/uint8."""
    if not t.is_cuda or t.dim() != 2:
        raise ValueError("output must be a 2-D GPU tensor")
    if t.shape[0] > 0 and t.shape[1] > 1 and t.stride(1) != 1:
        raise ValueError("output rows must be contiguous")
    code = {torch.float32: N.OUT_F32, torch.int16: N.OUT_I16, torch.uint8: N.OUT_U8}.get(t.dtype)
    if code is None:
        raise ValueError(f"unsupported output dtype {t.dtype}")
    o = N.fsg_output()
    o.data = t.data_ptr()
    o.ld = int(t.stride(0) if t.shape[0] > 1 else max(1, t.shape[1]))
    o.dtype = code
    o.clip01 = int(bool(clip01))
    if code != N.OUT_F32:
        if qp is None:
            raise ValueError("integer output needs quantisation params")
        o.a_coef = float(np.float32(qp["a_coef"]))
        o.b_coef = float(np.float32(qp["b_coef"]))
        o.dn_min, o.dn_max = int(qp["dn_min"]), int(qp["dn_max"])
    o.post_kind = int(post)
    o.host_stats[0] = float(host_stats[0]) if len(host_stats) > 0 else 0.0
    o.host_stats[1] = float(host_stats[1]) if len(host_stats) > 1 else 0.0
    if dev_stats is not None:
        if dev_stats.dtype != torch.float64 or not dev_stats.is_cuda:
            raise ValueError("dev_stats must be a float64 GPU tensor")
        o.dev_stats = dev_stats.data_ptr()
    return o


def empty_out(h: int, w: int, dtype: str = "float32", device=None) -> torch.Tensor:
    return torch.empty((h, w), dtype=_TORCH_DTYPE[dtype], device=device or "cuda")


def _lib():
    return N.lib()


# ------------------------------------------------------------------ stencils

def _rows_ptr(step_x_rows: Optional[torch.Tensor]):
    if step_x_rows is None:
        return None
    if step_x_rows.dtype != torch.float64 or not step_x_rows.is_contiguous():
        raise ValueError("step_x_rows must be a contiguous float64 tensor")
    return step_x_rows.data_ptr()


def shade_params(*, step_x: float, step_y: float, sign_x: float = 1.0, sign_y: float = 1.0,
                 z_factor: float = 1.0, azimuth: float = 315.0, altitude: float = 45.0,
                 unit: str = "degree", nan_free: bool = False,
                 step_x_rows: Optional[torch.Tensor] = None) -> N.fsg_shade_params:
    """step_x_rows: per window row signed x steps (float64, on the device), kept
    alive by the caller while the parameters are in use."""
    return N.fsg_shade_params(float(step_x), float(step_y), float(sign_x), float(sign_y),
                              float(z_factor), float(azimuth), float(altitude),
                              {"degree": 0, "percent": 1}.get(unit, 2), 0 if nan_free else 1,
                              _rows_ptr(step_x_rows))


def curvature_params(*, step_x: float, step_y: float, curvature_type: str,
                     nan_free: bool = False,
                     step_x_rows: Optional[torch.Tensor] = None) -> N.fsg_curvature_params:
    ct = {"mean": 0, "gaussian": 1, "planform": 2}.get(curvature_type, 3)
    return N.fsg_curvature_params(float(step_x), float(step_y), ct, 0 if nan_free else 1,
                                  _rows_ptr(step_x_rows))


def hillshade_slope(win_t: torch.Tensor, reg: N.fsg_region, *, step_x: float, step_y: float,
                    sign_x: float = 1.0, sign_y: float = 1.0, z_factor: float = 1.0,
                    azimuth: float = 315.0, altitude: float = 45.0, unit: str = "degree",
                    hs_out: Optional[N.fsg_output] = None,
                    sl_out: Optional[N.fsg_output] = None, nan_free: bool = False,
                    step_x_rows: Optional[torch.Tensor] = None) -> None:
    p = shade_params(step_x=step_x, step_y=step_y, sign_x=sign_x, sign_y=sign_y,
                     z_factor=z_factor, azimuth=azimuth, altitude=altitude, unit=unit,
                     nan_free=nan_free, step_x_rows=step_x_rows)
    w = window(win_t)
    N.check(_lib().fsg_hillshade_slope(C.byref(w), C.byref(reg), C.byref(p),
                                       C.byref(hs_out) if hs_out is not None else None,
                                       C.byref(sl_out) if sl_out is not None else None,
                                       _stream()))


def curvature(win_t, reg, *, step_x: float, step_y: float, curvature_type: str,
              out: N.fsg_output, nan_free: bool = False,
              step_x_rows: Optional[torch.Tensor] = None) -> None:
    p = curvature_params(step_x=step_x, step_y=step_y, curvature_type=curvature_type,
                         nan_free=nan_free, step_x_rows=step_x_rows)
    w = window(win_t)
    N.check(_lib().fsg_curvature(C.byref(w), C.byref(reg), C.byref(p), C.byref(out), _stream()))


# ------------------------------------------------------------------ rays

def openness_params(table, openness_type: str, raw_degrees: bool = False):
    """(params, keep-alive array) for a ray table: list over directions of
    lists over distances of (ox, oy, phys)|None."""
    n_dirs = len(table)
    n_dist = max((len(r) for r in table), default=0)
    arr = (N.fsg_ray_sample * max(1, n_dirs * n_dist))()
    for d, row in enumerate(table):
        for k in range(n_dist):
            e = arr[d * n_dist + k]
            entry = row[k] if k < len(row) else None
            if entry is None:
                e.skip = 1
            else:
                e.dx, e.dy, e.phys = int(entry[0]), int(entry[1]), float(entry[2])
    p = N.fsg_openness_params(0 if openness_type == "positive" else 1, n_dirs, n_dist,
                              1 if raw_degrees else 0, C.cast(arr, C.POINTER(N.fsg_ray_sample)))
    return p, arr


def pad_reflect(win_t: torch.Tensor, top: int, bottom: int, left: int,
                right: int) -> torch.Tensor:
    """The window padded per side with the symmetric ('reflect' in dask) rule."""
    h, w = win_t.shape
    out = torch.empty((h + top + bottom, w + left + right), dtype=torch.float32,
                      device=win_t.device)
    wd = window(win_t)
    N.check(_lib().fsg_pad_reflect(C.byref(wd), int(top), int(bottom), int(left), int(right),
                                   out.data_ptr(), out.stride(0), _stream()))
    return out


def openness(win_t, reg, *, table, openness_type: str, out: N.fsg_output,
             raw_degrees: bool = False) -> None:
    p, _arr = openness_params(table, openness_type, raw_degrees)
    w = window(win_t)
    N.check(_lib().fsg_openness(C.byref(w), C.byref(reg), C.byref(p), C.byref(out), _stream()))


def openness_points(win_t, points: torch.Tensor, *, table, openness_type: str,
                    raw_degrees: bool = True) -> torch.Tensor:
    """Openness at window pixels points (int32 [n, 2] = row, col, on the GPU)."""
    if points.dtype != torch.int32 or points.dim() != 2 or points.shape[1] != 2:
        raise ValueError("points must be an int32 [n, 2] tensor")
    pts = points.contiguous()
    out = torch.empty(pts.shape[0], dtype=torch.float32, device=win_t.device)
    p, _arr = openness_params(table, openness_type, raw_degrees)
    w = window(win_t)
    N.check(_lib().fsg_openness_points(C.byref(w), pts.data_ptr(), int(pts.shape[0]),
                                       C.byref(p), out.data_ptr(), _stream()))
    return out


def splitmix64(x: int) -> int:
    """splitmix64 (Steele, Lea & Flood 2014) on Python ints, 64-bit wrap."""
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def judge_segments(h: int, w: int, tile_h: int, tile_w: int, sample_frac: float, seed: int,
                   seg_len: int):
    """The judge's sampled segments, tile by tile: [((y0, x0, th, tw), [(row, col0), ...])]
    (the positions fsg_openness_judge draws)."""
    import math
    out = []
    n_tx = -(-w // tile_w)
    for t in range(-(-h // tile_h) * n_tx):
        y0, x0 = (t // n_tx) * tile_h, (t % n_tx) * tile_w
        th, tw = min(tile_h, h - y0), min(tile_w, w - x0)
        n = max(1, math.ceil(sample_frac * th * tw / seg_len))
        key = splitmix64(seed ^ splitmix64(t))
        segs = []
        for k in range(n):
            hh = splitmix64((key + k) & ((1 << 64) - 1))
            segs.append((y0 + (hh >> 32) % th, x0 + (hh & 0xFFFFFFFF) % (tw - seg_len + 1)))
        out.append(((y0, x0, th, tw), segs))
    return out


def openness_judge(win_t, approx: torch.Tensor, *, tile_h: int, tile_w: int, sample_frac: float,
                   seed: int, seg_len: int, table, openness_type: str = "positive",
                   raw_degrees: bool = True) -> torch.Tensor:
    """Per-tile mean |exact openness - approx| over sampled row segments (a
    float32 [n_tiles] tensor, tiles row-major)."""
    h, w = win_t.shape
    n_tiles = -(-h // tile_h) * -(-w // tile_w)
    out = torch.empty(n_tiles, dtype=torch.float32, device=win_t.device)
    wd = window(win_t)
    nbytes = _lib().fsg_openness_judge_workspace_size(C.byref(wd), tile_h, tile_w)
    ws = torch.empty(max(1, nbytes), dtype=torch.uint8, device=win_t.device)
    p, _arr = openness_params(table, openness_type, raw_degrees)
    N.check(_lib().fsg_openness_judge(C.byref(wd), C.byref(p), approx.data_ptr(), approx.stride(0),
                                      tile_h, tile_w, float(sample_frac), int(seed), int(seg_len),
                                      out.data_ptr(), ws.data_ptr(), nbytes, _stream()))
    return out


def ambient_occlusion(win_t, reg, *, table, intensity: float, pad_d: int,
                      out: N.fsg_output) -> None:
    """table: list of (dx, dy, phys, dist_factor)|None in reference order."""
    arr = (N.fsg_ao_sample * max(1, len(table)))()
    for i, entry in enumerate(table):
        if entry is None:
            arr[i].skip = 1
        else:
            arr[i].dx, arr[i].dy = int(entry[0]), int(entry[1])
            arr[i].phys, arr[i].dist_factor = float(entry[2]), float(entry[3])
    p = N.fsg_ao_params(len(table), int(pad_d), float(intensity),
                        C.cast(arr, C.POINTER(N.fsg_ao_sample)))
    w = window(win_t)
    L = _lib()
    nbytes = L.fsg_ambient_occlusion_workspace_size(C.byref(w), C.byref(reg))
    ws = torch.empty(max(1, nbytes), dtype=torch.uint8, device=win_t.device)
    N.check(L.fsg_ambient_occlusion(C.byref(w), C.byref(reg), C.byref(p), C.byref(out),
                                    ws.data_ptr(), nbytes, _stream()))


# ------------------------------------------------------------------ topousm

def topousm_params(radii: Sequence[int], weights: Sequence[float], ds: Sequence[int],
                   field: Optional[torch.Tensor] = None, full_shape=None, win_origin=(0, 0),
                   w_large: float = 0.0) -> N.fsg_topousm_params:
    if len(radii) > N.MAX_RADII:
        raise ValueError(f"at most {N.MAX_RADII} radii are supported")
    if len(weights) != len(radii):
        raise ValueError(
            f"Length of weights ({len(weights)}) must match length of radii ({len(radii)})")
    p = N.fsg_topousm_params()
    p.nan_hint = 1
    p.n_radii = len(radii)
    for i, (r, wt, d) in enumerate(zip(radii, weights, ds)):
        p.radii[i] = int(r)
        p.weights[i] = float(np.float32(wt))
        p.ds[i] = int(d)
    if field is not None:
        _check_window(field, "field")
        if not field.is_contiguous():
            raise ValueError("field must be contiguous")
        p.field = field.data_ptr()
        p.field_h, p.field_w = int(field.shape[0]), int(field.shape[1])
        p.full_h, p.full_w = int(full_shape[0]), int(full_shape[1])
        p.win_r0, p.win_c0 = int(win_origin[0]), int(win_origin[1])
        p.w_large = float(np.float32(w_large))
    return p


_SIDE_STREAMS = {}


class StreamRing:
    """Independent launches spread round-robin over the caller's current stream
    and n - 1 side streams (default priority, one set per device, reused), so
    that one launch's last workgroups overlap the next launch's first.  The
    side streams start after everything already queued on the caller's
    stream; join() makes the caller's stream wait for all of them.  Work on
    stream k must use resources of its own (workspaces) or read-only ones."""

    def __init__(self, n: int = 2):
        cur = torch.cuda.current_stream()
        side = _SIDE_STREAMS.setdefault(cur.device, [])
        while len(side) < n - 1:
            side.append(torch.cuda.Stream(device=cur.device))
        self.streams = [cur] + side[:n - 1]
        ev = torch.cuda.Event()
        ev.record(cur)
        for st in self.streams[1:]:
            st.wait_event(ev)

    def __len__(self):
        return len(self.streams)

    def stream(self, k: int):
        return self.streams[k % len(self.streams)]

    def join(self) -> None:
        cur = self.streams[0]
        for st in self.streams[1:]:
            ev = torch.cuda.Event()
            ev.record(st)
            cur.wait_event(ev)


class Workspace:
    """Grow-only device scratch buffer (one per caller/stream)."""

    def __init__(self):
        self.buf = None

    def get(self, nbytes: int, device) -> torch.Tensor:
        if self.buf is None or self.buf.numel() < nbytes or self.buf.device != torch.device(device):
            self.buf = torch.empty(max(256, int(nbytes)), dtype=torch.uint8, device=device)
        return self.buf


def topousm(win_t, reg, params: N.fsg_topousm_params, out: N.fsg_output,
            ws: Optional[Workspace] = None, out2: Optional[N.fsg_output] = None) -> None:
    w = window(win_t)
    L = _lib()
    nbytes = L.fsg_topousm_workspace_size(C.byref(w), C.byref(params))
    buf = (ws or Workspace()).get(nbytes, win_t.device)
    N.check(L.fsg_topousm(C.byref(w), C.byref(reg), C.byref(params), C.byref(out),
                          C.byref(out2) if out2 is not None else None,
                          buf.data_ptr(), buf.numel(), _stream()))


def set_topousm_variant(variant: int) -> None:
    """Main-pass kernel choice of fsg_topousm for this process: 0 (default)
    the single-wavefront kernel wherever it applies, 1 the workgroup kernels
    everywhere, 8..1024 the single-wavefront kernel with that many rows per
    block (identical results; for benchmarks, launch-shape sweeps and
    equality tests)."""
    N.check(_lib().fsg_topousm_set_variant(int(variant)))


def set_stencil_split_min(px: int) -> int:
    """Smallest interior (pixels) that hillshade/slope and curvature give to
    the row-walking interior kernels (the rest of the region, and smaller
    interiors, take the general kernel); px < 0 restores the default.  A
    process-wide setting for equality tests and tuning (identical results);
    returns the previous value."""
    return int(_lib().fsg_stencil_set_split_min(int(px)))


def topousm_noalloc(win_t, reg, params: N.fsg_topousm_params, out: N.fsg_output,
                    out2: Optional[N.fsg_output] = None) -> None:
    """topousm without a workspace: the library's fixed per-GPU scratch array
    holds the coarse grids (no allocation; overlapping calls on one GPU race)."""
    w = window(win_t)
    N.check(_lib().fsg_topousm_noalloc(C.byref(w), C.byref(reg), C.byref(params), C.byref(out),
                                       C.byref(out2) if out2 is not None else None, _stream()))


def block_mean(src_t: torch.Tensor, factor: int) -> torch.Tensor:
    h, w = src_t.shape
    oh, ow = max(1, -(-h // factor)), max(1, -(-w // factor))
    out = torch.empty((oh, ow), dtype=torch.float32, device=src_t.device)
    win = window(src_t)
    N.check(_lib().fsg_block_mean(C.byref(win), int(factor), out.data_ptr(), ow, _stream()))
    return out


def overview_average(src_t: torch.Tensor, oh: int, ow: int) -> torch.Tensor:
    """Area-weighted NaN-aware average of the raster onto an oh x ow grid
    (Resampling.average at any ratio; float64 sums and division)."""
    out = torch.empty((int(oh), int(ow)), dtype=torch.float32, device=src_t.device)
    win = window(src_t)
    N.check(_lib().fsg_overview_average(C.byref(win), int(oh), int(ow), out.data_ptr(),
                                        int(ow), _stream()))
    return out


def topousm_coarse_means(win_t: torch.Tensor, ds: int, radii_coarse: Sequence[int]) -> torch.Tensor:
    """Coarse means of one decimation level over a whole window (typically the
    raster): a (len(radii_coarse), ceil(h/ds), ceil(w/ds)) float32 tensor."""
    h, w = win_t.shape
    hc, wc = -(-h // ds), -(-w // ds)
    n = len(radii_coarse)
    out = torch.empty((n, hc, wc), dtype=torch.float32, device=win_t.device)
    L = _lib()
    win = window(win_t)
    nbytes = L.fsg_topousm_coarse_means_workspace_size(C.byref(win), int(ds))
    ws = torch.empty(max(1, nbytes), dtype=torch.uint8, device=win_t.device)
    rr = (C.c_int32 * n)(*[int(r) for r in radii_coarse])
    outs = (C.c_void_p * n)(*[out[i].data_ptr() for i in range(n)])
    N.check(L.fsg_topousm_coarse_means(C.byref(win), int(ds), n, rr, outs, wc, ws.data_ptr(),
                                       nbytes, _stream()))
    return out


def raster_scan(raster: torch.Tensor, factors: Sequence[int]):
    """One read of the raster: (device int64 NaN count, [NaN-aware block mean
    per factor]); factors are 4, 8 or 16, at most three."""
    h, w = raster.shape
    factors = [int(f) for f in factors]
    outs = [torch.empty((-(-h // f), -(-w // f)), dtype=torch.float32, device=raster.device)
            for f in factors]
    arr = (N.fsg_scan_out * max(1, len(factors)))()
    for k, (f, o) in enumerate(zip(factors, outs)):
        arr[k] = N.fsg_scan_out(f, o.data_ptr(), int(o.stride(0)))
    cnt = torch.zeros(1, dtype=torch.int64, device=raster.device)
    win = window(raster)
    N.check(_lib().fsg_raster_scan(C.byref(win), len(factors), arr, cnt.data_ptr(), _stream()))
    return cnt, outs


def topousm_coarse_means_batch(jobs: Sequence[Tuple[torch.Tensor, int, Sequence[int]]]
                               ) -> List[torch.Tensor]:
    """topousm_coarse_means for many NaN-free windows in a few launches.

    jobs: (window tensor, ds in {1, 4, 8, 16}, coarse radii >= 2, at most 4);
    returns one (len(radii), ceil(h/ds), ceil(w/ds)) float32 tensor per job."""
    n = len(jobs)
    if n == 0:
        return []
    arr = (N.fsg_coarse_job * n)()
    outs = []
    for k, (win_t, ds, radii) in enumerate(jobs):
        radii = [int(r) for r in radii]
        if not 1 <= len(radii) <= N.COARSE_JOB_MAX_RADII:
            raise ValueError("1 to 4 coarse radii per job")
        h, w = win_t.shape
        hc, wc = -(-h // int(ds)), -(-w // int(ds))
        o = torch.empty((len(radii), hc, wc), dtype=torch.float32, device=win_t.device)
        outs.append(o)
        J = arr[k]
        J.win = window(win_t)
        J.ds = int(ds)
        J.n_radii = len(radii)
        for i, r in enumerate(radii):
            J.radii[i] = r
            J.outs[i] = o[i].data_ptr()
        J.ld_out = wc
    L = _lib()
    nbytes = int(L.fsg_topousm_coarse_means_batch_workspace_size(arr, n))
    if nbytes <= 0:
        raise RuntimeError("fsg_topousm_coarse_means_batch_workspace_size failed")
    ws = torch.empty(nbytes, dtype=torch.uint8, device=jobs[0][0].device)
    N.check(L.fsg_topousm_coarse_means_batch(arr, n, ws.data_ptr(), nbytes, _stream()))
    return outs


def topousm_coarse_field(coarse_t: torch.Tensor, radii_coarse: Sequence[int],
                         weights: Sequence[float]) -> torch.Tensor:
    h, w = coarse_t.shape
    out = torch.empty((h, w), dtype=torch.float32, device=coarse_t.device)
    L = _lib()
    nbytes = L.fsg_topousm_field_workspace_size(h, w)
    ws = torch.empty(max(1, nbytes), dtype=torch.uint8, device=coarse_t.device)
    rr = (C.c_int32 * len(radii_coarse))(*[int(r) for r in radii_coarse])
    ww = (C.c_float * len(weights))(*[float(np.float32(x)) for x in weights])
    win = window(coarse_t)
    N.check(L.fsg_topousm_coarse_field(C.byref(win), len(radii_coarse), rr, ww, out.data_ptr(),
                                       w, ws.data_ptr(), nbytes, _stream()))
    return out


def topousm_field_from_means(means: Sequence[torch.Tensor], weights: Sequence[float]
                             ) -> torch.Tensor:
    """sum_i f32(w_i) * means[i] with topousm_coarse_field's arithmetic."""
    n = len(means)
    h, w = means[0].shape
    out = torch.empty((h, w), dtype=torch.float32, device=means[0].device)
    ptrs = (C.c_void_p * n)(*[m.data_ptr() for m in means])
    lds = (C.c_int64 * n)(*[int(m.stride(0)) for m in means])
    ww = (C.c_float * n)(*[float(np.float32(x)) for x in weights])
    N.check(_lib().fsg_topousm_field_from_means(ptrs, lds, ww, n, h, w, out.data_ptr(), w,
                                                _stream()))
    return out


# ------------------------------------------------------------------ stats

PCT_ABS, PCT_FINITE = 1, 2


def percentile(values: torch.Tensor, qs: Sequence[float], flags: int):
    """Exact percentiles on the GPU.  Returns device tensors (pct[nq] f64,
    n_valid[1] i64, moments[2] f64) without synchronising."""
    if values.dtype != torch.float32 or not values.is_cuda:
        raise ValueError("values must be a float32 GPU tensor")
    v = values.reshape(-1)
    if not v.is_contiguous():
        v = v.contiguous()
    dev = values.device
    q = torch.tensor(list(qs), dtype=torch.float64, device=dev)
    out = torch.empty(len(qs), dtype=torch.float64, device=dev)
    nval = torch.empty(1, dtype=torch.int64, device=dev)
    mom = torch.empty(2, dtype=torch.float64, device=dev)
    L = _lib()
    nbytes = L.fsg_percentile_workspace_size(v.numel())
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    N.check(L.fsg_percentile(v.data_ptr(), v.numel(), int(flags),
                             C.cast(q.data_ptr(), C.POINTER(C.c_double)), len(qs),
                             out.data_ptr(), nval.data_ptr(), mom.data_ptr(), ws.data_ptr(),
                             nbytes, _stream()))
    # keep q alive until the kernels that read it have run
    out._fsg_keep = (q, ws)
    return out, nval, mom


class PercentileBuffers:
    """Everything fsg_percentile reads and writes for n values and the
    percentiles qs, allocated up front, so that ``run`` neither allocates
    nor copies from the host (it can be captured in a graph)."""

    def __init__(self, n: int, qs: Sequence[float], device):
        self.n = int(n)
        self.q = torch.tensor(list(qs), dtype=torch.float64, device=device)
        self.pct = torch.empty(len(qs), dtype=torch.float64, device=device)
        self.nval = torch.empty(1, dtype=torch.int64, device=device)
        self.mom = torch.empty(2, dtype=torch.float64, device=device)
        self.nbytes = int(_lib().fsg_percentile_workspace_size(self.n))
        self.ws = torch.empty(self.nbytes, dtype=torch.uint8, device=device)
        self.stats = torch.empty(2, dtype=torch.float64, device=device)

    def run(self, values: torch.Tensor, flags: int, kind: int) -> torch.Tensor:
        """The percentiles of ``values`` (n contiguous float32) and their
        finalised statistics (fsg_finalize_stats of ``kind``) in ``stats``."""
        if values.dtype != torch.float32 or values.numel() != self.n or not values.is_contiguous():
            raise ValueError(f"values must be {self.n} contiguous float32 values")
        L = _lib()
        N.check(L.fsg_percentile(values.data_ptr(), self.n, int(flags),
                                 C.cast(self.q.data_ptr(), C.POINTER(C.c_double)),
                                 int(self.q.numel()), self.pct.data_ptr(), self.nval.data_ptr(),
                                 self.mom.data_ptr(), self.ws.data_ptr(), self.nbytes, _stream()))
        N.check(L.fsg_finalize_stats(int(kind), self.pct.data_ptr(), self.mom.data_ptr(),
                                     self.nval.data_ptr(), self.stats.data_ptr(), _stream()))
        return self.stats


def finalize_stats(kind: int, pct, moments, n_valid) -> torch.Tensor:
    stats = torch.empty(2, dtype=torch.float64, device=pct.device)
    N.check(_lib().fsg_finalize_stats(int(kind), pct.data_ptr(), moments.data_ptr(),
                                      n_valid.data_ptr(), stats.data_ptr(), _stream()))
    return stats


def copy_region(win_t, reg, dst: torch.Tensor) -> None:
    w = window(win_t)
    N.check(_lib().fsg_copy_region(C.byref(w), C.byref(reg), dst.data_ptr(), _stream()))


def encode(src_t: torch.Tensor, reg, out: N.fsg_output) -> None:
    w = window(src_t)
    N.check(_lib().fsg_encode(C.byref(w), C.byref(reg), C.byref(out), _stream()))


def quantize(t: torch.Tensor, qp: dict, dtype: str) -> torch.Tensor:
    """quantize_array on the GPU for a 1-D or 2-D float32 tensor."""
    src = t.to(torch.float32)
    shape = src.shape
    src2 = src.reshape(1, -1) if src.dim() != 2 else src
    src2 = src2.contiguous()
    dst = torch.empty(src2.shape, dtype=_TORCH_DTYPE[dtype], device=src2.device)
    encode(src2, full_region(src2), output(dst, qp=qp))
    return dst.reshape(shape)


def nodata_to_nan(raster: torch.Tensor, nodata: Optional[float], windows=()) -> torch.Tensor:
    """NoData sentinel -> NaN in place; returns per-window NoData counts (device)."""
    _check_window(raster, "raster")
    h, w = raster.shape
    counts = torch.zeros(max(1, len(windows)), dtype=torch.int64, device=raster.device)
    flat = (C.c_int32 * max(1, 4 * len(windows)))(*[int(v) for wd in windows for v in wd])
    has = nodata is not None
    N.check(_lib().fsg_nodata_to_nan(raster.data_ptr(), int(raster.stride(0)), h, w, int(has),
                                     float(nodata) if has else 0.0, flat, len(windows),
                                     counts.data_ptr(), _stream()))
    return counts

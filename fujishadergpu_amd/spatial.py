"""Spatial mode: multi-radius hillshade, slope, curvature, openness and
ambient occlusion on the GPU.

The reference's tile backend runs a spatial-mode tile through Dask as one
chunk (algorithms/tile/dask_bridge.py:395-405).  For every radius r,
multiscale_response_fields (algorithms/_nan_utils.py:455-524) either

* pads the window symmetrically by the algorithm's depth (map_overlap with
  boundary='reflect'), smooths it with _smooth_for_radius
  (_nan_utils.py:527-552: a NaN-aware Gaussian of sigma r / 2, or of r / f /
  2 on a NaN-aware f-decimated copy that is upsampled back), runs the local
  block function on it and trims the halo; or
* for radii above large_radius_threshold (_nan_utils.py:219-228), runs the
  same block function once on the raster's overview with the radius and
  the pixel scales scaled by the overview decimation, and samples that
  response bilinearly at the window's raster positions
  (coarse_large_radius_response, _nan_utils.py:328-427);

then the per-radius fields are combined (_impl_hillshade.py:110-131 for
hillshade, _combine_multiscale_dask _nan_utils.py:182-213 otherwise).

Openness and ambient occlusion take no Gaussian: their block functions
(compute_openness_spatial_block, _impl_openness.py:135-164;
compute_ambient_occlusion_spatial_block, _impl_ambient_occlusion.py:123-157)
run the local ray kernel on the block, or on its NaN-aware f-decimated copy
with the radius and the pixel scales scaled by f, upsampled back to the
block's shape.

Here each step is a kernel of csrc/spatial.hip (or the local stencil kernels on
the smoothed, padded window); only the region the caller asks for is formed.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _native as N
from . import ops
from .params import radius_to_downsample_factor

SPATIAL_AUTO_RADII = (2, 8, 32, 128, 512, 2048)   # common/spatial_mode.py


def _s():
    return torch.cuda.current_stream().cuda_stream


def clean_normalized_weights(weights) -> Optional[List[float]]:
    """_clean_normalized_weights (_nan_utils.py:148-161)."""
    cleaned = []
    for w in weights:
        try:
            fv = float(w)
        except (TypeError, ValueError):
            return None
        cleaned.append(fv if math.isfinite(fv) and fv > 0 else 0.0)
    s = float(sum(cleaned))
    if s <= 0:
        return None
    return [v / s for v in cleaned]


def resolve_radii_weights(radii, weights) -> Tuple[List[int], List[float]]:
    """_resolve_spatial_radii_weights / _normalize_spatial_radii
    (_nan_utils.py:77-129) for explicit radii; the 2**n profile when the
    weights do not fit."""
    if radii is None:
        radii = list(SPATIAL_AUTO_RADII)
    out: List[int] = []
    for r in radii:
        try:
            rv = int(round(float(r)))
        except (TypeError, ValueError):
            continue
        if rv > 0 and rv not in out:
            out.append(rv)
    if not out:
        out = list(SPATIAL_AUTO_RADII)
    if weights is not None and not isinstance(weights, (str, bytes)):
        try:
            fits = len(weights) == len(out)
        except TypeError:
            fits = False
        if fits:
            user = clean_normalized_weights(weights)
            if user is not None:
                return out, user
    raw = [2.0 ** (len(out) - 1 - i) for i in range(len(out))]
    tot = sum(raw)
    return out, [v / tot for v in raw]


RAY_ALGORITHMS = ("openness", "ambient_occlusion")
SPATIAL_ALGORITHMS = ("hillshade", "slope", "curvature") + RAY_ALGORITHMS


def depth_for(algorithm: str, r: float) -> int:
    """depth_for_scale (_impl_hillshade.py:102, _impl_slope.py:64,
    _impl_curvature.py:84, _impl_openness.py:195, _impl_ambient_occlusion.py:180)."""
    if algorithm in RAY_ALGORITHMS:
        return int(r) + 1
    if algorithm == "curvature":
        return max(3, int(float(r) * 2 + 2))
    return max(2, int(float(r) * 2 + 1))


def coarsen_factor_for_shape(h: int, w: int, coarse_max: int = 2048) -> int:
    """_nan_utils.py:231-236."""
    longest = max(int(h), int(w))
    if longest <= coarse_max:
        return 1
    return 1 << int(math.ceil(math.log2(longest / float(coarse_max))))


# ------------------------------------------------------------------ kernels

def gaussian_nan(win: torch.Tensor, pad: int, sigma: float) -> torch.Tensor:
    """NaN-aware Gaussian of the window symmetrically padded by pad."""
    h, w = win.shape
    out = torch.empty((h + 2 * pad, w + 2 * pad), dtype=torch.float32, device=win.device)
    nb = N.lib().fsg_gaussian_nan_workspace_size(h, w, pad)
    ws = torch.empty(max(1, nb), dtype=torch.uint8, device=win.device)
    wd = ops.window(win)
    N.check(N.lib().fsg_gaussian_nan(C.byref(wd), pad, float(sigma), out.data_ptr(),
                                     out.stride(0), ws.data_ptr(), nb, _s()))
    return out


def pad_symmetric(win: torch.Tensor, pad: int) -> torch.Tensor:
    h, w = win.shape
    out = torch.empty((h + 2 * pad, w + 2 * pad), dtype=torch.float32, device=win.device)
    wd = ops.window(win)
    N.check(N.lib().fsg_pad_symmetric(C.byref(wd), pad, out.data_ptr(), out.stride(0), _s()))
    return out


def downsample_nan_aware(win: torch.Tensor, factor: int) -> torch.Tensor:
    h, w = win.shape
    hc, wc = -(-h // factor), -(-w // factor)
    out = torch.empty((hc, wc), dtype=torch.float32, device=win.device)
    nb = N.lib().fsg_downsample_workspace_size(h, w, factor)
    ws = torch.empty(max(1, nb), dtype=torch.uint8, device=win.device)
    wd = ops.window(win)
    N.check(N.lib().fsg_downsample_nan_aware(C.byref(wd), factor, out.data_ptr(), wc,
                                             ws.data_ptr(), nb, _s()))
    return out


def upsample_to_shape(src: torch.Tensor, h: int, w: int) -> torch.Tensor:
    out = torch.empty((h, w), dtype=torch.float32, device=src.device)
    N.check(N.lib().fsg_upsample_to_shape(src.data_ptr(), src.stride(0), src.shape[0],
                                          src.shape[1], out.data_ptr(), out.stride(0), h, w,
                                          _s()))
    return out


def smooth_for_radius(block: torch.Tensor, pad: int, radius: float, pixel_size: float,
                      algorithm: str) -> Tuple[torch.Tensor, bool]:
    """_smooth_for_radius of the block padded by pad; also whether the result
    may hold NaN (only the decimated branch's upsampling can leave some)."""
    r = max(1.0, float(radius))
    if r <= 1.0:
        return (pad_symmetric(block, pad) if pad else block.contiguous()), True
    f = radius_to_downsample_factor(r, pixel_size=pixel_size, algorithm_name=algorithm)
    if f <= 1:
        return gaussian_nan(block, pad, max(0.5, r / 2.0)), False
    full = pad_symmetric(block, pad) if pad else block.contiguous()
    small = gaussian_nan(downsample_nan_aware(full, f), 0, max(0.5, (r / f) / 2.0))
    return upsample_to_shape(small, full.shape[0], full.shape[1]), True


def _local(algorithm: str, src: torch.Tensor, reg: N.fsg_region, out: torch.Tensor, p: dict,
           nan_free: bool) -> None:
    """The local block function on (smoothed) src, region reg into out."""
    from .algorithms import curvature_steps, gradient_steps
    px = p.get("pixel_size", 1.0)
    psx, psy = p.get("pixel_scale_x"), p.get("pixel_scale_y")
    if algorithm in ("hillshade", "slope"):
        sx, sy = gradient_steps(px, psx, psy)
        z = p.get("z_factor", 1.0)
        kw = dict(step_x=sx, step_y=sy, nan_free=nan_free)
        if algorithm == "hillshade":
            kw.update(sign_x=1.0 if (psx is None or float(psx) >= 0.0) else -1.0,
                      sign_y=1.0 if (psy is None or float(psy) >= 0.0) else -1.0,
                      z_factor=1.0 if z is None else float(z),
                      azimuth=float(p.get("azimuth", 315.0)),
                      altitude=float(p.get("altitude", 45.0)), hs_out=ops.output(out))
        else:
            kw.update(unit=p.get("unit", "degree"), sl_out=ops.output(out))
        ops.hillshade_slope(src, reg, **kw)
    else:
        sx, sy = curvature_steps(px, psx, psy)
        ops.curvature(src, reg, step_x=sx, step_y=sy,
                      curvature_type=p.get("curvature_type", "mean"), out=ops.output(out),
                      nan_free=nan_free)


def _ray_local(algorithm: str, src: torch.Tensor, reg: N.fsg_region, out: torch.Tensor,
               r: float, p: dict) -> None:
    """compute_openness_vectorized / compute_ambient_occlusion_block of src at
    radius r (max_distance / radius), region reg into out (no display stretch)."""
    from .algorithms import ao_table, openness_table
    px = p.get("pixel_size", 1.0)
    psx, psy = p.get("pixel_scale_x"), p.get("pixel_scale_y")
    o = ops.output(out)
    if algorithm == "openness":
        table = openness_table(p.get("num_directions", 16), r, px, psx, psy)
        ops.openness(src, reg, table=table, openness_type=p.get("openness_type", "positive"),
                     out=o)
    else:
        table = ao_table(p.get("num_samples", 16), r, px, psx, psy)
        ops.ambient_occlusion(src, reg, table=table, intensity=p.get("intensity", 1.0),
                              pad_d=max(1, int(round(float(r)))), out=o)


def _ray_block(algorithm: str, src: torch.Tensor, reg: N.fsg_region, out: torch.Tensor,
               r: float, p: dict) -> None:
    """compute_openness_spatial_block (_impl_openness.py:135-164) /
    compute_ambient_occlusion_spatial_block (_impl_ambient_occlusion.py:123-157)
    of the whole array src, region reg into out."""
    f = radius_to_downsample_factor(r, pixel_size=p.get("pixel_size", 1.0),
                                    algorithm_name=algorithm)
    if f <= 1:
        _ray_local(algorithm, src, reg, out, r, p)
        return
    small = downsample_nan_aware(src, f)
    if algorithm == "openness":
        r_s = float(max(2, int(round(float(r) / float(f)))))
    else:
        r_s = max(1.0, float(r) / float(f))
    resp = torch.empty_like(small)
    _ray_local(algorithm, small, ops.full_region(small), resp, r_s, _scaled(p, float(f)))
    up = upsample_to_shape(resp, src.shape[0], src.shape[1])
    out.copy_(up[reg.r0:reg.r0 + reg.h, reg.c0:reg.c0 + reg.w])


def _scaled(p: dict, fac: float) -> dict:
    kw = dict(p)
    kw["pixel_size"] = float(p.get("pixel_size", 1.0)) * fac
    for k in ("pixel_scale_x", "pixel_scale_y"):
        if kw.get(k) is not None:
            kw[k] = float(kw[k]) * fac
    return kw


def _coarse_response(algorithm: str, coarse: torch.Tensor, r_c: int, p: dict) -> torch.Tensor:
    """block_fn on a whole coarse array (no halo) at radius r_c."""
    if algorithm in RAY_ALGORITHMS:
        resp = torch.empty_like(coarse)
        _ray_block(algorithm, coarse, ops.full_region(coarse), resp, float(r_c), p)
        return resp
    sm, may_nan = smooth_for_radius(coarse, 0, r_c, p.get("pixel_size", 1.0), algorithm)
    resp = torch.empty_like(sm)
    _local(algorithm, sm, ops.full_region(sm), resp, p, nan_free=not may_nan)
    return resp


def _sample(resp: torch.Tensor, r0: int, c0: int, full_h: int, full_w: int,
            win: torch.Tensor, out: torch.Tensor) -> None:
    N.check(N.lib().fsg_sample_coarse(resp.data_ptr(), resp.stride(0), resp.shape[0],
                                      resp.shape[1], r0, c0, full_h, full_w, win.data_ptr(),
                                      win.stride(0), out.data_ptr(), out.stride(0),
                                      out.shape[0], out.shape[1], _s()))


def field_radii(algorithm: str, p: dict, direct: bool = False):
    """(radii, weights) of the per-radius fields, in the reference's order:
    _resolve_spatial_radii_weights, then openness's max(2, r) and AO's
    max(1, r) (_impl_openness.py:193, _impl_ambient_occlusion.py:178); the
    direct hillshade path keeps max(1, r) as given."""
    user_w = p.get("weights")
    if direct:
        return [max(1.0, float(r)) for r in (p.get("radii") or [1])], user_w
    radii, resolved_w = resolve_radii_weights(p.get("radii"), user_w)
    weights = (resolved_w if user_w is None else user_w) if algorithm == "hillshade" \
        else resolved_w
    if algorithm == "openness":
        radii = [int(max(2, round(float(r)))) for r in radii]
    elif algorithm == "ambient_occlusion":
        radii = [max(1, int(round(float(r)))) for r in radii]
    return radii, weights


def n_fields(algorithm: str, p: dict, direct: bool = False) -> int:
    """How many per-radius fields (the bands of agg='stack') ``p`` makes."""
    return len(field_radii(algorithm, p, direct)[0])


def run(algorithm: str, win: torch.Tensor, reg: N.fsg_region, out_descs, p: dict,
        direct: bool = False, bands=None) -> None:
    """The spatial-mode result of ``algorithm`` on the window, over region
    reg, written through the output descriptors.

    direct: hillshade's local-mode multi-radius path of the tile backend
    (_direct_hillshade, tile/dask_bridge.py:72-111): the spatial block function
    on the window itself (no halo, no overview), radii max(1, r) as given, and
    _combine_direct (tile/dask_bridge.py:28-69).

    bands: for agg='stack', one descriptor list per band (per radius, in
    radius order); band i receives field i unchanged, through its output
    epilogue (the band-first C x H x W stack of _nan_utils.py:191-194 and
    tile/dask_bridge.py:32-36)."""
    if algorithm not in SPATIAL_ALGORITHMS or (direct and algorithm != "hillshade"):
        raise ValueError(f"no spatial path for {algorithm}")
    if p.get("pixel_scale_x_rows") is not None:
        raise ValueError("per-row x scales are not defined for the spatial mode")
    h, w = win.shape
    px = float(p.get("pixel_size", 1.0) or 1.0)
    agg = str(p.get("agg", "mean") or "mean").lower()
    radii, weights = field_radii(algorithm, p, direct)
    if agg == "stack":
        if bands is None:
            raise ValueError("agg='stack' returns one band per radius: give a band-first "
                             "C x H x W output")
        if len(bands) != len(radii):
            raise ValueError(f"agg='stack' makes {len(radii)} bands; the output has "
                             f"{len(bands)}")
    elif bands is not None:
        raise ValueError(f"a band stack output needs agg='stack' (agg={agg!r})")
    thr = int(max(256, min(h, w) // 16))
    F = coarsen_factor_for_shape(h, w)
    overview = p.get("_overview_coarse_dem")
    origin = p.get("_tile_origin")
    coarse_ok = overview is not None and origin is not None
    chunk_cap = max(1, min(h, w) - 1)
    # a NaN-free window keeps every smoothed field NaN-free (all-valid block
    # means, Gaussian weights > 0, upsampling denominators of 1)
    nan_free = bool(p.get("_window_nan_free", False))
    base = {k: v for k, v in p.items() if k not in ("radii", "weights", "mode", "agg")}
    ray = algorithm in RAY_ALGORITHMS
    fields: List[torch.Tensor] = []
    for r in [float(v) for v in radii]:
        field = torch.empty((reg.h, reg.w), dtype=torch.float32, device=win.device)
        wsub = win[reg.r0:reg.r0 + reg.h, reg.c0:reg.c0 + reg.w]
        if direct:
            sm, may_nan = smooth_for_radius(win, 0, r, px, algorithm)
            _local(algorithm, sm, reg, field, base, nan_free=nan_free or not may_nan)
        elif int(round(r)) > thr and (F > 1 or coarse_ok):
            if coarse_ok:
                fac = float(p["_overview_decimation"])
                cache: Dict = p.setdefault("_overview_cache", {})
                key = (algorithm, int(round(r)), id(overview))
                if key not in cache:
                    resp = _coarse_response(algorithm, overview,
                                            max(1, int(round(r / fac))), _scaled(base, fac))
                    ready = torch.cuda.Event()
                    ready.record()
                    cache[key] = (resp, ready)
                resp, ready = cache[key]
                # tiles may run on another stream than the one that built it
                torch.cuda.current_stream().wait_event(ready)
                fh, fw = p.get("_tile_full_shape") or (h, w)
                _sample(resp, int(origin[0]) + reg.r0, int(origin[1]) + reg.c0, int(fh),
                        int(fw), wsub, field)
            else:
                # da.coarsen(nanmean, trim_excess) of the window, the block
                # function on it with the reflect halo, sampled back over the window
                hh, ww = (h // F) * F, (w // F) * F
                coarse = ops.block_mean(win[:hh, :ww], F)
                r_c = max(1, int(round(r / F)))
                kw = _scaled(base, float(F))
                d = max(1, min(depth_for(algorithm, r_c), max(1, min(coarse.shape) - 1)))
                creg = ops.region(d, d, coarse.shape[0], coarse.shape[1])
                resp = torch.empty_like(coarse)
                if ray:
                    _ray_block(algorithm, pad_symmetric(coarse, d), creg, resp, float(r_c), kw)
                else:
                    sm, may_nan = smooth_for_radius(coarse, d, r_c, kw["pixel_size"], algorithm)
                    _local(algorithm, sm, creg, resp, kw, nan_free=nan_free or not may_nan)
                _sample(resp, reg.r0, reg.c0, h, w, wsub, field)
        elif ray:
            d = max(1, min(depth_for(algorithm, r), chunk_cap))
            _ray_block(algorithm, pad_symmetric(win, d),
                       ops.region(d + reg.r0, d + reg.c0, reg.h, reg.w), field, r, base)
        else:
            d = max(1, min(depth_for(algorithm, r), chunk_cap))
            sm, may_nan = smooth_for_radius(win, d, r, px, algorithm)
            _local(algorithm, sm, ops.region(d + reg.r0, d + reg.c0, reg.h, reg.w), field, base,
                   nan_free=nan_free or not may_nan)
        fields.append(field)
    n = len(fields)
    if agg == "stack":
        one = (C.c_float * 1)(1.0)
        for f, descs in zip(fields, bands):
            ptr = (C.c_void_p * 1)(f.data_ptr())
            for o in descs:      # a one-field sum is the field itself
                N.check(N.lib().fsg_combine_fields(ptr, 1, f.stride(0), one, N.AGG_SUM, reg.h,
                                                   reg.w, C.byref(o), _s()))
        return
    # combine (the scalar weights are float32 in the products)
    wv = None
    kind = N.AGG_MEAN
    if direct:
        # _combine_direct: f0 alone; max / min; 0 + sum; f32-normalised weights;
        # else 0 + sum of f * f32(1 / n)
        if n == 1:
            wv, kind = np.ones(1, np.float32), N.AGG_SUM
        elif agg == "max":
            kind = N.AGG_MAX
        elif agg == "min":
            kind = N.AGG_MIN
        elif agg == "sum":
            kind = N.AGG_SUM
        else:
            kind = N.AGG_WEIGHTED
            wv = np.full(n, np.float32(1.0 / float(n)), np.float32)
            if isinstance(weights, (list, tuple)) and len(weights) == n:
                w32 = np.asarray(weights, dtype=np.float32)
                if np.isfinite(w32).all() and float(w32.sum()) > 0:
                    wv = (w32 / np.float32(float(w32.sum()))).astype(np.float32)
    elif algorithm == "hillshade":
        if agg == "max":
            kind = N.AGG_MAX
        elif agg == "min":
            kind = N.AGG_MIN
        elif (agg == "mean" and isinstance(weights, (list, tuple)) and len(weights) == n):
            w32 = np.asarray(weights, dtype=np.float32)
            if np.isfinite(w32).all() and float(w32.sum()) > 0:
                wv = (w32 / w32.sum()).astype(np.float32)
                kind = N.AGG_WEIGHTED
    else:
        if n == 1:
            wv, kind = np.ones(1, np.float32), N.AGG_SUM
        elif agg == "max":
            kind = N.AGG_MAX
        elif agg == "min":
            kind = N.AGG_MIN
        elif agg == "sum":
            kind = N.AGG_SUM
        elif weights is not None and len(weights) == n:
            clean = clean_normalized_weights(weights)
            if clean is not None:
                wv = np.asarray(clean, dtype=np.float32)
                kind = N.AGG_WEIGHTED
    ptrs = (C.c_void_p * n)(*[f.data_ptr() for f in fields])
    wts = (C.c_float * n)(*([float(x) for x in wv] if wv is not None else [0.0] * n))
    for o in out_descs:
        N.check(N.lib().fsg_combine_fields(ptrs, n, fields[0].stride(0), wts, kind, reg.h,
                                           reg.w, C.byref(o), _s()))

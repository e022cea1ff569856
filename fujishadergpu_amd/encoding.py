"""Output value ranges and integer quantisation parameters.

Restates io/output_encoding.py:40-175 (ranges, NoData sentinel, quantisation
coefficients).  The quantisation itself runs on the GPU in every output
epilogue (``fsg_output``); ``quantize_array`` is the drop-in for the
reference's array-level call and also runs on the GPU.
"""
from __future__ import annotations

from typing import Dict, Optional, Tuple

import numpy as np

_NORM_HEADROOM = 1.176
OUTPUT_VALUE_RANGES: Dict[str, Tuple[float, float]] = {
    "topousm_fast": (-_NORM_HEADROOM, _NORM_HEADROOM),
    "fractal_anomaly": (-_NORM_HEADROOM, _NORM_HEADROOM),
    "hillshade": (0.0, 1.0),
    "specular": (0.0, 1.0),
    "atmospheric_scattering": (0.0, 1.0),
    "curvature": (0.0, 1.0),
    "multi_light_uncertainty": (0.0, 1.0),
    "ambient_occlusion": (0.0, 1.0),
    "openness": (0.0, 1.0),
    "multiscale_terrain": (0.0, _NORM_HEADROOM),
    "visual_saliency": (0.0, _NORM_HEADROOM),
    "scale_space_surprise": (0.0, _NORM_HEADROOM),
    "npr_edges": (0.2, 1.0),
    "slope": (0.0, 90.0),
    "structure_tensor": (0.0, _NORM_HEADROOM),
    "frangi": (0.0, 1.0),
    "lic": (0.0, 1.0),
    "phase_congruency": (0.0, 1.0),
    "tv_decomposition": (0.0, 1.0),
    "scale_drift": (0.0, _NORM_HEADROOM),
}

SUPPORTED_OUTPUT_DTYPES = ("float32", "int16", "uint8")
_INT_MAXPOS: Dict[str, int] = {"int16": 32767, "uint8": 255}


def output_nodata_for_dtype(dtype) -> float:
    """io/output_encoding.py:81-86."""
    return float("nan") if np.dtype(dtype).kind == "f" else 0.0


def resolve_output_range(algorithm: str, *, params: Optional[dict] = None,
                         override: Optional[Tuple[float, float]] = None):
    """io/output_encoding.py:89-127."""
    if override is not None:
        lo, hi = float(override[0]), float(override[1])
        if hi > lo:
            return lo, hi
        raise ValueError(f"output range must satisfy high > low, got ({lo!r}, {hi!r})")
    algo = str(algorithm).lower()
    if algo == "slope" and params is not None:
        unit = str(params.get("unit", "degree")).lower()
        if unit == "radian":
            return (0.0, float(np.pi / 2.0))
        if unit != "degree":
            return None
    if params is not None:
        if algo == "structure_tensor" and str(params.get("st_output", "coherence")).lower() == "orientation":
            return (0.0, 1.0)
        if algo == "scale_drift" and str(params.get("drift_output", "magnitude")).lower() in (
                "direction", "divergence"):
            return (0.0, 1.0)
        if algo == "tv_decomposition" and str(params.get("component", "texture")).lower() == "structure":
            return None
    return OUTPUT_VALUE_RANGES.get(algo)


def quantize_params(lo: float, hi: float, dtype: str) -> Dict[str, float]:
    """io/output_encoding.py:130-175."""
    lo, hi = float(lo), float(hi)
    dt = str(dtype).lower()
    maxpos = _INT_MAXPOS[dt]
    signed = lo < 0.0 < hi
    if signed:
        a = max(abs(lo), abs(hi))
        a = a if a > 0 else 1.0
        if dt == "int16":
            a_coef, b_coef, dn_min, dn_max = maxpos / a, 0.0, -maxpos, maxpos
        else:
            a_coef, b_coef, dn_min, dn_max = (maxpos - 1) / 2.0 / a, (maxpos + 1) / 2.0, 1, maxpos
    else:
        span = (hi - lo) if (hi - lo) > 0 else 1.0
        a_coef = (maxpos - 1) / span
        b_coef = 1.0 - a_coef * lo
        dn_min, dn_max = 1, maxpos
    return {
        "a_coef": float(a_coef), "b_coef": float(b_coef),
        "dn_min": int(dn_min), "dn_max": int(dn_max),
        "scale": float(1.0 / a_coef), "offset": float(-b_coef / a_coef),
        "nodata": 0.0, "signed": bool(signed),
    }


def quantize_array(arr, qp: Dict[str, float], dtype: str):
    """quantize_array (io/output_encoding.py:178-190) on the GPU.

    ``arr`` may be a torch CUDA tensor (result stays on the device) or a NumPy
    array (copied to the device, quantised, copied back)."""
    import torch

    from . import ops
    host = isinstance(arr, np.ndarray)
    t = torch.as_tensor(np.asarray(arr, dtype=np.float32), device="cuda") if host else arr
    out = ops.quantize(t, qp, dtype)
    return out.cpu().numpy() if host else out

"""ctypes binding of libfsg.so (include/fsg.h).

The library is the product: there is no CPU fallback.  ``lib()`` raises
``NativeLibraryError`` when libfsg.so is missing or lacks a symbol, and every
wrapper raises ``FsgError`` on a non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FSG_LIBRARY", os.path.join(_PKG, "libfsg.so"))

FSG_OK, FSG_E_INVALID, FSG_E_HIP, FSG_E_UNSUPPORTED, FSG_E_WORKSPACE = 0, 1, 2, 3, 4
OUT_F32, OUT_I16, OUT_U8 = 0, 1, 2
POST_NONE, POST_DIV, POST_STRETCH = 0, 1, 2
MAX_RADII = 16
ABI_VERSION = 20  # FSG_ABI_VERSION of include/fsg.h


class NativeLibraryError(RuntimeError):
    """libfsg.so is missing, stale or incomplete."""


class FsgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[fsg status {code}] {msg}")
        self.code = code


class fsg_window(C.Structure):
    _fields_ = [("data", C.c_void_p), ("ld", C.c_int64), ("h", C.c_int32), ("w", C.c_int32)]


class fsg_region(C.Structure):
    _fields_ = [("r0", C.c_int32), ("c0", C.c_int32), ("h", C.c_int32), ("w", C.c_int32)]


class fsg_output(C.Structure):
    _fields_ = [
        ("data", C.c_void_p), ("ld", C.c_int64), ("dtype", C.c_int32), ("clip01", C.c_int32),
        ("a_coef", C.c_float), ("b_coef", C.c_float), ("dn_min", C.c_int32), ("dn_max", C.c_int32),
        ("post_kind", C.c_int32), ("pad0", C.c_int32), ("host_stats", C.c_double * 2),
        ("dev_stats", C.c_void_p),
    ]


class fsg_shade_params(C.Structure):
    _fields_ = [
        ("step_x", C.c_double), ("step_y", C.c_double), ("sign_x", C.c_double),
        ("sign_y", C.c_double), ("z_factor", C.c_double), ("azimuth_deg", C.c_double),
        ("altitude_deg", C.c_double), ("slope_unit", C.c_int32), ("nan_hint", C.c_int32),
        ("step_x_rows", C.c_void_p),
    ]


class fsg_curvature_params(C.Structure):
    _fields_ = [("step_x", C.c_double), ("step_y", C.c_double),
                ("curvature_type", C.c_int32), ("nan_hint", C.c_int32),
                ("step_x_rows", C.c_void_p)]


class fsg_ray_sample(C.Structure):
    _fields_ = [("dx", C.c_int32), ("dy", C.c_int32), ("skip", C.c_int32), ("pad0", C.c_int32),
                ("phys", C.c_double)]


class fsg_openness_params(C.Structure):
    _fields_ = [("openness_type", C.c_int32), ("n_dirs", C.c_int32), ("n_dist", C.c_int32),
                ("raw_degrees", C.c_int32), ("samples", C.POINTER(fsg_ray_sample))]


class fsg_ao_sample(C.Structure):
    _fields_ = [("dx", C.c_int32), ("dy", C.c_int32), ("skip", C.c_int32), ("pad0", C.c_int32),
                ("phys", C.c_double), ("dist_factor", C.c_double)]


class fsg_ao_params(C.Structure):
    _fields_ = [("n_samples", C.c_int32), ("pad_d", C.c_int32), ("intensity", C.c_double),
                ("samples", C.POINTER(fsg_ao_sample))]


class fsg_topousm_params(C.Structure):
    _fields_ = [
        ("n_radii", C.c_int32), ("radii", C.c_int32 * MAX_RADII), ("weights", C.c_float * MAX_RADII),
        ("ds", C.c_int32 * MAX_RADII), ("field", C.c_void_p), ("field_h", C.c_int32),
        ("field_w", C.c_int32), ("full_h", C.c_int32), ("full_w", C.c_int32),
        ("win_r0", C.c_int32), ("win_c0", C.c_int32), ("w_large", C.c_float), ("nan_hint", C.c_int32),
        ("mean", C.c_void_p * MAX_RADII), ("mean_ld", C.c_int64 * MAX_RADII),
    ]


class fsg_scan_out(C.Structure):
    _fields_ = [("factor", C.c_int32), ("out", C.c_void_p), ("ld_out", C.c_int64)]


COARSE_JOB_MAX_RADII = 4


class fsg_coarse_job(C.Structure):
    _fields_ = [
        ("win", fsg_window), ("ds", C.c_int32), ("n_radii", C.c_int32),
        ("radii", C.c_int32 * COARSE_JOB_MAX_RADII), ("outs", C.c_void_p * COARSE_JOB_MAX_RADII),
        ("ld_out", C.c_int64),
    ]


class fsg_batch_tile(C.Structure):
    _fields_ = [
        ("win", fsg_window), ("reg", fsg_region), ("hillshade", fsg_output), ("slope", fsg_output),
        ("curvature", fsg_output), ("topousm", fsg_output), ("topousm2", fsg_output),
        ("shade", fsg_shade_params), ("curv", fsg_curvature_params), ("usm", fsg_topousm_params),
        ("openness", fsg_output), ("open", fsg_openness_params),
    ]


class fsg_cnn_net(C.Structure):
    _fields_ = [("wpack", C.c_void_p * 6), ("bias", C.c_void_p * 6), ("cin", C.c_int32),
                ("cout", C.c_int32)]


class fsg_cnn_conv(C.Structure):
    _fields_ = [("x", C.c_void_p), ("wpack", C.c_void_p), ("bias", C.c_void_p),
                ("pre_in", C.c_void_p), ("pre_out", C.c_void_p), ("y", C.c_void_p),
                ("yf", C.c_void_p), ("batch", C.c_int32), ("h", C.c_int32), ("w", C.c_int32),
                ("cin", C.c_int32), ("kpad", C.c_int32), ("cout", C.c_int32),
                ("epilogue", C.c_int32), ("pad0", C.c_int32)]


CNN_GELU, CNN_LINEAR, CNN_GELU_GRAD = 0, 1, 2
AGG_WEIGHTED, AGG_MEAN, AGG_SUM, AGG_MAX, AGG_MIN = 0, 1, 2, 3, 4


class fsg_mode_stats(C.Structure):
    _fields_ = [("mode", C.c_double), ("count", C.c_int64), ("n_valid", C.c_int64),
                ("vmin", C.c_double), ("vmax", C.c_double), ("n_distinct", C.c_int64)]


class fsg_cog_raster(C.Structure):
    _fields_ = [("data", C.c_void_p), ("ld", C.c_int64), ("h", C.c_int32), ("w", C.c_int32),
                ("dtype", C.c_int32), ("has_geotransform", C.c_int32),
                ("geotransform", C.c_double * 6), ("epsg", C.c_int32),
                ("epsg_geographic", C.c_int32), ("has_nodata", C.c_int32), ("pad0", C.c_int32),
                ("nodata", C.c_double)]


class fsg_cog_options(C.Structure):
    _fields_ = [("block", C.c_int32), ("zstd_level", C.c_int32), ("overview_count", C.c_int32),
                ("threads", C.c_int32)]


class fsg_cog_report(C.Structure):
    _fields_ = [("n_levels", C.c_int32), ("pad0", C.c_int32), ("data_start", C.c_uint64),
                ("compressed_bytes", C.c_uint64), ("raw_bytes", C.c_uint64)]


class fsg_cog_info_t(C.Structure):
    _fields_ = [("n_levels", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("tile_w", C.c_int32), ("tile_h", C.c_int32), ("dtype", C.c_int32),
                ("compression", C.c_int32), ("predictor", C.c_int32),
                ("has_geotransform", C.c_int32), ("pad0", C.c_int32),
                ("geotransform", C.c_double * 6), ("epsg", C.c_int32),
                ("epsg_geographic", C.c_int32), ("has_nodata", C.c_int32), ("pad1", C.c_int32),
                ("nodata", C.c_double)]


FILL_NONE, FILL_ENCLOSED, FILL_ALL = 0, 1, 2
PREPARE_MAX_NODATA = 16


class fsg_batch_config(C.Structure):
    _fields_ = [("wg_per_cu", C.c_int32), ("order", C.c_int32), ("n_cu", C.c_int32),
                ("chunk", C.c_int32)]


P = C.POINTER
vp = C.c_void_p
_SIGS = {
    "fsg_abi_version": (C.c_int, []),
    "fsg_last_error": (C.c_char_p, []),
    "fsg_launch_count": (C.c_int64, []),
    "fsg_timing_enable": (C.c_int, [C.c_int32]),
    "fsg_timing_read": (C.c_int, [C.c_char_p, P(C.c_double), P(C.c_int64)]),
    "fsg_timing_read_union": (C.c_int, [C.c_char_p, P(C.c_double), P(C.c_int64),
                                        P(C.c_double)]),
    "fsg_nodata_to_nan": (C.c_int, [vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                    P(C.c_int32), C.c_int32, vp, vp]),
    "fsg_hillshade_slope": (C.c_int, [P(fsg_window), P(fsg_region), P(fsg_shade_params),
                                      P(fsg_output), P(fsg_output), vp]),
    "fsg_curvature": (C.c_int, [P(fsg_window), P(fsg_region), P(fsg_curvature_params),
                                P(fsg_output), vp]),
    "fsg_openness": (C.c_int, [P(fsg_window), P(fsg_region), P(fsg_openness_params),
                               P(fsg_output), vp]),
    "fsg_ambient_occlusion_workspace_size": (C.c_size_t, [P(fsg_window), P(fsg_region)]),
    "fsg_ambient_occlusion": (C.c_int, [P(fsg_window), P(fsg_region), P(fsg_ao_params),
                                        P(fsg_output), vp, C.c_size_t, vp]),
    "fsg_topousm_workspace_size": (C.c_size_t, [P(fsg_window), P(fsg_topousm_params)]),
    "fsg_topousm": (C.c_int, [P(fsg_window), P(fsg_region), P(fsg_topousm_params),
                              P(fsg_output), P(fsg_output), vp, C.c_size_t, vp]),
    "fsg_topousm_noalloc": (C.c_int, [P(fsg_window), P(fsg_region), P(fsg_topousm_params),
                                      P(fsg_output), P(fsg_output), vp]),
    "fsg_topousm_set_variant": (C.c_int, [C.c_int32]),
    "fsg_stencil_set_split_min": (C.c_int64, [C.c_int64]),
    "fsg_block_mean": (C.c_int, [P(fsg_window), C.c_int32, vp, C.c_int64, vp]),
    "fsg_overview_average": (C.c_int, [P(fsg_window), C.c_int32, C.c_int32, vp, C.c_int64, vp]),
    "fsg_topousm_coarse_means_workspace_size": (C.c_size_t, [P(fsg_window), C.c_int32]),
    "fsg_topousm_coarse_means": (C.c_int, [P(fsg_window), C.c_int32, C.c_int32, P(C.c_int32),
                                           P(vp), C.c_int64, vp, C.c_size_t, vp]),
    "fsg_topousm_field_from_means": (C.c_int, [P(vp), P(C.c_int64), P(C.c_float), C.c_int32,
                                               C.c_int32, C.c_int32, vp, C.c_int64, vp]),
    "fsg_raster_scan": (C.c_int, [P(fsg_window), C.c_int32, P(fsg_scan_out), vp, vp]),
    "fsg_topousm_coarse_means_batch_workspace_size": (C.c_size_t, [P(fsg_coarse_job),
                                                                    C.c_int32]),
    "fsg_topousm_coarse_means_batch": (C.c_int, [P(fsg_coarse_job), C.c_int32, vp, C.c_size_t,
                                                 vp]),
    "fsg_openness_points": (C.c_int, [P(fsg_window), vp, C.c_int32, P(fsg_openness_params), vp,
                                      vp]),
    "fsg_openness_judge_workspace_size": (C.c_size_t, [P(fsg_window), C.c_int32, C.c_int32]),
    "fsg_openness_judge": (C.c_int, [P(fsg_window), P(fsg_openness_params), vp, C.c_int64,
                                     C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.c_int32,
                                     vp, vp, C.c_size_t, vp]),
    "fsg_pad_symmetric": (C.c_int, [P(fsg_window), C.c_int32, vp, C.c_int64, vp]),
    "fsg_pad_reflect": (C.c_int, [P(fsg_window), C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp,
                                  C.c_int64, vp]),
    "fsg_gaussian_nan_workspace_size": (C.c_size_t, [C.c_int32, C.c_int32, C.c_int32]),
    "fsg_gaussian_nan": (C.c_int, [P(fsg_window), C.c_int32, C.c_double, vp, C.c_int64, vp,
                                   C.c_size_t, vp]),
    "fsg_downsample_workspace_size": (C.c_size_t, [C.c_int32, C.c_int32, C.c_int32]),
    "fsg_downsample_nan_aware": (C.c_int, [P(fsg_window), C.c_int32, vp, C.c_int64, vp,
                                           C.c_size_t, vp]),
    "fsg_upsample_to_shape": (C.c_int, [vp, C.c_int64, C.c_int32, C.c_int32, vp, C.c_int64,
                                        C.c_int32, C.c_int32, vp]),
    "fsg_sample_coarse": (C.c_int, [vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_int32, vp, C.c_int64, vp, C.c_int64,
                                    C.c_int32, C.c_int32, vp]),
    "fsg_combine_fields": (C.c_int, [P(vp), C.c_int32, C.c_int64, P(C.c_float), C.c_int32,
                                     C.c_int32, C.c_int32, P(fsg_output), vp]),
    "fsg_cog_write": (C.c_int, [C.c_char_p, P(fsg_cog_raster), P(fsg_cog_options),
                                P(fsg_cog_report), vp]),
    "fsg_cog_open": (C.c_int, [C.c_char_p, P(vp)]),
    "fsg_cog_info": (C.c_int, [vp, C.c_int32, P(fsg_cog_info_t)]),
    "fsg_cog_read": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp,
                               C.c_int64, C.c_int32, vp]),
    "fsg_cog_close": (None, [vp]),
    "fsg_prepare_sample": (C.c_int, [P(fsg_window), C.c_int32, C.c_int32, vp, C.c_int64, vp]),
    "fsg_value_mode_workspace_size": (C.c_size_t, [C.c_int64]),
    "fsg_value_mode": (C.c_int, [vp, C.c_int64, C.c_int32, C.c_float, vp, vp, C.c_size_t, vp]),
    "fsg_count_equal": (C.c_int, [vp, C.c_int64, C.c_float, vp, vp]),
    "fsg_coarse_average": (C.c_int, [vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                     P(C.c_float), C.c_int32, vp, C.c_int64, vp, vp, vp]),
    "fsg_pushpull_workspace_size": (C.c_size_t, [C.c_int32, C.c_int32]),
    "fsg_pushpull_fill": (C.c_int, [vp, C.c_int64, vp, C.c_int32, C.c_int32, C.c_int32, vp,
                                    C.c_int64, vp, C.c_size_t, vp]),
    "fsg_edge_connected": (C.c_int, [vp, C.c_int32, C.c_int32, vp, vp, P(C.c_int32), vp]),
    "fsg_prepare_fill": (C.c_int, [P(fsg_window), P(C.c_float), C.c_int32, C.c_int32, C.c_int32,
                                   vp, vp, C.c_int32, C.c_int32, vp, C.c_int64, vp]),
    "fsg_cnn_conv3x3": (C.c_int, [P(fsg_cnn_conv), vp]),
    "fsg_cnn_conv3x3_wgrad": (C.c_int, [vp, vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp,
                                        vp, vp]),
    "fsg_cnn_pack": (C.c_int, [vp, C.c_int32, C.c_int32, vp, C.c_int32, vp, C.c_int32, vp]),
    "fsg_cnn_adam": (C.c_int, [vp, vp, vp, vp, C.c_int32, C.c_float, C.c_float, C.c_float,
                               C.c_float, C.c_int32, vp]),
    "fsg_cnn_to_bf16": (C.c_int, [vp, vp, C.c_int64, vp]),
    "fsg_cnn_features": (C.c_int, [vp, C.c_int64, P(vp), C.c_int64, P(C.c_float), C.c_int32,
                                   C.c_int32, vp, vp]),
    "fsg_cnn_upsample": (C.c_int, [vp, vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32, vp,
                                   C.c_int64, C.c_int32, C.c_int32, vp]),
    "fsg_cnn_mse": (C.c_int, [vp, C.c_int32, vp, vp, vp, vp, C.c_int64, C.c_float, vp, vp, vp]),
    "fsg_cnn_forward": (C.c_int, [P(fsg_cnn_net), vp, C.c_int32, C.c_int32, C.c_int32, vp,
                                  C.c_int32, vp]),
    "fsg_cnn_compose": (C.c_int, [vp, C.c_int32, vp, vp, C.c_int64, C.c_int32, C.c_int32,
                                  C.c_int32, vp, C.c_int64, C.c_float, vp, C.c_int64, C.c_int32,
                                  C.c_int32, vp]),
    "fsg_cnn_detail_label": (C.c_int, [vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       vp, vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32, vp,
                                       C.c_int64, C.c_float, vp, vp, vp]),
    "fsg_batch_create": (C.c_int, [P(fsg_batch_tile), C.c_int32, P(vp)]),
    "fsg_batch_run": (C.c_int, [vp, P(fsg_batch_config), vp]),
    "fsg_batch_capture": (C.c_int, [vp, P(fsg_batch_config), vp]),
    "fsg_batch_replay": (C.c_int, [vp, vp]),
    "fsg_batch_info": (C.c_int, [vp, P(C.c_int64)]),
    "fsg_batch_destroy": (None, [vp]),
    "fsg_topousm_field_workspace_size": (C.c_size_t, [C.c_int32, C.c_int32]),
    "fsg_topousm_coarse_field": (C.c_int, [P(fsg_window), C.c_int32, P(C.c_int32), P(C.c_float),
                                           vp, C.c_int64, vp, C.c_size_t, vp]),
    "fsg_percentile_workspace_size": (C.c_size_t, [C.c_int64]),
    "fsg_percentile": (C.c_int, [vp, C.c_int64, C.c_int32, P(C.c_double), C.c_int32, vp, vp, vp,
                                 vp, C.c_size_t, vp]),
    "fsg_finalize_stats": (C.c_int, [C.c_int32, vp, vp, vp, vp, vp]),
    "fsg_copy_region": (C.c_int, [P(fsg_window), P(fsg_region), vp, vp]),
    "fsg_encode": (C.c_int, [P(fsg_window), P(fsg_region), P(fsg_output), vp]),
}

_lock = threading.Lock()
_lib = None


def declared_symbols():
    """Names of every entry point the C header declares (parsed from include/fsg.h)."""
    import re
    hdr = os.path.join(os.path.dirname(_PKG), "include", "fsg.h")
    text = open(hdr, encoding="utf-8").read()
    return sorted(set(re.findall(r"\b(fsg_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load libfsg.so once; raise NativeLibraryError if it is missing/incomplete."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryError(
                f"{LIB_PATH} not found: build it with `python -m fujishadergpu_amd.build` "
                "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as exc:
            raise NativeLibraryError(f"cannot load {LIB_PATH}: {exc}") from exc
        for name, (res, args) in _SIGS.items():
            try:
                fn = getattr(L, name)
            except AttributeError as exc:
                raise NativeLibraryError(f"{LIB_PATH} lacks symbol {name}") from exc
            fn.restype = res
            fn.argtypes = args
        if L.fsg_abi_version() != ABI_VERSION:
            raise NativeLibraryError("libfsg.so ABI version mismatch")
        _lib = L
        return _lib


def check(code: int) -> None:
    if code != FSG_OK:
        msg = lib().fsg_last_error()
        raise FsgError(code, msg.decode("utf-8", "replace") if msg else "")

// Spatial-mode building blocks (multi-radius hillshade, slope and curvature):
// the NaN-aware Gaussian of a symmetrically padded window, the padded copy,
// the corner-aligned upsampling to a shape, the bilinear sample of a global
// coarse response at a window's raster positions, and the weighted combine of
// per-radius fields.
//
// The reference runs a spatial-mode tile through Dask as one chunk
// (algorithms/tile/dask_bridge.py:395-405): map_overlap with boundary
// 'reflect' pads the window symmetrically ([d c b a | a b c d | d c b a],
// dask.array.overlap.reflect), the block function smooths the padded block
// (_smooth_for_radius, _nan_utils.py:527-552) and runs the local block
// function on it, and the result is trimmed.  Large radii come from one
// global overview (coarse_large_radius_response, _nan_utils.py:328-427).
#include <algorithm>
#include <cmath>
#include <vector>

#include "fsg_common.hpp"

namespace fsg {

constexpr int GN_MAX_R = 160;   // Gaussian taps per side (sigma <= 40 at truncate 4)

struct GaussW {
  int r;
  double g[GN_MAX_R + 1];      // centre, then distance 1 .. r
};

// symmetric ("reflect" in dask) index of i in [0, n), for |excess| <= n
__device__ __forceinline__ int sym(int i, int n) {
  return i < 0 ? -1 - i : (i >= n ? 2 * n - 1 - i : i);
}

// Value of the virtually padded window at padded coordinates, 'nearest'
// beyond the padded extent [0, n + 2 pad).
struct PadSrc {
  const float* p;
  int64_t ld;
  int h, w, pad;
  __device__ __forceinline__ float at(int pr, int pc) const {
    const int hp = h + 2 * pad, wp = w + 2 * pad;
    pr = min(max(pr, 0), hp - 1);
    pc = min(max(pc, 0), wp - 1);
    return p[(int64_t)sym(pr - pad, h) * ld + sym(pc - pad, w)];
  }
};

// Pass 0 (axis 0) of the NaN-aware Gaussian over the padded extent:
// tv = f32(sum g * where(nan, 0, x)), tw = f32(sum g * valid), scipy's
// symmetric order (centre, then pairs at distance r .. 1) in float64.
__global__ void k_gauss_nan_rows(PadSrc S, GaussW G, float* __restrict__ tv,
                                 float* __restrict__ tw, int64_t ldt) {
  const int pc = blockIdx.x * blockDim.x + threadIdx.x;
  const int hp = S.h + 2 * S.pad, wp = S.w + 2 * S.pad;
  if (pc >= wp) return;
  for (int pr = blockIdx.y; pr < hp; pr += gridDim.y) {
    const float x0 = S.at(pr, pc);
    double sv = (isnan(x0) ? 0.0 : (double)x0) * G.g[0];
    double sw = (isnan(x0) ? 0.0 : 1.0) * G.g[0];
    for (int d = G.r; d >= 1; d--) {
      const float a = S.at(pr - d, pc), b = S.at(pr + d, pc);
      sv += ((isnan(a) ? 0.0 : (double)a) + (isnan(b) ? 0.0 : (double)b)) * G.g[d];
      sw += ((isnan(a) ? 0.0 : 1.0) + (isnan(b) ? 0.0 : 1.0)) * G.g[d];
    }
    tv[(int64_t)pr * ldt + pc] = (float)sv;
    tw[(int64_t)pr * ldt + pc] = (float)sw;
  }
}

// Pass 1 (axis 1) and where(sw > 0, sv / sw, 0) in float32 ('nearest' along
// the padded rows).
__global__ void k_gauss_nan_cols(const float* __restrict__ tv, const float* __restrict__ tw,
                                 int64_t ldt, int hp, int wp, GaussW G, float* __restrict__ out,
                                 int64_t ldo) {
  const int pc = blockIdx.x * blockDim.x + threadIdx.x;
  if (pc >= wp) return;
  for (int pr = blockIdx.y; pr < hp; pr += gridDim.y) {
    const float* rv = tv + (int64_t)pr * ldt;
    const float* rw = tw + (int64_t)pr * ldt;
    double sv = (double)rv[pc] * G.g[0];
    double sw = (double)rw[pc] * G.g[0];
    for (int d = G.r; d >= 1; d--) {
      const int a = max(pc - d, 0), b = min(pc + d, wp - 1);
      sv += ((double)rv[a] + (double)rv[b]) * G.g[d];
      sw += ((double)rw[a] + (double)rw[b]) * G.g[d];
    }
    const float fv = (float)sv, fw = (float)sw;
    out[(int64_t)pr * ldo + pc] = fw > 0.0f ? fv / fw : 0.0f;
  }
}

__global__ void k_pad_sym(PadSrc S, float* __restrict__ out, int64_t ldo) {
  const int pc = blockIdx.x * blockDim.x + threadIdx.x;
  const int hp = S.h + 2 * S.pad, wp = S.w + 2 * S.pad;
  if (pc >= wp) return;
  for (int pr = blockIdx.y; pr < hp; pr += gridDim.y) out[(int64_t)pr * ldo + pc] = S.at(pr, pc);
}

// upsample_to_shape (_nan_utils.py:671-698): scipy zoom order 1, corner
// aligned (in = out * (n_in - 1) / (n_out - 1)), of where(nan, 0, x) and of
// valid, num / max(den, 1e-6) where den > 1e-3 else NaN; the zoom output
// shape round(n_in * (N / n_in)) is edge-padded (or cropped) to N.  A
// NaN-free source gives den = 1 and the plain zoom bit for bit.
__global__ void k_upsample_shape(const float* __restrict__ src, int64_t lds, int hs, int ws,
                                 int zh, int zw, float* __restrict__ out, int64_t ldo, int h,
                                 int w) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= w) return;
  const double zx = zw > 1 ? (double)(ws - 1) / (double)(zw - 1) : 0.0;
  const double zy = zh > 1 ? (double)(hs - 1) / (double)(zh - 1) : 0.0;
  const int jj = min(j, zw - 1);   // edge padding past the zoom output
  double cx = fmin(fmax((double)jj * zx, 0.0), (double)(ws - 1));
  const int x0 = (int)floor(cx), x1 = min(x0 + 1, ws - 1);
  const double tx = cx - (double)x0;
  for (int i = blockIdx.y; i < h; i += gridDim.y) {
    const int ii = min(i, zh - 1);
    double cy = fmin(fmax((double)ii * zy, 0.0), (double)(hs - 1));
    const int y0 = (int)floor(cy), y1 = min(y0 + 1, hs - 1);
    const double ty = cy - (double)y0;
    const float a = src[(int64_t)y0 * lds + x0], b = src[(int64_t)y0 * lds + x1];
    const float c = src[(int64_t)y1 * lds + x0], d = src[(int64_t)y1 * lds + x1];
    auto fz = [](float v) { return isnan(v) ? 0.0 : (double)v; };
    auto vz = [](float v) { return isnan(v) ? 0.0 : 1.0; };
    const double num = (fz(a) * (1.0 - tx) + fz(b) * tx) * (1.0 - ty) +
                       (fz(c) * (1.0 - tx) + fz(d) * tx) * ty;
    const double den = (vz(a) * (1.0 - tx) + vz(b) * tx) * (1.0 - ty) +
                       (vz(c) * (1.0 - tx) + vz(d) * tx) * ty;
    const float fn = (float)num, fd = (float)den;
    out[(int64_t)i * ldo + j] = fd > 1e-3f ? fn / fmaxf(fd, 1e-6f) : __builtin_nanf("");
  }
}

// _bilinear_sample_coarse (_nan_utils.py:255-281): map_coordinates order 1,
// mode 'nearest', at float32 coordinates (g + 0.5) * f32(n_c / N) - 0.5 of the
// window's raster rows r0 + i and columns c0 + j; NaN where the window is NaN.
__global__ void k_sample_coarse(const float* __restrict__ coarse, int64_t ldc, int hc, int wc,
                                int r0, int c0, float fy, float fx, const float* __restrict__ win,
                                int64_t ldw, float* __restrict__ out, int64_t ldo, int h, int w) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= w) return;
  const float fxc = ((float)(c0 + j) + 0.5f) * fx - 0.5f;
  const double cx = fmin(fmax((double)fxc, 0.0), (double)(wc - 1));
  const int x0 = (int)floor(cx), x1 = min(x0 + 1, wc - 1);
  const double tx = cx - (double)x0;
  for (int i = blockIdx.y; i < h; i += gridDim.y) {
    const float fyc = ((float)(r0 + i) + 0.5f) * fy - 0.5f;
    const double cy = fmin(fmax((double)fyc, 0.0), (double)(hc - 1));
    const int y0 = (int)floor(cy), y1 = min(y0 + 1, hc - 1);
    const double ty = cy - (double)y0;
    const double top = (double)coarse[(int64_t)y0 * ldc + x0] * (1.0 - tx) +
                       (double)coarse[(int64_t)y0 * ldc + x1] * tx;
    const double bot = (double)coarse[(int64_t)y1 * ldc + x0] * (1.0 - tx) +
                       (double)coarse[(int64_t)y1 * ldc + x1] * tx;
    float v = (float)(top * (1.0 - ty) + bot * ty);
    if (win && isnan(win[(int64_t)i * ldw + j])) v = __builtin_nanf("");
    out[(int64_t)i * ldo + j] = v;
  }
}

constexpr int CMB_MAX = 16;
struct CombineArgs {
  const float* f[CMB_MAX];
  int64_t ld;
  int n, agg;           // FSG_AGG_*
  float w[CMB_MAX];
  int h, w_;
};

// _combine_multiscale_dask (_nan_utils.py:182-213) and hillshade's inline
// variant: weighted sum out = f0 * w0 (+ f_i * w_i in order), or the float32
// mean / min / max / sum over the fields; then the output epilogue.
__global__ void k_combine(CombineArgs A, OutDev o) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= A.w_) return;
  const PostK pp = load_post(o);
  for (int i = blockIdx.y; i < A.h; i += gridDim.y) {
    const int64_t e = (int64_t)i * A.ld + j;
    float v = A.f[0][e];
    if (A.agg == FSG_AGG_WEIGHTED) {
      v = v * A.w[0];
      for (int k = 1; k < A.n; k++) v = v + A.f[k][e] * A.w[k];
    } else if (A.agg == FSG_AGG_MEAN) {
      for (int k = 1; k < A.n; k++) v = v + A.f[k][e];
      v = v / (float)A.n;
    } else if (A.agg == FSG_AGG_SUM) {
      for (int k = 1; k < A.n; k++) v = v + A.f[k][e];
    } else if (A.agg == FSG_AGG_MAX) {
      for (int k = 1; k < A.n; k++) {
        const float x = A.f[k][e];
        v = (isnan(v) || isnan(x)) ? __builtin_nanf("") : fmaxf(v, x);
      }
    } else {
      for (int k = 1; k < A.n; k++) {
        const float x = A.f[k][e];
        v = (isnan(v) || isnan(x)) ? __builtin_nanf("") : fminf(v, x);
      }
    }
    store_out(o, i, j, apply_post(pp, v));
  }
}

static dim3 grid2(int w, int h) {
  return dim3((unsigned)((w + 255) / 256), (unsigned)std::min(h, 16384));
}

}  // namespace fsg

using namespace fsg;

extern "C" size_t fsg_gaussian_nan_workspace_size(int32_t h, int32_t w, int32_t pad) {
  if (h <= 0 || w <= 0 || pad < 0) return 0;
  return 2 * ((sizeof(float) * (size_t)(h + 2 * pad) * (w + 2 * pad) + 255) & ~(size_t)255);
}

extern "C" int fsg_gaussian_nan(const fsg_window* src, int32_t pad, double sigma, float* out,
                                int64_t ldo, void* workspace, size_t workspace_bytes,
                                void* stream) {
  int rc = validate_window(src, "fsg_gaussian_nan");
  if (rc) return rc;
  if (src->h <= 0 || src->w <= 0 || pad < 0 || pad > std::min(src->h, src->w) || !out ||
      !(sigma > 0.0))
    return fail(FSG_E_INVALID, "fsg_gaussian_nan: bad arguments (pad must be <= the window)");
  const int hp = src->h + 2 * pad, wp = src->w + 2 * pad;
  if (ldo < wp) return fail(FSG_E_INVALID, "fsg_gaussian_nan: ldo smaller than the padded width");
  const size_t need = fsg_gaussian_nan_workspace_size(src->h, src->w, pad);
  if (!workspace || workspace_bytes < need)
    return fail(FSG_E_WORKSPACE, "fsg_gaussian_nan: workspace too small");
  GaussW G{};
  G.r = (int)(4.0 * sigma + 0.5);   // scipy truncate = 4.0
  if (G.r > GN_MAX_R) return fail(FSG_E_UNSUPPORTED, "fsg_gaussian_nan: sigma above 40");
  std::vector<double> full(2 * G.r + 1);
  gaussian_weights(sigma, G.r, full.data());
  for (int d = 0; d <= G.r; d++) G.g[d] = full[G.r + d];
  float* tv = (float*)workspace;
  float* tw = (float*)((char*)workspace + need / 2);
  hipStream_t s = as_stream(stream);
  PadSrc S{src->data, src->ld, src->h, src->w, pad};
  TimedScope ts("k_gauss_nan", s);
  hipLaunchKernelGGL(k_gauss_nan_rows, grid2(wp, hp), dim3(256), 0, s, S, G, tv, tw,
                     (int64_t)wp);
  FSG_CHECK_LAUNCH("k_gauss_nan_rows");
  hipLaunchKernelGGL(k_gauss_nan_cols, grid2(wp, hp), dim3(256), 0, s, tv, tw, (int64_t)wp, hp,
                     wp, G, out, ldo);
  FSG_CHECK_LAUNCH("k_gauss_nan_cols");
  return FSG_OK;
}

extern "C" int fsg_pad_symmetric(const fsg_window* src, int32_t pad, float* out, int64_t ldo,
                                 void* stream) {
  int rc = validate_window(src, "fsg_pad_symmetric");
  if (rc) return rc;
  if (src->h <= 0 || src->w <= 0 || pad < 0 || pad > std::min(src->h, src->w) || !out ||
      ldo < src->w + 2 * pad)
    return fail(FSG_E_INVALID, "fsg_pad_symmetric: bad arguments (pad must be <= the window)");
  const int hp = src->h + 2 * pad, wp = src->w + 2 * pad;
  PadSrc S{src->data, src->ld, src->h, src->w, pad};
  hipLaunchKernelGGL(k_pad_sym, grid2(wp, hp), dim3(256), 0, as_stream(stream), S, out, ldo);
  FSG_CHECK_LAUNCH("k_pad_sym");
  return FSG_OK;
}

extern "C" int fsg_upsample_to_shape(const float* src, int64_t lds, int32_t hs, int32_t ws,
                                     float* out, int64_t ldo, int32_t h, int32_t w,
                                     void* stream) {
  if (!src || !out || hs <= 0 || ws <= 0 || h <= 0 || w <= 0 || lds < ws || ldo < w)
    return fail(FSG_E_INVALID, "fsg_upsample_to_shape: bad arguments");
  // scipy zoom output shape: round(n_in * (N / n_in))
  const int zh = (int)std::nearbyint((double)hs * ((double)h / (double)hs));
  const int zw = (int)std::nearbyint((double)ws * ((double)w / (double)ws));
  hipLaunchKernelGGL(k_upsample_shape, grid2(w, h), dim3(256), 0, as_stream(stream), src, lds,
                     hs, ws, std::max(1, zh), std::max(1, zw), out, ldo, h, w);
  FSG_CHECK_LAUNCH("k_upsample_shape");
  return FSG_OK;
}

extern "C" int fsg_sample_coarse(const float* coarse, int64_t ldc, int32_t hc, int32_t wc,
                                 int32_t r0, int32_t c0, int32_t full_h, int32_t full_w,
                                 const float* nan_src, int64_t ld_nan, float* out, int64_t ldo,
                                 int32_t h, int32_t w, void* stream) {
  if (!coarse || !out || hc <= 0 || wc <= 0 || full_h <= 0 || full_w <= 0 || h <= 0 || w <= 0 ||
      ldc < wc || ldo < w || (nan_src && ld_nan < w))
    return fail(FSG_E_INVALID, "fsg_sample_coarse: bad arguments");
  const float fy = (float)((double)hc / (double)full_h), fx = (float)((double)wc / (double)full_w);
  hipLaunchKernelGGL(k_sample_coarse, grid2(w, h), dim3(256), 0, as_stream(stream), coarse, ldc,
                     hc, wc, r0, c0, fy, fx, nan_src, ld_nan, out, ldo, h, w);
  FSG_CHECK_LAUNCH("k_sample_coarse");
  return FSG_OK;
}

extern "C" int fsg_combine_fields(const float* const* fields, int32_t n, int64_t ld,
                                  const float* weights, int32_t agg, int32_t h, int32_t w,
                                  const fsg_output* out, void* stream) {
  if (!fields || n <= 0 || n > CMB_MAX || h <= 0 || w <= 0 || ld < w || !out ||
      agg < FSG_AGG_WEIGHTED || agg > FSG_AGG_MIN || (agg == FSG_AGG_WEIGHTED && !weights))
    return fail(FSG_E_INVALID, "fsg_combine_fields: bad arguments");
  fsg_region reg{0, 0, h, w};
  int rc = validate_output(out, &reg, "fsg_combine_fields");
  if (rc) return rc;
  CombineArgs A{};
  for (int k = 0; k < n; k++) {
    if (!fields[k]) return fail(FSG_E_INVALID, "fsg_combine_fields: field NULL");
    A.f[k] = fields[k];
    A.w[k] = weights ? weights[k] : 0.0f;
  }
  A.ld = ld;
  A.n = n;
  A.agg = agg;
  A.h = h;
  A.w_ = w;
  hipLaunchKernelGGL(k_combine, grid2(w, h), dim3(256), 0, as_stream(stream), A, make_out(out));
  FSG_CHECK_LAUNCH("k_combine");
  return FSG_OK;
}

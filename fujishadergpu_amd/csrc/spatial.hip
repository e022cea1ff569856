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

// The NaN-aware Gaussian runs as two LDS-tiled passes over 64 x 64 output
// tiles.  Each of the 4 waves of a block owns GN_M consecutive outputs along
// the filter axis for 64 lines across it, so one sample loaded from LDS into
// a register serves GN_M outputs: the pair (x[-d], x[+d]) of output m at
// distance d is x[m - d], x[m + d], and a block of GN_M distances reads two
// register windows of 2 GN_M - 1 samples that slide by GN_M per block.  Every
// output still accumulates in scipy's symmetric order, in float64: the
// centre, then acc + (x[-d] + x[+d]) * g[d] for d = r .. 1, each operation
// rounded on its own (the library builds with -ffp-contract=off).
constexpr int GN_T = 64;   // output tile edge
constexpr int GN_M = 16;   // outputs per lane along the filter axis (4 waves x 16 = 64)

// sample as the accumulation sees it: 0 plain, 1 NaN -> 0, 2 validity (0 / 1)
template <int MODE>
__device__ __forceinline__ double gn_val(float f) {
  if constexpr (MODE == 0) return (double)f;
  else if constexpr (MODE == 1) return isnan(f) ? 0.0 : (double)f;
  else return isnan(f) ? 0.0 : 1.0;
}

// acc[m] = sum_k g[|k|] v(s[(m + k) ST]), k in [-r, r], in the order above.
// s points at the lane's first output; the LDS line holds r samples before
// it and r + GN_M - 1 after it.
template <int MODE, int ST>
__device__ __forceinline__ void gn_line(const float* s, const GaussW& G, double (&acc)[GN_M]) {
  const double g0 = G.g[0];
#pragma unroll
  for (int m = 0; m < GN_M; m++) acc[m] = gn_val<MODE>(s[m * ST]) * g0;
  int d = G.r;
  for (; d % GN_M != 0; d--) {   // the top r mod GN_M distances one at a time
    const double g = G.g[d];
#pragma unroll
    for (int m = 0; m < GN_M; m++)
      acc[m] = acc[m] + (gn_val<MODE>(s[(m - d) * ST]) + gn_val<MODE>(s[(m + d) * ST])) * g;
  }
  if (d == 0) return;
  // distances d .. d - GN_M + 1: L[k] = x[k - d], R[k] = x[k + d - GN_M + 1]
  double L[2 * GN_M - 1], R[2 * GN_M - 1];
#pragma unroll
  for (int k = 0; k < 2 * GN_M - 1; k++) {
    L[k] = gn_val<MODE>(s[(k - d) * ST]);
    R[k] = gn_val<MODE>(s[(k + d - (GN_M - 1)) * ST]);
  }
  for (;;) {
#pragma unroll
    for (int dd = 0; dd < GN_M; dd++) {
      const double g = G.g[d - dd];
#pragma unroll
      for (int m = 0; m < GN_M; m++) acc[m] = acc[m] + (L[m + dd] + R[m - dd + GN_M - 1]) * g;
    }
    d -= GN_M;
    if (d == 0) break;
#pragma unroll
    for (int k = 0; k < GN_M - 1; k++) L[k] = L[k + GN_M];
#pragma unroll
    for (int k = GN_M - 1; k < 2 * GN_M - 1; k++) L[k] = gn_val<MODE>(s[(k - d) * ST]);
#pragma unroll
    for (int k = 2 * GN_M - 2; k >= GN_M; k--) R[k] = R[k - GN_M];
#pragma unroll
    for (int k = 0; k < GN_M; k++) R[k] = gn_val<MODE>(s[(k + d - (GN_M - 1)) * ST]);
  }
}

// Pass 0 (axis 0) over the padded extent: tv = f32(sum g * where(nan, 0, x)),
// tw = f32(sum g * valid).  A tile whose input span holds no NaN has tw =
// f32(wsum) everywhere, wsum being the same accumulation over all-valid
// samples (computed on the host by the same operation sequence).
// LDS: (64 + 2r) rows x 64 columns.
__global__ __launch_bounds__(256, 2) void k_gauss_nan_rows(PadSrc S, GaussW G,
                                                           float* __restrict__ tv,
                                                           float* __restrict__ tw, int64_t ldt,
                                                           float wconst) {
  extern __shared__ float gl[];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int hp = S.h + 2 * S.pad, wp = S.w + 2 * S.pad;
  const int tx0 = blockIdx.x * GN_T, ty0 = blockIdx.y * GN_T;
  const int nrow = GN_T + 2 * G.r;
  const int pc = min(tx0 + lane, wp - 1);
  bool nan = false;
  // staging in batches of 8 rows per wave (8 loads in flight per lane)
  for (int i0 = q; i0 < nrow; i0 += 32) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + 4 * u;
      v[u] = i < nrow ? S.at(ty0 - G.r + i, pc) : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + 4 * u;
      if (i < nrow) {
        gl[i * GN_T + lane] = v[u];
        nan |= isnan(v[u]);
      }
    }
  }
  const bool any_nan = __syncthreads_or(nan);
  const float* s = gl + (q * GN_M + G.r) * GN_T + lane;
  const int x = tx0 + lane;
  const int y0 = ty0 + q * GN_M;
  const int nm = x < wp ? min(GN_M, hp - y0) : 0;   // rows this lane stores
  double acc[GN_M];
  if (any_nan) gn_line<1, GN_T>(s, G, acc);
  else gn_line<0, GN_T>(s, G, acc);
#pragma unroll
  for (int m = 0; m < GN_M; m++)
    if (m < nm) tv[(int64_t)(y0 + m) * ldt + x] = (float)acc[m];
  if (!any_nan) {
#pragma unroll
    for (int m = 0; m < GN_M; m++)
      if (m < nm) tw[(int64_t)(y0 + m) * ldt + x] = wconst;
    return;
  }
  gn_line<2, GN_T>(s, G, acc);
#pragma unroll
  for (int m = 0; m < GN_M; m++)
    if (m < nm) tw[(int64_t)(y0 + m) * ldt + x] = (float)acc[m];
}

// Pass 1 (axis 1, 'nearest' along the padded rows) and where(sw > 0, sv / sw,
// 0) in float32.  A tile whose tw span is all f32(wsum) divides by the
// constant f32 of that accumulation (wdiv, from the host).  LDS: 64 rows x
// (64 + 2r + 1) columns (odd pitch: lanes walk rows without bank conflicts),
// reused for tw, tv and the transposed output tile.
__global__ __launch_bounds__(256, 2) void k_gauss_nan_cols(const float* __restrict__ tv,
                                                        const float* __restrict__ tw, int64_t ldt,
                                                        int hp, int wp, GaussW G,
                                                        float* __restrict__ out, int64_t ldo,
                                                        float wconst, float wdiv) {
  extern __shared__ float gl[];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int tx0 = blockIdx.x * GN_T, ty0 = blockIdx.y * GN_T;
  const int ncol = GN_T + 2 * G.r, P = ncol + 1;
  // staging: column chunks of 64, each in batches of 8 rows per wave (8 loads in flight)
  auto stage = [&](const float* src) {
    bool off = false;
    for (int c = 0; c < ncol; c += 64) {
      const int j = c + lane;
      const bool jok = j < ncol;
      const int col = min(max(tx0 - G.r + j, 0), wp - 1);
#pragma unroll
      for (int k0 = 0; k0 < GN_T / 4; k0 += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int i = q + 4 * (k0 + u);
          v[u] = jok ? src[(int64_t)min(ty0 + i, hp - 1) * ldt + col] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int i = q + 4 * (k0 + u);
          if (jok) {
            gl[i * P + j] = v[u];
            off |= __float_as_uint(v[u]) != __float_as_uint(wconst);
          }
        }
      }
    }
    return off;
  };
  const float* s = gl + lane * P + q * GN_M + G.r;
  double acc[GN_M];
  float fw[GN_M], fo[GN_M];
  const bool varied = __syncthreads_or(stage(tw));
  if (varied) {
    gn_line<0, 1>(s, G, acc);
#pragma unroll
    for (int m = 0; m < GN_M; m++) fw[m] = (float)acc[m];
  } else {
#pragma unroll
    for (int m = 0; m < GN_M; m++) fw[m] = wdiv;
  }
  __syncthreads();
  stage(tv);
  __syncthreads();
  gn_line<0, 1>(s, G, acc);
#pragma unroll
  for (int m = 0; m < GN_M; m++) {
    const float v = (float)acc[m];
    fo[m] = fw[m] > 0.0f ? v / fw[m] : 0.0f;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < GN_M; m++) gl[lane * (GN_T + 1) + q * GN_M + m] = fo[m];
  __syncthreads();
  const int x = tx0 + lane;
  if (x >= wp) return;
  for (int i = q; i < GN_T; i += 4) {
    const int y = ty0 + i;
    if (y < hp) out[(int64_t)y * ldo + x] = gl[i * (GN_T + 1) + lane];
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
  // the all-valid accumulations: pass 0 over validity 1, pass 1 over f32 of that
  double w0 = 1.0 * G.g[0];
  for (int d = G.r; d >= 1; d--) w0 = w0 + (1.0 + 1.0) * G.g[d];
  const float wconst = (float)w0;
  double w1 = (double)wconst * G.g[0];
  for (int d = G.r; d >= 1; d--) w1 = w1 + ((double)wconst + (double)wconst) * G.g[d];
  const float wdiv = (float)w1;
  float* tv = (float*)workspace;
  float* tw = (float*)((char*)workspace + need / 2);
  hipStream_t s = as_stream(stream);
  PadSrc S{src->data, src->ld, src->h, src->w, pad};
  const dim3 grid((wp + GN_T - 1) / GN_T, (hp + GN_T - 1) / GN_T);
  const size_t lds_r = sizeof(float) * (GN_T + 2 * G.r) * GN_T;
  const size_t lds_c = sizeof(float) * GN_T * (GN_T + 2 * G.r + 1);
  TimedScope ts("k_gauss_nan", s);
  hipLaunchKernelGGL(k_gauss_nan_rows, grid, dim3(256), lds_r, s, S, G, tv, tw, (int64_t)wp,
                     wconst);
  FSG_CHECK_LAUNCH("k_gauss_nan_rows");
  hipLaunchKernelGGL(k_gauss_nan_cols, grid, dim3(256), lds_c, s, tv, tw, (int64_t)wp, hp, wp, G,
                     out, ldo, wconst, wdiv);
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

// dask map_overlap boundary='reflect' at the raster edges of a window: each
// side padded by its own amount with numpy's 'symmetric' rule
__global__ void k_pad_sides(const float* __restrict__ src, int64_t lds, int h, int w, int top,
                            int left, float* __restrict__ out, int64_t ldo, int ho, int wo) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= wo) return;
  const int sc = sym(c - left, w);
  for (int r = blockIdx.y; r < ho; r += gridDim.y)
    out[(int64_t)r * ldo + c] = src[(int64_t)sym(r - top, h) * lds + sc];
}

extern "C" int fsg_pad_reflect(const fsg_window* src, int32_t top, int32_t bottom, int32_t left,
                               int32_t right, float* out, int64_t ldo, void* stream) {
  int rc = validate_window(src, "fsg_pad_reflect");
  if (rc) return rc;
  if (src->h <= 0 || src->w <= 0 || !out || top < 0 || bottom < 0 || left < 0 || right < 0 ||
      std::max(top, bottom) > src->h || std::max(left, right) > src->w ||
      ldo < (int64_t)src->w + left + right)
    return fail(FSG_E_INVALID,
                "fsg_pad_reflect: bad arguments (each pad at most the window extent)");
  const int ho = src->h + top + bottom, wo = src->w + left + right;
  hipLaunchKernelGGL(k_pad_sides, grid2(wo, ho), dim3(256), 0, as_stream(stream), src->data,
                     src->ld, src->h, src->w, top, left, out, ldo, ho, wo);
  FSG_CHECK_LAUNCH("k_pad_sides");
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

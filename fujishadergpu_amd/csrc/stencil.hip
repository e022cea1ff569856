// Gradient stencils: fused hillshade + slope, and curvature.
//
// Derivatives are numpy/cupy gradient(f, h_row, h_col, edge_order=2) on the
// window: interior (f[i+1] - f[i-1]) / f32(2h), window edges
// a*f0 + b*f1 + c*f2 with f32 coefficients (-1.5/h, 2/h, -0.5/h) and the
// mirror (0.5/h, -2/h, 1.5/h), every operation rounded in f32.  NaN cells are
// replaced by the NaN-aware sigma=1 Gaussian fill of handle_nan_for_gradient
// (_nan_utils.py:50-74), evaluated on demand in scipy's two-pass order.
#include <algorithm>
#include <cmath>

#include "fsg_common.hpp"

namespace fsg {

struct WinDev {
  const float* p;
  int64_t ld;
  int h, w;
};

__device__ __forceinline__ float raw(const WinDev& W, int r, int c) {
  return W.p[(int64_t)r * W.ld + c];
}

// Gaussian sigma=1, truncate=4 weights: g[0] centre, g[d] distance d.
struct Gauss9 {
  double g[5];
};

// NaN-aware local fill at (r, c): scipy gaussian_filter(mode='nearest') of
// filled*valid and of valid, axis 0 then axis 1 with float32 intermediates,
// symmetric accumulation order (centre, then distance 4..1), then
// where(sw > 0, sv / sw, 0) in float32 (_nan_utils.py:18-31).
__device__ __noinline__ float nan_fill(const WinDev W, int r, int c, const Gauss9 G) {
  float vsv[9], vsw[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int cc = clampi(c + k - 4, W.w);
    const float x0 = raw(W, r, cc);
    const bool v0 = !isnan(x0);
    double sv = (v0 ? (double)x0 : 0.0) * G.g[0];
    double sw = (v0 ? 1.0 : 0.0) * G.g[0];
#pragma unroll
    for (int d = 4; d >= 1; d--) {
      const float a = raw(W, clampi(r - d, W.h), cc);
      const float b = raw(W, clampi(r + d, W.h), cc);
      const double fa = isnan(a) ? 0.0 : (double)a, fb = isnan(b) ? 0.0 : (double)b;
      const double wa = isnan(a) ? 0.0 : 1.0, wb = isnan(b) ? 0.0 : 1.0;
      sv += (fa + fb) * G.g[d];
      sw += (wa + wb) * G.g[d];
    }
    vsv[k] = (float)sv;
    vsw[k] = (float)sw;
  }
  double sv = (double)vsv[4] * G.g[0];
  double sw = (double)vsw[4] * G.g[0];
#pragma unroll
  for (int d = 4; d >= 1; d--) {
    sv += ((double)vsv[4 - d] + (double)vsv[4 + d]) * G.g[d];
    sw += ((double)vsw[4 - d] + (double)vsw[4 + d]) * G.g[d];
  }
  const float fsv = (float)sv, fsw = (float)sw;
  return fsw > 0.0f ? fsv / fsw : 0.0f;
}

// Filled value (NaN replaced by the local fill).  NANAW = false is the variant
// for windows known to hold no NaN: no fill call, identical results there.
template <bool NANAW>
__device__ __forceinline__ float filled(const WinDev& W, int r, int c, const Gauss9& G) {
  const float v = raw(W, r, c);
  if (!NANAW) return v;
  return isnan(v) ? nan_fill(W, r, c, G) : v;
}

// numpy gradient coefficients for one axis (all float32, from a Python step).
struct Axis {
  float two_h;       // f32(2*h)
  float a0, b0, c0;  // f32(-1.5/h), f32(2/h), f32(-0.5/h)
  float a1, b1, c1;  // f32(0.5/h),  f32(-2/h), f32(1.5/h)
};

inline Axis make_axis(double h) {
  Axis a;
  a.two_h = (float)(2. * h);
  a.a0 = (float)(-1.5 / h);
  a.b0 = (float)(2. / h);
  a.c0 = (float)(-0.5 / h);
  a.a1 = (float)(0.5 / h);
  a.b1 = (float)(-2. / h);
  a.c1 = (float)(1.5 / h);
  return a;
}

// One derivative along an axis of length n at index i from accessor f(k).
template <class F>
__device__ __forceinline__ float deriv(const Axis& A, int i, int n, F f) {
  if (i == 0) {
    const float t0 = A.a0 * f(0);
    const float t1 = A.b0 * f(1);
    const float s = t0 + t1;
    const float t2 = A.c0 * f(2);
    return s + t2;
  }
  if (i == n - 1) {
    const float t0 = A.a1 * f(n - 3);
    const float t1 = A.b1 * f(n - 2);
    const float s = t0 + t1;
    const float t2 = A.c1 * f(n - 1);
    return s + t2;
  }
  const float d = f(i + 1) - f(i - 1);
  return d / A.two_h;
}

// ------------------------------------------------------------ hillshade+slope

struct ShadeK {
  Axis ax, ay;    // |step| derivative coefficients
  float zf;       // f32(z_factor) (hillshade only)
  float sgx, sgy; // f32 orientation signs
  double lx, ly, lz;
  int unit;
  int need_hs, need_sl;
  int z_is_one;
};

template <bool NANAW>
__global__ void __launch_bounds__(256)
k_hillshade_slope(WinDev W, int r0, int c0, int rh, int rw, ShadeK K, Gauss9 G, OutDev ohs,
                  OutDev osl) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (c >= rw || r >= rh) return;
  const int wr = r0 + r, wc = c0 + c;
  const float zc = raw(W, wr, wc);
  const PostK phs = load_post(ohs), psl = load_post(osl);
  if (isnan(zc)) {
    const float qn = __builtin_nanf("");
    if (K.need_hs) store_out(ohs, r, c, apply_post(phs, qn));
    if (K.need_sl) store_out(osl, r, c, apply_post(psl, qn));
    return;
  }
  // Unscaled filled neighbours along each axis.
  auto fr = [&](int k) { return filled<NANAW>(W, k, wc, G); };
  auto fcl = [&](int k) { return filled<NANAW>(W, wr, k, G); };
  const float dy = deriv(K.ay, wr, W.h, fr);
  const float dx = deriv(K.ax, wc, W.w, fcl);
  if (K.need_sl) {
    const float m = sqrtf(dx * dx + dy * dy);
    const float rad = atanf(m);
    float s;
    if (K.unit == 0) {
      s = (float)((double)rad * (180.0 / M_PI));
    } else if (K.unit == 1) {
      s = tanf(rad) * 100.0f;
    } else {
      s = rad;
    }
    store_out(osl, r, c, apply_post(psl, s));
  }
  if (K.need_hs) {
    float hdy = dy, hdx = dx;
    if (!K.z_is_one) {
      auto frz = [&](int k) { return filled<NANAW>(W, k, wc, G) * K.zf; };
      auto fcz = [&](int k) { return filled<NANAW>(W, wr, k, G) * K.zf; };
      hdy = deriv(K.ay, wr, W.h, frz);
      hdx = deriv(K.ax, wc, W.w, fcz);
    }
    const float e = hdx * K.sgx;
    const float n = hdy * K.sgy;
    const float norm = sqrtf(e * e + n * n + 1.0f);
    // float64 from here: the light vector is a 0-d float64 array.
    double h = (double)(-e) * K.lx;
    h = h + (double)(-n) * K.ly;
    h = h + K.lz;
    h = h / (double)norm;
    h = fmin(fmax(h, 0.0), 1.0);
    store_out(ohs, r, c, apply_post(phs, (float)h));
  }
}

// ------------------------------------------------------------------ curvature

struct CurvK {
  Axis ax, ay;    // |step| (first derivatives)
  Axis sx, sy;    // signed steps (second derivatives)
  int type;
  float gamma;
};

template <bool NANAW>
__device__ __forceinline__ float d_col(const WinDev& W, const Axis& A, int i, int j,
                                       const Gauss9& G) {
  auto f = [&](int k) { return filled<NANAW>(W, i, k, G); };
  return deriv(A, j, W.w, f);
}
template <bool NANAW>
__device__ __forceinline__ float d_row(const WinDev& W, const Axis& A, int i, int j,
                                       const Gauss9& G) {
  auto f = [&](int k) { return filled<NANAW>(W, k, j, G); };
  return deriv(A, i, W.h, f);
}

template <bool NANAW>
__global__ void __launch_bounds__(256)
k_curvature(WinDev W, int r0, int c0, int rh, int rw, CurvK K, Gauss9 G, OutDev out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (c >= rw || r >= rh) return;
  const int i = r0 + r, j = c0 + c;
  const PostK pp = load_post(out);
  const float zc = raw(W, i, j);
  if (isnan(zc)) {
    store_out(out, r, c, apply_post(pp, __builtin_nanf("")));
    return;
  }
  // first derivatives (|step|)
  const float p = d_col<NANAW>(W, K.ax, i, j, G);  // dx
  const float q = d_row<NANAW>(W, K.ay, i, j, G);  // dy
  // second derivatives (signed steps) of dx and dy
  // dyy, dyx = gradient(dy, step_y, step_x); dxy, dxx = gradient(dx, ...)
  auto dx_along_col = [&](int k) { return d_col<NANAW>(W, K.ax, i, k, G); };
  auto dx_along_row = [&](int k) { return d_col<NANAW>(W, K.ax, k, j, G); };
  auto dy_along_row = [&](int k) { return d_row<NANAW>(W, K.ay, k, j, G); };
  auto dy_along_col = [&](int k) { return d_row<NANAW>(W, K.ay, i, k, G); };
  const float dxx = deriv(K.sx, j, W.w, dx_along_col);
  const float dxy = deriv(K.sy, i, W.h, dx_along_row);
  const float dyy = deriv(K.sy, i, W.h, dy_along_row);
  const float dyx = deriv(K.sx, j, W.w, dy_along_col);
  float cv;
  const float one = 1.0f, two = 2.0f;
  if (K.type == 0) {
    const float s = (dxy + dyx) / two;
    const float den = powf(one + p * p + q * q, 1.5f);
    const float num = (one + q * q) * dxx - two * p * q * s + (one + p * p) * dyy;
    cv = -num / (two * den + 1e-10f);
  } else if (K.type == 1) {
    const float b = one + p * p + q * q;
    cv = (dxx * dyy - dxy * dxy) / (b * b);
  } else if (K.type == 2) {
    cv = (-(q * q * dxx - two * p * q * dxy + p * p * dyy)) /
         (powf(p * p + q * q, 1.5f) + 1e-10f);
  } else {
    cv = (-(p * p * dxx + two * p * q * dxy + q * q * dyy)) /
         ((p * p + q * q) * powf(one + p * p + q * q, 1.5f) + 1e-10f);
  }
  const float cn = tanhf(cv * 100.0f);
  float res = (cn + one) / two;
  res = powf(res, K.gamma);
  store_out(out, r, c, apply_post(pp, res));
}

Gauss9 make_gauss9() {
  double w[9];
  gaussian_weights(1.0, 4, w);
  Gauss9 G;
  for (int d = 0; d <= 4; d++) G.g[d] = w[4 + d];
  return G;
}

}  // namespace fsg

using namespace fsg;

extern "C" int fsg_hillshade_slope(const fsg_window* win, const fsg_region* reg,
                                   const fsg_shade_params* p, const fsg_output* hillshade,
                                   const fsg_output* slope, void* stream) {
  int rc = validate_window(win, "fsg_hillshade_slope");
  if (rc) return rc;
  rc = validate_region(win, reg, "fsg_hillshade_slope");
  if (rc) return rc;
  if (!p) return fail(FSG_E_INVALID, "fsg_hillshade_slope: params NULL");
  if ((rc = validate_output(hillshade, reg, "fsg_hillshade_slope"))) return rc;
  if ((rc = validate_output(slope, reg, "fsg_hillshade_slope"))) return rc;
  if (win->h < 3 || win->w < 3)
    return fail(FSG_E_INVALID,
                "fsg_hillshade_slope: Shape of array too small to calculate a numerical "
                "gradient, at least (edge_order + 1) elements are required.");
  if (!(p->step_x > 0) || !(p->step_y > 0))
    return fail(FSG_E_INVALID, "fsg_hillshade_slope: steps must be positive");
  if (!hillshade && !slope) return FSG_OK;
  if (reg->h == 0 || reg->w == 0) return FSG_OK;
  ShadeK K;
  K.ax = make_axis(p->step_x);
  K.ay = make_axis(p->step_y);
  K.zf = (float)p->z_factor;
  K.z_is_one = (K.zf == 1.0f);
  K.sgx = (float)p->sign_x;
  K.sgy = (float)p->sign_y;
  const double alt = p->altitude_deg * (M_PI / 180.0);
  const double az = p->azimuth_deg * (M_PI / 180.0);
  K.lx = std::sin(az) * std::cos(alt);
  K.ly = std::cos(az) * std::cos(alt);
  K.lz = std::sin(alt);
  K.unit = p->slope_unit;
  K.need_hs = hillshade != nullptr;
  K.need_sl = slope != nullptr;
  WinDev W{win->data, win->ld, win->h, win->w};
  dim3 blk(64, 4), grd((reg->w + 63) / 64, (reg->h + 3) / 4);
  TimedScope ts("k_hillshade_slope", as_stream(stream));
  if (p->nan_hint != 0)
    hipLaunchKernelGGL(k_hillshade_slope<true>, grd, blk, 0, as_stream(stream), W, reg->r0,
                       reg->c0, reg->h, reg->w, K, make_gauss9(), make_out(hillshade),
                       make_out(slope));
  else
    hipLaunchKernelGGL(k_hillshade_slope<false>, grd, blk, 0, as_stream(stream), W, reg->r0,
                       reg->c0, reg->h, reg->w, K, make_gauss9(), make_out(hillshade),
                       make_out(slope));
  FSG_CHECK_LAUNCH("k_hillshade_slope");
  return FSG_OK;
}

extern "C" int fsg_curvature(const fsg_window* win, const fsg_region* reg,
                             const fsg_curvature_params* p, const fsg_output* out,
                             void* stream) {
  int rc = validate_window(win, "fsg_curvature");
  if (rc) return rc;
  rc = validate_region(win, reg, "fsg_curvature");
  if (rc) return rc;
  if (!p || !out) return fail(FSG_E_INVALID, "fsg_curvature: params/output NULL");
  if ((rc = validate_output(out, reg, "fsg_curvature"))) return rc;
  if (win->h < 3 || win->w < 3)
    return fail(FSG_E_INVALID,
                "fsg_curvature: Shape of array too small to calculate a numerical "
                "gradient, at least (edge_order + 1) elements are required.");
  if (p->curvature_type < 0 || p->curvature_type > 3)
    return fail(FSG_E_INVALID, "fsg_curvature: bad curvature_type");
  if (p->step_x == 0.0 || p->step_y == 0.0)
    return fail(FSG_E_INVALID, "fsg_curvature: zero step");
  if (reg->h == 0 || reg->w == 0) return FSG_OK;
  CurvK K;
  K.ax = make_axis(std::fabs(p->step_x));
  K.ay = make_axis(std::fabs(p->step_y));
  K.sx = make_axis(p->step_x);
  K.sy = make_axis(p->step_y);
  K.type = p->curvature_type;
  K.gamma = (float)(1.0 / 2.2);
  WinDev W{win->data, win->ld, win->h, win->w};
  dim3 blk(64, 4), grd((reg->w + 63) / 64, (reg->h + 3) / 4);
  TimedScope ts("k_curvature", as_stream(stream));
  if (p->nan_hint != 0)
    hipLaunchKernelGGL(k_curvature<true>, grd, blk, 0, as_stream(stream), W, reg->r0, reg->c0,
                       reg->h, reg->w, K, make_gauss9(), make_out(out));
  else
    hipLaunchKernelGGL(k_curvature<false>, grd, blk, 0, as_stream(stream), W, reg->r0, reg->c0,
                       reg->h, reg->w, K, make_gauss9(), make_out(out));
  FSG_CHECK_LAUNCH("k_curvature");
  return FSG_OK;
}

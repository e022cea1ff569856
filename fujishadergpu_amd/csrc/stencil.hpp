// Gradient-stencil device code shared by the per-window kernels
// (stencil.hip) and the batched fused launch (batch.hip).
//
// Derivatives are numpy/cupy gradient(f, h_row, h_col, edge_order=2) on the
// window: interior (f[i+1] - f[i-1]) / f32(2h), window edges
// a*f0 + b*f1 + c*f2 with f32 coefficients (-1.5/h, 2/h, -0.5/h) and the
// mirror (0.5/h, -2/h, 1.5/h), every operation rounded in f32.  NaN cells are
// replaced by the NaN-aware sigma=1 Gaussian fill of handle_nan_for_gradient
// (_nan_utils.py:50-74), evaluated on demand in scipy's two-pass order.
//
// The per-pixel functions take an accessor f(r, c) returning the filled
// value at window coordinates, so a kernel reading global memory and one
// reading an LDS tile run the same float32 operations in the same order.
#pragma once
#include <cmath>

#include "fsg_common.hpp"

namespace fsg {

__device__ __forceinline__ float raw(const WinDev& W, int r, int c) {
  return W.p[(int64_t)r * W.ld + c];
}

// NaN-aware local fill at (r, c): scipy gaussian_filter(mode='nearest') of
// filled*valid and of valid, axis 0 then axis 1 with float32 intermediates,
// symmetric accumulation order (centre, then distance 4..1), then
// where(sw > 0, sv / sw, 0) in float32 (_nan_utils.py:18-31).
// axis-0 pass of that fill at column cc: (sv, sw) rounded to float32.
// ROLLED keeps the nine loads in a loop (the per-cell fill is rare, and its
// registers would count against every kernel that inlines it); the unrolled
// form issues them together.
template <bool ROLLED = true>
__device__ __forceinline__ void nan_fill_col(const WinDev& W, int r, int cc, const Gauss9& G,
                                             float& fv, float& fw) {
  const float x0 = raw(W, r, cc);
  const bool v0 = !isnan(x0);
  double sv = (v0 ? (double)x0 : 0.0) * G.g[0];
  double sw = (v0 ? 1.0 : 0.0) * G.g[0];
  auto step = [&](int d) {
    const float a = raw(W, clampi(r - d, W.h), cc);
    const float b = raw(W, clampi(r + d, W.h), cc);
    const double fa = isnan(a) ? 0.0 : (double)a, fb = isnan(b) ? 0.0 : (double)b;
    const double wa = isnan(a) ? 0.0 : 1.0, wb = isnan(b) ? 0.0 : 1.0;
    sv += (fa + fb) * G.g[d];
    sw += (wa + wb) * G.g[d];
  };
  if (ROLLED) {
#pragma unroll 1
    for (int d = 4; d >= 1; d--) step(d);
  } else {
#pragma unroll
    for (int d = 4; d >= 1; d--) step(d);
  }
  fv = (float)sv;
  fw = (float)sw;
}

// NaN-aware local fill at (r, c): scipy gaussian_filter(mode='nearest') of
// filled*valid and of valid, axis 0 then axis 1 with float32 intermediates,
// symmetric accumulation order (centre, then distance 4..1), then
// where(sw > 0, sv / sw, 0) in float32 (_nan_utils.py:18-31).  The axis-1
// pass takes the column pairs in that order, so only two columns are live.
static __device__ __forceinline__ float nan_fill(const WinDev& W, int r, int c, const Gauss9& G) {
  float cv, cw;
  nan_fill_col(W, r, clampi(c, W.w), G, cv, cw);
  double sv = (double)cv * G.g[0];
  double sw = (double)cw * G.g[0];
#pragma unroll 1
  for (int d = 4; d >= 1; d--) {
    float av, aw, bv, bw;
    nan_fill_col(W, r, clampi(c - d, W.w), G, av, aw);
    nan_fill_col(W, r, clampi(c + d, W.w), G, bv, bw);
    sv += ((double)av + (double)bv) * G.g[d];
    sw += ((double)aw + (double)bw) * G.g[d];
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

__host__ __device__ inline Axis make_axis(double h) {
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

struct CurvK {
  Axis ax, ay;    // |step| (first derivatives)
  Axis sx, sy;    // signed steps (second derivatives)
  int type;
  float gamma;
};

// Kernel constants from the C-ABI parameters (host side).
inline ShadeK make_shadek(const fsg_shade_params* p, bool need_hs, bool need_sl) {
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
  K.need_hs = need_hs;
  K.need_sl = need_sl;
  return K;
}

inline CurvK make_curvk(const fsg_curvature_params* p) {
  CurvK K;
  K.ax = make_axis(std::fabs(p->step_x));
  K.ay = make_axis(std::fabs(p->step_y));
  K.sx = make_axis(p->step_x);
  K.sy = make_axis(p->step_y);
  K.type = p->curvature_type;
  K.gamma = (float)(1.0 / 2.2);
  return K;
}

// Gradient (d/dcol, d/drow) at window (i, j) of an h x w window.
template <class F>
__device__ __forceinline__ void grad_at(const Axis& ax, const Axis& ay, int i, int j, int h,
                                        int w, F f, float& dx, float& dy) {
  dy = deriv(ay, i, h, [&](int k) { return f(k, j); });
  dx = deriv(ax, j, w, [&](int k) { return f(i, k); });
}

// compute_slope_block's unit conversion (_impl_slope.py:23-33).  The
// gradient magnitude takes the hardware square root (one ulp; slope is
// checked to a tolerance, and the correctly rounded expansion costs ~15
// instructions more per pixel).
__device__ __forceinline__ float slope_from(const ShadeK& K, float dx, float dy) {
  const float m = __builtin_amdgcn_sqrtf(dx * dx + dy * dy);
  const float rad = atanf(m);
  if (K.unit == 0) return (float)((double)rad * (180.0 / M_PI));
  if (K.unit == 1) return tanf(rad) * 100.0f;
  return rad;
}

// The shading from the z-scaled gradient (hdx, hdy): surface normal against
// the float64 light vector, clipped to [0, 1] (_impl_hillshade.py:20-54).
__device__ __forceinline__ float shade_from_grad(const ShadeK& K, float hdx, float hdy) {
  const float e = hdx * K.sgx;
  const float n = hdy * K.sgy;
  const float norm = sqrtf(e * e + n * n + 1.0f);
  // float64 from here: the light vector is a 0-d float64 array.
  double hs = (double)(-e) * K.lx;
  hs = hs + (double)(-n) * K.ly;
  hs = hs + K.lz;
  hs = hs / (double)norm;
  hs = fmin(fmax(hs, 0.0), 1.0);
  return (float)hs;
}

// compute_hillshade_block's shading from the (unscaled) gradient; a z factor
// other than 1 re-derives from the scaled values as the reference does.
template <class F>
__device__ __forceinline__ float hillshade_from(const ShadeK& K, const Axis& axr, int i, int j,
                                                int h, int w, F f, float dx, float dy) {
  float hdy = dy, hdx = dx;
  if (!K.z_is_one) {
    hdy = deriv(K.ay, i, h, [&](int k) { return f(k, j) * K.zf; });
    hdx = deriv(axr, j, w, [&](int k) { return f(i, k) * K.zf; });
  }
  return shade_from_grad(K, hdx, hdy);
}

__device__ __forceinline__ float curvature_from(const CurvK& K, float p, float q, float dxx,
                                                float dxy, float dyy, float dyx);

// compute_curvature_block (_impl_curvature.py:19-57) at window (i, j):
// first derivatives with |step|, second derivatives of those with the signed
// steps, the curvature formula, tanh(100 c), (x + 1) / 2, gamma.
template <class F>
__device__ __forceinline__ float curvature_at(const CurvK& K, int i, int j, int h, int w, F f) {
  const float p = deriv(K.ax, j, w, [&](int k) { return f(i, k); });  // dx
  const float q = deriv(K.ay, i, h, [&](int k) { return f(k, j); });  // dy
  // dyy, dyx = gradient(dy, step_y, step_x); dxy, dxx = gradient(dx, ...)
  const float dxx = deriv(K.sx, j, w, [&](int k) {
    return deriv(K.ax, k, w, [&](int m) { return f(i, m); });
  });
  const float dxy = deriv(K.sy, i, h, [&](int k) {
    return deriv(K.ax, j, w, [&](int m) { return f(k, m); });
  });
  const float dyy = deriv(K.sy, i, h, [&](int k) {
    return deriv(K.ay, k, h, [&](int m) { return f(m, j); });
  });
  const float dyx = deriv(K.sx, j, w, [&](int k) {
    return deriv(K.ay, i, h, [&](int m) { return f(m, k); });
  });
  return curvature_from(K, p, q, dxx, dxy, dyy, dyx);
}

// x^1.5 and a / b from the hardware square root and reciprocal (about one
// ulp each).
__device__ __forceinline__ float pow15(float x) { return x * __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// The curvature formula, tanh(100 c), (x + 1) / 2 and gamma from the first
// (p = d/dcol, q = d/drow) and second derivatives (compute_curvature_block,
// _impl_curvature.py:35-57).  The arithmetic is float32 like the reference's,
// with hardware transcendentals: x^1.5 = x sqrt(x); tanh(y) as
// 2 / (1 + e^(-2y)) - 1; res^gamma as 2^(gamma log2 res).  A few float32 ulps
// from the reference's libm values (the parity tests' display-value bound is
// 4e-4, one saturated tanh ulp).
__device__ __forceinline__ float curvature_from(const CurvK& K, float p, float q, float dxx,
                                                float dxy, float dyy, float dyx) {
  float cv;
  const float one = 1.0f, two = 2.0f;
  if (K.type == 0) {
    const float s = (dxy + dyx) / two;
    const float den = pow15(one + p * p + q * q);
    const float num = (one + q * q) * dxx - two * p * q * s + (one + p * p) * dyy;
    cv = fdiv(-num, two * den + 1e-10f);
  } else if (K.type == 1) {
    const float b = one + p * p + q * q;
    cv = fdiv(dxx * dyy - dxy * dxy, b * b);
  } else if (K.type == 2) {
    cv = fdiv(-(q * q * dxx - two * p * q * dxy + p * p * dyy), pow15(p * p + q * q) + 1e-10f);
  } else {
    cv = fdiv(-(p * p * dxx + two * p * q * dxy + q * q * dyy),
              (p * p + q * q) * pow15(one + p * p + q * q) + 1e-10f);
  }
  // tanh(100 cv) = 2 / (1 + e^(-200 cv)) - 1, rounded to float32 like the
  // reference's tanh (its saturated tail rounds to -1 + k ulp), then
  // (tanh + 1) / 2 as the reference
  const float e = __expf(-200.0f * cv);
  const float cn = two * __builtin_amdgcn_rcpf(one + e) - one;
  const float res = (cn + one) / two;
  return __builtin_amdgcn_exp2f(K.gamma * __builtin_amdgcn_logf(res));
}

// ------------------------------------------------------------ tiled block
// One 64 x 16 block of a window region: hillshade, slope and curvature from
// one staged tile.  Pass 1 stages the filled values (3-pixel halo, clipped to
// the window); pass 2 computes the first derivatives P = d/dcol and Q = d/drow
// once per position over the block plus a 2-pixel ring; pass 3 forms the
// second derivatives from those tables.  Every value equals the one the
// nested per-pixel formulas give (a first derivative at a position does not
// depend on which pixel asks for it), so the outputs are the same bits.
constexpr int ST_TW = 64, ST_TH = 16, ST_HALO = 3, ST_RING = 2;
constexpr int ST_LT = ST_TW + 2 * ST_HALO + 1;      // staged row stride
constexpr int ST_ROWS = ST_TH + 2 * ST_HALO;
constexpr int ST_PW = ST_TW + 2 * ST_RING;          // derivative table width
constexpr int ST_PH = ST_TH + 2 * ST_RING;
constexpr int ST_AXIS_FLOATS = 8;                    // one Axis, padded
constexpr int ST_LDS_FLOATS = ST_ROWS * ST_LT + 2 * ST_PH * ST_PW;
// with per-row x coefficient tables (|step| and signed) for the staged rows
constexpr int ST_LDS_FLOATS_ROWS = ST_LDS_FLOATS + 2 * ST_ROWS * ST_AXIS_FLOATS;

struct StencilTile {
  WinDev W;
  const double* xrows;  // per-row signed x steps of the window, or NULL
  int r0, c0, rh, rw;
  int nbx;         // blocks per block row of the region
  int nanaw;
  int want_hs, want_sl, want_cv;
  int shade_from_tables;  // with curvature: hillshade/slope gradient equals the tables' (same |steps|)
  ShadeK shade;
  CurvK curv;
  OutDev hs, sl, cv;
};

inline bool same_axis(const Axis& a, const Axis& b) {
  return a.two_h == b.two_h && a.a0 == b.a0 && a.b0 == b.b0 && a.c0 == b.c0 && a.a1 == b.a1 &&
         a.b1 == b.b1 && a.c1 == b.c1;
}

// ROWS: the window has per-row x steps (S.xrows); the variant without them
// keeps the constant coefficients in registers.
template <bool NANAW, bool ROWS>
__device__ __forceinline__ void stencil_tile_block_v(const StencilTile& S, int by, int bx,
                                                     const Gauss9& G, float* lds) {
  const int rl = S.r0 + by * ST_TH, cl = S.c0 + bx * ST_TW;  // window coords of the block
  const int H = S.W.h, Wd = S.W.w;
  const int tr0 = max(rl - ST_HALO, 0), tr1 = min(rl + ST_TH + ST_HALO, H);
  const int tc0 = max(cl - ST_HALO, 0), tc1 = min(cl + ST_TW + ST_HALO, Wd);
  const int nr = tr1 - tr0, nc = tc1 - tc0;
  float* F = lds;
  float* Pt = lds + ST_ROWS * ST_LT;
  float* Qt = Pt + ST_PH * ST_PW;
  Axis* Xa = (Axis*)(Qt + ST_PH * ST_PW);     // per staged row: |x step| coefficients
  Axis* Xs = (Axis*)((float*)Xa + ST_ROWS * ST_AXIS_FLOATS);  // signed x step
  const int t = threadIdx.x;
  {
    // lanes 0..63 of each wave walk columns, the 4 waves walk rows; all of a
    // lane's loads are issued before any is used
    constexpr int NRW = (ST_ROWS + 3) / 4, NCW = (ST_TW + 2 * ST_HALO + 63) / 64;
    float v[NRW][NCW];
#pragma unroll
    for (int i = 0; i < NRW; i++)
#pragma unroll
      for (int j = 0; j < NCW; j++) {
        const int rr = (t >> 6) + 4 * i, cc = (t & 63) + 64 * j;
        v[i][j] = (rr < nr && cc < nc) ? raw(S.W, tr0 + rr, tc0 + cc) : 0.0f;
      }
#pragma unroll
    for (int i = 0; i < NRW; i++)
#pragma unroll
      for (int j = 0; j < NCW; j++) {
        const int rr = (t >> 6) + 4 * i, cc = (t & 63) + 64 * j;
        if (rr < nr && cc < nc) F[rr * ST_LT + cc] = v[i][j];
      }
    if (NANAW) {
      // NaN cells (rare) take the local fill; one call site, the lane's own cells
#pragma unroll 1
      for (int e = 0; e < NRW * NCW; e++) {
        const int rr = (t >> 6) + 4 * (e / NCW), cc = (t & 63) + 64 * (e % NCW);
        if (rr < nr && cc < nc && isnan(F[rr * ST_LT + cc]))
          F[rr * ST_LT + cc] = nan_fill(S.W, tr0 + rr, tc0 + cc, G);
      }
    }
  }
  if (ROWS && t < nr) {
    const double hx = S.xrows[tr0 + t];
    *(Axis*)((float*)Xa + t * ST_AXIS_FLOATS) = make_axis(fabs(hx));
    *(Axis*)((float*)Xs + t * ST_AXIS_FLOATS) = make_axis(hx);
  }
  __syncthreads();
  auto f = [&](int r, int c) { return F[(r - tr0) * ST_LT + (c - tc0)]; };
  // first-derivative tables over rows rl-2 .. rl+TH+1, cols cl-2 .. cl+TW+1
  // the tables pay off for curvature's second derivatives; hillshade/slope
  // alone take their gradient straight from the staged tile
  const bool need_tab = S.want_cv;
  const Axis ax = S.want_cv ? S.curv.ax : S.shade.ax;
  const Axis ay = S.want_cv ? S.curv.ay : S.shade.ay;
  // x-derivative coefficients of window row i: the row's own step, or the constant one
  auto axr = [&](int i, const Axis& c) -> Axis {
    return ROWS ? *(const Axis*)((const float*)Xa + (i - tr0) * ST_AXIS_FLOATS) : c;
  };
  auto sxr = [&](int i, const Axis& c) -> Axis {
    return ROWS ? *(const Axis*)((const float*)Xs + (i - tr0) * ST_AXIS_FLOATS) : c;
  };
  if (need_tab) {
    for (int e = t; e < ST_PH * ST_PW; e += 256) {
      const int pr = e / ST_PW, pc = e - pr * ST_PW;
      const int i = rl - ST_RING + pr, j = cl - ST_RING + pc;
      if (i >= 0 && i < H && j >= 0 && j < Wd) {
        Pt[e] = deriv(axr(i, ax), j, Wd, [&](int k) { return f(i, k); });
        Qt[e] = deriv(ay, i, H, [&](int k) { return f(k, j); });
      }
    }
  }
  __syncthreads();
  auto P = [&](int i, int j) { return Pt[(i - rl + ST_RING) * ST_PW + (j - cl + ST_RING)]; };
  auto Q = [&](int i, int j) { return Qt[(i - rl + ST_RING) * ST_PW + (j - cl + ST_RING)]; };
  const int c = t & 63;
  const PostK phs = load_post(S.hs), psl = load_post(S.sl), pcv = load_post(S.cv);
  for (int k = t >> 6; k < ST_TH; k += 4) {
    const int rr = by * ST_TH + k, cc = bx * ST_TW + c;  // region coordinates
    if (rr < S.rh && cc < S.rw) {
      const int wr = S.r0 + rr, wc = S.c0 + cc;
      const float z = NANAW ? raw(S.W, wr, wc) : f(wr, wc);
      if (isnan(z)) {
        const float qn = __builtin_nanf("");
        if (S.want_hs) store_out(S.hs, rr, cc, apply_post(phs, qn));
        if (S.want_sl) store_out(S.sl, rr, cc, apply_post(psl, qn));
        if (S.want_cv) store_out(S.cv, rr, cc, apply_post(pcv, qn));
        continue;
      }
      if (S.want_hs || S.want_sl) {
        float dx, dy;
        const Axis sax = axr(wr, S.shade.ax);
        if (S.want_cv && S.shade_from_tables) {
          dx = P(wr, wc);
          dy = Q(wr, wc);
        } else {
          grad_at(sax, S.shade.ay, wr, wc, H, Wd, f, dx, dy);
        }
        if (S.want_sl) store_out(S.sl, rr, cc, apply_post(psl, slope_from(S.shade, dx, dy)));
        if (S.want_hs)
          store_out(S.hs, rr, cc,
                    apply_post(phs, hillshade_from(S.shade, sax, wr, wc, H, Wd, f, dx, dy)));
      }
      if (S.want_cv) {
        const CurvK& K = S.curv;
        const Axis ksx = sxr(wr, K.sx);
        const float p = P(wr, wc), q = Q(wr, wc);
        const float dxx = deriv(ksx, wc, Wd, [&](int m) { return P(wr, m); });
        const float dxy = deriv(K.sy, wr, H, [&](int m) { return P(m, wc); });
        const float dyy = deriv(K.sy, wr, H, [&](int m) { return Q(m, wc); });
        const float dyx = deriv(ksx, wc, Wd, [&](int m) { return Q(wr, m); });
        store_out(S.cv, rr, cc, apply_post(pcv, curvature_from(K, p, q, dxx, dxy, dyy, dyx)));
      }
    }
  }
  __syncthreads();
}

template <bool NANAW>
__device__ __forceinline__ void stencil_tile_block(const StencilTile& S, int by, int bx,
                                                   const Gauss9& G, float* lds) {
  if (S.xrows)
    stencil_tile_block_v<NANAW, true>(S, by, bx, G, lds);
  else
    stencil_tile_block_v<NANAW, false>(S, by, bx, G, lds);
}

}  // namespace fsg

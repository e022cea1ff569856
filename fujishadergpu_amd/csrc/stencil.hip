// Gradient stencils: fused hillshade + slope, and curvature, one window per
// launch.  The arithmetic lives in stencil.hpp.
#include <algorithm>
#include <cmath>

#include "stencil.hpp"

namespace fsg {

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
  auto f = [&](int i, int j) { return filled<NANAW>(W, i, j, G); };
  float dx, dy;
  grad_at(K.ax, K.ay, wr, wc, W.h, W.w, f, dx, dy);
  if (K.need_sl) store_out(osl, r, c, apply_post(psl, slope_from(K, dx, dy)));
  if (K.need_hs)
    store_out(ohs, r, c, apply_post(phs, hillshade_from(K, wr, wc, W.h, W.w, f, dx, dy)));
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
  auto f = [&](int a, int b) { return filled<NANAW>(W, a, b, G); };
  store_out(out, r, c, apply_post(pp, curvature_at(K, i, j, W.h, W.w, f)));
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
  const ShadeK K = make_shadek(p, hillshade != nullptr, slope != nullptr);
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
  const CurvK K = make_curvk(p);
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

// Gradient stencils: fused hillshade + slope, and curvature, one window per
// launch, one workgroup per 64 x 16 block; the block function and the
// arithmetic live in stencil.hpp (shared with the batched launch).
#include <algorithm>
#include <cmath>

#include "stencil.hpp"

namespace fsg {

// One workgroup per 64 x 16 block of the region (stencil.hpp); ROWS: the
// window has per-row x steps (larger LDS for the coefficient tables).
template <bool NANAW, bool ROWS>
__global__ void __launch_bounds__(256) k_stencil_tiles(StencilTile S, Gauss9 G) {
  __shared__ float lds[ROWS ? ST_LDS_FLOATS_ROWS : ST_LDS_FLOATS];
  stencil_tile_block_v<NANAW, ROWS>(S, blockIdx.y, blockIdx.x, G, lds);
}

// Hillshade and/or slope without curvature and with constant steps: the
// gradient read straight from global memory (L1 serves the neighbours); the
// same per-pixel operations as the tiled block (deriv's central difference
// inside the window, deriv itself at its edges), so the same bits, without
// its LDS staging and barriers.
//
// NaN-free windows: each lane walks SHADE_R rows of one column, the column's
// values one row ahead in registers.
constexpr int SHADE_R = 8;

template <bool NANAW>
__global__ void __launch_bounds__(256) k_shade_rows(StencilTile S, Gauss9 G) {
  const int cc = blockIdx.x * 64 + (threadIdx.x & 63);
  const int rr0 = (blockIdx.y * 4 + (threadIdx.x >> 6)) * SHADE_R;
  if (rr0 >= S.rh || cc >= S.rw) return;
  const int H = S.W.h, Wd = S.W.w;
  const int wc = S.c0 + cc;
  const PostK phs = load_post(S.hs), psl = load_post(S.sl);
  auto f = [&](int r, int c) { return filled<NANAW>(S.W, r, c, G); };
  const int n = min(SHADE_R, S.rh - rr0);
  const int wr0 = S.r0 + rr0;
  const bool col_in = wc > 0 && wc < Wd - 1;
  float up = wr0 > 0 ? raw(S.W, wr0 - 1, wc) : 0.0f;
  float mid = raw(S.W, wr0, wc);
  for (int k = 0; k < n; k++) {
    const int rr = rr0 + k, wr = wr0 + k;
    const float down = wr + 1 < H ? raw(S.W, wr + 1, wc) : 0.0f;
    if (isnan(mid)) {
      const float qn = __builtin_nanf("");
      if (S.want_hs) store_out(S.hs, rr, cc, apply_post(phs, qn));
      if (S.want_sl) store_out(S.sl, rr, cc, apply_post(psl, qn));
    } else {
      float dx, dy;
      if (col_in && wr > 0 && wr < H - 1) {
        // deriv's interior branch on the filled values
        const float fu = (NANAW && isnan(up)) ? f(wr - 1, wc) : up;
        const float fd = (NANAW && isnan(down)) ? f(wr + 1, wc) : down;
        dy = (fd - fu) / S.shade.ay.two_h;
        const float fr = f(wr, wc + 1), fl = f(wr, wc - 1);
        dx = (fr - fl) / S.shade.ax.two_h;
      } else {
        grad_at(S.shade.ax, S.shade.ay, wr, wc, H, Wd, f, dx, dy);
      }
      if (S.want_sl) store_out(S.sl, rr, cc, apply_post(psl, slope_from(S.shade, dx, dy)));
      if (S.want_hs)
        store_out(S.hs, rr, cc,
                  apply_post(phs, hillshade_from(S.shade, S.shade.ax, wr, wc, H, Wd, f, dx, dy)));
    }
    up = mid;
    mid = down;
  }
}

// Windows with NaN: one lane per pixel (the fill inlined at each neighbour
// read takes the registers that a row walk would need).
template <bool NANAW>
__global__ void __launch_bounds__(256) k_shade_px(StencilTile S, Gauss9 G) {
  const int cc = blockIdx.x * 64 + (threadIdx.x & 63);
  const int rr = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (rr >= S.rh || cc >= S.rw) return;
  const int H = S.W.h, Wd = S.W.w;
  const int wr = S.r0 + rr, wc = S.c0 + cc;
  const float z = raw(S.W, wr, wc);
  const PostK phs = load_post(S.hs), psl = load_post(S.sl);
  if (isnan(z)) {
    const float qn = __builtin_nanf("");
    if (S.want_hs) store_out(S.hs, rr, cc, apply_post(phs, qn));
    if (S.want_sl) store_out(S.sl, rr, cc, apply_post(psl, qn));
    return;
  }
  auto f = [&](int r, int c) { return filled<NANAW>(S.W, r, c, G); };
  float dx, dy;
  grad_at(S.shade.ax, S.shade.ay, wr, wc, H, Wd, f, dx, dy);
  if (S.want_sl) store_out(S.sl, rr, cc, apply_post(psl, slope_from(S.shade, dx, dy)));
  if (S.want_hs)
    store_out(S.hs, rr, cc,
              apply_post(phs, hillshade_from(S.shade, S.shade.ax, wr, wc, H, Wd, f, dx, dy)));
}

static int launch_stencil(const StencilTile& S, const char* name, hipStream_t s) {
  dim3 grd(S.nbx, (S.rh + ST_TH - 1) / ST_TH);
  TimedScope ts(name, s);
  if (S.xrows) {
    if (S.nanaw)
      hipLaunchKernelGGL((k_stencil_tiles<true, true>), grd, dim3(256), 0, s, S, make_gauss9());
    else
      hipLaunchKernelGGL((k_stencil_tiles<false, true>), grd, dim3(256), 0, s, S, make_gauss9());
  } else {
    if (S.nanaw)
      hipLaunchKernelGGL((k_stencil_tiles<true, false>), grd, dim3(256), 0, s, S, make_gauss9());
    else
      hipLaunchKernelGGL((k_stencil_tiles<false, false>), grd, dim3(256), 0, s, S, make_gauss9());
  }
  FSG_CHECK_LAUNCH(name);
  return FSG_OK;
}

static StencilTile stencil_tile(const fsg_window* win, const fsg_region* reg) {
  StencilTile S{};
  S.W = WinDev{win->data, win->ld, win->h, win->w};
  S.r0 = reg->r0;
  S.c0 = reg->c0;
  S.rh = reg->h;
  S.rw = reg->w;
  S.nbx = (reg->w + ST_TW - 1) / ST_TW;
  return S;
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
  StencilTile S = stencil_tile(win, reg);
  S.shade = make_shadek(p, hillshade != nullptr, slope != nullptr);
  S.want_hs = hillshade != nullptr;
  S.want_sl = slope != nullptr;
  S.shade_from_tables = 1;
  S.xrows = p->step_x_rows;
  S.nanaw = p->nan_hint != 0;
  if (hillshade) S.hs = make_out(hillshade);
  if (slope) S.sl = make_out(slope);
  if (!S.xrows) {
    hipStream_t s = as_stream(stream);
    TimedScope ts("k_hillshade_slope", s);
    if (S.nanaw) {
      dim3 grd((reg->w + 63) / 64, (reg->h + 3) / 4);
      hipLaunchKernelGGL(k_shade_px<true>, grd, dim3(256), 0, s, S, make_gauss9());
    } else {
      dim3 grd((reg->w + 63) / 64, (reg->h + 4 * SHADE_R - 1) / (4 * SHADE_R));
      hipLaunchKernelGGL(k_shade_rows<false>, grd, dim3(256), 0, s, S, make_gauss9());
    }
    FSG_CHECK_LAUNCH("k_shade_direct");
    return FSG_OK;
  }
  return launch_stencil(S, "k_hillshade_slope", as_stream(stream));
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
  StencilTile S = stencil_tile(win, reg);
  S.curv = make_curvk(p);
  S.want_cv = 1;
  S.xrows = p->step_x_rows;
  S.nanaw = p->nan_hint != 0;
  S.cv = make_out(out);
  return launch_stencil(S, "k_curvature", as_stream(stream));
}

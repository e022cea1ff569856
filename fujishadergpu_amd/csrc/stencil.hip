// Gradient stencils: fused hillshade + slope, and curvature, one window per
// launch, one workgroup per 64 x 16 block; the block function and the
// arithmetic live in stencil.hpp (shared with the batched launch).
#include <algorithm>
#include <atomic>
#include <cmath>

#include "stencil.hpp"
#include "stencil_plan.hpp"

namespace fsg {

// The general kernels run over up to six rectangles of the region in one
// launch (the region itself, or the window's border strips and the interior
// columns the four-column kernel leaves): a 1-D grid, blocks numbered rect by
// rect.
constexpr int RS_MAX = 6;
struct RectSet {
  int n;
  int r0[RS_MAX], c0[RS_MAX], h[RS_MAX], w[RS_MAX];  // window coordinates
  int nbx[RS_MAX];
  int pre[RS_MAX + 1];  // blocks before each rectangle
};

// This block's rectangle as a stencil tile of its own, and the block's
// position in it; false past the last rectangle's blocks.
__device__ __forceinline__ bool rect_block(const StencilTile& S, const RectSet& R,
                                           StencilTile& T, int& bx, int& by) {
  int b = blockIdx.x;
  if (b >= R.pre[R.n]) return false;
  int k = 0;
  while (k + 1 < R.n && b >= R.pre[k + 1]) k++;
  b -= R.pre[k];
  by = b / R.nbx[k];
  bx = b - by * R.nbx[k];
  T = S;
  T.r0 = R.r0[k];
  T.c0 = R.c0[k];
  T.rh = R.h[k];
  T.rw = R.w[k];
  T.nbx = R.nbx[k];
  const int dr = R.r0[k] - S.r0, dc = R.c0[k] - S.c0;
  T.hs = out_shift(S.hs, dr, dc);
  T.sl = out_shift(S.sl, dr, dc);
  T.cv = out_shift(S.cv, dr, dc);
  return true;
}

// One workgroup per 64 x 16 block (stencil.hpp); ROWS: the window has
// per-row x steps (larger LDS for the coefficient tables).
template <bool NANAW, bool ROWS>
__global__ void __launch_bounds__(256) k_stencil_tiles(StencilTile S, RectSet R, Gauss9 G) {
  __shared__ float lds[ROWS ? ST_LDS_FLOATS_ROWS : ST_LDS_FLOATS];
  StencilTile T;
  int bx, by;
  if (!rect_block(S, R, T, bx, by)) return;  // the whole block
  stencil_tile_block_v<NANAW, ROWS>(T, by, bx, G, lds);
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
__global__ void __launch_bounds__(256) k_shade_rows(StencilTile S0, RectSet R, Gauss9 G) {
  StencilTile S;
  int bx, by;
  if (!rect_block(S0, R, S, bx, by)) return;
  const int cc = bx * 64 + (threadIdx.x & 63);
  const int rr0 = (by * 4 + (threadIdx.x >> 6)) * SHADE_R;
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
__global__ void __launch_bounds__(256) k_shade_px(StencilTile S0, RectSet R, Gauss9 G) {
  StencilTile S;
  int bx, by;
  if (!rect_block(S0, R, S, bx, by)) return;
  const int cc = bx * 64 + (threadIdx.x & 63);
  const int rr = by * 4 + (threadIdx.x >> 6);
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

// The rectangle set of up to RS_MAX rectangles with blocks of bh rows.
static RectSet rect_set(const Rect* q, int n, int bh) {
  RectSet R{};
  R.n = 0;
  R.pre[0] = 0;
  for (int k = 0; k < n; k++) {
    if (q[k].h <= 0 || q[k].w <= 0) continue;
    const int i = R.n++;
    R.r0[i] = q[k].r0;
    R.c0[i] = q[k].c0;
    R.h[i] = q[k].h;
    R.w[i] = q[k].w;
    R.nbx[i] = (q[k].w + 63) / 64;
    R.pre[i + 1] = R.pre[i] + R.nbx[i] * ((q[k].h + bh - 1) / bh);
  }
  return R;
}

static int launch_stencil(const StencilTile& S, const char* name, hipStream_t s) {
  const Rect q{S.r0, S.c0, S.rh, S.rw};
  const RectSet R = rect_set(&q, 1, ST_TH);
  TimedScope ts(name, s);
  if (R.n == 0) return FSG_OK;
  dim3 grd(R.pre[R.n]);
  if (S.xrows) {
    if (S.nanaw)
      hipLaunchKernelGGL((k_stencil_tiles<true, true>), grd, dim3(256), 0, s, S, R, make_gauss9());
    else
      hipLaunchKernelGGL((k_stencil_tiles<false, true>), grd, dim3(256), 0, s, S, R, make_gauss9());
  } else {
    if (S.nanaw)
      hipLaunchKernelGGL((k_stencil_tiles<true, false>), grd, dim3(256), 0, s, S, R, make_gauss9());
    else
      hipLaunchKernelGGL((k_stencil_tiles<false, false>), grd, dim3(256), 0, s, S, R, make_gauss9());
  }
  FSG_CHECK_LAUNCH(name);
  return FSG_OK;
}

// The one-column launch over the interior rectangle: one wave per `cols`
// columns and `seg` rows, four waves (segments) per workgroup.
static bool rows_grid(const Rect& in, int cols, int seg, dim3& grd) {
  const int64_t nseg = ((int64_t)in.h + seg - 1) / seg;
  const int64_t gy = (nseg + 3) / 4, gx = ((int64_t)in.w + cols - 1) / cols;
  if (gy > 65535 || gx > 0x7fffffff) return false;
  grd = dim3((unsigned)gx, (unsigned)gy);
  return true;
}

static RowsTile rows_tile(const StencilTile& S, const fsg_region* reg, const Rect& in) {
  RowsTile R{};
  R.W = S.W;
  R.r0 = reg->r0;
  R.c0 = reg->c0;
  R.ir0 = in.r0;
  R.ir1 = in.r0 + in.h;
  R.ic0 = in.c0;
  R.ic1 = in.c0 + in.w;
  return R;
}

// The general kernels (every position, numpy's edge formulas included) over
// up to RS_MAX rectangles of the region, one launch: hillshade/slope (one
// lane per pixel with NaN, a row walk without) or curvature (tiled block).
static void launch_general(const StencilTile& S, const Rect* q, int n, bool curv,
                           hipStream_t s) {
  const int bh = curv ? ST_TH : (S.nanaw ? 4 : 4 * SHADE_R);
  const RectSet R = rect_set(q, n, bh);
  if (R.n == 0) return;
  const dim3 grd((unsigned)R.pre[R.n]);
  if (curv) {
    if (S.nanaw)
      hipLaunchKernelGGL((k_stencil_tiles<true, false>), grd, dim3(256), 0, s, S, R, make_gauss9());
    else
      hipLaunchKernelGGL((k_stencil_tiles<false, false>), grd, dim3(256), 0, s, S, R,
                         make_gauss9());
  } else if (S.nanaw) {
    hipLaunchKernelGGL(k_shade_px<true>, grd, dim3(256), 0, s, S, R, make_gauss9());
  } else {
    hipLaunchKernelGGL(k_shade_rows<false>, grd, dim3(256), 0, s, S, R, make_gauss9());
  }
}

// Interiors smaller than this take the general kernel alone: one launch
// instead of two, and no row-walking launch with too few rows per wave to
// hide its loads (2048^2 windows ran faster as one general launch).
constexpr int64_t SPLIT_MIN_DEFAULT = int64_t(1) << 23;
static std::atomic<int64_t> g_split_min{SPLIT_MIN_DEFAULT};
// NaN-aware windows: the general kernel's fill is far slower than the row
// kernels' wave fill, so their interiors take the row kernels from 2^20
// pixels (or the process-wide setting, when smaller).
constexpr int64_t SPLIT_MIN_NANAW = int64_t(1) << 20;

// Constant-step hillshade/slope or curvature: the row-walking kernel over
// the interior, then ONE general launch over the border strips (and the
// interior columns the four-column kernel leaves at the sides).  Small
// interiors, and divisors outside div_rn's range, take the general kernel
// over the whole region.
static void launch_split(const StencilTile& S, const fsg_window* win, const fsg_region* reg,
                         bool curv, hipStream_t s) {
  const Rect in = interior(win, reg, curv ? 2 : 1);
  RowsTile R = rows_tile(S, reg, in);
  const int cols1 = curv ? SR_CURV_COLS : SR_SHADE_COLS;
  R.seg = seg_rows(in.h, (in.w + cols1 - 1) / cols1);
  dim3 grd;
  const int64_t split_all = g_split_min.load(std::memory_order_relaxed);
  const int64_t split_min = S.nanaw ? std::min(SPLIT_MIN_NANAW, split_all) : split_all;
  bool ok = (int64_t)in.h * in.w >= split_min &&
            in.h > 0 && rows_grid(in, cols1, R.seg, grd);
  if (curv) {
    R.curv = S.curv;
    R.dx = make_divk(S.curv.ax.two_h);
    R.dy = make_divk(S.curv.ay.two_h);
    R.sx = make_divk(S.curv.sx.two_h);
    R.sy = make_divk(S.curv.sy.two_h);
    R.o0 = S.cv;
    ok = ok && fdiv_ok(R.sx.b) && fdiv_ok(R.sy.b);
  } else {
    R.shade = S.shade;
    R.dx = make_divk(S.shade.ax.two_h);
    R.dy = make_divk(S.shade.ay.two_h);
    R.want_hs = S.want_hs;
    R.want_sl = S.want_sl;
    R.o0 = S.hs;
    R.o1 = S.sl;
  }
  ok = ok && fdiv_ok(R.dx.b) && fdiv_ok(R.dy.b);
  Rect gen[RS_MAX];
  int ng = 0;
  if (ok) {
    // four columns per lane where whole waves fit (the interior columns left
    // at the sides go to the general launch), else one column per lane
    RowsTile4 S4;
    dim3 grd4;
    int oc0 = 0, oc1 = 0;
    const OutDev outs[2] = {R.o0, R.o1};
    const int ic1 = in.c0 + in.w;
    if (rows4_setup(R, in, outs, curv ? 1 : 2, curv ? SR4_LEFT_CURV : SR4_LEFT_SHADE, S4, grd4,
                    oc0, oc1)) {
      if (curv) {
        if (S.nanaw)
          hipLaunchKernelGGL(k_curv_inner4<true>, grd4, dim3(256), 0, s, S4, make_gauss9());
        else
          hipLaunchKernelGGL(k_curv_inner4<false>, grd4, dim3(256), 0, s, S4, make_gauss9());
      } else {
        if (S.nanaw)
          hipLaunchKernelGGL(k_shade_inner4<true>, grd4, dim3(256), 0, s, S4, make_gauss9());
        else
          hipLaunchKernelGGL(k_shade_inner4<false>, grd4, dim3(256), 0, s, S4, make_gauss9());
      }
      if (in.c0 < oc0) gen[ng++] = Rect{in.r0, in.c0, in.h, oc0 - in.c0};
      if (oc1 < ic1) gen[ng++] = Rect{in.r0, oc1, in.h, ic1 - oc1};
    } else if (curv) {
      if (S.nanaw)
        hipLaunchKernelGGL(k_curv_inner<true>, grd, dim3(256), 0, s, R, make_gauss9());
      else
        hipLaunchKernelGGL(k_curv_inner<false>, grd, dim3(256), 0, s, R, make_gauss9());
    } else {
      if (S.nanaw)
        hipLaunchKernelGGL(k_shade_inner<true>, grd, dim3(256), 0, s, R, make_gauss9());
      else
        hipLaunchKernelGGL(k_shade_inner<false>, grd, dim3(256), 0, s, R, make_gauss9());
    }
    ng += border_strips(reg, in, gen + ng);
  } else {
    gen[ng++] = Rect{reg->r0, reg->c0, reg->h, reg->w};
  }
  launch_general(S, gen, ng, curv, s);
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
    launch_split(S, win, reg, false, s);
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
  if (!S.xrows) {
    hipStream_t s = as_stream(stream);
    TimedScope ts("k_curvature", s);
    launch_split(S, win, reg, true, s);
    FSG_CHECK_LAUNCH("k_curvature");
    return FSG_OK;
  }
  return launch_stencil(S, "k_curvature", as_stream(stream));
}

extern "C" int64_t fsg_stencil_set_split_min(int64_t px) {
  return g_split_min.exchange(px < 0 ? SPLIT_MIN_DEFAULT : px, std::memory_order_relaxed);
}

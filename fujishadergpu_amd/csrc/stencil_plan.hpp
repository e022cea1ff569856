// Host-side planning of the stencils' interior split, shared by the
// per-window entry points (stencil.hip) and the batched launch (batch.hip):
// the window interior where every derivative is the central one, the border
// strips around it, and the four-columns-per-lane wave geometry.
#pragma once
#include <algorithm>
#include <cstdint>

#include "stencil_rows.hpp"

namespace fsg {

// The output descriptor moved to region coordinates (dr, dc).
__host__ __device__ inline OutDev out_shift(OutDev o, int dr, int dc) {
  if (!o.data) return o;
  const int64_t es = o.dtype == FSG_OUT_F32 ? 4 : (o.dtype == FSG_OUT_I16 ? 2 : 1);
  o.data = (char*)o.data + ((int64_t)dr * o.ld + dc) * es;
  return o;
}

struct Rect {
  int r0, c0, h, w;  // window coordinates
};

// The interior rectangle of the region: window rows [m, H-m) and columns
// [m, W-m), where every derivative the stencil takes is the central one.
inline Rect interior(int H, int W, int rr0, int rc0, int rh, int rw, int m) {
  const int r0 = std::max(rr0, m), r1 = std::min(rr0 + rh, H - m);
  const int c0 = std::max(rc0, m), c1 = std::min(rc0 + rw, W - m);
  if (r0 >= r1 || c0 >= c1) return Rect{0, 0, 0, 0};
  return Rect{r0, c0, r1 - r0, c1 - c0};
}

inline Rect interior(const fsg_window* win, const fsg_region* reg, int m) {
  return interior(win->h, win->w, reg->r0, reg->c0, reg->h, reg->w, m);
}

// The region less the interior rectangle, as up to four strips.
inline int border_strips(int R0, int C0, int rh, int rw, const Rect& in, Rect* out) {
  const int R1 = R0 + rh, C1 = C0 + rw;
  int n = 0;
  if (in.h == 0 || in.w == 0) {
    out[n++] = Rect{R0, C0, rh, rw};
    return n;
  }
  const int ir1 = in.r0 + in.h, ic1 = in.c0 + in.w;
  if (in.r0 > R0) out[n++] = Rect{R0, C0, in.r0 - R0, rw};
  if (R1 > ir1) out[n++] = Rect{ir1, C0, R1 - ir1, rw};
  if (in.c0 > C0) out[n++] = Rect{in.r0, C0, in.h, in.c0 - C0};
  if (C1 > ic1) out[n++] = Rect{in.r0, ic1, in.h, C1 - ic1};
  return n;
}

inline int border_strips(const fsg_region* reg, const Rect& in, Rect* out) {
  return border_strips(reg->r0, reg->c0, reg->h, reg->w, in, out);
}

// Rows per wave for an interior of h rows and `strips` waves across.
inline int seg_rows(int64_t h, int64_t strips) {
  const int64_t s = h * strips / SR_MIN_WAVES;
  return (int)std::max<int64_t>(SR_SEG_MIN, std::min<int64_t>(SR_SEG_MAX, s));
}

// Element offset of a pointer within its 16-byte group, or -1 when it is
// not 4-byte aligned.
inline int quad_phase(const void* p) {
  const uintptr_t u = (uintptr_t)p;
  return (u & 3) ? -1 : (int)((u >> 2) & 3);
}

inline int mod4(int x) { return ((x % 4) + 4) % 4; }

// The four-columns-per-lane launch over the interior columns [oc0, oc1) that
// whole waves cover: needs a 4-byte aligned window; each wave's 256 columns
// lie inside the window; float32 outputs take 16-byte (8-byte) stores.  seg > 0
// fixes the rows per wave (else seg_rows).  False when no wave fits.
inline bool rows4_setup(const RowsTile& R, const Rect& in, const OutDev* outs, int n_out,
                        int left, RowsTile4& S4, dim3& grd, int& oc0, int& oc1, int seg = 0) {
  const int cols = 256 - 2 * left;
  const int a = quad_phase(R.W.p);
  if (a < 0 || in.w == 0) return false;
  // wave starts on 16-byte boundaries when every row keeps the same phase
  // (row stride a multiple of 4 floats), else anywhere (dword-aligned loads)
  const int q = R.W.ld % 4 == 0 ? 4 : 1;
  auto modq = [q](int x) { return q == 4 ? mod4(x) : 0; };
  int cb0 = (in.c0 - left) - modq(a + in.c0 - left);  // aligned, cb0 + left <= in.c0
  if (cb0 < 0) cb0 += q * ((-cb0 + q - 1) / q);  // first aligned start inside the window
  const int cb_max = R.W.w - 256;                // a wave's last column <= W - 1
  const int in_c1 = in.c0 + in.w;
  if (cb0 > cb_max) return false;
  const int cb_last = cb_max - modq(a + cb_max);  // aligned like cb0, <= cb_max
  const int64_t n_strips = ((int64_t)in_c1 - (cb0 + left) + cols - 1) / cols;
  oc0 = std::max(in.c0, cb0 + left);
  oc1 = std::min(in_c1, cb_last + left + cols);  // the last wave's last output + 1
  if (seg <= 0) seg = seg_rows(in.h, n_strips);
  const int64_t nseg = ((int64_t)in.h + seg - 1) / seg;
  if (n_strips <= 0 || oc0 >= oc1 || nseg > 65535) return false;
  S4.T = R;
  S4.T.ic0 = oc0;
  S4.T.ic1 = oc1;
  S4.T.seg = seg;
  S4.cb0 = cb0;
  S4.cb_last = cb_last;
  S4.n_strips = (int)n_strips;
  int vec = 1;
  for (int k = 0; k < n_out; k++) {
    const OutDev& o = outs[k];
    if (!o.data) continue;
    vec = vec && o.dtype == FSG_OUT_F32 && quad_phase(o.data) >= 0;
  }
  S4.vec_out = vec;
  grd = dim3((unsigned)((n_strips + 3) / 4), (unsigned)nseg);
  return true;
}

}  // namespace fsg

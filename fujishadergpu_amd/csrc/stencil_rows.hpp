// Row-walking gradient kernels for the interior of a window.
//
// Inside the window every derivative is numpy's central difference, so a
// lane can own one column and walk down its rows: it loads each row's value
// once (the NaN fill applied at the load), takes its left and right
// neighbours' values from the adjacent lanes with whole-wave DPP shifts, and
// keeps the rows above and below in registers.  Curvature also keeps the
// first derivatives P = d/dcol and Q = d/drow of the rows above and below, so
// every first derivative is formed once per position, like the tiled block's
// tables.  A wave covers 64 columns and writes the middle 62 (hillshade,
// slope) or 60 (curvature); the window's outer rows and columns, where
// numpy's one-sided formulas apply, go to the general kernels.
//
// Divisions by f32(2h) use div_rn (the correctly rounded quotient from the
// host's RN(1/2h)), so every value is the same float32 operation as the
// general kernels' and the outputs are the same bits.  Curvature drops the
// range guard (div_rn_unguarded): the same quotients for every dividend in
// [2^-75, 2^75] and zero, which real elevation data stays within.
#pragma once
#include "fdiv.hpp"
#include "stencil.hpp"

namespace fsg {

// The value of lane L-1 (lane 0 gets 0) and of lane L+1 (lane 63 gets 0):
// DPP wave_shr:1 and wave_shl:1 of a 32-bit value.
__device__ __forceinline__ float lane_left(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_right(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// A divisor f32(2h) and its reciprocal RN(1 / f32(2h)) (host computed).
// ok: b is in div_rn's range (else the IEEE division).
struct DivK {
  float b, y;
  int ok;
};

__host__ __device__ inline DivK make_divk(float two_h) {
  return DivK{two_h, 1.0f / two_h, fdiv_ok(two_h) ? 1 : 0};
}

__device__ __forceinline__ float dq(float a, const DivK& d) {
  return d.ok ? div_rn(a, d.b, d.y) : a / d.b;
}
// curvature's divisions: the unguarded form (its output is checked to a
// tolerance, and every elevation difference of real data is in its range)
__device__ __forceinline__ float dqu(float a, const DivK& d) {
  return d.ok ? div_rn_unguarded(a, d.b, d.y) : a / d.b;
}

// Columns a wave writes: 64 lanes less the neighbours the stencil reads.
constexpr int SR_SHADE_COLS = 62, SR_CURV_COLS = 60;
// Rows one wave walks: at most 64, fewer on small windows so that the launch
// still has SR_MIN_WAVES waves (a wave walks its rows one after another),
// at least 8.  16384 waves (about two rounds of resident waves) measured
// about 7 % faster than 8192 on 8192^2 windows, NaN-free and NaN-aware.
constexpr int SR_SEG_MAX = 64, SR_SEG_MIN = 8, SR_MIN_WAVES = 16384;

// One interior launch: the window, the region's origin (output coordinates
// are relative to it) and the interior rectangle [ir0, ir1) x [ic0, ic1) in
// window coordinates.
struct RowsTile {
  WinDev W;
  int r0, c0;
  int ir0, ir1, ic0, ic1;
  int seg;      // rows per wave
  DivK dx, dy;  // f32(2|step_x|), f32(2|step_y|): first derivatives
  DivK sx, sy;  // f32(2 step_x), f32(2 step_y): curvature's second derivatives
  ShadeK shade;
  CurvK curv;
  int want_hs, want_sl;
  OutDev o0, o1;  // hillshade and slope, or curvature in o0
  // the batch's combined walk (stencil4_wave): curvature's output, and the
  // window's per-row signed x steps (NULL: the constant dx / sx)
  int want_cv;
  OutDev o2;
  const double* xrows;
};

// The window value at (r, c) with the NaN fill applied; z is the raw value.
template <bool NANAW>
__device__ __forceinline__ float rows_fill(const WinDev& W, int r, int c, float z,
                                           const Gauss9& G) {
  if (NANAW && isnan(z)) return nan_fill(W, r, c, G);
  return z;
}

// Rows a lane has in flight ahead of the row it uses: one 256-byte load per
// wave and row is too little to cover the memory latency on its own.
constexpr int SR_AHEAD = 4;

// Whether lane L's value is valid in any of lanes L-2 .. L+2 (a lane outside
// the wave counts as valid).
__device__ __forceinline__ bool near_valid2(bool v) {
  const int x = v ? 1 : 0;
  const int l1 = __builtin_amdgcn_update_dpp(1, x, 0x138, 0xf, 0xf, false);
  const int r1 = __builtin_amdgcn_update_dpp(1, x, 0x130, 0xf, 0xf, false);
  const int l2 = __builtin_amdgcn_update_dpp(1, l1, 0x138, 0xf, 0xf, false);
  const int r2 = __builtin_amdgcn_update_dpp(1, r1, 0x130, 0xf, 0xf, false);
  return (x | l1 | r1 | l2 | r2) != 0;
}

// Column sums the wave fill keeps per wave: its 64 columns and 4 either side.
constexpr int SR_FILL_COLS = 72;

// The NaN fill of row x for the lanes with `want`, computed by the whole wave
// at once: every lane forms the axis-0 pass of its own column (nan_fill_col,
// nine loads issued together), lanes 0..7 also the 4 columns either side of
// the wave, and the axis-1 pass reads the nine column sums from LDS in
// nan_fill's order (centre, then distance 4..1), so the value is nan_fill's
// bits.  c is the lane's column (cb + lane), cb the wave's first column.
// Every lane of the wave must call it (it exchanges through LDS).
__device__ __forceinline__ float wave_fill(const WinDev& W, int x, int c, int cb, bool want,
                                           const Gauss9& G, float2* cols) {
  const int lane = threadIdx.x & 63;
  float cv, cw;
  nan_fill_col<false>(W, x, clampi(c, W.w), G, cv, cw);
  cols[4 + lane] = make_float2(cv, cw);
  if (lane < 8) {
    const int col = lane < 4 ? cb - 4 + lane : cb + 64 + (lane - 4);
    float ev, ew;
    nan_fill_col<false>(W, x, clampi(col, W.w), G, ev, ew);
    cols[lane < 4 ? lane : 64 + lane] = make_float2(ev, ew);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float f = 0.0f;
  if (want) {
    double sv = (double)cv * G.g[0];
    double sw = (double)cw * G.g[0];
#pragma unroll
    for (int d = 4; d >= 1; d--) {
      const float2 a = cols[4 + lane - d], b = cols[4 + lane + d];
      sv += ((double)a.x + (double)b.x) * G.g[d];
      sw += ((double)a.y + (double)b.y) * G.g[d];
    }
    const float fsv = (float)sv, fsw = (float)sw;
    f = fsw > 0.0f ? fsv / fsw : 0.0f;
  }
  // the next call overwrites the sums
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return f;
}

// Hillshade and/or slope over the interior rows [1, H-2] x cols [1, W-2].
// Rows arrive one per iteration from ra - 1 on (the first two iterations
// only fill).  A NaN cell takes the fill only when an output that reads it,
// (r - 1, c), (r + 1, c) or (r, c -+ 1), is valid; the others feed NaN outputs.
template <bool NANAW>
__global__ void __launch_bounds__(256) k_shade_inner(RowsTile S, Gauss9 G) {
  const int lane = threadIdx.x & 63;
  const int c = S.ic0 - 1 + blockIdx.x * SR_SHADE_COLS + lane;
  const int ra = S.ir0 + (blockIdx.y * 4 + (threadIdx.x >> 6)) * S.seg;
  if (ra >= S.ir1) return;  // the whole wave
  const int rb = min(ra + S.seg, S.ir1);
  const int H = S.W.h;
  const int cl = min(c, S.W.w - 1);  // lanes past the window read its last column
  const bool out = lane >= 1 && lane <= SR_SHADE_COLS && c < S.ic1;
  const int64_t ld = S.W.ld;
  const float* p = S.W.p + (int64_t)(ra - 1) * ld + cl;  // row ra - 1 (ra >= 1)
  const PostK phs = load_post(S.o0), psl = load_post(S.o1);
  const ShadeK& K = S.shade;
  // iteration r holds zd = row r + 1 and zq[k] = row r + 2 + k
  float zd = p[0], zq[SR_AHEAD];
#pragma unroll
  for (int k = 0; k < SR_AHEAD; k++) zq[k] = ra + k < H ? p[(1 + k) * ld] : 0.0f;
  p += (1 + SR_AHEAD) * ld;  // row ra + SR_AHEAD
  float fu = 0.0f, fm = 0.0f, zm = 0.0f;
  __shared__ float2 fill_cols[4][SR_FILL_COLS];
  for (int r = ra - 2; r < rb; r++) {
    float fd = zd;
    if (NANAW) {
      // every lane takes part in the shifts, before any branch
      const float zl = lane_left(zd), zr = lane_right(zd);
      const bool want = isnan(zd) & (!isnan(zm) | !isnan(zq[0]) | !isnan(zl) | !isnan(zr));
      if (__ballot(want)) {  // the whole wave
        const float f = wave_fill(S.W, r + 1, c, c - lane, want, G, fill_cols[threadIdx.x >> 6]);
        if (want) fd = f;
      }
      if (isnan(fd)) fd = 0.0f;  // NaN cells no output reads
    }
    const float fl = lane_left(fm), fr = lane_right(fm);
    if (r >= ra && out) {
      const int orr = r - S.r0, occ = c - S.c0;
      if (isnan(zm)) {
        const float qn = __builtin_nanf("");
        if (S.want_hs) store_out(S.o0, orr, occ, apply_post(phs, qn));
        if (S.want_sl) store_out(S.o1, orr, occ, apply_post(psl, qn));
      } else {
        const float dy = dq(fd - fu, S.dy);
        const float dx = dq(fr - fl, S.dx);
        if (S.want_sl) store_out(S.o1, orr, occ, apply_post(psl, slope_from(K, dx, dy)));
        if (S.want_hs) {
          float hdx = dx, hdy = dy;
          if (!K.z_is_one) {
            hdy = dq(fd * K.zf - fu * K.zf, S.dy);
            hdx = dq(fr * K.zf - fl * K.zf, S.dx);
          }
          store_out(S.o0, orr, occ, apply_post(phs, shade_from_grad(K, hdx, hdy)));
        }
      }
    }
    fu = fm;
    fm = fd;
    zm = zd;
    zd = zq[0];
#pragma unroll
    for (int k = 0; k + 1 < SR_AHEAD; k++) zq[k] = zq[k + 1];
    zq[SR_AHEAD - 1] = r + 2 + SR_AHEAD < H ? p[0] : 0.0f;  // row r + 2 + SR_AHEAD
    p += ld;
  }
}

// Curvature over the interior rows [2, H-3] x cols [2, W-3].  Rows arrive one
// per iteration from ra - 2 on; the first four iterations only build the rows'
// values and first derivatives.  A NaN cell takes the fill only when a valid
// cell lies within two rows and two columns of it (every output that reads
// it does); the others feed NaN outputs only.
template <bool NANAW>
__global__ void __launch_bounds__(256) k_curv_inner(RowsTile S, Gauss9 G) {
  const int lane = threadIdx.x & 63;
  const int c = S.ic0 - 2 + blockIdx.x * SR_CURV_COLS + lane;
  const int ra = S.ir0 + (blockIdx.y * 4 + (threadIdx.x >> 6)) * S.seg;
  if (ra >= S.ir1) return;  // the whole wave
  const int rb = min(ra + S.seg, S.ir1);
  const int H = S.W.h;
  const int cl = min(c, S.W.w - 1);
  const bool out = lane >= 2 && lane < 2 + SR_CURV_COLS && c < S.ic1;
  const int64_t ld = S.W.ld;
  const float* p = S.W.p + (int64_t)(ra - 2) * ld + cl;  // row ra - 2 (ra >= 2)
  const PostK pcv = load_post(S.o0);
  const CurvK& K = S.curv;
  // iteration r holds zn = row r + 2 and zq[k] = row r + 3 + k
  float zn = p[0], zq[SR_AHEAD];
#pragma unroll
  for (int k = 0; k < SR_AHEAD; k++) zq[k] = ra - 1 + k < H ? p[(1 + k) * ld] : 0.0f;
  p += (1 + SR_AHEAD) * ld;  // row ra - 1 + SR_AHEAD
  // after iteration r: fi, fi1 = rows r + 1, r + 2 (filled); pm, p0 = P(r),
  // P(r + 1); qm, q0 = Q(r), Q(r + 1); zc, zc1 = raw rows r + 1, r + 2
  float fi = 0.0f, fi1 = 0.0f, pm = 0.0f, p0 = 0.0f, qm = 0.0f, q0 = 0.0f;
  float zc = 0.0f, zc1 = 0.0f;
  __shared__ float2 fill_cols[4][SR_FILL_COLS];
  for (int r = ra - 4; r < rb; r++) {
    const float z4 = zn;  // row r + 2
    float f4 = z4;
    if (NANAW) {
      const bool col = !isnan(zc) | !isnan(zc1) | !isnan(z4) | !isnan(zq[0]) |
                       !isnan(zq[1]);  // rows r .. r + 4 of this column
      const bool want = isnan(z4) & near_valid2(col);
      if (__ballot(want)) {  // the whole wave
        const float f = wave_fill(S.W, r + 2, c, c - lane, want, G, fill_cols[threadIdx.x >> 6]);
        if (want) f4 = f;
      }
      if (isnan(f4)) f4 = 0.0f;  // NaN cells no output reads
    }
    // every lane takes part in the shifts, before any branch
    const float l1 = lane_left(fi1), r1 = lane_right(fi1);
    const float lp = lane_left(p0), rp = lane_right(p0);
    const float lq = lane_left(q0), rq = lane_right(q0);
    const float pp = dqu(r1 - l1, S.dx);  // P(r + 1)
    const float qp = dqu(f4 - fi, S.dy);  // Q(r + 1)
    if (r >= ra && out) {
      const float dxx = dqu(rp - lp, S.sx);
      const float dyx = dqu(rq - lq, S.sx);
      const float dxy = dqu(pp - pm, S.sy);
      const float dyy = dqu(qp - qm, S.sy);
      const int orr = r - S.r0, occ = c - S.c0;
      const float v = isnan(zc) ? __builtin_nanf("")
                                : curvature_from(K, p0, q0, dxx, dxy, dyy, dyx);
      store_out(S.o0, orr, occ, apply_post(pcv, v));
    }
    pm = p0;
    p0 = pp;
    qm = q0;
    q0 = qp;
    fi = fi1;
    fi1 = f4;
    zc = zc1;
    zc1 = z4;
    zn = zq[0];
#pragma unroll
    for (int k = 0; k + 1 < SR_AHEAD; k++) zq[k] = zq[k + 1];
    zq[SR_AHEAD - 1] = r + 3 + SR_AHEAD < H ? p[0] : 0.0f;  // row r + 3 + SR_AHEAD
    p += ld;
  }
}

// ---------------------------------------------------------------- 4 columns
// The same walks with four adjacent columns per lane: 16-byte loads (1 KB per
// wave and row instead of 256 bytes, three 128-byte lines for 240 useful bytes
// before), and only the group's outer neighbours cross lanes.  A wave covers
// 256 columns and writes 256 - 2 LEFT of them from column LEFT on; the next
// wave starts that many columns later.  With a row stride that is a multiple
// of 4 floats every wave starts at a 16-byte boundary in every row; other
// strides take dword-aligned 16-byte accesses.  LEFT = 4: lanes 1..62 write
// whole groups (16-byte stores), lanes 0 and 63 only supply neighbours;
// LEFT = 2: every lane writes, the first and last lanes half a group (8-byte
// stores).  Measured: hillshade/slope (two outputs) run faster with LEFT = 4,
// curvature with LEFT = 2.  Waves go only where their 256 columns lie inside
// the window (the last one moved left to the last start that fits, aligned
// like the others, overlapping its neighbour), and the few interior columns
// left at the sides take the general kernel.
template <int LEFT>
constexpr int sr4_cols() {
  return 256 - 2 * LEFT;
}
constexpr int SR4_LEFT_SHADE = 4, SR4_LEFT_CURV = 2;
constexpr int SR4_AHEAD = 1;  // rows in flight ahead of the one used (more cost occupancy)

struct RowsTile4 {
  RowsTile T;
  int cb0;      // first column of wave 0 (a 16-byte aligned address in every
                // row when the row stride is a multiple of 4 floats)
  int cb_last;  // the last start (aligned like cb0) whose 256 columns fit the
                // window: later waves start there (overlapping writes of equal values)
  int n_strips; // waves across the interior
  int vec_out;  // o0 (and o1) are float32: 16-byte (8-byte) stores per group
};

// The raw values of columns c4 .. c4+3 of the row at rowp, as one 16-byte
// load: the launch only places waves whose 256 columns lie inside the window.
// The address is 4-byte aligned; gfx9 and later take multi-dword global
// accesses at dword alignment, a 16-byte aligned one (row stride a multiple of
// 4 floats) touches one fewer cache line per wave and row.
__device__ __forceinline__ void load_group(const float* rowp, int c4, float z[4]) {
  float4 t;
  __builtin_memcpy(&t, rowp + c4, sizeof(t));
  z[0] = t.x;
  z[1] = t.y;
  z[2] = t.z;
  z[3] = t.w;
}

// Column sums the four-column wave fill keeps per wave: 256 columns and 4
// either side.
constexpr int SR4_FILL_COLS = 264;

// wave_fill for four columns per lane: the lane's four column sums (nine
// loads each) and lanes 0..7 the four columns either side of the wave go to
// LDS, then each wanted cell reads its nine in nan_fill's order, so the value
// is nan_fill's bits.  z holds the raw row, want which cells need the fill;
// the group's values come back in f (other NaN cells as 0: no output reads
// them).  Every lane of the wave must call it.
__device__ __forceinline__ void wave_fill4(const WinDev& W, int x, int c4, int cb,
                                           const float z[4], const bool want[4],
                                           const Gauss9& G, float2* cols, float f[4]) {
  const int lane = threadIdx.x & 63;
#pragma unroll 1
  for (int k = 0; k < 4; k++) {
    float cv, cw;
    nan_fill_col<false>(W, x, clampi(c4 + k, W.w), G, cv, cw);
    cols[4 + 4 * lane + k] = make_float2(cv, cw);
  }
  if (lane < 8) {
    const int col = lane < 4 ? cb - 4 + lane : cb + 256 + (lane - 4);
    float ev, ew;
    nan_fill_col<false>(W, x, clampi(col, W.w), G, ev, ew);
    cols[lane < 4 ? lane : 256 + lane] = make_float2(ev, ew);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int k = 0; k < 4; k++) {
    float v = isnan(z[k]) ? 0.0f : z[k];
    if (want[k]) {
      const int e = 4 + 4 * lane + k;
      const float2 ce = cols[e];
      double sv = (double)ce.x * G.g[0];
      double sw = (double)ce.y * G.g[0];
#pragma unroll
      for (int d = 4; d >= 1; d--) {
        const float2 a = cols[e - d], b = cols[e + d];
        sv += ((double)a.x + (double)b.x) * G.g[d];
        sw += ((double)a.y + (double)b.y) * G.g[d];
      }
      const float fsv = (float)sv, fsw = (float)sw;
      v = fsw > 0.0f ? fsv / fsw : 0.0f;
    }
    f[k] = v;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Writes v[0..3] at output row orr, columns occ .. occ+3 where ok[k] is set.
// vec: a float32 output (4-byte aligned, like the input row).  HALVES stores
// each half as one 8-byte store (a wave's first and last lanes write one half
// each), else a whole group as one 16-byte store; groups cut by the region's
// first or last column, and other outputs, store per column.
template <bool HALVES>
__device__ __forceinline__ void store_group(const OutDev& o, bool vec, int orr, int occ,
                                            const bool ok[4], const float v[4]) {
  if (vec) {
    float* d = (float*)o.data + (int64_t)orr * o.ld + occ;
    if (!HALVES && ok[0] && ok[3]) {
      const float4 t = make_float4(v[0], v[1], v[2], v[3]);
      __builtin_memcpy(d, &t, sizeof(t));
      return;
    }
    if (HALVES) {
      const float2 lo = make_float2(v[0], v[1]), hi = make_float2(v[2], v[3]);
      if (ok[0] && ok[1]) __builtin_memcpy(d, &lo, sizeof(lo));
      if (ok[2] && ok[3]) __builtin_memcpy(d + 2, &hi, sizeof(hi));
      if (ok[0] == ok[1] && ok[2] == ok[3]) return;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (ok[k]) store_out(o, orr, occ + k, v[k]);
}

// Wave `strip` of row segment `sg`: its columns and rows; false when the wave
// has no work.
template <int LEFT>
__device__ __forceinline__ bool rows4_at(const RowsTile4& S, int strip, int sg, int& c4, int& cb,
                                         int& ra, int& rb) {
  const int lane = threadIdx.x & 63;
  ra = S.T.ir0 + sg * S.T.seg;
  if (strip >= S.n_strips || ra >= S.T.ir1) return false;
  rb = min(ra + S.T.seg, S.T.ir1);
  cb = min(S.cb0 + strip * sr4_cols<LEFT>(), S.cb_last);
  c4 = cb + 4 * lane;
  return true;
}

// The wave's strip and row segment; false when the wave has no work.
template <int LEFT>
__device__ __forceinline__ bool rows4_wave(const RowsTile4& S, int& c4, int& cb, int& ra,
                                           int& rb) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int strip = blockIdx.x * 4 + wave;
  ra = S.T.ir0 + blockIdx.y * S.T.seg;
  if (strip >= S.n_strips || ra >= S.T.ir1) return false;
  rb = min(ra + S.T.seg, S.T.ir1);
  cb = min(S.cb0 + strip * sr4_cols<LEFT>(), S.cb_last);
  c4 = cb + 4 * lane;
  return true;
}

// Hillshade and/or slope, four columns per lane.  Rows arrive one per
// iteration from ra - 1 on; the first two iterations only fill (so the NaN
// fill has one call site).
template <bool NANAW>
__global__ void __launch_bounds__(256) k_shade_inner4(RowsTile4 S4, Gauss9 G) {
  constexpr int LEFT = SR4_LEFT_SHADE;
  int c4, cb, ra, rb;
  if (!rows4_wave<LEFT>(S4, c4, cb, ra, rb)) return;  // the whole wave
  const RowsTile& S = S4.T;
  const int H = S.W.h;
  const int lo = max(cb + LEFT, S.ic0), hi = min(cb + LEFT + sr4_cols<LEFT>(), S.ic1);
  bool ok[4];
#pragma unroll
  for (int k = 0; k < 4; k++) ok[k] = c4 + k >= lo && c4 + k < hi;
  const bool any = ok[0] || ok[1] || ok[2] || ok[3];
  const int64_t ld = S.W.ld;
  const float* rowp = S.W.p + (int64_t)(ra - 1) * ld;  // row ra - 1 (ra >= 1)
  const PostK phs = load_post(S.o0), psl = load_post(S.o1);
  const ShadeK& K = S.shade;
  constexpr int AHEAD = SR4_AHEAD;
  // iteration r holds zd = row r + 1 and zq[a] = row r + 2 + a
  float zd[4], zq[AHEAD][4];
  load_group(rowp, c4, zd);
#pragma unroll
  for (int a = 0; a < AHEAD; a++) {
    if (ra + a < H)
      load_group(rowp + (1 + a) * ld, c4, zq[a]);
    else
      zq[a][0] = zq[a][1] = zq[a][2] = zq[a][3] = 0.0f;
  }
  rowp += (1 + AHEAD) * ld;  // row ra + AHEAD
  float fu[4] = {0.f, 0.f, 0.f, 0.f}, fm[4] = {0.f, 0.f, 0.f, 0.f};
  float zm[4] = {0.f, 0.f, 0.f, 0.f};
  __shared__ float2 fill_cols[4][SR4_FILL_COLS];
  for (int r = ra - 2; r < rb; r++) {
    float fd[4] = {zd[0], zd[1], zd[2], zd[3]};
    if (NANAW) {
      // a NaN cell of row r + 1 is filled when an output reading it is valid:
      // (r, c), (r + 2, c), (r + 1, c -+ 1); every lane takes part in the shifts
      const float zl = lane_left(zd[3]), zr = lane_right(zd[0]);
      bool want[4];
      bool any_want = false;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float left = k == 0 ? zl : zd[k - 1], right = k == 3 ? zr : zd[k + 1];
        want[k] = isnan(zd[k]) & (!isnan(zm[k]) | !isnan(zq[0][k]) | !isnan(left) |
                                  !isnan(right));
        any_want |= want[k];
      }
      if (__ballot(any_want))  // the whole wave
        wave_fill4(S.W, r + 1, c4, cb, zd, want, G, fill_cols[threadIdx.x >> 6], fd);
      else
#pragma unroll
        for (int k = 0; k < 4; k++) fd[k] = isnan(zd[k]) ? 0.0f : zd[k];
    }
    const float fl = lane_left(fm[3]), fr = lane_right(fm[0]);
    if (r >= ra && any) {
      float hs[4], sl[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float left = k == 0 ? fl : fm[k - 1];
        const float right = k == 3 ? fr : fm[k + 1];
        hs[k] = sl[k] = __builtin_nanf("");
        if (!isnan(zm[k])) {
          const float dy = dq(fd[k] - fu[k], S.dy);
          const float dx = dq(right - left, S.dx);
          if (S.want_sl) sl[k] = slope_from(K, dx, dy);
          if (S.want_hs) {
            float hdx = dx, hdy = dy;
            if (!K.z_is_one) {
              hdy = dq(fd[k] * K.zf - fu[k] * K.zf, S.dy);
              hdx = dq(right * K.zf - left * K.zf, S.dx);
            }
            hs[k] = shade_from_grad(K, hdx, hdy);
          }
        }
        hs[k] = apply_post(phs, hs[k]);
        sl[k] = apply_post(psl, sl[k]);
      }
      const int orr = r - S.r0, occ = c4 - S.c0;
      if (S.want_hs) store_group<LEFT == 2>(S.o0, S4.vec_out, orr, occ, ok, hs);
      if (S.want_sl) store_group<LEFT == 2>(S.o1, S4.vec_out, orr, occ, ok, sl);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      fu[k] = fm[k];
      fm[k] = fd[k];
      zm[k] = zd[k];
      zd[k] = zq[0][k];
    }
#pragma unroll
    for (int a = 0; a + 1 < AHEAD; a++)
#pragma unroll
      for (int k = 0; k < 4; k++) zq[a][k] = zq[a + 1][k];
    if (r + 2 + AHEAD < H)  // row r + 2 + AHEAD
      load_group(rowp, c4, zq[AHEAD - 1]);
    rowp += ld;
  }
}

// Curvature, four columns per lane.  Rows arrive one per iteration from
// ra - 2 on; the first four iterations only build the rows' values and first
// derivatives.
template <bool NANAW>
__global__ void __launch_bounds__(256) k_curv_inner4(RowsTile4 S4, Gauss9 G) {
  constexpr int LEFT = SR4_LEFT_CURV;
  int c4, cb, ra, rb;
  if (!rows4_wave<LEFT>(S4, c4, cb, ra, rb)) return;  // the whole wave
  const RowsTile& S = S4.T;
  const int H = S.W.h;
  const int lo = max(cb + LEFT, S.ic0), hi = min(cb + LEFT + sr4_cols<LEFT>(), S.ic1);
  bool ok[4];
#pragma unroll
  for (int k = 0; k < 4; k++) ok[k] = c4 + k >= lo && c4 + k < hi;
  const bool any = ok[0] || ok[1] || ok[2] || ok[3];
  const int64_t ld = S.W.ld;
  const float* rowp = S.W.p + (int64_t)(ra - 2) * ld;  // row ra - 2 (ra >= 2)
  const PostK pcv = load_post(S.o0);
  const CurvK& K = S.curv;
  // the NaN-aware variant looks two rows ahead: the fill test reads rows
  // r .. r + 4
  constexpr int AHEAD = NANAW ? 2 : SR4_AHEAD;
  // iteration r holds zn = row r + 2 and zq[a] = row r + 3 + a
  float zn[4], zq[AHEAD][4];
  load_group(rowp, c4, zn);
#pragma unroll
  for (int a = 0; a < AHEAD; a++) {
    if (ra - 1 + a < H)
      load_group(rowp + (1 + a) * ld, c4, zq[a]);
    else
      zq[a][0] = zq[a][1] = zq[a][2] = zq[a][3] = 0.0f;
  }
  rowp += (1 + AHEAD) * ld;  // row ra - 1 + AHEAD
  // after iteration r: fi, fi1 = rows r + 1, r + 2 (filled); pm, p0 = P(r),
  // P(r + 1); qm, q0 = Q(r), Q(r + 1); zc, zc1 = raw rows r + 1, r + 2
  float fi[4] = {0.f, 0.f, 0.f, 0.f}, fi1[4] = {0.f, 0.f, 0.f, 0.f};
  float pm[4] = {0.f, 0.f, 0.f, 0.f}, p0[4] = {0.f, 0.f, 0.f, 0.f};
  float qm[4] = {0.f, 0.f, 0.f, 0.f}, q0[4] = {0.f, 0.f, 0.f, 0.f};
  float zc[4] = {0.f, 0.f, 0.f, 0.f}, zc1[4] = {0.f, 0.f, 0.f, 0.f};
  __shared__ float2 fill_cols[4][SR4_FILL_COLS];
  for (int r = ra - 4; r < rb; r++) {
    float f4[4] = {zn[0], zn[1], zn[2], zn[3]}, pp[4], qp[4];
    if (NANAW) {
      // a NaN cell of row r + 2 is filled when a valid cell lies within two
      // rows and two columns of it (the reach of the derivatives reading it)
      int colv = 0;
#pragma unroll
      for (int k = 0; k < 4; k++)
        colv |= (!isnan(zc[k]) | !isnan(zc1[k]) | !isnan(zn[k]) | !isnan(zq[0][k]) |
                 !isnan(zq[1][k])) << k;
      // the neighbour lanes' masks; outside the wave counts as valid
      const int ml = __builtin_amdgcn_update_dpp(0xf, colv, 0x138, 0xf, 0xf, false);
      const int mr = __builtin_amdgcn_update_dpp(0xf, colv, 0x130, 0xf, 0xf, false);
      const int w12 = ml | (colv << 4) | (mr << 8);  // columns c4 - 4 .. c4 + 7
      bool want[4];
      bool any_want = false;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        want[k] = isnan(zn[k]) & (((w12 >> (k + 2)) & 0x1f) != 0);
        any_want |= want[k];
      }
      if (__ballot(any_want))  // the whole wave
        wave_fill4(S.W, r + 2, c4, cb, zn, want, G, fill_cols[threadIdx.x >> 6], f4);
      else
#pragma unroll
        for (int k = 0; k < 4; k++) f4[k] = isnan(zn[k]) ? 0.0f : zn[k];
    }
    const float l1 = lane_left(fi1[3]), r1 = lane_right(fi1[0]);
    const float lp = lane_left(p0[3]), rp = lane_right(p0[0]);
    const float lq = lane_left(q0[3]), rq = lane_right(q0[0]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      pp[k] = dqu((k == 3 ? r1 : fi1[k + 1]) - (k == 0 ? l1 : fi1[k - 1]), S.dx);  // P(r + 1)
      qp[k] = dqu(f4[k] - fi[k], S.dy);                                            // Q(r + 1)
    }
    if (r >= ra && any) {
      float v[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float dxx = dqu((k == 3 ? rp : p0[k + 1]) - (k == 0 ? lp : p0[k - 1]), S.sx);
        const float dyx = dqu((k == 3 ? rq : q0[k + 1]) - (k == 0 ? lq : q0[k - 1]), S.sx);
        const float dxy = dqu(pp[k] - pm[k], S.sy);
        const float dyy = dqu(qp[k] - qm[k], S.sy);
        v[k] = isnan(zc[k]) ? __builtin_nanf("")
                            : curvature_from(K, p0[k], q0[k], dxx, dxy, dyy, dyx);
        v[k] = apply_post(pcv, v[k]);
      }
      store_group<LEFT == 2>(S.o0, S4.vec_out, r - S.r0, c4 - S.c0, ok, v);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      pm[k] = p0[k];
      p0[k] = pp[k];
      qm[k] = q0[k];
      q0[k] = qp[k];
      fi[k] = fi1[k];
      fi1[k] = f4[k];
      zc[k] = zc1[k];
      zc1[k] = zn[k];
      zn[k] = zq[0][k];
    }
#pragma unroll
    for (int a = 0; a + 1 < AHEAD; a++)
#pragma unroll
      for (int k = 0; k < 4; k++) zq[a][k] = zq[a + 1][k];
    if (r + 3 + AHEAD < H)  // row r + 3 + AHEAD
      load_group(rowp, c4, zq[AHEAD - 1]);
    rowp += ld;
  }
}


// ---------------------------------------------------------- batch items
// Hillshade, slope and curvature of a NaN-free window's interior in one walk
// (the batch's stencil items; the window interior is rows [2, H-3] x cols
// [2, W-3]): curvature's first derivatives P(r) and Q(r) are the gradient the
// tiled block gives hillshade and slope when all three share the steps.  With
// ROWS, the x divisors are row r's own: f32(2|hx|) for P and hillshade,
// f32(2 hx) for curvature's second derivatives, RN(1/b) by an IEEE division
// per row (the same values the tiled block's per-row coefficient tables
// hold).  Curvature geometry: 252 columns per wave, half-group stores.
__device__ __forceinline__ DivK row_divk(double h) { return make_divk((float)(2. * h)); }

template <bool ROWS>
__device__ __forceinline__ void stencil4_wave(const RowsTile4& S4, int strip, int sg) {
  constexpr int LEFT = SR4_LEFT_CURV;
  int c4, cb, ra, rb;
  if (!rows4_at<LEFT>(S4, strip, sg, c4, cb, ra, rb)) return;  // the whole wave
  const RowsTile& S = S4.T;
  const int H = S.W.h;
  const int lo = max(cb + LEFT, S.ic0), hi = min(cb + LEFT + sr4_cols<LEFT>(), S.ic1);
  bool ok[4];
#pragma unroll
  for (int k = 0; k < 4; k++) ok[k] = c4 + k >= lo && c4 + k < hi;
  const bool any = ok[0] || ok[1] || ok[2] || ok[3];
  const int64_t ld = S.W.ld;
  const float* rowp = S.W.p + (int64_t)(ra - 2) * ld;  // row ra - 2 (ra >= 2)
  const PostK phs = load_post(S.o0), psl = load_post(S.o1), pcv = load_post(S.o2);
  const CurvK& K = S.curv;
  const ShadeK& KS = S.shade;
  constexpr int AHEAD = SR4_AHEAD;
  float zn[4], zq[AHEAD][4];
  load_group(rowp, c4, zn);
#pragma unroll
  for (int a = 0; a < AHEAD; a++) {
    if (ra - 1 + a < H)
      load_group(rowp + (1 + a) * ld, c4, zq[a]);
    else
      zq[a][0] = zq[a][1] = zq[a][2] = zq[a][3] = 0.0f;
  }
  rowp += (1 + AHEAD) * ld;
  float fi[4] = {0.f, 0.f, 0.f, 0.f}, fi1[4] = {0.f, 0.f, 0.f, 0.f};
  float pm[4] = {0.f, 0.f, 0.f, 0.f}, p0[4] = {0.f, 0.f, 0.f, 0.f};
  float qm[4] = {0.f, 0.f, 0.f, 0.f}, q0[4] = {0.f, 0.f, 0.f, 0.f};
  float zc[4] = {0.f, 0.f, 0.f, 0.f}, zc1[4] = {0.f, 0.f, 0.f, 0.f};
  for (int r = ra - 4; r < rb; r++) {
    const float* f4 = zn;  // NaN-free window: row r + 2 as loaded
    float pp[4], qp[4];
    // P(r + 1) takes row r + 1's |x| step
    const DivK dxp = ROWS ? row_divk(fabs(S.xrows[min(max(r + 1, 0), H - 1)])) : S.dx;
    const float l1 = lane_left(fi1[3]), r1 = lane_right(fi1[0]);
    const float lp = lane_left(p0[3]), rp = lane_right(p0[0]);
    const float lq = lane_left(q0[3]), rq = lane_right(q0[0]);
    // first derivatives: the guarded division when hillshade or slope reads
    // them (hillshade's bits), the unguarded one for curvature alone
    const bool exact = S.want_hs || S.want_sl;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const float dp = (k == 3 ? r1 : fi1[k + 1]) - (k == 0 ? l1 : fi1[k - 1]);
      pp[k] = exact ? dq(dp, dxp) : dqu(dp, dxp);
      qp[k] = exact ? dq(f4[k] - fi[k], S.dy) : dqu(f4[k] - fi[k], S.dy);
    }
    if (r >= ra && any) {
      const int orr = r - S.r0, occ = c4 - S.c0;
      if (S.want_cv) {
        const DivK sxr = ROWS ? row_divk(S.xrows[r]) : S.sx;
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const float dxx = dqu((k == 3 ? rp : p0[k + 1]) - (k == 0 ? lp : p0[k - 1]), sxr);
          const float dyx = dqu((k == 3 ? rq : q0[k + 1]) - (k == 0 ? lq : q0[k - 1]), sxr);
          const float dxy = dqu(pp[k] - pm[k], S.sy);
          const float dyy = dqu(qp[k] - qm[k], S.sy);
          v[k] = isnan(zc[k]) ? __builtin_nanf("")
                              : curvature_from(K, p0[k], q0[k], dxx, dxy, dyy, dyx);
          v[k] = apply_post(pcv, v[k]);
        }
        store_group<true>(S.o2, S4.vec_out, orr, occ, ok, v);
      }
      if (S.want_hs || S.want_sl) {
        const DivK dxr = ROWS ? row_divk(fabs(S.xrows[r])) : S.dx;
        float hs[4], sl[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          hs[k] = sl[k] = __builtin_nanf("");
          if (!isnan(zc[k])) {
            // the gradient of row r: P(r), Q(r)
            const float dx = p0[k], dy = q0[k];
            if (S.want_sl) sl[k] = slope_from(KS, dx, dy);
            // (z factor 1 only: the batch keeps other hillshade tiles on
            // the tiled block, which re-derives from the scaled values)
            if (S.want_hs) hs[k] = shade_from_grad(KS, dx, dy);
          }
          hs[k] = apply_post(phs, hs[k]);
          sl[k] = apply_post(psl, sl[k]);
        }
        if (S.want_hs) store_group<true>(S.o0, S4.vec_out, orr, occ, ok, hs);
        if (S.want_sl) store_group<true>(S.o1, S4.vec_out, orr, occ, ok, sl);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      pm[k] = p0[k];
      p0[k] = pp[k];
      qm[k] = q0[k];
      q0[k] = qp[k];
      fi[k] = fi1[k];
      fi1[k] = f4[k];
      zc[k] = zc1[k];
      zc1[k] = zn[k];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) zn[k] = zq[0][k];
#pragma unroll
    for (int a = 0; a + 1 < AHEAD; a++)
#pragma unroll
      for (int k = 0; k < 4; k++) zq[a][k] = zq[a + 1][k];
    if (r + 3 + AHEAD < H)
      load_group(rowp, c4, zq[AHEAD - 1]);
    rowp += ld;
  }
}

}  // namespace fsg

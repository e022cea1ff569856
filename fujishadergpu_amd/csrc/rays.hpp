// Sparse Yokoyama openness: the per-pixel device functions and the ray
// table, shared by the per-window kernels (rays.hip) and the persistent batch
// kernel (batch.hip), so both produce the same bits.
#pragma once
#include <algorithm>
#include <cmath>
#include <string>

#include "fsg_common.hpp"

namespace fsg {

struct WinDevR {
  const float* p;
  int64_t ld;
  int h, w;
};

constexpr int OPEN_MAX = 320;  // n_dirs * n_dist entries passed by value

struct OpenTab {
  int n_dirs, n_dist;
  int negative;
  int raw;         // 1: mean angle in degrees instead of the display value
  int dmax_y, dmax_x;  // largest |offset| on each axis (interior test)
  float half_pi;   // f32(pi/2)
  float gamma;     // f32(1/2.2)
  short2 off[OPEN_MAX];
  float inv[OPEN_MAX];
};

// Mean openness angle (radians) at window (i, j): per direction the extreme
// elevation ratio over its samples, one atan, pi/2 -/+ that angle, averaged
// over the directions that saw a valid sample (_impl_openness.py:78-126).
__device__ __forceinline__ float openness_angle(const WinDevR& W, int i, int j, float z,
                                                const OpenTab& T) {
  float angle_sum = 0.0f, dir_count = 0.0f;
  for (int d = 0; d < T.n_dirs; d++) {
    float ext = T.negative ? INFINITY : -INFINITY;
    bool any = false;
    for (int k = 0; k < T.n_dist; k++) {
      const int e = d * T.n_dist + k;
      const short2 o = T.off[e];
      if (o.x == 0 && o.y == 0) continue;
      const int si = i + o.y, sj = j + o.x;
      if (si < 0 || si >= W.h || sj < 0 || sj >= W.w) continue;
      const float s = W.p[(int64_t)si * W.ld + sj];
      if (isnan(s)) continue;
      const float q = (s - z) * T.inv[e];
      ext = T.negative ? fminf(ext, q) : fmaxf(ext, q);
      any = true;
    }
    if (any) {
      const float a = atanf(ext);
      const float da = T.negative ? (T.half_pi + a) : (T.half_pi - a);
      angle_sum = angle_sum + da;
      dir_count = dir_count + 1.0f;
    }
  }
  return angle_sum / fmaxf(dir_count, 1.0f);
}

// The same angle with the distances of a direction unrolled (n_dist <= 10,
// the reference's table), all loads of a direction issued before any is
// used.  The extreme starts as NaN and takes fmaxf / fminf with every ratio:
// IEEE maxNum / minNum drop a NaN operand, so a NaN sample, a skipped table
// entry (its inverse distance is NaN) and, with CHECK, a sample outside the
// window (loaded from a clamped address and replaced by NaN) change nothing,
// and the extreme stays NaN exactly when no sample of the direction is valid.
// That is openness_angle's "extreme over the valid samples, any valid?" with
// no per-sample mask (max and min are exact, so the order of the valid
// ratios does not matter): the same bits.  CHECK = false is the interior
// variant (every sample inside the window); OFF32 (interior only): the rows
// a pixel's rays reach, i - dmax_y .. i + dmax_y, span fewer than 2^31
// elements, so a sample's address is a 32-bit element offset from the first
// of them (one vector add per sample).
constexpr int OPEN_UNROLL = 10;

template <int ND, bool CHECK, bool OFF32, bool NEG>
__device__ __forceinline__ float openness_angle_s(const WinDevR& W, int i, int j, float z,
                                                  const OpenTab& T) {
  // ND = 10: the table row of a direction is a fixed 10 entries (wide scalar
  // loads, no guards); ND = 0: guarded by the runtime n_dist (<= 10)
  constexpr int NU = ND > 0 ? ND : OPEN_UNROLL;
  const int nd = ND > 0 ? ND : T.n_dist;
  const float* pc = W.p + (int64_t)i * W.ld + j;   // this pixel
  // OFF32: the first row the rays reach, and the pixel's offset from it
  const float* pr = W.p + (int64_t)(i - T.dmax_y) * W.ld;
  int pix = T.dmax_y * (int)W.ld + j;
  // opaque to the optimiser: a sample's offset stays pix + (a scalar), one
  // vector add, instead of being refactored into a per-sample multiply
  asm volatile("" : "+v"(pix));
  const int ld32 = (int)W.ld;
  float angle_sum = 0.0f, dir_count = 0.0f;
  for (int d = 0; d < T.n_dirs; d++) {
    float q[NU];
#pragma unroll
    for (int k = 0; k < NU; k++) {
      q[k] = __builtin_nanf("");
      if (k < nd) {
        const short2 o = T.off[d * nd + k];
        float v;
        if (CHECK) {
          const int si = i + o.y, sj = j + o.x;
          const bool in = si >= 0 && si < W.h && sj >= 0 && sj < W.w;
          v = W.p[(int64_t)min(max(si, 0), W.h - 1) * W.ld + min(max(sj, 0), W.w - 1)];
          v = in ? v : __builtin_nanf("");
        } else if (OFF32) {
          v = pr[(unsigned)(pix + o.y * ld32 + o.x)];
        } else {
          // wave-uniform element offset: scalar arithmetic, one address add per lane
          v = pc[(int64_t)o.y * W.ld + o.x];
        }
        q[k] = (v - z) * T.inv[d * nd + k];
      }
    }
    float ext = __builtin_nanf("");
#pragma unroll
    for (int k = 0; k < NU; k++) ext = NEG ? fminf(ext, q[k]) : fmaxf(ext, q[k]);
    if (!isnan(ext)) {
      const float a = atanf(ext);
      const float da = NEG ? (T.half_pi + a) : (T.half_pi - a);
      angle_sum = angle_sum + da;
      dir_count = dir_count + 1.0f;
    }
  }
  return angle_sum / fmaxf(dir_count, 1.0f);
}

template <int ND, bool CHECK, bool OFF32 = false>
__device__ __forceinline__ float openness_angle_u(const WinDevR& W, int i, int j, float z,
                                                  const OpenTab& T) {
  return T.negative ? openness_angle_s<ND, CHECK, OFF32, true>(W, i, j, z, T)
                    : openness_angle_s<ND, CHECK, OFF32, false>(W, i, j, z, T);
}

// the rows any pixel's rays reach (2 dmax_y + 1 of them) span fewer than
// 2^31 elements: 32-bit sample offsets on interior blocks
__device__ __forceinline__ bool open_off32(const WinDevR& W, const OpenTab& T) {
  return (int64_t)(2 * T.dmax_y + 1) * W.ld < ((int64_t)1 << 31);
}

// Mean angle at (i, j): the unrolled variant when the table allows it.
// interior: the caller knows every sample of (i, j) lies inside the window.
__device__ __forceinline__ float openness_angle_at(const WinDevR& W, int i, int j, float z,
                                                   const OpenTab& T, bool interior) {
  if (T.n_dist == OPEN_UNROLL) {
    if (!interior) return openness_angle_u<OPEN_UNROLL, true>(W, i, j, z, T);
    return open_off32(W, T) ? openness_angle_u<OPEN_UNROLL, false, true>(W, i, j, z, T)
                         : openness_angle_u<OPEN_UNROLL, false>(W, i, j, z, T);
  }
  if (T.n_dist > OPEN_UNROLL) return openness_angle(W, i, j, z, T);
  return interior ? openness_angle_u<0, false>(W, i, j, z, T)
                  : openness_angle_u<0, true>(W, i, j, z, T);
}

// every sample of pixels in rows [i0, i1] x cols [j0, j1] lies inside the window
__device__ __forceinline__ bool openness_interior(const WinDevR& W, const OpenTab& T, int i0,
                                                  int i1, int j0, int j1) {
  return i0 - T.dmax_y >= 0 && i1 + T.dmax_y < W.h && j0 - T.dmax_x >= 0 && j1 + T.dmax_x < W.w;
}

// display value (op / (pi/2), clipped, gamma) or, with T.raw, the angle in degrees
__device__ __forceinline__ float openness_value(const OpenTab& T, float angle) {
  if (T.raw) return (float)((double)angle * (180.0 / M_PI));
  float op = angle / T.half_pi;
  op = fminf(fmaxf(op, 0.0f), 1.0f);
  return powf(op, T.gamma);
}

// The device ray table of an fsg_openness_params (offsets and f32(1/phys)).
inline int openness_table(const fsg_openness_params* p, OpenTab& T, const char* who) {
  if (!p) return fail(FSG_E_INVALID, std::string(who) + ": params NULL");
  if (p->n_dirs < 0 || p->n_dist < 0 || (p->n_dirs * p->n_dist > 0 && !p->samples))
    return fail(FSG_E_INVALID, std::string(who) + ": bad ray table");
  if (p->n_dirs * p->n_dist > OPEN_MAX)
    return fail(FSG_E_UNSUPPORTED, std::string(who) + ": ray table larger than 320 samples");
  T = OpenTab{};
  T.n_dirs = p->n_dirs;
  T.n_dist = p->n_dist;
  T.negative = p->openness_type != 0;
  T.raw = p->raw_degrees != 0;
  T.half_pi = (float)(M_PI / 2);
  T.gamma = (float)(1.0 / 2.2);
  for (int e = 0; e < p->n_dirs * p->n_dist; e++) {
    const fsg_ray_sample& s = p->samples[e];
    if (s.skip || (s.dx == 0 && s.dy == 0)) {
      // never a sample: openness_angle skips (0, 0); the unrolled variants'
      // NaN ratio drops out of their fmaxf / fminf
      T.off[e] = make_short2(0, 0);
      T.inv[e] = __builtin_nanf("");
      continue;
    }
    if (std::abs(s.dx) > 32767 || std::abs(s.dy) > 32767)
      return fail(FSG_E_UNSUPPORTED, std::string(who) + ": offset beyond int16");
    T.off[e] = make_short2((short)s.dx, (short)s.dy);
    T.inv[e] = 1.0f / (float)s.phys;
    T.dmax_x = std::max(T.dmax_x, std::abs(s.dx));
    T.dmax_y = std::max(T.dmax_y, std::abs(s.dy));
  }
  return FSG_OK;
}


}  // namespace fsg

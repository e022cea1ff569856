// TopoUSM main pass, single-wavefront blocks (the default for NaN-free
// windows whose terms run [1..3 box radii <= 32][<= 3 bilinear sources]).
//
// One 64-lane wavefront owns a strip of WAVE_NC = 256 window columns (four
// per lane) and walks down SH rows of it; the output columns are the middle
// 256 - 2R.  Per row:
//   1. vertical: float64 running window sums per radius and column, kept in
//      registers (uniform_filter's axis-0 pass), rounded to float32 means;
//   2. horizontal: the row prefix of each radius's float32 means, built in
//      float64 by a per-lane 4-column prefix and a DPP wavefront scan, goes to
//      a per-wave LDS buffer; box sum = P[c + r + 1] - P[c - r];
//   3. the decimated terms and the overview field from float64 bilinear
//      values: the two horizontally interpolated coarse rows of the current
//      row pair sit in the wave's LDS (each lane reads back only its own
//      columns) and are rebuilt when the pair changes (row coefficients from
//      a table computed once per launch, k_wave_tables);
//   then the normalisation and both encodings.
// Every partial sum of the float32 inputs is exact in float64, so the prefix
// differences equal the running sums of the workgroup kernel bit for bit, and
// so do all results (same rounding points, same term order).  No workgroup
// barrier: the LDS buffers belong to the wave, whose LDS operations execute
// in program order.
#pragma once
#include "usm.hpp"

namespace fsg {

constexpr int WAVE_NC = 256;   // window columns per wavefront strip
constexpr int WAVE_PB = 81;    // prefix entries per slot: blocks -8 .. 72 of 4 columns
constexpr int WAVE_MAX_R = 32;
constexpr int WAVE_MAX_SRC = 3;
constexpr int WAVE_MAX_NR = 3;

// What the wavefront kernel reads: a compact copy of UsmArgs.
struct WaveArgs {
  const float* win;
  int64_t ld;
  int H, W;                 // window
  int r0, c0, rh, rw;       // region in window coordinates
  int R;
  int rad[WAVE_MAX_NR];
  double inv[WAVE_MAX_NR];
  float wt[WAVE_MAX_NR + WAVE_MAX_SRC];
  float wl;
  int nz;                   // decimated sources (the field, if any, is the last)
  int has_field;
  const float* sp[WAVE_MAX_SRC];
  int64_t sld[WAVE_MAX_SRC];
  // per source: rowT[s * rh + y] = (a0, a1, ty) of region row y, the
  // float64 weight as its two 32-bit halves; the column mapping itself
  SrcRef src[WAVE_MAX_SRC];
  const int4* rowT;
  OutDev o1, o2;
  // block geometry
  int SW, SH, n_strips, n_blocks, per_xcd;
};

// v from lane (lane - shift) under a DPP control; lanes without a source get 0.
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffull), CTRL,
                                             ROW_MASK, BANK_MASK, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK,
                                             BANK_MASK, true);
  return __builtin_bit_cast(double,
                            ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Inclusive prefix sum over the 64 lanes: Hillis-Steele within each 16-lane
// row (row_shr 1, 2, 4, 8), then row 0's total into row 1 and row 2's into
// row 3 (row_bcast:15), then lanes 0-31's total into rows 2-3 (row_bcast:31).
__device__ __forceinline__ double wave_scan_incl(double x) {
  x = x + dpp_f64<0x111>(x);
  x = x + dpp_f64<0x112>(x);
  x = x + dpp_f64<0x114>(x);
  x = x + dpp_f64<0x118>(x);
  x = x + dpp_f64<0x142, 0xa>(x);
  x = x + dpp_f64<0x143, 0xc>(x);
  return x;
}

// Prefix entry k (k in [-32, 292]) of the slot-major buffer: entry k lives in
// slot k & 3 at block (k >> 2) + 8, so the entries a lane needs at a fixed
// offset from its own 4 columns sit at consecutive addresses across lanes.
__device__ __forceinline__ int wave_pidx(int k) { return (k & 3) * WAVE_PB + (k >> 2) + 8; }

__device__ __forceinline__ double bits_f64(int lo, int hi) {
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// The epilogue of four consecutive outputs, branch-free per value (the
// uniform choices are made once per call).  post: apply_post's arithmetic;
// store: store_out's (clip, quantisation, NaN -> 0 for integer outputs).
__device__ __forceinline__ float post_value(const PostK& p, float v) {
  if (p.kind == 1) {
    const float q = v / p.sc;                     // NaN stays NaN
    return p.zero ? (isnan(v) ? v : 0.0f) : q;
  }
  if (p.kind == 2) {
    const float q = (v - p.lo) / p.sc;
    return isnan(q) ? q : fmaxf(q, 0.0f);
  }
  return v;
}

__device__ __forceinline__ float dn_value(const OutDev& o, float v) {
  float t = o.a * v;
  t = t + o.b;
  t = rintf(t);
  t = fminf(fmaxf(t, o.dn_min), o.dn_max);
  return isfinite(v) ? t : 0.0f;
}

// Values m_lo .. m_hi - 1 of four consecutive output columns starting at
// region column c, row r.  full4: every lane's group is all-or-nothing
// (uniform), so a lane with outputs writes one vector.
__device__ __forceinline__ void store_out4(const OutDev& o, int r, int c, const float v[4],
                                           int m_lo, int m_hi, bool full4) {
  float w[4];
#pragma unroll
  for (int m = 0; m < 4; m++) w[m] = v[m];
  if (o.clip01) {
#pragma unroll
    for (int m = 0; m < 4; m++) w[m] = isnan(w[m]) ? w[m] : fminf(fmaxf(w[m], 0.0f), 1.0f);
  }
  const int64_t off = (int64_t)r * o.ld + c;
  if (o.dtype == FSG_OUT_F32) {
    float* p = (float*)o.data + off;
    if (full4) {  // one 16-byte store (dword-aligned)
      const float4 q = make_float4(w[0], w[1], w[2], w[3]);
      __builtin_memcpy(p, &q, sizeof(q));
    } else {
#pragma unroll
      for (int m = 0; m < 4; m++)
        if (m >= m_lo && m < m_hi) p[m] = w[m];
    }
    return;
  }
  float d[4];
#pragma unroll
  for (int m = 0; m < 4; m++) d[m] = dn_value(o, w[m]);
  if (o.dtype == FSG_OUT_I16) {
    int16_t* p = (int16_t*)o.data + off;
#pragma unroll
    for (int m = 0; m < 4; m++)
      if (m >= m_lo && m < m_hi) p[m] = (int16_t)(int)d[m];
  } else {
    uint8_t* p = (uint8_t*)o.data + off;
    if (full4 && (((uintptr_t)o.data | (uintptr_t)o.ld) & 3) == 0) {
      *(uint32_t*)p = (uint32_t)(int)d[0] | ((uint32_t)(int)d[1] << 8) |
                      ((uint32_t)(int)d[2] << 16) | ((uint32_t)(int)d[3] << 24);
    } else {
#pragma unroll
      for (int m = 0; m < 4; m++)
        if (m >= m_lo && m < m_hi) p[m] = (uint8_t)(int)d[m];
    }
  }
}

// One SW x SH block of the region.  P: this wave's LDS prefix buffer (4 *
// WAVE_PB doubles, one radius at a time).  Hc: its LDS copy of the float64
// coarse-row values of each source, [NS][2][WAVE_NC] (lane t uses columns
// 4t .. 4t + 3 only).  FAST: every row and column the block reads lies inside
// the window (no reflection); a lane's four columns are one 16-byte load,
// dword-aligned (gfx9 and later take it as one access at any column).
// KR = 1: the box radii are {2, 8, 32} (every prefix-entry offset an
// immediate); KR = 0: the radii of A.
template <int KR>
__device__ __forceinline__ int box_radius(const WaveArgs& A, int i) {
  if (KR == 1) return i == 0 ? 2 : (i == 1 ? 8 : 32);
  return A.rad[i];
}

template <int NR, int NS, bool FAST, int KR>
__device__ __forceinline__ void usm_wave_block(const WaveArgs& A, int tile, double* P,
                                               double* Hc) {
  const int t = threadIdx.x & 63;
  const int strip = tile % A.n_strips;
  const int seg = tile / A.n_strips;
  const int x0 = strip * A.SW;
  const int y0 = seg * A.SH;
  const int rows = min(A.SH, A.rh - y0);
  const int cols = min(A.SW, A.rw - x0);
  const int R = KR == 1 ? 32 : A.R;
  const int wy0 = A.r0 + y0;
  const int cs = A.c0 + x0 - R;   // window column of strip column 0
  const int cw0 = cs + 4 * t;     // this lane's first window column
  const int H = A.H, W = A.W;
  const int64_t ld = A.ld;
  const float* Wp = A.win;
  // byte offset of this lane's four columns from a row start (FAST loads)
  const unsigned loff = (unsigned)cw0 * 4u;

  auto load4 = [&](int wrow, float v[4]) {
    if (FAST) {
      const char* rp = (const char*)(Wp + (int64_t)wrow * ld);
      float4 q;
      __builtin_memcpy(&q, rp + loff, sizeof(q));
      v[0] = q.x;
      v[1] = q.y;
      v[2] = q.z;
      v[3] = q.w;
    } else {
      const int rr = __builtin_amdgcn_readfirstlane(refl(wrow, H));
      const float* rp = Wp + (int64_t)rr * ld;
#pragma unroll
      for (int m = 0; m < 4; m++) v[m] = rp[refl(cw0 + m, W)];
    }
  };

  // this lane's columns m_lo .. m_hi - 1 are outputs, at region columns xr0 + m
  const int xr0 = x0 + 4 * t - R;
  const int m_lo = max(0, R - 4 * t);
  const int m_hi = min(4, R + cols - 4 * t);

  // vertical sums for the row before the first output row, rows in
  // ascending order (eight rows' loads in flight at a time)
  double S[NR][4];
#pragma unroll
  for (int i = 0; i < NR; i++)
#pragma unroll
    for (int m = 0; m < 4; m++) S[i][m] = 0.0;
  auto add_row = [&](int d, const float v[4]) {
#pragma unroll
    for (int i = 0; i < NR; i++) {
      const int r = box_radius<KR>(A, i);
      if (d >= -r && d <= r) {
#pragma unroll
        for (int m = 0; m < 4; m++) S[i][m] += (double)v[m];
      }
    }
  };
  int d = -R;
#pragma unroll 1
  for (; d + 7 <= R; d += 8) {
    float v[8][4];
#pragma unroll
    for (int k = 0; k < 8; k++) load4(wy0 - 1 + d + k, v[k]);
#pragma unroll
    for (int k = 0; k < 8; k++) add_row(d + k, v[k]);
  }
#pragma unroll 1
  for (; d <= R; d++) {
    float v[4];
    load4(wy0 - 1 + d, v);
    add_row(d, v);
  }

  if (t == 0) P[wave_pidx(0)] = 0.0;   // prefix entry k = 0 (the empty sum)
  double* const Pt = P + t;            // block index t of every slot
  int curA[NS > 0 ? NS : 1];
#pragma unroll
  for (int s = 0; s < NS; s++) curA[s] = -1;
  // this lane's region columns for the bilinear tables (clamped to the output span)
  int xc[4];
#pragma unroll
  for (int m = 0; m < 4; m++) xc[m] = min(max(xr0 + m, x0), x0 + cols - 1);

  const PostK p1 = load_post(A.o1);
  const bool two = A.o2.data != nullptr;
  // every lane's four columns are all outputs or none (uniform)
  const bool full4 = (R & 3) == 0 && (cols & 3) == 0 && (x0 & 3) == 0;

  // the rows an output row reads: its own, and the entering and leaving row
  // of each radius; loaded one output row ahead
  float zq[4], vq[NR][2][4];
  auto load_rows = [&](int wy) {
    load4(wy, zq);
#pragma unroll
    for (int i = 0; i < NR; i++) {
      const int r = box_radius<KR>(A, i);
      load4(wy + r, vq[i][0]);
      load4(wy - r - 1, vq[i][1]);
    }
  };
  load_rows(wy0);

  for (int yy = 0; yy < rows; yy++) {
    float z[4], vn[NR][4], vo[NR][4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
      z[m] = zq[m];
#pragma unroll
      for (int i = 0; i < NR; i++) {
        vn[i][m] = vq[i][0][m];
        vo[i][m] = vq[i][1][m];
      }
    }
    if (yy + 1 < rows) load_rows(wy0 + yy + 1);
    float acc[4];
#pragma unroll
    for (int i = 0; i < NR; i++) {
      // 1. vertical running sums -> float32 means
      const int r = box_radius<KR>(A, i);
      const double inv = A.inv[i];
      float V[4];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        S[i][m] += (double)vn[i][m];
        S[i][m] -= (double)vo[i][m];
        V[m] = (float)(S[i][m] * inv);
      }
      // 2. row prefix of the means -> LDS, then the box sums
      const double q0 = (double)V[0];
      const double q1 = q0 + (double)V[1];
      const double q2 = q1 + (double)V[2];
      const double q3 = q2 + (double)V[3];
      const double g = wave_scan_incl(q3);
      const double e = g - q3;
      Pt[1 * WAVE_PB + 8] = e + q0;   // entry 4t + 1
      Pt[2 * WAVE_PB + 8] = e + q1;   // entry 4t + 2
      Pt[3 * WAVE_PB + 8] = e + q2;   // entry 4t + 3
      Pt[0 * WAVE_PB + 9] = g;        // entry 4t + 4
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const float w = A.wt[i];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        // entries 4t + m + r + 1 and 4t + m - r: slot and block offsets
        // relative to lane t depend on m and r only
        const int kp = m + r + 1, km = m - r;
        const double box = Pt[(kp & 3) * WAVE_PB + (kp >> 2) + 8] -
                           Pt[(km & 3) * WAVE_PB + (km >> 2) + 8];
        const float bm = (float)(box * inv);
        const float diff = w * (z[m] - bm);
        acc[m] = (i == 0) ? diff : acc[m] + diff;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // 3. decimated terms, field, post, store
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int4 rt = A.rowT[s * A.rh + y0 + yy];
      const int a0 = __builtin_amdgcn_readfirstlane(rt.x);
      double* hA = Hc + (s * 2 + 0) * WAVE_NC + 4 * t;
      double* hB = Hc + (s * 2 + 1) * WAVE_NC + 4 * t;
      if (a0 != curA[s]) {
        // once per coarse row pair: the horizontal interpolation of the two
        // coarse rows at this lane's four columns (src_axis's arithmetic);
        // two columns at a time keep it off the steady rows' registers
        const float* ra = A.sp[s] + (int64_t)a0 * A.sld[s];
        const float* rb = A.sp[s] + (int64_t)rt.y * A.sld[s];
#pragma unroll 2
        for (int m = 0; m < 4; m++) {
          int i0, i1;
          double tx;
          src_axis(A.src[s], 0, A.c0 + xc[m], i0, i1, tx);
          const float va = ra[i0], vb = ra[i1], vc = rb[i0], vd = rb[i1];
          const double mx = 1.0 - tx;
          hA[m] = (double)va * mx + (double)vb * tx;
          hB[m] = (double)vc * mx + (double)vd * tx;
        }
        curA[s] = a0;
      }
      const double ty = bits_f64(rt.z, rt.w);
      const double my = 1.0 - ty;
      if (s < A.nz) {
        const float wt = A.wt[NR + s];
#pragma unroll
        for (int m = 0; m < 4; m++) {
          const float mean = (float)(hA[m] * my + hB[m] * ty);
          const float diff = wt * (z[m] - mean);
          acc[m] = acc[m] + diff;
        }
      } else {
#pragma unroll
        for (int m = 0; m < 4; m++) {
          const float up = (float)(hA[m] * my + hB[m] * ty);
          const float large = A.wl * z[m] - up;
          acc[m] = (isnan(z[m]) ? __builtin_nanf("") : acc[m]) + large;
        }
      }
    }
    if (!A.has_field) {
#pragma unroll
      for (int m = 0; m < 4; m++) acc[m] = isnan(z[m]) ? __builtin_nanf("") : acc[m];
    }
    if (m_lo < m_hi) {
      float v[4];
#pragma unroll
      for (int m = 0; m < 4; m++) v[m] = post_value(p1, acc[m]);
      const int yr = y0 + yy;
      store_out4(A.o1, yr, xr0, v, m_lo, m_hi, full4);
      if (two) store_out4(A.o2, yr, xr0, v, m_lo, m_hi, full4);
    }
  }
}

}  // namespace fsg

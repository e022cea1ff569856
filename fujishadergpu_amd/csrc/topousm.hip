// TopoUSM Fast: multiscale box-mean unsharp mask.
//
// Per window (the reference's tile window), for radii r_i with weights w_i
// (compute_topousm_fast_efficient_block, _impl_topousm_fast.py:49-100):
//   ds_i == 1, r_i >= 2 : NaN-aware box mean of size 2r+1, mode 'reflect'
//   ds_i == 1, r_i <= 1 : NaN-aware Gaussian sigma=1, mode 'nearest'
//   ds_i  > 1           : NaN-aware ds x ds block mean (+ enclosed void fill),
//                         box (or sigma=1 Gaussian) on that grid, then a
//                         corner-aligned bilinear zoom back to the window
//   sum_i w_i * (z - mean_i), NaN restored, + (w_large * z - bilinear(field))
//   for the large-radius overview field, then / f32(p99 scale).
//
// Box means follow scipy.ndimage.uniform_filter's structure: axis-0 means
// rounded to float32, then axis-1 means of those, with exact (float64)
// window sums; NaN-aware means are where(sw > 0, sv / sw, 0) in float32.
//
// Main kernel layout (k_usm_main): one workgroup of 256 threads owns a strip
// of SW output columns x SH output rows.  Threads first walk down their column
// (SW + 2R columns, R = largest full-resolution box radius) keeping running
// float64 window sums per radius; every K rows the axis-0 means land in LDS,
// then 256 threads re-partition as K rows x S segments and slide running sums
// along each row segment.  Results are staged in LDS and written coalesced.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "usm.hpp"
#include "usm_wave.hpp"

namespace fsg {

__device__ __forceinline__ float rawv(const float* p, int64_t ld, int r, int c) {
  return p[(int64_t)r * ld + c];
}

// NaN-aware sigma=1 Gaussian mean at (r, c) of an h x w grid, scipy order,
// mode 'nearest': where(sw > 0, sv / sw, 0)  (_nan_utils.py:18-31).  Without
// NaN this equals gaussian_filter(block) (sw rounds to 1.0f).
__device__ __noinline__ float gauss9_mean(const float* p, int64_t ld, int h, int w, int r, int c,
                                          Gauss9 G) {
  float vsv[9], vsw[9];
  for (int k = 0; k < 9; k++) {
    const int cc = clampi(c + k - 4, w);
    const float x0 = rawv(p, ld, r, cc);
    const bool v0 = !isnan(x0);
    double sv = (v0 ? (double)x0 : 0.0) * G.g[0];
    double sw = (v0 ? 1.0 : 0.0) * G.g[0];
    for (int d = 4; d >= 1; d--) {
      const float a = rawv(p, ld, clampi(r - d, h), cc);
      const float b = rawv(p, ld, clampi(r + d, h), cc);
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
  for (int d = 4; d >= 1; d--) {
    sv += ((double)vsv[4 - d] + (double)vsv[4 + d]) * G.g[d];
    sw += ((double)vsw[4 - d] + (double)vsw[4 + d]) * G.g[d];
  }
  const float fsv = (float)sv, fsw = (float)sw;
  return fsw > 0.0f ? fsv / fsw : 0.0f;
}

// ------------------------------------------------------------ coarse kernels

// _downsample_nan_aware cell means + "any NaN cell" flag.  One block covers
// 256 consecutive source columns of one coarse row (f source rows): each lane
// sums its column, then groups of f lanes reduce by shuffles (f is a power of
// two <= 16).  Cell sums are exact in float64, so the order does not matter.
__device__ __forceinline__ void decimate_block(const WinDev& W, int f, float* __restrict__ out,
                                               int oh, int ow, int* __restrict__ nan_flag,
                                               int bx, int by) {
  const int orow = by;
  const int c = bx * 256 + threadIdx.x;  // source column
  const int r0 = orow * f, r1 = min(r0 + f, W.h);
  double tot = 0.0;
  int cnt = 0;
  if (c < W.w) {
    for (int r = r0; r < r1; r++) {
      const float v = W.p[(int64_t)r * W.ld + c];
      if (isfinite(v)) {
        tot += (double)v;
        cnt++;
      }
    }
  }
  for (int off = 1; off < f; off <<= 1) {
    tot += __shfl_down(tot, off, 64);
    cnt += __shfl_down(cnt, off, 64);
  }
  if ((threadIdx.x & (f - 1)) == 0) {
    const int oc = c / f;
    if (oc < ow && orow < oh) {
      float res;
      if (cnt > 0) {
        res = (float)tot / (float)cnt;
      } else {
        res = __builtin_nanf("");
        if (nan_flag) atomicOr(nan_flag, 1);
      }
      out[(int64_t)orow * ow + oc] = res;
    }
  }
}

__global__ void __launch_bounds__(256)
k_decimate(WinDev W, int f, float* __restrict__ out, int oh, int ow, int* __restrict__ nan_flag) {
  for (int orow = blockIdx.y; orow < oh; orow += gridDim.y)   // grid-stride over rows
    decimate_block(W, f, out, oh, ow, nan_flag, blockIdx.x, orow);
}

// Columns c..c+3 of row r (c < w): one 16-byte load where all four lie inside
// (dword-aligned: windows start at any column), else per column with the
// columns past w absent (NaN).
__device__ __forceinline__ void row4(const WinDev& W, int r, int c, float v[4]) {
  const float* p = W.p + (int64_t)r * W.ld + c;
  if (c + 3 < W.w) {
    float4 q;
    __builtin_memcpy(&q, p, sizeof(q));
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = c + k < W.w ? p[k] : __builtin_nanf("");
  }
}

// float64 sum and count of the finite values of four columns c..c+3 (c < w)
// over rows r0..r1-1, added row by row and column by column.  The loads of
// four rows are issued before their values are added (a runtime-length loop
// would otherwise wait for each row's load in turn).
__device__ __forceinline__ void dec_rows4(const WinDev& W, int c, int r0, int r1, double& tot,
                                          int& cnt) {
  int r = r0;
  for (; r + 4 <= r1; r += 4) {
    float v[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++) row4(W, r + i, c, v[i]);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (isfinite(v[i][k])) {
          tot += (double)v[i][k];
          cnt++;
        }
  }
  for (; r < r1; r++) {
    float v[4];
    row4(W, r, c, v);
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (isfinite(v[k])) {
        tot += (double)v[k];
        cnt++;
      }
  }
}

// The same cell means for f >= 4 (a power of two <= 16) with four source
// columns per lane: one 16-byte load per row (dword-aligned, which gfx9 and
// later take as one access), then groups of f/4 lanes reduce by shuffles.
// The sums are exact in float64, so the cells equal k_decimate's bit for bit.
// Block bx covers source columns bx*1024 .. bx*1024 + 1023 of output row orow;
// every lane of the block must call it (the shuffles).
__device__ __forceinline__ void decimate4_block(const WinDev& W, int f, float* __restrict__ out,
                                                int oh, int ow, int* __restrict__ nan_flag,
                                                int bx, int orow) {
  const int c = (bx * 256 + threadIdx.x) * 4;  // first of this lane's source columns
  const int r0 = orow * f, r1 = min(r0 + f, W.h);
  double tot = 0.0;
  int cnt = 0;
  if (c < W.w) dec_rows4(W, c, r0, r1, tot, cnt);
  for (int off = 1; off < f / 4; off <<= 1) {
    tot += __shfl_down(tot, off, 64);
    cnt += __shfl_down(cnt, off, 64);
  }
  if ((threadIdx.x & (f / 4 - 1)) == 0) {
    const int oc = c / f;
    if (oc < ow && orow < oh) {
      float res;
      if (cnt > 0) {
        res = (float)tot / (float)cnt;
      } else {
        res = __builtin_nanf("");
        if (nan_flag) atomicOr(nan_flag, 1);
      }
      out[(int64_t)orow * ow + oc] = res;
    }
  }
}

__global__ void __launch_bounds__(256)
k_decimate4(WinDev W, int f, float* __restrict__ out, int oh, int ow, int* __restrict__ nan_flag) {
  // grid-stride over output rows (block-uniform, so the shuffles see every lane)
  for (int orow = blockIdx.y; orow < oh; orow += gridDim.y)
    decimate4_block(W, f, out, oh, ow, nan_flag, blockIdx.x, orow);
}

// Several independent decimations in one launch (the coarse preparation of a
// batch of NaN-free windows): flat block index -> job by binary search over
// the jobs' block-count prefix, then the per-window kernel's block.
__device__ __forceinline__ int batch_job_of(const int* __restrict__ pre, int n, int g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256)
k_decimate_batch(const DecJob* __restrict__ jobs, const int* __restrict__ pre, int n) {
  const int j = batch_job_of(pre, n, blockIdx.x);
  const DecJob J = jobs[j];
  const int local = blockIdx.x - pre[j];
  if (J.f >= 4)  // job-uniform: nbx counts 1024-column blocks
    decimate4_block(J.W, J.f, J.out, J.oh, J.ow, nullptr, local % J.nbx, local / J.nbx);
  else
    decimate_block(J.W, J.f, J.out, J.oh, J.ow, nullptr, local % J.nbx, local / J.nbx);
}

// Separable Gaussian pass of the enclosed-void fill, mode 'nearest', scipy
// symmetric order.  Pass 0: axis 0 of (where(nan,0,x), valid) -> (tv, tw);
// pass 1: axis 1 of (tv, tw) -> (sv, sw).  Gated on the NaN flag.
constexpr int VF_MAX_RAD = 320;
struct VFWeights {
  int rad;
  int by_value;    // 1: g[] holds the scipy weights; 0: evaluate on device
  double sigma2, inv_sum;
  double g[VF_MAX_RAD + 1];
};

__device__ __forceinline__ double vf_w(const VFWeights& V, int d) {
  return V.by_value ? V.g[d] : exp(-0.5 / V.sigma2 * (double)(d * d)) * V.inv_sum;
}

__global__ void k_voidfill_pass(const float* __restrict__ a, const float* __restrict__ b,
                                float* __restrict__ oa, float* __restrict__ ob, int h, int w,
                                VFWeights V, int axis, const int* __restrict__ nan_flag) {
  if (*nan_flag == 0) return;
  const int rad = V.rad;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)h * w;
       e += (int64_t)gridDim.x * blockDim.x) {
  const int r = (int)(e / w), c = (int)(e - (int64_t)r * w);
  double sv, sw;
  if (axis == 0) {
    // a = coarse (NaN -> value 0, weight 0)
    const float x0 = a[(int64_t)r * w + c];
    sv = (isnan(x0) ? 0.0 : (double)x0) * vf_w(V, 0);
    sw = (isnan(x0) ? 0.0 : 1.0) * vf_w(V, 0);
    for (int d = rad; d >= 1; d--) {
      const float xa = a[(int64_t)clampi(r - d, h) * w + c];
      const float xb = a[(int64_t)clampi(r + d, h) * w + c];
      sv += ((isnan(xa) ? 0.0 : (double)xa) + (isnan(xb) ? 0.0 : (double)xb)) * vf_w(V, d);
      sw += ((isnan(xa) ? 0.0 : 1.0) + (isnan(xb) ? 0.0 : 1.0)) * vf_w(V, d);
    }
  } else {
    const int64_t row = (int64_t)r * w;
    sv = (double)a[row + c] * vf_w(V, 0);
    sw = (double)b[row + c] * vf_w(V, 0);
    for (int d = rad; d >= 1; d--) {
      sv += ((double)a[row + clampi(c - d, w)] + (double)a[row + clampi(c + d, w)]) * vf_w(V, d);
      sw += ((double)b[row + clampi(c - d, w)] + (double)b[row + clampi(c + d, w)]) * vf_w(V, d);
    }
  }
  oa[(int64_t)r * w + c] = (float)sv;
  ob[(int64_t)r * w + c] = (float)sw;
  }
}

__global__ void k_zero_int(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// coarse = where(isnan & sw > 0.5, sv / max(sw, 1e-6), coarse)
__global__ void k_voidfill_apply(float* __restrict__ coarse, const float* __restrict__ sv,
                                 const float* __restrict__ sw, int n,
                                 const int* __restrict__ nan_flag) {
  if (*nan_flag == 0) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float x = coarse[i];
    if (isnan(x) && sw[i] > 0.5f) coarse[i] = sv[i] / fmaxf(sw[i], 1e-6f);
  }
}

// NaN-aware box mean on a small grid as two transposing axis-0 passes.  A
// 64-lane block owns 64 columns x 64 rows: each lane keeps a running float64
// window sum down its column, the 64x64 results go through LDS and are written
// transposed, so both passes read and write coalesced rows.
//   pass 0: x (h x w, NaN-aware)  -> tv^T, tw^T (w x h): float32 axis-0 means
//   pass 1: tv^T, tw^T (w x h)    -> out (h x w): axis-1 means of those and
//           where(sw > 0, sv / sw, 0)   (_nan_utils.py:34-47 structure)
template <int PASS>
__device__ __forceinline__ void box_t_block(const float* __restrict__ a,
                                            const float* __restrict__ b, int64_t lda, int h,
                                            int w, int r, double inv, float* __restrict__ oa,
                                            float* __restrict__ ob, int64_t ldo, int bx, int by,
                                            float (*ta)[65], float (*tb)[65]) {
  const int lane = threadIdx.x;
  const int c = bx * 64 + lane;  // column of the (h x w) input
  const int y0 = by * 64;
  const int y1 = min(y0 + 64, h);
  if (c < w) {
    const float* pa = a + c;
    const float* pb = PASS ? b + c : nullptr;
    double s = 0.0, sw = 0.0;
    // window rows y0-r .. y0+r, 16 independent loads per batch
    for (int d0 = -r; d0 <= r; d0 += 16) {
      float va[16], vb[16];
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int64_t o = (int64_t)refl(y0 + min(d0 + e, r), h) * lda;
        va[e] = pa[o];
        if (PASS) vb[e] = pb[o];
      }
#pragma unroll
      for (int e = 0; e < 16; e++) {
        if (d0 + e > r) break;
        if (PASS == 0) {
          if (!isnan(va[e])) { s += (double)va[e]; sw += 1.0; }
        } else {
          s += (double)va[e];
          sw += (double)vb[e];
        }
      }
    }
    // running sums down the block's rows, 8 rows of loads per batch
    for (int yb = y0; yb < y1; yb += 8) {
      float na[8], oa_[8], nb[8], obv[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int y = min(yb + e, y1 - 1);
        const int64_t on = (int64_t)refl(y + r, h) * lda, oo = (int64_t)refl(y - r - 1, h) * lda;
        na[e] = pa[on];
        oa_[e] = pa[oo];
        if (PASS) {
          nb[e] = pb[on];
          obv[e] = pb[oo];
        }
      }
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int y = yb + e;
        if (y >= y1) break;
        if (y > y0) {
          if (PASS == 0) {
            if (!isnan(na[e])) { s += (double)na[e]; sw += 1.0; }
            if (!isnan(oa_[e])) { s -= (double)oa_[e]; sw -= 1.0; }
          } else {
            s += (double)na[e] - (double)oa_[e];
            sw += (double)nb[e] - (double)obv[e];
          }
        }
        const float fv = (float)(s * inv), fw = (float)(sw * inv);
        if (PASS == 0) {
          ta[y - y0][lane] = fv;
          tb[y - y0][lane] = fw;
        } else {
          ta[y - y0][lane] = fw > 0.0f ? fv / fw : 0.0f;
        }
      }
    }
  }
  __syncthreads();
  // transposed write: output row = input column, output column = input row
  const int ncol = min(64, w - bx * 64);
  for (int cc = 0; cc < ncol; cc++) {
    const int y = y0 + lane;
    if (y < y1) {
      const int64_t o = (int64_t)(bx * 64 + cc) * ldo + y;
      oa[o] = ta[lane][cc];
      if (PASS == 0) ob[o] = tb[lane][cc];
    }
  }
}

template <int PASS>
__global__ void __launch_bounds__(64)
k_box_t(const float* __restrict__ a, const float* __restrict__ b, int64_t lda, int h, int w,
        int r, double inv, float* __restrict__ oa, float* __restrict__ ob, int64_t ldo) {
  __shared__ float ta[64][65];
  __shared__ float tb[64][65];
  box_t_block<PASS>(a, b, lda, h, w, r, inv, oa, ob, ldo, blockIdx.x, blockIdx.y, ta, tb);
}

// Box passes of several independent grids in one launch (see k_decimate_batch).
template <int PASS>
__global__ void __launch_bounds__(64)
k_box_t_batch(const BoxJob* __restrict__ jobs, const int* __restrict__ pre, int n) {
  __shared__ float ta[64][65];
  __shared__ float tb[64][65];
  const int j = batch_job_of(pre, n, blockIdx.x);
  const BoxJob J = jobs[j];
  const int local = blockIdx.x - pre[j];
  box_t_block<PASS>(J.a, J.b, J.lda, J.h, J.w, J.r, J.inv, J.oa, J.ob, J.ldo, local % J.nbx,
                    local / J.nbx, ta, tb);
}

// ---- coarse means of many NaN-free windows (fsg_topousm_coarse_means_batch)
//
// Job tables travel as kernel arguments.  A decimation job reads its window
// once for one factor fa, or for two factors fa < fb (fb = 2 fa or 4 fa, both
// powers of two in 4..16) when two jobs share a window.
constexpr int CMB_MAX_DEC = 16;
constexpr int CMB_MAX_BOX = 24;

struct DualDecJob {
  WinDev W;
  float* oa;       // fa grid (ceil(h/fa) x ceil(w/fa), dense)
  float* ob;       // fb grid, or nullptr
  int fa, fb;      // fb = 0: one factor
  int oha, owa, ohb, owb;
  int nbx;         // blocks per band (1024 source columns each)
};
struct DecTab {
  DualDecJob j[CMB_MAX_DEC];
  int pre[CMB_MAX_DEC + 1];
  int n;
};
struct BoxTab {
  BoxJob j[CMB_MAX_BOX];
  int pre[CMB_MAX_BOX + 1];
  int n;
};

__device__ __forceinline__ int tab_job_of(const int* pre, int n, int g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// reduce over groups of g lanes (g a power of two <= 4), result in the group's first lane
__device__ __forceinline__ void dec_reduce(double& tot, int& cnt, int g) {
  for (int off = 1; off < g; off <<= 1) {
    tot += __shfl_down(tot, off, 64);
    cnt += __shfl_down(cnt, off, 64);
  }
}

// Block (bx, by): source columns bx*1024 + 4 lane .. +3; source rows of coarse
// row `by` of the largest factor.  Cell sums are exact in float64, so a cell's
// value does not depend on how its rows and columns are grouped: the cells
// equal k_decimate4's bit for bit.
__global__ void __launch_bounds__(256) k_decimate_tab(DecTab T) {
  const int j = tab_job_of(T.pre, T.n, blockIdx.x);
  const DualDecJob& J = T.j[j];
  const int local = blockIdx.x - T.pre[j];
  const int bx = local % J.nbx, by = local / J.nbx;
  const int c = (bx * 256 + threadIdx.x) * 4;
  if (J.fb == 0) {
    const int f = J.fa;
    const int r0 = by * f, r1 = min(r0 + f, J.W.h);
    double tot = 0.0;
    int cnt = 0;
    if (c < J.W.w) dec_rows4(J.W, c, r0, r1, tot, cnt);
    dec_reduce(tot, cnt, f / 4);
    if ((threadIdx.x & (f / 4 - 1)) == 0) {
      const int oc = c / f;
      if (oc < J.owa && by < J.oha)
        J.oa[(int64_t)by * J.owa + oc] = cnt > 0 ? (float)tot / (float)cnt : __builtin_nanf("");
    }
    return;
  }
  const int fa = J.fa, fb = J.fb, m = fb / fa;
  double tb = 0.0;
  int cb = 0;
  for (int g = 0; g < m; g++) {
    const int r0 = by * fb + g * fa, r1 = min(r0 + fa, J.W.h);
    double ta = 0.0;
    int ca = 0;
    if (c < J.W.w && r0 < r1) dec_rows4(J.W, c, r0, r1, ta, ca);
    tb += ta;
    cb += ca;
    dec_reduce(ta, ca, fa / 4);
    const int ra = by * m + g;
    if ((threadIdx.x & (fa / 4 - 1)) == 0) {
      const int oc = c / fa;
      if (oc < J.owa && ra < J.oha)
        J.oa[(int64_t)ra * J.owa + oc] = ca > 0 ? (float)ta / (float)ca : __builtin_nanf("");
    }
  }
  dec_reduce(tb, cb, fb / 4);
  if ((threadIdx.x & (fb / 4 - 1)) == 0) {
    const int oc = c / fb;
    if (oc < J.owb && by < J.ohb)
      J.ob[(int64_t)by * J.owb + oc] = cb > 0 ? (float)tb / (float)cb : __builtin_nanf("");
  }
}

// Box means along axis 0 of a grid WITHOUT NaN, written transposed: one
// 64-column x 64-row tile per one-wave block.  Both passes of a NaN-free grid
// are this operation (pass 0 on the grid, pass 1 on its transposed means):
// with no NaN the NaN-aware passes' count planes are exactly 1.0f, so
// fv / fw = fv and the results equal box_t_block's bit for bit (each pass
// updates its running sum in the order box_t_block's pass does), from half the
// traffic and half the LDS.
template <int PASS>
__device__ __forceinline__ void box_plain_block(const float* __restrict__ a, int64_t lda, int h,
                                                int w, int r, double inv, float* __restrict__ o,
                                                int64_t ldo, int nsub, int bx, int by,
                                                float (*ta)[65]) {
  // a band of nsub 64-row sub-bands: the window sum is built once per band
  // (2r + 1 rows) and runs on across its sub-bands, each transposed via LDS
  const int lane = threadIdx.x;
  const int c = bx * 64 + lane;
  const int yb0 = by * 64 * nsub;
  const float* pa = a + c;
  double s = 0.0;
  if (c < w) {
    // window rows yb0-r .. yb0+r, 64 independent loads per batch
    for (int d0 = -r; d0 <= r; d0 += 64) {
      float va[64];
#pragma unroll
      for (int e = 0; e < 64; e++) va[e] = pa[(int64_t)refl(yb0 + min(d0 + e, r), h) * lda];
#pragma unroll
      for (int e = 0; e < 64; e++)
        if (d0 + e <= r) s += (double)va[e];
    }
  }
  for (int sb = 0; sb < nsub; sb++) {
    const int y0 = yb0 + sb * 64;
    if (y0 >= h) break;  // block-uniform
    const int y1 = min(y0 + 64, h);
    if (c < w) {
      for (int yb = y0; yb < y1; yb += 32) {
        float na[32], ov[32];
#pragma unroll
        for (int e = 0; e < 32; e++) {
          const int y = min(yb + e, y1 - 1);
          na[e] = pa[(int64_t)refl(y + r, h) * lda];
          ov[e] = pa[(int64_t)refl(y - r - 1, h) * lda];
        }
#pragma unroll
        for (int e = 0; e < 32; e++) {
          const int y = yb + e;
          if (y < y1) {
            if (y > yb0) {
              if (PASS == 0) {
                s += (double)na[e];
                s -= (double)ov[e];
              } else {
                s += (double)na[e] - (double)ov[e];
              }
            }
            ta[y - y0][lane] = (float)(s * inv);
          }
        }
      }
    }
    __syncthreads();
    const int ncol = min(64, w - bx * 64);
    for (int cc = 0; cc < ncol; cc++) {
      const int y = y0 + lane;
      if (y < y1) o[(int64_t)(bx * 64 + cc) * ldo + y] = ta[lane][cc];
    }
    __syncthreads();
  }
}

template <int PASS>
__global__ void __launch_bounds__(64) k_box_plain_tab(BoxTab T) {
  __shared__ float ta[64][65];
  const int j = tab_job_of(T.pre, T.n, blockIdx.x);
  const BoxJob& J = T.j[j];
  const int local = blockIdx.x - T.pre[j];
  box_plain_block<PASS>(J.a, J.lda, J.h, J.w, J.r, J.inv, J.oa, J.ldo, J.nsub > 1 ? J.nsub : 1,
                        local % J.nbx, local / J.nbx, ta);
}

__global__ void k_gauss9_grid(const float* __restrict__ x, int64_t ldx, int h, int w,
                              float* __restrict__ out, int64_t ldo, Gauss9 G) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= w) return;
  for (int r = blockIdx.y; r < h; r += gridDim.y)
    out[(int64_t)r * ldo + c] = gauss9_mean(x, ldx, h, w, r, c, G);
}

// field = (first ? 0 : field) + f32(w) * mean, in radius order
__global__ void k_field_accum(float* __restrict__ field, int64_t ldf, const float* __restrict__ m,
                              int64_t ldm, int h, int w, float wt, int first) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= w) return;
  for (int r = blockIdx.y; r < h; r += gridDim.y) {
    const float t = wt * m[(int64_t)r * ldm + c];
    float* p = field + (int64_t)r * ldf + c;
    *p = first ? t : (*p + t);
  }
}

// ------------------------------------------------------------ main kernel

// bilinear value from a staged tile (rows from r_lo, cols from c_lo)
__device__ __forceinline__ float src_sample(const SrcRef& S, const float* tile, int r_lo, int c_lo,
                                            int yw, int xw) {
  int a0, a1, b0, b1;
  double ty, tx;
  src_axis(S, 1, yw, a0, a1, ty);
  src_axis(S, 0, xw, b0, b1, tx);
  a0 -= r_lo; a1 -= r_lo; b0 -= c_lo; b1 -= c_lo;
  const double a = tile[a0 * S.cols_max + b0], b = tile[a0 * S.cols_max + b1];
  const double c = tile[a1 * S.cols_max + b0], d = tile[a1 * S.cols_max + b1];
  const double top = a * (1.0 - tx) + b * tx;
  const double bot = c * (1.0 - tx) + d * tx;
  return (float)(top * (1.0 - ty) + bot * ty);
}

template <int NR, bool NANAW, int K, bool GAUSS>
__global__ void __launch_bounds__(256) k_usm_main(UsmArgs A) {
  constexpr int KS = 8;           // rows per batch of column loads
  constexpr int NRA = NR > 0 ? NR : 1;
  extern __shared__ float lds[];
  const int NC = A.SW + 2 * A.R;
  const int LDC = NC + 1;         // padded LDS row stride
  float* Vs = lds;                                   // [NR][K][LDC] axis-0 means
  float* Ws = Vs + NR * K * LDC;                     // [NR][K][LDC] validity means
  float* Zs = NANAW ? (Ws + NR * K * LDC) : Ws;      // [K][LDC] centre values / results
  float* Src = Zs + K * LDC;                         // staging tiles of bilinear sources
  // bilinear coefficient tables (8-byte entries from a 16-byte boundary):
  // per source, K row entries and SW column entries
  int2* RowI = (int2*)(lds + ((NR * K * LDC * (NANAW ? 2 : 1) + K * LDC + A.src_floats + 3) & ~3));
  double* RowT = (double*)(RowI + MAX_SRC * K);
  int2* ColI = (int2*)(RowT + MAX_SRC * K);
  double* ColT = (double*)(ColI + A.n_src * A.SW);

  __shared__ int s_clo[MAX_SRC], s_ncol[MAX_SRC];

  const int t = threadIdx.x;
  const int strip = blockIdx.x % A.n_strips;
  const int seg = blockIdx.x / A.n_strips;
  const int x0 = strip * A.SW;           // first output column (region coordinates)
  const int y0 = seg * A.SH;             // first output row (region coordinates)
  const int rows = min(A.SH, A.rh - y0);
  const int cols = min(A.SW, A.rw - x0);
  if (rows <= 0 || cols <= 0) return;
  const int wx0 = A.c0 + x0, wy0 = A.r0 + y0;  // window coordinates
  const int H = A.W.h;
  const int64_t ld = A.W.ld;

  // per-source column ranges and column coefficient tables (once per block)
  for (int s = 0; s < A.n_src; s++) {
    const SrcRef S = A.src[s];
    int a0, a1, b0, b1;
    double tt;
    src_axis(S, 0, wx0, a0, a1, tt);
    src_axis(S, 0, wx0 + cols - 1, b0, b1, tt);
    if (t == 0) {
      s_clo[s] = a0;
      s_ncol[s] = min(b1 - a0 + 1, S.cols_max);
    }
    // column indices relative to the staged tile's first column a0
    for (int c = t; c < A.SW; c += 256) {
      int i0, i1;
      double tx;
      src_axis(S, 0, wx0 + min(c, cols - 1), i0, i1, tx);
      ColI[s * A.SW + c] = make_int2(i0 - a0, i1 - a0);
      ColT[s * A.SW + c] = tx;
    }
  }

  // column owned in the vertical stage; every row address is block-uniform,
  // so loads are (uniform row pointer)[per-lane column]
  const bool has_col = t < NC;
  const int jw = has_col ? refl(wx0 - A.R + t, A.W.w) : 0;
  const float* Wp = A.W.p;

  double S[NRA];
  int Cn[NRA];
#pragma unroll
  for (int i = 0; i < NR; i++) {
    double s = 0.0;
    int n = 0;
    const int r = A.rad[i];
    // rows wy0-1-r .. wy0-1+r in batches of 8 independent loads (exact sums)
    for (int d0 = -r; d0 <= r; d0 += 8) {
      float v[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int d = min(d0 + e, r);
        v[e] = urow(Wp, reflf(wy0 - 1 + d, H), ld)[jw];
      }
#pragma unroll
      for (int e = 0; e < 8; e++) {
        if (d0 + e > r) break;
        if (NANAW) {
          if (!isnan(v[e])) { s += (double)v[e]; n++; }
        } else {
          s += (double)v[e];
        }
      }
    }
    S[i] = s;
    Cn[i] = n;
  }
  __syncthreads();

  const int n_chunks = (rows + K - 1) / K;
  const int segs = 256 / K;
  const int L = A.SW / segs;

  for (int ch = 0; ch < n_chunks; ch++) {
    const int wyc = wy0 + ch * K;
    const int nrow = min(K, rows - ch * K);
    // ---------------- axis-0 stage: batched column loads, running sums
    for (int kb = 0; kb < K; kb += KS) {
      float vz[KS], vn[NRA][KS], vo[NRA][KS];
#pragma unroll
      for (int k = 0; k < KS; k++) {
        const int yw = wyc + kb + k;
        vz[k] = urow(Wp, min(yw, H - 1), ld)[jw];
#pragma unroll
        for (int i = 0; i < NR; i++) {
          vn[i][k] = urow(Wp, reflf(yw + A.rad[i], H), ld)[jw];
          vo[i][k] = urow(Wp, reflf(yw - A.rad[i] - 1, H), ld)[jw];
        }
      }
      if (has_col) {
#pragma unroll
        for (int k = 0; k < KS; k++) {
#pragma unroll
          for (int i = 0; i < NR; i++) {
            if (NANAW) {
              if (!isnan(vn[i][k])) { S[i] += (double)vn[i][k]; Cn[i]++; }
              if (!isnan(vo[i][k])) { S[i] -= (double)vo[i][k]; Cn[i]--; }
              Ws[(i * K + kb + k) * LDC + t] = (float)((double)Cn[i] * A.inv[i]);
            } else {
              S[i] += (double)vn[i][k];
              S[i] -= (double)vo[i][k];
            }
            Vs[(i * K + kb + k) * LDC + t] = (float)(S[i] * A.inv[i]);
          }
          Zs[(kb + k) * LDC + t] = vz[k];
        }
      }
    }
    // ---------------- stage the bilinear sources and their row coefficients
    for (int s = 0; s < A.n_src; s++) {
      const SrcRef& SR = A.src[s];
      int a0, a1, b0, b1;
      double tt;
      src_axis(SR, 1, wyc, a0, a1, tt);
      src_axis(SR, 1, wyc + nrow - 1, b0, b1, tt);
      const int nr = min(b1 - a0 + 1, SR.rows_max);
      const int nc = s_ncol[s], clo = s_clo[s];
      float* tile = Src + SR.lds_off;
      for (int e = t; e < nr * nc; e += 256) {
        const int rr = e / nc, cc = e - rr * nc;
        tile[rr * SR.cols_max + cc] = SR.p[(int64_t)(a0 + rr) * SR.ld + clo + cc];
      }
      if (t < K) {
        int i0, i1;
        double ty;
        src_axis(SR, 1, wyc + min(t, nrow - 1), i0, i1, ty);
        RowI[s * K + t] = make_int2((i0 - a0) * SR.cols_max, (i1 - a0) * SR.cols_max);
        RowT[s * K + t] = ty;
      }
    }
    __syncthreads();

    // ---------------- axis-1 stage + combine
    {
      const int k = t / segs;
      const int sg = t % segs;
      if (k < nrow) {
        const int cbase = A.R + sg * L;  // LDS column of this segment's first output
        double Hs[NRA], HW[NRA];
#pragma unroll
        for (int i = 0; i < NR; i++) {
          const float* vrow = Vs + (i * K + k) * LDC;
          const float* wrow = Ws + (i * K + k) * LDC;
          double h = 0.0, hw = 0.0;
#pragma unroll 4
          for (int d = -A.rad[i]; d <= A.rad[i]; d++) {
            h += (double)vrow[cbase + d];
            if (NANAW) hw += (double)wrow[cbase + d];
          }
          Hs[i] = h;
          HW[i] = hw;
        }
        for (int m = 0; m < L; m++) {
          const int lc = cbase + m;
          float bm[NRA];
#pragma unroll
          for (int i = 0; i < NR; i++) {
            if (m > 0) {
              const float* vrow = Vs + (i * K + k) * LDC;
              const int r = A.rad[i];
              Hs[i] += (double)vrow[lc + r] - (double)vrow[lc - r - 1];
              if (NANAW) {
                const float* wrow = Ws + (i * K + k) * LDC;
                HW[i] += (double)wrow[lc + r] - (double)wrow[lc - r - 1];
              }
            }
            const float fv = (float)(Hs[i] * A.inv[i]);
            if (NANAW) {
              const float fw = (float)(HW[i] * A.inv[i]);
              bm[i] = fw > 0.0f ? fv / fw : 0.0f;
            } else {
              bm[i] = fv;
            }
          }
          const int cl = sg * L + m;      // strip column
          const float z = Zs[k * LDC + lc];
          float acc = 0.0f;
          // term table unrolled with compile-time indices: kind/slot/weight
          // come from the kernel arguments (scalar registers)
#pragma unroll
          for (int q = 0; q < FSG_MAX_RADII; q++) {
            if (q < A.n_terms) {
              float mean = 0.0f;
              const int kd = A.kind[q];
              if (kd == TERM_BOX) {
#pragma unroll
                for (int i = 0; i < NR; i++)
                  if (i == A.slot[q]) mean = bm[i];
              } else if (GAUSS && kd == TERM_GAUSS) {
                mean = gauss9_mean(A.W.p, A.W.ld, A.W.h, A.W.w, wyc + k, A.c0 + x0 + cl, A.G);
              } else {
                const int s = A.slot[q];
                const int2 ri = RowI[s * K + k];
                const double ty = RowT[s * K + k];
                const int2 ci = ColI[s * A.SW + cl];
                const double tx = ColT[s * A.SW + cl];
                const float* tile = Src + A.src[s].lds_off;
                const double a = tile[ri.x + ci.x], b = tile[ri.x + ci.y];
                const double c = tile[ri.y + ci.x], d = tile[ri.y + ci.y];
                const double top = a * (1.0 - tx) + b * tx;
                const double bot = c * (1.0 - tx) + d * tx;
                mean = (float)(top * (1.0 - ty) + bot * ty);
              }
              const float diff = A.wt[q] * (z - mean);
              acc = (q == 0) ? diff : acc + diff;
            }
          }
          if (isnan(z)) acc = __builtin_nanf("");
          if (A.field_src >= 0) {
            const int s = A.field_src;
            const int2 ri = RowI[s * K + k];
            const double ty = RowT[s * K + k];
            const int2 ci = ColI[s * A.SW + cl];
            const double tx = ColT[s * A.SW + cl];
            const float* tile = Src + A.src[s].lds_off;
            const double a = tile[ri.x + ci.x], b = tile[ri.x + ci.y];
            const double c = tile[ri.y + ci.x], d = tile[ri.y + ci.y];
            const double top = a * (1.0 - tx) + b * tx;
            const double bot = c * (1.0 - tx) + d * tx;
            const float up = (float)(top * (1.0 - ty) + bot * ty);
            const float large = A.wl * z - up;
            acc = (A.n_terms > 0) ? acc + large : large;
          }
          // the thread's own Zs cell is read above only by this thread
          Zs[k * LDC + lc] = acc;
        }
      }
    }
    __syncthreads();

    // ---------------- coalesced write: one region row per iteration
    {
      const PostK p1 = load_post(A.o1);
      const bool two = A.o2.data != nullptr;
      for (int k = 0; k < nrow; k++) {
        if (t < cols) {
          const float v = apply_post(p1, Zs[k * LDC + A.R + t]);
          const int yr = y0 + ch * K + k, xr = x0 + t;
          store_out(A.o1, yr, xr, v);
          if (two) store_out(A.o2, yr, xr, v);
        }
      }
    }
    __syncthreads();
  }
}

template <int NR, bool NANAW>
static int launch_usm(const UsmArgs& A, int n_blocks, hipStream_t s) {
  constexpr int K = (NR <= 3) ? 16 : 8;
  const int LDC = A.SW + 2 * A.R + 1;
  size_t lds = sizeof(float) * ((size_t)NR * K * LDC * (NANAW ? 2 : 1) + (size_t)K * LDC +
                                (size_t)A.src_floats);
  lds = (lds + 15) & ~(size_t)15;
  lds += (sizeof(int2) + sizeof(double)) * ((size_t)MAX_SRC * K + (size_t)A.n_src * A.SW);
  if (lds > 150 * 1024) return fail(FSG_E_UNSUPPORTED, "k_usm_main: LDS budget exceeded");
  bool gauss = false;
  for (int q = 0; q < A.n_terms; q++) gauss |= (A.kind[q] == TERM_GAUSS);
  TimedScope ts("k_usm_main", s);
  if (gauss)
    hipLaunchKernelGGL((k_usm_main<NR, NANAW, K, true>), dim3(n_blocks), dim3(256), lds, s, A);
  else
    hipLaunchKernelGGL((k_usm_main<NR, NANAW, K, false>), dim3(n_blocks), dim3(256), lds, s, A);
  FSG_CHECK_LAUNCH("k_usm_main");
  return FSG_OK;
}

static int dispatch_usm(const UsmArgs& A, bool nanaw, int n_blocks, hipStream_t s) {
#define FSG_USM_CASE(N)                                                   \
  case N:                                                                 \
    return nanaw ? launch_usm<N, true>(A, n_blocks, s) : launch_usm<N, false>(A, n_blocks, s);
  switch (A.nr) {
    FSG_USM_CASE(0)
    FSG_USM_CASE(1)
    FSG_USM_CASE(2)
    FSG_USM_CASE(3)
    FSG_USM_CASE(4)
    FSG_USM_CASE(5)
    FSG_USM_CASE(6)
    default:
      return fail(FSG_E_UNSUPPORTED, "fsg_topousm: more than 6 full-resolution box radii");
  }
#undef FSG_USM_CASE
}

// ------------------------------------------------------------ fast main kernel

// One launch per window: XCD-aware order, consecutive blocks of one band of
// the region land on one XCD (L2).
template <int NR, bool NANAW>
__global__ void __launch_bounds__(256) k_usm_fast(UsmArgs A, int n_tiles, int per_xcd) {
  extern __shared__ float lds[];
  const int b = blockIdx.x;
  const int tile = (b & 7) * per_xcd + (b >> 3);
  if (tile >= n_tiles) return;
  usm_fast_block<NR, NANAW>(A, tile, lds);
}

template <int NR, bool NANAW>
static int launch_fast(const UsmArgs& A, int n_tiles, hipStream_t s) {
  const size_t lds = fast_lds_bytes(A, NANAW);
  if (lds > 150 * 1024) return fail(FSG_E_UNSUPPORTED, "k_usm_fast: LDS budget exceeded");
  const int per_xcd = (n_tiles + 7) / 8;
  TimedScope ts("k_usm_main", s);
  hipLaunchKernelGGL((k_usm_fast<NR, NANAW>), dim3(per_xcd * 8), dim3(256), lds, s, A, n_tiles,
                     per_xcd);
  FSG_CHECK_LAUNCH("k_usm_fast");
  return FSG_OK;
}

// ------------------------------------------------- single-wavefront kernel

// 0: the wavefront kernel wherever it applies; 1: the workgroup kernels only
static std::atomic<int> g_usm_variant{0};

// One launch: the blocks whose rows and columns all lie inside the window
// (no reflection) take the FAST = true code, the others the reflecting loads.
template <int NR, int NS, int KR>
__device__ __forceinline__ void wave_tile(const WaveArgs& A, int tile, double* P, double* Hc) {
  const int x0 = (tile % A.n_strips) * A.SW;
  const int y0 = (tile / A.n_strips) * A.SH;
  const int rows = min(A.SH, A.rh - y0);
  const int cs = A.c0 + x0 - A.R;
  const int wy0 = A.r0 + y0;
  const bool fast = cs >= 0 && cs + WAVE_NC <= A.W && wy0 - A.R - 1 >= 0 &&
                    wy0 + rows - 1 + A.R <= A.H - 1;
  if (fast)
    usm_wave_block<NR, NS, true, KR>(A, tile, P, Hc);
  else
    usm_wave_block<NR, NS, false, KR>(A, tile, P, Hc);
}

// Block b takes tile (b & 7) * per_xcd + (b >> 3): consecutive tiles of a
// strip (vertical neighbours, sharing halo rows) land on one XCD and its L2.
template <int NR, int NS, int KR>
__global__ void __launch_bounds__(64) k_usm_wave(WaveArgs A) {
  __shared__ double P[4 * WAVE_PB];
  __shared__ double Hc[(NS > 0 ? NS : 1) * 2 * WAVE_NC];
  const int b = blockIdx.x;
  const int tile = (b & 7) * A.per_xcd + (b >> 3);
  if (tile >= A.n_blocks) return;
  wave_tile<NR, NS, KR>(A, tile, P, Hc);
}

// Many windows' wavefront main passes in one launch (fsg_batch): block b
// finds its window by binary search over the block prefix (each window's
// count a multiple of 8, so b & 7 is also the window-local XCD slot), then
// runs k_usm_wave's block on that window's arguments, read through the
// constant address space.
template <int NR, int NS, int KR>
__global__ void __launch_bounds__(64) k_usm_wave_batch(const WaveArgs* args, const int* pre,
                                                       int n) {
  __shared__ double P[4 * WAVE_PB];
  __shared__ double Hc[(NS > 0 ? NS : 1) * 2 * WAVE_NC];
  const int b = blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (desc(pre, mid) <= b) lo = mid; else hi = mid - 1;
  }
  const WaveArgs& A = desc(args, lo);
  const int local = b - desc(pre, lo);
  const int tile = (local & 7) * A.per_xcd + (local >> 3);
  if (tile >= A.n_blocks) return;
  wave_tile<NR, NS, KR>(A, tile, P, Hc);
}

// Bilinear row coefficients of every source at every region row (the same
// src_axis arithmetic as the workgroup kernel's staging).
struct WaveTabArgs {
  SrcRef src[WAVE_MAX_SRC];
  int n_src, r0, rh;
  int4* rowT;
};

__global__ void __launch_bounds__(256) k_wave_tables(WaveTabArgs T) {
  for (int e = blockIdx.x * 256 + threadIdx.x; e < T.n_src * T.rh; e += gridDim.x * 256) {
    const int s = e / T.rh, k = e - s * T.rh;
    int a, b;
    double w;
    src_axis(T.src[s], 1, T.r0 + k, a, b, w);
    const unsigned long long u = __builtin_bit_cast(unsigned long long, w);
    T.rowT[e] = make_int4(a, b, (int)(unsigned)(u & 0xffffffffull), (int)(u >> 32));
  }
}

// Rows per wavefront block: the fewest that still give every strip's blocks
// a slot in the launch's first (and only) round of resident blocks -- the
// blocks of one launch run concurrently and each one's time grows with its
// rows (plus a 2R + 1 row start-up), so a second, partial round would cost
// a whole block time.
static int wave_rows(int rh, int n_strips, int capacity) {
  const int v = g_usm_variant.load(std::memory_order_relaxed);
  if (v >= 8) return v;
  const int segs = std::max(1, capacity / std::max(1, n_strips));
  return std::max(16, (rh + segs - 1) / segs);
}

// Resident blocks of one k_usm_wave instance on the current device (CUs x
// blocks per CU under its registers and LDS).
template <int NR, int NS, int KR>
static int wave_capacity() {
  static const int cap = [] {
    int dev = 0, cus = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_usm_wave<NR, NS, KR>, 64, 0) !=
            hipSuccess || nb <= 0) {
      (void)hipGetLastError();
      nb = 8;
    }
    return cus * nb;
  }();
  return cap;
}

template <int NR, int NS, int KR>
static int launch_wave(WaveArgs A, hipStream_t s) {
  A.SH = wave_rows(A.rh, A.n_strips, wave_capacity<NR, NS, KR>());
  A.n_blocks = A.n_strips * ((A.rh + A.SH - 1) / A.SH);
  A.per_xcd = (A.n_blocks + 7) / 8;
  TimedScope ts("k_usm_main", s);
  hipLaunchKernelGGL((k_usm_wave<NR, NS, KR>), dim3(A.per_xcd * 8), dim3(64), 0, s, A);
  FSG_CHECK_LAUNCH("k_usm_wave");
  return FSG_OK;
}

// The wavefront kernel handles NaN-free windows in the fast term order with
// 1..3 box radii up to 32 and at most 3 bilinear sources.
static bool wave_eligible(const UsmArgs& A, bool nanaw) {
  return g_usm_variant.load(std::memory_order_relaxed) != 1 && !nanaw && A.nr >= 1 &&
         A.nr <= WAVE_MAX_NR && A.R <= WAVE_MAX_R && A.n_src <= WAVE_MAX_SRC;
}

// The wavefront kernel's arguments for a fast-path call (coefficient tables
// at tab, the workspace's table area) and launches the tables' kernel.
static int wave_args(const UsmArgs& U, int4* tab, WaveArgs& A, hipStream_t s) {
  A = WaveArgs{};
  A.win = U.W.p;
  A.ld = U.W.ld;
  A.H = U.W.h;
  A.W = U.W.w;
  A.r0 = U.r0;
  A.c0 = U.c0;
  A.rh = U.rh;
  A.rw = U.rw;
  A.R = U.R;
  for (int i = 0; i < U.nr; i++) {
    A.rad[i] = U.rad[i];
    A.inv[i] = U.inv[i];
    A.wt[i] = U.wt[i];
  }
  A.nz = U.n_src - (U.field_src >= 0 ? 1 : 0);
  A.has_field = U.field_src >= 0 ? 1 : 0;
  for (int j = 0; j < A.nz; j++) A.wt[U.nr + j] = U.wt[U.nr + j];
  A.wl = U.wl;
  WaveTabArgs T{};
  for (int j = 0; j < U.n_src; j++) {
    A.sp[j] = U.src[j].p;
    A.sld[j] = U.src[j].ld;
    A.src[j] = U.src[j];
    T.src[j] = U.src[j];
  }
  A.rowT = tab;
  A.o1 = U.o1;
  A.o2 = U.o2;
  A.SW = WAVE_NC - 2 * U.R;
  A.n_strips = (U.rw + A.SW - 1) / A.SW;
  if (U.n_src > 0) {
    T.n_src = U.n_src;
    T.r0 = U.r0;
    T.rh = U.rh;
    T.rowT = tab;
    const int n = U.n_src * U.rh;
    hipLaunchKernelGGL(k_wave_tables, dim3(std::min((n + 255) / 256, 1024)), dim3(256), 0, s, T);
    FSG_CHECK_LAUNCH("k_wave_tables");
  }
  return FSG_OK;
}

// Builds the coefficient tables and launches the wavefront kernel for the
// arguments of a fast-path call.
static int dispatch_wave(const UsmArgs& U, int4* tab, hipStream_t s) {
  WaveArgs A;
  const int rc = wave_args(U, tab, A, s);
  if (rc) return rc;
  if (U.nr == 3 && U.rad[0] == 2 && U.rad[1] == 8 && U.rad[2] == 32) {
    if (U.n_src == 0) return launch_wave<3, 0, 1>(A, s);
    if (U.n_src == 1) return launch_wave<3, 1, 1>(A, s);
    if (U.n_src == 2) return launch_wave<3, 2, 1>(A, s);
    if (U.n_src == 3) return launch_wave<3, 3, 1>(A, s);
  }
#define FSG_WAVE_CASE(NRV, NSV) \
  if (U.nr == NRV && U.n_src == NSV) return launch_wave<NRV, NSV, 0>(A, s);
  FSG_WAVE_CASE(1, 0) FSG_WAVE_CASE(1, 1) FSG_WAVE_CASE(1, 2) FSG_WAVE_CASE(1, 3)
  FSG_WAVE_CASE(2, 0) FSG_WAVE_CASE(2, 1) FSG_WAVE_CASE(2, 2) FSG_WAVE_CASE(2, 3)
  FSG_WAVE_CASE(3, 0) FSG_WAVE_CASE(3, 1) FSG_WAVE_CASE(3, 2) FSG_WAVE_CASE(3, 3)
#undef FSG_WAVE_CASE
  return fail(FSG_E_UNSUPPORTED, "k_usm_wave: unsupported term counts");
}

static int dispatch_fast(const UsmArgs& A, bool nanaw, int n_tiles, hipStream_t s) {
#define FSG_FAST_CASE(N)                                                      \
  case N:                                                                     \
    return nanaw ? launch_fast<N, true>(A, n_tiles, s) : launch_fast<N, false>(A, n_tiles, s);
  switch (A.nr) {
    FSG_FAST_CASE(0)
    FSG_FAST_CASE(1)
    FSG_FAST_CASE(2)
    FSG_FAST_CASE(3)
    FSG_FAST_CASE(4)
    FSG_FAST_CASE(5)
    FSG_FAST_CASE(6)
    default:
      return fail(FSG_E_UNSUPPORTED, "fsg_topousm: more than 6 full-resolution box radii");
  }
#undef FSG_FAST_CASE
}

// Exact staging-tile bounds: rows spanned by any K-row chunk of the region and
// columns spanned by any SW-column strip (same arithmetic as the device).
static void src_bounds(SrcRef& S, int r0, int rh, int c0, int rw, int K, int SW, int SH) {
  int rmax = 1, cmax = 1, a0, a1, b0, b1;
  double tt;
  for (int y0 = 0; y0 < rh; y0 += SH) {
    const int rows = std::min(SH, rh - y0);
    for (int ch = 0; ch * K < rows; ch++) {
      const int first = r0 + y0 + ch * K, last = first + std::min(K, rows - ch * K) - 1;
      src_axis(S, 1, first, a0, a1, tt);
      src_axis(S, 1, last, b0, b1, tt);
      rmax = std::max(rmax, b1 - a0 + 1);
    }
  }
  for (int x0 = 0; x0 < rw; x0 += SW) {
    const int cols = std::min(SW, rw - x0);
    src_axis(S, 0, c0 + x0, a0, a1, tt);
    src_axis(S, 0, c0 + x0 + cols - 1, b0, b1, tt);
    cmax = std::max(cmax, b1 - a0 + 1);
  }
  S.rows_max = rmax;
  S.cols_max = cmax;
}

// ------------------------------------------------------------ host helpers

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct CoarsePlan {
  int n_ds = 0;
  int ds_val[FSG_MAX_RADII];
  int ds_of_radius[FSG_MAX_RADII];  // index into ds_val, or -1
  int hc[FSG_MAX_RADII], wc[FSG_MAX_RADII];
  size_t tab_off = 0;
  size_t grid_off[FSG_MAX_RADII];   // per distinct ds: decimated grid
  size_t mean_off[FSG_MAX_RADII];   // per radius: coarse mean grid
  size_t flag_off;
  size_t tmp_off[4];
  size_t total = 0;
  int max_cells = 0;
  int max_vrad = 0;
};

static int void_fill_radius(int hc, int wc, double* sigma_out) {
  const double sigma = std::max(1.0, (double)std::min(hc, wc) / 64.0);
  *sigma_out = sigma;
  return gaussian_radius(sigma);
}

static CoarsePlan plan_coarse(const fsg_window* win, const fsg_topousm_params* p) {
  CoarsePlan P;
  size_t off = 0;
  for (int i = 0; i < p->n_radii; i++) {
    P.ds_of_radius[i] = -1;
    const int ds = p->ds[i];
    if (ds <= 1 || p->mean[i]) continue;  // full resolution, or supplied by the caller
    int idx = -1;
    for (int j = 0; j < P.n_ds; j++)
      if (P.ds_val[j] == ds) idx = j;
    if (idx < 0) {
      idx = P.n_ds++;
      P.ds_val[idx] = ds;
      P.hc[idx] = std::max(1, (win->h + ds - 1) / ds);
      P.wc[idx] = std::max(1, (win->w + ds - 1) / ds);
      P.grid_off[idx] = off;
      off += align256(sizeof(float) * (size_t)P.hc[idx] * P.wc[idx]);
      P.max_cells = std::max(P.max_cells, P.hc[idx] * P.wc[idx]);
      double sg;
      P.max_vrad = std::max(P.max_vrad, void_fill_radius(P.hc[idx], P.wc[idx], &sg));
    }
    P.ds_of_radius[i] = idx;
    P.mean_off[i] = off;
    off += align256(sizeof(float) * (size_t)P.hc[idx] * P.wc[idx]);
  }
  P.flag_off = off;
  off += align256(sizeof(int) * FSG_MAX_RADII);
  for (int k = 0; k < 4; k++) {
    P.tmp_off[k] = off;
    off += align256(sizeof(float) * (size_t)P.max_cells);
  }
  // bilinear coefficient tables of the wavefront kernel (per source, one
  // entry per window row and per window column at most)
  P.tab_off = off;
  off += align256(sizeof(int4) * (size_t)MAX_SRC * ((size_t)win->h + (size_t)win->w));
  P.total = off;
  return P;
}

// Box (r >= 2) or sigma=1 Gaussian (r <= 1) NaN-aware mean of an h x w grid.
static int coarse_mean(const float* x, int64_t ldx, int h, int w, int r, float* out, int64_t ldo,
                       float* tv, float* tw, hipStream_t s) {
  if (r <= 1) {
    dim3 blk(128), grd((w + 127) / 128, std::min(h, 32768));
    hipLaunchKernelGGL(k_gauss9_grid, grd, blk, 0, s, x, ldx, h, w, out, ldo, make_gauss9());
    FSG_CHECK_LAUNCH("k_gauss9_grid");
    return FSG_OK;
  }
  const double inv = 1.0 / (double)(2 * r + 1);
  // pass 0: (h x w) -> transposed (w x h) axis-0 means
  hipLaunchKernelGGL(k_box_t<0>, dim3((w + 63) / 64, (h + 63) / 64), dim3(64), 0, s, x,
                     nullptr, ldx, h, w, r, inv, tv, tw, (int64_t)h);
  FSG_CHECK_LAUNCH("k_box_t");
  // pass 1: axis 0 of the transposed grid (= axis 1), transposed back
  hipLaunchKernelGGL(k_box_t<1>, dim3((h + 63) / 64, (w + 63) / 64), dim3(64), 0, s, tv, tw,
                     (int64_t)h, w, h, r, inv, out, nullptr, ldo);
  FSG_CHECK_LAUNCH("k_box_t");
  return FSG_OK;
}

// Decimated grid with the enclosed-void fill (_nan_utils.py:604-668).
static int decimate(const fsg_window* win, int ds, float* grid, int hc, int wc, int* flag,
                    float* t0, float* t1, float* t2, float* t3, bool nan_free, hipStream_t s) {
  if (ds > 16 || (ds & (ds - 1)) != 0)
    return fail(FSG_E_UNSUPPORTED, "fsg_topousm: downsample factor must be a power of two <= 16");
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, flag, 1);
  FSG_CHECK_LAUNCH("k_zero_int");
  WinDev W{win->data, win->ld, win->h, win->w};
  if (ds >= 4) {
    dim3 blk(256), grd((win->w + 1023) / 1024, std::min(hc, 32768));
    hipLaunchKernelGGL(k_decimate4, grd, blk, 0, s, W, ds, grid, hc, wc, flag);
    FSG_CHECK_LAUNCH("k_decimate4");
  } else {
    dim3 blk(256), grd((win->w + 255) / 256, std::min(hc, 32768));
    hipLaunchKernelGGL(k_decimate, grd, blk, 0, s, W, ds, grid, hc, wc, flag);
    FSG_CHECK_LAUNCH("k_decimate");
  }
  // a NaN-free window has no all-NaN cell: the enclosed-void fill is a no-op
  if (nan_free) return FSG_OK;
  double sigma;
  const int rad = void_fill_radius(hc, wc, &sigma);
  VFWeights V{};
  V.rad = rad;
  V.sigma2 = sigma * sigma;
  {
    std::vector<double> full(2 * rad + 1);
    gaussian_weights(sigma, rad, full.data());
    if (rad <= VF_MAX_RAD) {
      V.by_value = 1;
      for (int d = 0; d <= rad; d++) V.g[d] = full[rad + d];
    } else {
      std::vector<double> phi(2 * rad + 1);
      for (int i = 0; i <= 2 * rad; i++) {
        const double x = (double)(i - rad);
        phi[i] = std::exp(-0.5 / V.sigma2 * (x * x));
      }
      V.by_value = 0;
      V.inv_sum = 1.0 / np_pairwise_sum(phi.data(), 2 * rad + 1);
    }
  }
  const int n = hc * wc;
  dim3 blk(256), grd(std::min(1024, (n + 255) / 256));
  hipLaunchKernelGGL(k_voidfill_pass, grd, blk, 0, s, grid, grid, t0, t1, hc, wc, V, 0, flag);
  FSG_CHECK_LAUNCH("k_voidfill_pass");
  hipLaunchKernelGGL(k_voidfill_pass, grd, blk, 0, s, t0, t1, t2, t3, hc, wc, V, 1, flag);
  FSG_CHECK_LAUNCH("k_voidfill_pass");
  hipLaunchKernelGGL(k_voidfill_apply, grd, blk, 0, s, grid, t2, t3, n, flag);
  FSG_CHECK_LAUNCH("k_voidfill_apply");
  return FSG_OK;
}

}  // namespace fsg

using namespace fsg;

extern "C" size_t fsg_topousm_workspace_size(const fsg_window* win,
                                             const fsg_topousm_params* p) {
  if (!win || !p || p->n_radii < 0 || p->n_radii > FSG_MAX_RADII) return 0;
  return plan_coarse(win, p).total;
}

int fsg::usm_prepare(const fsg_window* win, const fsg_region* reg, const fsg_topousm_params* p,
                     const fsg_output* out, const fsg_output* out2, void* workspace,
                     size_t workspace_bytes, hipStream_t s, UsmArgs* A_out, bool* fast_out,
                     int* n_blocks_out, bool launch_prep) {
  *n_blocks_out = 0;
  *fast_out = false;
  int rc = validate_window(win, "fsg_topousm");
  if (rc) return rc;
  rc = validate_region(win, reg, "fsg_topousm");
  if (rc) return rc;
  if (!p || !out) return fail(FSG_E_INVALID, "fsg_topousm: params/output NULL");
  if ((rc = validate_output(out, reg, "fsg_topousm"))) return rc;
  if ((rc = validate_output(out2, reg, "fsg_topousm"))) return rc;
  if (p->n_radii < 0 || p->n_radii > FSG_MAX_RADII)
    return fail(FSG_E_INVALID, "fsg_topousm: n_radii out of range");
  if (p->n_radii == 0 && !p->field)
    return fail(FSG_E_INVALID, "fsg_topousm: no radii and no large field");
  if (win->h == 0 || win->w == 0) return fail(FSG_E_INVALID, "fsg_topousm: empty window");
  if (reg->h == 0 || reg->w == 0) return FSG_OK;

  const CoarsePlan P = plan_coarse(win, p);
  if (P.total > 0 && (!workspace || workspace_bytes < P.total))
    return fail(FSG_E_WORKSPACE, "fsg_topousm: workspace too small");
  char* ws = (char*)workspace;

  UsmArgs A{};
  A.W = WinDev{win->data, win->ld, win->h, win->w};
  A.r0 = reg->r0;
  A.c0 = reg->c0;
  A.rh = reg->h;
  A.rw = reg->w;
  A.n_terms = p->n_radii;
  A.nr = 0;
  A.R = 0;
  for (int i = 0; i < p->n_radii; i++) {
    const int r = p->radii[i];
    A.wt[i] = p->weights[i];
    if (p->ds[i] > 1) {
      A.kind[i] = TERM_COARSE;
      A.slot[i] = i;
    } else if (r <= 1) {
      A.kind[i] = TERM_GAUSS;
      A.slot[i] = 0;
    } else {
      int slot = -1;
      for (int j = 0; j < A.nr; j++)
        if (A.rad[j] == r) slot = j;
      if (slot < 0) {
        if (A.nr >= MAX_BOX)
          return fail(FSG_E_UNSUPPORTED, "fsg_topousm: more than 6 full-resolution box radii");
        slot = A.nr++;
        A.rad[slot] = r;
        A.inv[slot] = 1.0 / (double)(2 * r + 1);
        A.R = std::max(A.R, r);
      }
      A.kind[i] = TERM_BOX;
      A.slot[i] = slot;
    }
  }
  if (A.R > 96)
    return fail(FSG_E_UNSUPPORTED,
                "fsg_topousm: full-resolution box radius > 96 (the downsample rule caps it "
                "near 41)");

  // coarse grids for ds > 1 radii
  int* flags = (int*)(ws + P.flag_off);
  float* tmp[4];
  for (int k = 0; k < 4; k++) tmp[k] = (float*)(ws + P.tmp_off[k]);
  for (int j = 0; j < P.n_ds && launch_prep; j++) {
    rc = decimate(win, P.ds_val[j], (float*)(ws + P.grid_off[j]), P.hc[j], P.wc[j], flags + j,
                  tmp[0], tmp[1], tmp[2], tmp[3], p->nan_hint == 0, s);
    if (rc) return rc;
  }
  // Fast path when the terms run [distinct box radii in slot order][decimated
  // radii]: the order of every ascending radius list.  Other orders (a sigma=1
  // Gaussian term, a repeated radius, a decimated radius before a box radius)
  // use the general kernel.  Both accumulate in term order.
  // (single reflection of window rows: the last chunk's rows reach K - 1
  // past the region, so the window must hold at least R + K rows)
  bool fast = A.R + FAST_K <= win->h;
  for (int q = 0; q < A.n_terms; q++)
    fast = fast && (q < A.nr ? (A.kind[q] == TERM_BOX && A.slot[q] == q)
                             : A.kind[q] == TERM_COARSE);
  // strip geometry (needed for the staging bounds)
  const int K = fast ? FAST_K : (A.nr <= 3) ? 16 : 8;
  int SW = ((256 - 2 * A.R) / 32) * 32;
  if (SW < 32) return fail(FSG_E_UNSUPPORTED, "fsg_topousm: box radius too large");
  A.SW = SW;
  A.SH = fast ? FAST_SH : 256;
  A.n_strips = (reg->w + SW - 1) / SW;
  A.n_src = 0;
  A.field_src = -1;
  int src_floats = 0;
  for (int i = 0; i < p->n_radii; i++) {
    const int ds = p->ds[i];
    if (ds <= 1) continue;
    const int hc = std::max(1, (win->h + ds - 1) / ds), wc = std::max(1, (win->w + ds - 1) / ds);
    const float* mean = p->mean[i];
    int64_t mean_ld = p->mean_ld[i];
    if (!mean) {
      const int j = P.ds_of_radius[i];
      const int r_small = std::max(1, (int)std::nearbyint((double)p->radii[i] / (double)ds));
      float* m = (float*)(ws + P.mean_off[i]);
      if (launch_prep) {
        rc = coarse_mean((float*)(ws + P.grid_off[j]), P.wc[j], P.hc[j], P.wc[j], r_small, m,
                         P.wc[j], tmp[0], tmp[1], s);
        if (rc) return rc;
      }
      mean = m;
      mean_ld = P.wc[j];
    } else if (mean_ld < wc) {
      return fail(FSG_E_INVALID, "fsg_topousm: mean_ld smaller than the coarse grid width");
    }
    // scipy zoom output shape round(hc * (H / hc)) must be the window shape
    const double zoom_y = (double)win->h / (double)hc;
    const double zoom_x = (double)win->w / (double)wc;
    if ((int)std::nearbyint(hc * zoom_y) != win->h || (int)std::nearbyint(wc * zoom_x) != win->w)
      return fail(FSG_E_UNSUPPORTED, "fsg_topousm: zoom output shape differs from window");
    if (A.n_src >= MAX_SRC)
      return fail(FSG_E_UNSUPPORTED, "fsg_topousm: more than 5 decimated radii");
    SrcRef& S = A.src[A.n_src];
    S = SrcRef{};
    S.p = mean;
    S.ld = mean_ld;
    S.h = hc;
    S.w = wc;
    S.mode = 0;
    S.zy = (win->h != 1) ? (double)(hc - 1) / (double)(win->h - 1) : 1.0;
    S.zx = (win->w != 1) ? (double)(wc - 1) / (double)(win->w - 1) : 1.0;
    src_bounds(S, reg->r0, reg->h, reg->c0, reg->w, K, A.SW, A.SH);
    S.lds_off = src_floats;
    src_floats += S.rows_max * (fast ? A.SW : S.cols_max);
    A.slot[i] = A.n_src++;
  }
  if (p->field) {
    if (p->field_h <= 0 || p->field_w <= 0 || p->full_h <= 0 || p->full_w <= 0)
      return fail(FSG_E_INVALID, "fsg_topousm: bad field geometry");
    if (A.n_src >= MAX_SRC) return fail(FSG_E_UNSUPPORTED, "fsg_topousm: too many sources");
    SrcRef& S = A.src[A.n_src];
    S = SrcRef{};
    S.p = p->field;
    S.ld = p->field_w;
    S.h = p->field_h;
    S.w = p->field_w;
    S.mode = 1;
    S.oy = p->win_r0;
    S.ox = p->win_c0;
    S.fy = (float)((double)p->field_h / (double)p->full_h);
    S.fx = (float)((double)p->field_w / (double)p->full_w);
    src_bounds(S, reg->r0, reg->h, reg->c0, reg->w, K, A.SW, A.SH);
    S.lds_off = src_floats;
    src_floats += S.rows_max * (fast ? A.SW : S.cols_max);
    A.field_src = A.n_src++;
    A.wl = p->w_large;
  }
  A.src_floats = src_floats;
  A.G = make_gauss9();
  A.o1 = make_out(out);
  A.o2 = out2 ? make_out(out2) : OutDev{};
  const int n_segs = (reg->h + A.SH - 1) / A.SH;
  *A_out = A;
  *fast_out = fast;
  *n_blocks_out = A.n_strips * n_segs;
  return FSG_OK;
}

int fsg::usm_prep_jobs(const fsg_window* win, const fsg_topousm_params* p, void* workspace,
                       std::vector<DecJob>& dec, std::vector<std::vector<BoxJob>>& box0,
                       std::vector<std::vector<BoxJob>>& box1) {
  const CoarsePlan P = plan_coarse(win, p);
  char* ws = (char*)workspace;
  // same order and buffers as usm_prepare's launch_prep branches
  for (int j = 0; j < P.n_ds; j++)
    if (P.ds_val[j] > 16 || (P.ds_val[j] & (P.ds_val[j] - 1)) != 0) return FSG_E_UNSUPPORTED;
  for (int i = 0; i < p->n_radii; i++) {
    if (p->ds[i] <= 1 || p->mean[i]) continue;
    const int r_small = std::max(1, (int)std::nearbyint((double)p->radii[i] / (double)p->ds[i]));
    if (r_small <= 1) return FSG_E_UNSUPPORTED;
  }
  const WinDev W{win->data, win->ld, win->h, win->w};
  for (int j = 0; j < P.n_ds; j++) {
    const int cols = P.ds_val[j] >= 4 ? 1024 : 256;  // per block (decimate4_block / decimate_block)
    dec.push_back(DecJob{W, P.ds_val[j], (float*)(ws + P.grid_off[j]), P.hc[j], P.wc[j],
                         (win->w + cols - 1) / cols});
  }
  float* tv = (float*)(ws + P.tmp_off[0]);
  float* tw = (float*)(ws + P.tmp_off[1]);
  int level = 0;
  for (int i = 0; i < p->n_radii; i++) {
    if (p->ds[i] <= 1 || p->mean[i]) continue;
    const int j = P.ds_of_radius[i];
    const int r = std::max(1, (int)std::nearbyint((double)p->radii[i] / (double)p->ds[i]));
    const int h = P.hc[j], w = P.wc[j];
    const double inv = 1.0 / (double)(2 * r + 1);
    const float* grid = (const float*)(ws + P.grid_off[j]);
    float* m = (float*)(ws + P.mean_off[i]);
    if ((int)box0.size() <= level) {
      box0.emplace_back();
      box1.emplace_back();
    }
    box0[level].push_back(BoxJob{grid, nullptr, (int64_t)w, h, w, r, inv, tv, tw, (int64_t)h,
                                 (w + 63) / 64});
    box1[level].push_back(BoxJob{tv, tw, (int64_t)h, w, h, r, inv, m, nullptr, (int64_t)w,
                                 (h + 63) / 64});
    level++;
  }
  return FSG_OK;
}

// Resident blocks of one k_usm_wave_batch instance on the current device.
template <int NS>
static int wave_batch_capacity() {
  static const int cap = [] {
    int dev = 0, cus = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_usm_wave_batch<3, NS, 1>, 64, 0) !=
            hipSuccess || nb <= 0) {
      (void)hipGetLastError();
      nb = 8;
    }
    return cus * nb;
  }();
  return cap;
}

int fsg::wave_batch_create(const std::vector<WaveTileIn>& in, std::vector<char>& taken,
                           WaveBatch* wb) {
  *wb = WaveBatch{};
  taken.assign(in.size(), 0);
  if (g_usm_variant.load(std::memory_order_relaxed) == 1) return FSG_OK;
  // the instances batched here: the default radii, 0..3 bilinear sources
  auto inst_of = [](const UsmArgs& U) {
    if (U.nr != 3 || U.rad[0] != 2 || U.rad[1] != 8 || U.rad[2] != 32) return -1;
    if (U.R > WAVE_MAX_R || U.n_src > WAVE_MAX_SRC || U.rh <= 0 || U.rw <= 0) return -1;
    return U.n_src;
  };
  int count[WAVE_MAX_SRC + 1] = {0};
  for (const WaveTileIn& t : in) {
    const int k = inst_of(*t.args);
    if (k >= 0) count[k]++;
  }
  int inst = -1;
  for (int k = 0; k <= WAVE_MAX_SRC; k++)
    if (count[k] > 0 && (inst < 0 || count[k] > count[inst])) inst = k;
  if (inst < 0) return FSG_OK;
  std::vector<WaveArgs> A;
  std::vector<size_t> idx;
  int64_t strip_rows = 0;
  for (size_t i = 0; i < in.size(); i++) {
    if (inst_of(*in[i].args) != inst) continue;
    WaveArgs a;
    int4* tab = (int4*)((char*)in[i].workspace + plan_coarse(in[i].win, in[i].p).tab_off);
    if (int rc = wave_args(*in[i].args, tab, a, nullptr)) return rc;
    strip_rows += (int64_t)a.n_strips * a.rh;
    A.push_back(a);
    idx.push_back(i);
  }
  // rows per block: the batch's strip rows over 32 rounds of resident blocks
  // (at most a window's rows, at least 16), the windows with the longest
  // blocks first, so that the short ones fill in behind them.  Each block
  // pays a 2R + 1 row start-up, yet short blocks balance better: on 256
  // tiles of sides 256-4096 (about 44 rows per block) 32 rounds ran 11.9 ms,
  // 16 rounds 12.2, 4 rounds 14.9, 2 rounds 16.6 and 64 rounds 12.3.
  int cap = 0;
  switch (inst) {
    case 0: cap = wave_batch_capacity<0>(); break;
    case 1: cap = wave_batch_capacity<1>(); break;
    case 2: cap = wave_batch_capacity<2>(); break;
    default: cap = wave_batch_capacity<3>(); break;
  }
  const int64_t slots = 32 * (int64_t)std::max(1, cap);
  const int64_t sh_all = std::max<int64_t>(16, (strip_rows + slots - 1) / slots);
  for (WaveArgs& a : A) a.SH = (int)std::max<int64_t>(16, std::min<int64_t>(sh_all, a.rh));
  std::stable_sort(A.begin(), A.end(),
                   [](const WaveArgs& x, const WaveArgs& y) { return x.SH > y.SH; });
  std::vector<int> pre{0};
  for (WaveArgs& a : A) {
    const int64_t nb = (int64_t)a.n_strips * ((a.rh + a.SH - 1) / a.SH);
    a.n_blocks = (int)nb;
    a.per_xcd = (int)((nb + 7) / 8);
    const int64_t next = (int64_t)pre.back() + 8 * (int64_t)a.per_xcd;
    if (next > INT32_MAX) return fail(FSG_E_UNSUPPORTED, "wave_batch_create: too many blocks");
    pre.push_back((int)next);
  }
  void* d_args = nullptr;
  int* d_pre = nullptr;
  int rc = check_hip(hipMalloc(&d_args, sizeof(WaveArgs) * A.size()), "wave_batch_create");
  if (!rc) rc = check_hip(hipMalloc((void**)&d_pre, sizeof(int) * pre.size()), "wave_batch_create");
  if (!rc)
    rc = check_hip(hipMemcpy(d_args, A.data(), sizeof(WaveArgs) * A.size(), hipMemcpyHostToDevice),
                   "wave_batch_create");
  if (!rc)
    rc = check_hip(hipMemcpy(d_pre, pre.data(), sizeof(int) * pre.size(), hipMemcpyHostToDevice),
                   "wave_batch_create");
  // the coefficient tables were queued on the null stream
  if (!rc) rc = check_hip(hipStreamSynchronize(nullptr), "wave_batch_create");
  if (rc) {
    if (d_args) hipFree(d_args);
    if (d_pre) hipFree(d_pre);
    return rc;
  }
  wb->inst = inst;
  wb->n_tiles = (int)A.size();
  wb->n_blocks = pre.back();
  wb->d_args = d_args;
  wb->d_pre = d_pre;
  for (size_t i : idx) taken[i] = 1;
  return FSG_OK;
}

int fsg::wave_batch_launch(const WaveBatch& wb, hipStream_t s) {
  if (wb.n_tiles <= 0 || wb.n_blocks <= 0) return FSG_OK;
  const WaveArgs* a = (const WaveArgs*)wb.d_args;
  TimedScope ts("k_usm_wave_batch", s);
  switch (wb.inst) {
    case 0: hipLaunchKernelGGL((k_usm_wave_batch<3, 0, 1>), dim3(wb.n_blocks), dim3(64), 0, s, a, wb.d_pre, wb.n_tiles); break;
    case 1: hipLaunchKernelGGL((k_usm_wave_batch<3, 1, 1>), dim3(wb.n_blocks), dim3(64), 0, s, a, wb.d_pre, wb.n_tiles); break;
    case 2: hipLaunchKernelGGL((k_usm_wave_batch<3, 2, 1>), dim3(wb.n_blocks), dim3(64), 0, s, a, wb.d_pre, wb.n_tiles); break;
    case 3: hipLaunchKernelGGL((k_usm_wave_batch<3, 3, 1>), dim3(wb.n_blocks), dim3(64), 0, s, a, wb.d_pre, wb.n_tiles); break;
    default: return fail(FSG_E_INVALID, "wave_batch_launch: no instance");
  }
  FSG_CHECK_LAUNCH("k_usm_wave_batch");
  return FSG_OK;
}

void fsg::wave_batch_destroy(WaveBatch* wb) {
  if (wb->d_args) hipFree(wb->d_args);
  if (wb->d_pre) hipFree(wb->d_pre);
  *wb = WaveBatch{};
}

int fsg::launch_dec_batch(const DecJob* jobs, const int* pre, int n_jobs, int n_blocks,
                          hipStream_t s) {
  if (n_jobs <= 0 || n_blocks <= 0) return FSG_OK;
  hipLaunchKernelGGL(k_decimate_batch, dim3(n_blocks), dim3(256), 0, s, jobs, pre, n_jobs);
  FSG_CHECK_LAUNCH("k_decimate_batch");
  return FSG_OK;
}

int fsg::launch_box_batch(int pass, const BoxJob* jobs, const int* pre, int n_jobs, int n_blocks,
                          hipStream_t s) {
  if (n_jobs <= 0 || n_blocks <= 0) return FSG_OK;
  if (pass == 0)
    hipLaunchKernelGGL(k_box_t_batch<0>, dim3(n_blocks), dim3(64), 0, s, jobs, pre, n_jobs);
  else
    hipLaunchKernelGGL(k_box_t_batch<1>, dim3(n_blocks), dim3(64), 0, s, jobs, pre, n_jobs);
  FSG_CHECK_LAUNCH("k_box_t_batch");
  return FSG_OK;
}

extern "C" size_t fsg_downsample_workspace_size(int32_t h, int32_t w, int32_t factor) {
  if (h <= 0 || w <= 0 || factor < 1) return 0;
  const size_t cells = (size_t)((h + factor - 1) / factor) * (size_t)((w + factor - 1) / factor);
  return 4 * align256(sizeof(float) * cells) + align256(sizeof(int) * 4);
}

extern "C" int fsg_downsample_nan_aware(const fsg_window* src, int32_t factor, float* out,
                                        int64_t ldo, void* workspace, size_t workspace_bytes,
                                        void* stream) {
  int rc = validate_window(src, "fsg_downsample_nan_aware");
  if (rc) return rc;
  if (src->h <= 0 || src->w <= 0 || factor < 2 || !out)
    return fail(FSG_E_INVALID, "fsg_downsample_nan_aware: bad arguments");
  const int hc = (src->h + factor - 1) / factor, wc = (src->w + factor - 1) / factor;
  if (ldo != wc)
    return fail(FSG_E_INVALID, "fsg_downsample_nan_aware: out must be a dense ceil(h/f) x "
                               "ceil(w/f) grid (ldo = ceil(w/f))");
  const size_t need = fsg_downsample_workspace_size(src->h, src->w, factor);
  if (!workspace || workspace_bytes < need)
    return fail(FSG_E_WORKSPACE, "fsg_downsample_nan_aware: workspace too small");
  char* ws = (char*)workspace;
  const size_t g = align256(sizeof(float) * (size_t)hc * wc);
  float* t[4];
  for (int k = 0; k < 4; k++) t[k] = (float*)(ws + k * g);
  int* flag = (int*)(ws + 4 * g);
  return decimate(src, factor, out, hc, wc, flag, t[0], t[1], t[2], t[3], false,
                  as_stream(stream));
}

extern "C" int fsg_topousm(const fsg_window* win, const fsg_region* reg,
                           const fsg_topousm_params* p, const fsg_output* out,
                           const fsg_output* out2, void* workspace, size_t workspace_bytes,
                           void* stream) {
  hipStream_t s = as_stream(stream);
  UsmArgs A;
  bool fast = false;
  int n_blocks = 0;
  const int rc =
      usm_prepare(win, reg, p, out, out2, workspace, workspace_bytes, s, &A, &fast, &n_blocks);
  if (rc || n_blocks == 0) return rc;
  // The NaN-aware variant is exact for every window; nan_hint == 0 (the caller
  // knows the window holds no NaN) selects the cheaper plain-mean variant,
  // whose results are identical on NaN-free windows.
  if (fast && wave_eligible(A, p->nan_hint != 0))
    return dispatch_wave(A, (int4*)((char*)workspace + plan_coarse(win, p).tab_off), s);
  if (fast) return dispatch_fast(A, /*nanaw=*/p->nan_hint != 0, n_blocks, s);
  return dispatch_usm(A, /*nanaw=*/p->nan_hint != 0, n_blocks, s);
}

extern "C" int fsg_topousm_set_variant(int32_t variant) {
  if (variant != 0 && variant != 1 && !(variant >= 8 && variant <= 1024))
    return fail(FSG_E_INVALID,
                "fsg_topousm_set_variant: 0 (default), 1 (workgroup kernels) or 8..1024 "
                "(wavefront kernel rows per block)");
  g_usm_variant.store(variant, std::memory_order_relaxed);
  return FSG_OK;
}

// ------------------------------------------------ no-allocation entry point
// The workspace of fsg_topousm_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code never passes its name to a
// kernel: the name is a host-side stand-in, and the device address comes from
// hipGetSymbolAddress, once per device, kept in g_scratch_ptr.
static __device__ __align__(256) unsigned char scratch[FSG_TOPOUSM_SCRATCH_BYTES];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static int scratch_for_current_device(void** out) {
  int dev = -1;
  if (int rc = check_hip(hipGetDevice(&dev), "fsg_topousm_noalloc: hipGetDevice")) return rc;
  if (dev < 0 || dev >= kMaxDevices)
    return fail(FSG_E_UNSUPPORTED, "fsg_topousm_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (int rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)),
                           "fsg_topousm_noalloc: hipGetSymbolAddress"))
      return rc;
    if (!ptr) return fail(FSG_E_HIP, "fsg_topousm_noalloc: scratch has no device address");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return FSG_OK;
}

extern "C" int fsg_topousm_noalloc(const fsg_window* win, const fsg_region* reg,
                                   const fsg_topousm_params* p, const fsg_output* out,
                                   const fsg_output* out2, void* stream) {
  if (!win || !p) return fail(FSG_E_INVALID, "fsg_topousm_noalloc: window/params NULL");
  const size_t need = fsg_topousm_workspace_size(win, p);
  if (need > (size_t)FSG_TOPOUSM_SCRATCH_BYTES)
    return fail(FSG_E_WORKSPACE,
                "fsg_topousm_noalloc: this window needs " + std::to_string(need) +
                    " bytes of coarse-grid scratch, above the fixed maximum of " +
                    std::to_string((size_t)FSG_TOPOUSM_SCRATCH_BYTES) +
                    " (use fsg_topousm with a workspace)");
  void* ws = nullptr;
  if (int rc = scratch_for_current_device(&ws)) return rc;
  return fsg_topousm(win, reg, p, out, out2, ws, (size_t)FSG_TOPOUSM_SCRATCH_BYTES, stream);
}

extern "C" size_t fsg_topousm_coarse_means_workspace_size(const fsg_window* win, int32_t ds) {
  if (!win || win->h <= 0 || win->w <= 0 || ds < 1) return 0;
  const size_t cells = (size_t)((win->h + ds - 1) / ds) * (size_t)((win->w + ds - 1) / ds);
  // decimated grid, flag, four void-fill temporaries (two reused by the box passes)
  return 5 * align256(sizeof(float) * cells) + align256(sizeof(int) * 4);
}

extern "C" int fsg_topousm_coarse_means(const fsg_window* win, int32_t ds, int32_t n_radii,
                                        const int32_t* radii_coarse, float* const* outs,
                                        int64_t ld_out, void* workspace, size_t workspace_bytes,
                                        void* stream) {
  int rc = validate_window(win, "fsg_topousm_coarse_means");
  if (rc) return rc;
  if (win->h == 0 || win->w == 0) return fail(FSG_E_INVALID, "fsg_topousm_coarse_means: empty");
  if (ds < 1 || n_radii <= 0 || !radii_coarse || !outs)
    return fail(FSG_E_INVALID, "fsg_topousm_coarse_means: bad arguments");
  const int hc = (win->h + ds - 1) / ds, wc = (win->w + ds - 1) / ds;
  if (ld_out < wc) return fail(FSG_E_INVALID, "fsg_topousm_coarse_means: ld_out < grid width");
  for (int i = 0; i < n_radii; i++)
    if (!outs[i]) return fail(FSG_E_INVALID, "fsg_topousm_coarse_means: output NULL");
  const size_t need = fsg_topousm_coarse_means_workspace_size(win, ds);
  if (!workspace || workspace_bytes < need)
    return fail(FSG_E_WORKSPACE, "fsg_topousm_coarse_means: workspace too small");
  char* ws = (char*)workspace;
  const size_t g = align256(sizeof(float) * (size_t)hc * wc);
  float* grid = (float*)ws;
  float* tmp[4];
  for (int k = 0; k < 4; k++) tmp[k] = (float*)(ws + (1 + k) * g);
  int* flag = (int*)(ws + 5 * g);
  hipStream_t s = as_stream(stream);
  int64_t ldg = wc;
  if (ds == 1) {  // the window is already the grid
    grid = const_cast<float*>(win->data);
    ldg = win->ld;
  } else {
    rc = decimate(win, ds, grid, hc, wc, flag, tmp[0], tmp[1], tmp[2], tmp[3], false, s);
    if (rc) return rc;
  }
  for (int i = 0; i < n_radii; i++) {
    rc = coarse_mean(grid, ldg, hc, wc, std::max(1, (int)radii_coarse[i]), outs[i], ld_out, tmp[0],
                     tmp[1], s);
    if (rc) return rc;
  }
  return FSG_OK;
}

// Workspace of the batched coarse means: per job its decimated grid, per
// (job, radius) the transposed axis-0 means of box pass 0 (the passes of
// every job run in the same launches, so nothing is shared).
static size_t cmb_job_bytes(const fsg_coarse_job& J) {
  const size_t g = align256(sizeof(float) * (size_t)((J.win.h + J.ds - 1) / J.ds) *
                            (size_t)((J.win.w + J.ds - 1) / J.ds));
  return g * (1 + (size_t)std::max(0, (int)J.n_radii));
}

extern "C" size_t fsg_topousm_coarse_means_batch_workspace_size(const fsg_coarse_job* jobs,
                                                                int32_t n) {
  if (!jobs || n <= 0) return 0;
  size_t b = 0;
  for (int i = 0; i < n; i++) {
    if (jobs[i].win.h <= 0 || jobs[i].win.w <= 0 || jobs[i].ds < 1) return 0;
    b += cmb_job_bytes(jobs[i]);
  }
  return b;
}

static bool same_window(const fsg_window& a, const fsg_window& b) {
  return a.data == b.data && a.ld == b.ld && a.h == b.h && a.w == b.w;
}

extern "C" int fsg_topousm_coarse_means_batch(const fsg_coarse_job* jobs, int32_t n,
                                              void* workspace, size_t workspace_bytes,
                                              void* stream) {
  if (!jobs || n <= 0) return fail(FSG_E_INVALID, "fsg_topousm_coarse_means_batch: no jobs");
  for (int i = 0; i < n; i++) {
    const fsg_coarse_job& J = jobs[i];
    int rc = validate_window(&J.win, "fsg_topousm_coarse_means_batch");
    if (rc) return rc;
    if (J.win.h == 0 || J.win.w == 0)
      return fail(FSG_E_INVALID, "fsg_topousm_coarse_means_batch: empty window");
    if (J.ds != 1 && (J.ds < 4 || J.ds > 16 || (J.ds & (J.ds - 1)) != 0))
      return fail(FSG_E_UNSUPPORTED,
                  "fsg_topousm_coarse_means_batch: ds must be 1, 4, 8 or 16 (use "
                  "fsg_topousm_coarse_means)");
    if (J.n_radii < 1 || J.n_radii > FSG_COARSE_JOB_MAX_RADII)
      return fail(FSG_E_INVALID, "fsg_topousm_coarse_means_batch: n_radii out of range");
    const int wc = (J.win.w + J.ds - 1) / J.ds;
    if (J.ld_out < wc)
      return fail(FSG_E_INVALID, "fsg_topousm_coarse_means_batch: ld_out < grid width");
    for (int k = 0; k < J.n_radii; k++) {
      if (!J.outs[k]) return fail(FSG_E_INVALID, "fsg_topousm_coarse_means_batch: output NULL");
      if (J.radii[k] <= 1)
        return fail(FSG_E_UNSUPPORTED,
                    "fsg_topousm_coarse_means_batch: coarse radius <= 1 (the Gaussian branch; "
                    "use fsg_topousm_coarse_means)");
    }
  }
  const size_t need = fsg_topousm_coarse_means_batch_workspace_size(jobs, n);
  if (!workspace || workspace_bytes < need)
    return fail(FSG_E_WORKSPACE, "fsg_topousm_coarse_means_batch: workspace too small");
  hipStream_t s = as_stream(stream);
  // workspace carve-up
  std::vector<float*> grid(n);
  std::vector<int64_t> ldg(n);
  std::vector<std::vector<float*>> tv(n);
  {
    char* p = (char*)workspace;
    for (int i = 0; i < n; i++) {
      const fsg_coarse_job& J = jobs[i];
      const size_t g = align256(sizeof(float) * (size_t)((J.win.h + J.ds - 1) / J.ds) *
                                (size_t)((J.win.w + J.ds - 1) / J.ds));
      grid[i] = (float*)p;
      ldg[i] = (J.win.w + J.ds - 1) / J.ds;
      p += g;
      if (J.ds == 1) {  // the window already is the grid
        grid[i] = const_cast<float*>(J.win.data);
        ldg[i] = J.win.ld;
      }
      for (int k = 0; k < J.n_radii; k++) {
        tv[i].push_back((float*)p);
        p += g;
      }
    }
  }
  // decimations: a job whose window another job shares at 2x or 4x its factor
  // rides along with that job's read
  std::vector<DualDecJob> dec;
  std::vector<int> paired(n, 0);
  for (int i = 0; i < n; i++)
    if (jobs[i].ds == 1) paired[i] = 1;  // nothing to decimate
  for (int i = 0; i < n; i++) {
    if (paired[i]) continue;
    const fsg_coarse_job& A = jobs[i];
    int partner = -1;
    for (int k = 0; k < n && partner < 0; k++) {
      if (k == i || paired[k] || !same_window(A.win, jobs[k].win)) continue;
      const int fa = A.ds, fb = jobs[k].ds;
      if (fb == 2 * fa || fb == 4 * fa) partner = k;
    }
    // a larger factor paired with a smaller one later is handled from that side
    if (partner < 0) {
      bool later = false;
      for (int k = 0; k < n; k++)
        if (k != i && !paired[k] && same_window(A.win, jobs[k].win) &&
            (A.ds == 2 * jobs[k].ds || A.ds == 4 * jobs[k].ds))
          later = true;
      if (later) continue;
    }
    DualDecJob D{};
    D.W = WinDev{A.win.data, A.win.ld, A.win.h, A.win.w};
    D.oa = grid[i];
    D.fa = A.ds;
    D.oha = (A.win.h + A.ds - 1) / A.ds;
    D.owa = (A.win.w + A.ds - 1) / A.ds;
    D.nbx = (A.win.w + 1023) / 1024;
    paired[i] = 1;
    if (partner >= 0) {
      const fsg_coarse_job& B = jobs[partner];
      D.ob = grid[partner];
      D.fb = B.ds;
      D.ohb = (B.win.h + B.ds - 1) / B.ds;
      D.owb = (B.win.w + B.ds - 1) / B.ds;
      paired[partner] = 1;
    }
    dec.push_back(D);
  }
  for (int i = 0; i < n; i++)
    if (!paired[i]) return fail(FSG_E_INVALID, "fsg_topousm_coarse_means_batch: unplanned job");
  for (size_t a = 0; a < dec.size(); a += CMB_MAX_DEC) {
    DecTab T{};
    T.n = (int)std::min<size_t>(CMB_MAX_DEC, dec.size() - a);
    T.pre[0] = 0;
    for (int k = 0; k < T.n; k++) {
      T.j[k] = dec[a + k];
      const int rows = T.j[k].fb ? T.j[k].ohb : T.j[k].oha;  // bands of the largest factor
      T.pre[k + 1] = T.pre[k] + T.j[k].nbx * rows;
    }
    hipLaunchKernelGGL(k_decimate_tab, dim3(T.pre[T.n]), dim3(256), 0, s, T);
    FSG_CHECK_LAUNCH("k_decimate_tab");
  }
  // box passes (NaN-free: means only): every (job, radius) in the same two launches
  std::vector<BoxJob> b0, b1;
  for (int i = 0; i < n; i++) {
    const fsg_coarse_job& J = jobs[i];
    const int h = (J.win.h + J.ds - 1) / J.ds, w = (J.win.w + J.ds - 1) / J.ds;
    for (int k = 0; k < J.n_radii; k++) {
      const int r = J.radii[k];
      const double inv = 1.0 / (double)(2 * r + 1);
      // one 64-row sub-band per band: the launch lasts as long as its
      // slowest block's chain of loads, which taller bands lengthen
      const int nsub = 1;
      b0.push_back(BoxJob{grid[i], nullptr, ldg[i], h, w, r, inv, tv[i][k], nullptr,
                          (int64_t)h, (w + 63) / 64, nsub});
      b1.push_back(BoxJob{tv[i][k], nullptr, (int64_t)h, w, h, r, inv, J.outs[k], nullptr,
                          (int64_t)J.ld_out, (h + 63) / 64, nsub});
    }
  }
  for (int pass = 0; pass < 2; pass++) {
    const std::vector<BoxJob>& v = pass ? b1 : b0;
    for (size_t a = 0; a < v.size(); a += CMB_MAX_BOX) {
      BoxTab T{};
      T.n = (int)std::min<size_t>(CMB_MAX_BOX, v.size() - a);
      T.pre[0] = 0;
      for (int k = 0; k < T.n; k++) {
        T.j[k] = v[a + k];
        T.pre[k + 1] = T.pre[k] + box_blocks(T.j[k]);
      }
      if (pass == 0)
        hipLaunchKernelGGL(k_box_plain_tab<0>, dim3(T.pre[T.n]), dim3(64), 0, s, T);
      else
        hipLaunchKernelGGL(k_box_plain_tab<1>, dim3(T.pre[T.n]), dim3(64), 0, s, T);
      FSG_CHECK_LAUNCH("k_box_plain_tab");
    }
  }
  return FSG_OK;
}

// field = sum_i f32(w_i) * mean_i in radius order, the first term without an
// addition (k_field_accum's arithmetic, all terms in one pass)
constexpr int FIELD_MAX_TERMS = 8;
struct FieldTerms {
  const float* m[FIELD_MAX_TERMS];
  int64_t ld[FIELD_MAX_TERMS];
  float wt[FIELD_MAX_TERMS];
  int n;
};

__global__ void k_field_sum(FieldTerms T, float* __restrict__ out, int64_t ldo, int h, int w) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= w) return;
  for (int r = blockIdx.y; r < h; r += gridDim.y) {
    float acc = T.wt[0] * T.m[0][(int64_t)r * T.ld[0] + c];
    for (int i = 1; i < T.n; i++) {
      const float t = T.wt[i] * T.m[i][(int64_t)r * T.ld[i] + c];
      acc = acc + t;
    }
    out[(int64_t)r * ldo + c] = acc;
  }
}

extern "C" int fsg_topousm_field_from_means(const float* const* means, const int64_t* ld_means,
                                            const float* weights, int32_t n, int32_t h,
                                            int32_t w, float* out, int64_t ld_out,
                                            void* stream) {
  if (n < 1 || n > FIELD_MAX_TERMS || !means || !ld_means || !weights || !out || h <= 0 ||
      w <= 0 || ld_out < w)
    return fail(FSG_E_INVALID, "fsg_topousm_field_from_means: bad arguments");
  FieldTerms T{};
  T.n = n;
  for (int i = 0; i < n; i++) {
    if (!means[i] || ld_means[i] < w)
      return fail(FSG_E_INVALID, "fsg_topousm_field_from_means: bad mean grid");
    T.m[i] = means[i];
    T.ld[i] = ld_means[i];
    T.wt[i] = weights[i];
  }
  dim3 blk(256), grd((w + 255) / 256, std::min(h, 32768));
  hipLaunchKernelGGL(k_field_sum, grd, blk, 0, as_stream(stream), T, out, ld_out, h, w);
  FSG_CHECK_LAUNCH("k_field_sum");
  return FSG_OK;
}

extern "C" size_t fsg_topousm_field_workspace_size(int32_t h, int32_t w) {
  if (h <= 0 || w <= 0) return 0;
  return 3 * align256(sizeof(float) * (size_t)h * w);
}

extern "C" int fsg_topousm_coarse_field(const fsg_window* coarse, int32_t n_radii,
                                        const int32_t* radii_coarse, const float* weights,
                                        float* out, int64_t ld_out, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  int rc = validate_window(coarse, "fsg_topousm_coarse_field");
  if (rc) return rc;
  if (n_radii <= 0 || !radii_coarse || !weights || !out || ld_out < coarse->w)
    return fail(FSG_E_INVALID, "fsg_topousm_coarse_field: bad arguments");
  const int h = coarse->h, w = coarse->w;
  if (h == 0 || w == 0) return fail(FSG_E_INVALID, "fsg_topousm_coarse_field: empty grid");
  const size_t need = fsg_topousm_field_workspace_size(h, w);
  if (!workspace || workspace_bytes < need)
    return fail(FSG_E_WORKSPACE, "fsg_topousm_coarse_field: workspace too small");
  char* ws = (char*)workspace;
  float* tv = (float*)ws;
  float* tw = (float*)(ws + align256(sizeof(float) * (size_t)h * w));
  float* mean = (float*)(ws + 2 * align256(sizeof(float) * (size_t)h * w));
  hipStream_t s = as_stream(stream);
  for (int i = 0; i < n_radii; i++) {
    const int r = std::max(1, (int)radii_coarse[i]);
    rc = coarse_mean(coarse->data, coarse->ld, h, w, r, mean, w, tv, tw, s);
    if (rc) return rc;
    dim3 blk(128), grd((w + 127) / 128, std::min(h, 32768));
    hipLaunchKernelGGL(k_field_accum, grd, blk, 0, s, out, ld_out, mean, (int64_t)w, h, w,
                       weights[i], i == 0 ? 1 : 0);
    FSG_CHECK_LAUNCH("k_field_accum");
  }
  return FSG_OK;
}

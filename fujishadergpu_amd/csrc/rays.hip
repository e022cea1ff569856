// Ray marches: sparse Yokoyama openness and stylised ambient occlusion.
//
// Openness (compute_openness_vectorized, _impl_openness.py:31-132): per
// direction the horizon extreme of atan((z_s - z) / d) over the valid
// in-window samples.  atan is monotone, so the kernel takes the extreme of the
// slope ratio and evaluates one atanf per direction instead of one per sample.
// The ratio uses f32(1/phys) multiplication (within 1.5 ulp of the reference's
// division by f32(phys)).
//
// Ambient occlusion (compute_ambient_occlusion_block,
// _impl_ambient_occlusion.py:32-121): edge-padded samples, clamp(atan,0,pi/4),
// ring weighting, mean over valid samples, then a sigma=1 'nearest' Gaussian
// (scipy two-pass order) of where(nan, 1, ao), gamma, NaN restored.
#include <algorithm>
#include <cmath>

#include <string>

#include "rays.hpp"

namespace fsg {

// Block order of the ray kernels: 64 x 4 pixel blocks on a 1-D grid.  With
// at least 8 block columns the region is cut into 8 column panels, one per
// XCD (workgroup b runs on XCD b mod 8), walked row by row inside the panel:
// the rows an XCD's resident blocks sample (their 4 rows plus the rays'
// vertical reach) then span one panel's width, not the whole region's, and
// stay in that XCD's L2.  Narrower regions keep the plain row-major order.
struct RayGrid {
  int nbx, nby;   // 64-column and 4-row blocks of the region
  int pw;         // blocks per panel row (0: row-major)
};

inline RayGrid ray_grid(int h, int w) {
  RayGrid G;
  G.nbx = (w + 63) / 64;
  G.nby = (h + 3) / 4;
  G.pw = G.nbx >= 8 ? (G.nbx + 7) / 8 : 0;
  return G;
}

inline int64_t ray_grid_blocks(const RayGrid& G) {
  return G.pw ? (int64_t)8 * G.pw * G.nby : (int64_t)G.nbx * G.nby;
}

// (bx, by) of block b; false for the padding blocks of the last panel
__device__ __forceinline__ bool ray_block(const RayGrid& G, int b, int& bx, int& by) {
  if (G.pw == 0) {
    by = b / G.nbx;
    bx = b - by * G.nbx;
    return true;
  }
  const int xcd = b & 7, k = b >> 3;
  by = k / G.pw;
  bx = xcd * G.pw + (k - by * G.pw);
  return bx < G.nbx && by < G.nby;
}

// ND: the table's distances per direction when it is 1 .. 10 (fixed-length
// rows: wide scalar table loads, no guards), 0 for any other count.
template <int ND>
__global__ void __launch_bounds__(256)
k_openness(WinDevR W, int r0, int c0, int rh, int rw, OpenTab T, OutDev out, RayGrid G) {
  int bx, by;
  if (!ray_block(G, blockIdx.x, bx, by)) return;
  const int c = bx * 64 + threadIdx.x;
  const int r = by * 4 + threadIdx.y;
  // block-uniform: do all samples of the block's pixels stay inside the window?
  const int bi0 = r0 + by * 4, bj0 = c0 + bx * 64;
  const bool interior = openness_interior(W, T, bi0, bi0 + 3, bj0, bj0 + 63);
  if (c >= rw || r >= rh) return;
  const int i = r0 + r, j = c0 + c;
  const PostK pp = load_post(out);
  const float z = W.p[(int64_t)i * W.ld + j];
  if (isnan(z)) {
    store_out(out, r, c, apply_post(pp, __builtin_nanf("")));
    return;
  }
  float a;
  if constexpr (ND > 0)
    a = !interior ? openness_angle_u<ND, true>(W, i, j, z, T)
        : open_off32(W, T) ? openness_angle_u<ND, false, true>(W, i, j, z, T)
                        : openness_angle_u<ND, false>(W, i, j, z, T);
  else
    a = openness_angle_at(W, i, j, z, T, interior);
  store_out(out, r, c, apply_post(pp, openness_value(T, a)));
}

// The same at a list of window pixels (row, col), float32 out[n].
__global__ void __launch_bounds__(256)
k_openness_points(WinDevR W, const int2* __restrict__ pts, int n, OpenTab T,
                  float* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int2 q = pts[e];
  // a point outside the window reads nothing and gets NaN
  if (q.x < 0 || q.x >= W.h || q.y < 0 || q.y >= W.w) {
    out[e] = __builtin_nanf("");
    return;
  }
  const float z = W.p[(int64_t)q.x * W.ld + q.y];
  out[e] = isnan(z) ? __builtin_nanf("")
                    : openness_value(T, openness_angle_at(W, q.x, q.y, z, T, false));
}

// Sampled exact check of an approximate openness-angle raster (the learned
// surrogate's judge; no reference counterpart).  The window is cut into
// tile_h x tile_w tiles (ragged at the far edges); tile t (row-major) gets
// n_t = max(1, ceil(frac * th * tw / L)) segments of L pixels along one row,
// segment k at
//   h = splitmix64(splitmix64(seed ^ splitmix64(t)) + k),
//   row = y0 + (h >> 32) % th,  col0 = x0 + (h & 0xffffffff) % (tw - L + 1).
// One wave per segment, lane = pixel, so the ray gathers coalesce as in the
// dense kernel.  Block (t, split) takes segments split*4 + wave, +
// JUDGE_SPLIT*4, ... and writes the sum and count of |exact - approx| over
// pixels where both are numbers; the reduction order is fixed, so the result
// is deterministic.
constexpr int JUDGE_SPLIT = 64;

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct JudgeGeo {
  int tile_h, tile_w, n_tx, L;
  double frac;
  uint64_t seed;
};

__device__ __forceinline__ int judge_count(const JudgeGeo& G, int th, int tw) {
  return max(1, (int)ceil(G.frac * (double)th * (double)tw / (double)G.L));
}

__global__ void __launch_bounds__(256)
k_openness_judge(WinDevR W, OpenTab T, const float* __restrict__ approx, int64_t lda, JudgeGeo G,
                 double* __restrict__ part) {
  const int tile = blockIdx.x / JUDGE_SPLIT, split = blockIdx.x % JUDGE_SPLIT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int y0 = (tile / G.n_tx) * G.tile_h, x0 = (tile % G.n_tx) * G.tile_w;
  const int th = min(G.tile_h, W.h - y0), tw = min(G.tile_w, W.w - x0);
  const int n_seg = judge_count(G, th, tw);
  const uint64_t key = splitmix64(G.seed ^ splitmix64((uint64_t)tile));
  float sum = 0.0f, cnt = 0.0f;
  for (int s = split * 4 + wave; s < n_seg; s += JUDGE_SPLIT * 4) {
    const uint64_t h = splitmix64(key + (uint64_t)s);
    const int i = y0 + (int)((h >> 32) % (uint64_t)th);
    const int jseg = x0 + (int)((h & 0xffffffffull) % (uint64_t)(tw - G.L + 1));
    const int j = jseg + lane;
    const bool interior = openness_interior(W, T, i, i, jseg, jseg + G.L - 1);
    if (lane < G.L) {
      const float z = W.p[(int64_t)i * W.ld + j];
      const float a = approx[(int64_t)i * lda + j];
      if (!isnan(z) && !isnan(a)) {
        const float v = openness_value(T, openness_angle_at(W, i, j, z, T, interior));
        sum += fabsf(v - a);
        cnt += 1.0f;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    sum += __shfl_down(sum, o);
    cnt += __shfl_down(cnt, o);
  }
  __shared__ double red[4][2];
  if (lane == 0) {
    red[wave][0] = sum;
    red[wave][1] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ps = 0.0, pc = 0.0;
    for (int w = 0; w < 4; w++) {
      ps += red[w][0];
      pc += red[w][1];
    }
    part[(int64_t)blockIdx.x * 2] = ps;
    part[(int64_t)blockIdx.x * 2 + 1] = pc;
  }
}

// mean absolute error per tile (NaN for a tile without a valid sample)
__global__ void k_openness_judge_finish(const double* __restrict__ part, int n_tiles,
                                        float* __restrict__ tile_mae) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles) return;
  double s = 0.0, c = 0.0;
  for (int k = 0; k < JUDGE_SPLIT; k++) {
    s += part[((int64_t)t * JUDGE_SPLIT + k) * 2];
    c += part[((int64_t)t * JUDGE_SPLIT + k) * 2 + 1];
  }
  tile_mae[t] = c > 0.0 ? (float)(s / c) : __builtin_nanf("");
}

constexpr int AO_MAX = 256;

struct AoTab {
  int n;
  float intensity;  // f32(intensity)
  float quarter_pi; // f32(pi/4)
  float gamma;
  int dmax_y, dmax_x;  // largest |offset| on each axis (interior test)
  short2 off[AO_MAX];
  float inv[AO_MAX];
  float df[AO_MAX];
};

// Occlusion sum and valid-sample count at window (i, j), samples in table
// order, 8 loads issued before any is used.  CHECK = false: every sample
// lies inside the window (wave-uniform element offsets, no clamp); with CHECK
// the sample coordinates are clamped to the window ('edge' padding).  The
// float32 operations and their order are those of the per-sample loop.
template <bool CHECK>
__device__ __forceinline__ void ao_sum(const WinDevR& W, int i, int j, float z, const AoTab& T,
                                       float& total, float& count) {
  const float* pc = W.p + (int64_t)i * W.ld + j;
  total = 0.0f;
  count = 0.0f;
  for (int e0 = 0; e0 < T.n; e0 += 8) {
    float v[8];
    bool ok[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int e = e0 + k;
      ok[k] = false;
      v[k] = 0.0f;
      if (e < T.n) {
        const short2 o = T.off[e];
        ok[k] = !(o.x == 0 && o.y == 0);
        if (CHECK)
          v[k] = W.p[(int64_t)clampi(i + o.y, W.h) * W.ld + clampi(j + o.x, W.w)];
        else
          v[k] = pc[(int64_t)o.y * W.ld + o.x];
      }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int e = e0 + k;
      if (e < T.n) {
        const bool valid = ok[k] && !isnan(v[k]);
        const float a = atanf((v[k] - z) * T.inv[e]);
        float occ = fmaxf(0.0f, a) / T.quarter_pi;
        occ = fminf(occ, 1.0f);
        const float nt = total + occ * T.df[e];
        total = valid ? nt : total;
        count = valid ? count + 1.0f : count;
      }
    }
  }
}

// Pass 1: AO before the Gaussian on the expanded region, stored as
// where(isnan(z), 1, ao) -- the array the reference smooths.
__global__ void __launch_bounds__(256)
k_ao_raw(WinDevR W, int er0, int ec0, int eh, int ew, AoTab T, float* __restrict__ dst) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (c >= ew || r >= eh) return;
  const int i = er0 + r, j = ec0 + c;
  // block-uniform: do all samples of the block's pixels stay inside the window?
  const int bi0 = er0 + blockIdx.y * blockDim.y, bj0 = ec0 + blockIdx.x * blockDim.x;
  const bool interior = bi0 - T.dmax_y >= 0 && bi0 + (int)blockDim.y - 1 + T.dmax_y < W.h &&
                        bj0 - T.dmax_x >= 0 && bj0 + (int)blockDim.x - 1 + T.dmax_x < W.w;
  const float z = W.p[(int64_t)i * W.ld + j];
  float res;
  if (isnan(z)) {
    res = 1.0f;
  } else {
    float total, count;
    if (interior)
      ao_sum<false>(W, i, j, z, T, total, count);
    else
      ao_sum<true>(W, i, j, z, T, total, count);
    count = fmaxf(count, 1.0f);
    const float mean = total / count;
    float ao = 1.0f - mean * T.intensity;
    res = fminf(fmaxf(ao, 0.0f), 1.0f);
  }
  dst[(int64_t)r * ew + c] = res;
}

struct G9 {
  double g[5];
};

// Pass 2: sigma=1 Gaussian ('nearest' at the window edge), gamma, NaN restore.
__global__ void __launch_bounds__(256)
k_ao_smooth(WinDevR W, int r0, int c0, int rh, int rw, int er0, int ec0, int eh, int ew,
            const float* __restrict__ src, G9 G, float gamma, OutDev out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (c >= rw || r >= rh) return;
  const int i = r0 + r, j = c0 + c;
  const PostK pp = load_post(out);
  const float z = W.p[(int64_t)i * W.ld + j];
  if (isnan(z)) {
    store_out(out, r, c, apply_post(pp, __builtin_nanf("")));
    return;
  }
  float v[9];
  for (int k = 0; k < 9; k++) {
    const int jj = clampi(j + k - 4, W.w) - ec0;
    const float* colp = src + jj;
    double s = (double)colp[(int64_t)(i - er0) * ew] * G.g[0];
    for (int d = 4; d >= 1; d--) {
      const int ia = clampi(i - d, W.h) - er0, ib = clampi(i + d, W.h) - er0;
      s += ((double)colp[(int64_t)ia * ew] + (double)colp[(int64_t)ib * ew]) * G.g[d];
    }
    v[k] = (float)s;
  }
  double s = (double)v[4] * G.g[0];
  for (int d = 4; d >= 1; d--) s += ((double)v[4 - d] + (double)v[4 + d]) * G.g[d];
  const float ao = (float)s;
  store_out(out, r, c, apply_post(pp, powf(ao, gamma)));
}

static void expanded(const fsg_window* win, const fsg_region* reg, int* er0, int* ec0, int* eh,
                     int* ew) {
  *er0 = std::max(0, reg->r0 - 4);
  *ec0 = std::max(0, reg->c0 - 4);
  const int er1 = std::min(win->h, reg->r0 + reg->h + 4);
  const int ec1 = std::min(win->w, reg->c0 + reg->w + 4);
  *eh = er1 - *er0;
  *ew = ec1 - *ec0;
}

}  // namespace fsg

using namespace fsg;

extern "C" int fsg_openness(const fsg_window* win, const fsg_region* reg,
                            const fsg_openness_params* p, const fsg_output* out, void* stream) {
  int rc = validate_window(win, "fsg_openness");
  if (rc) return rc;
  rc = validate_region(win, reg, "fsg_openness");
  if (rc) return rc;
  if (!out) return fail(FSG_E_INVALID, "fsg_openness: output NULL");
  if ((rc = validate_output(out, reg, "fsg_openness"))) return rc;
  OpenTab T;
  if ((rc = openness_table(p, T, "fsg_openness"))) return rc;
  if (reg->h == 0 || reg->w == 0) return FSG_OK;
  WinDevR W{win->data, win->ld, win->h, win->w};
  const RayGrid G = ray_grid(reg->h, reg->w);
  if (ray_grid_blocks(G) > INT32_MAX)
    return fail(FSG_E_UNSUPPORTED, "fsg_openness: region too large for one launch");
  dim3 blk(64, 4), grd((unsigned)ray_grid_blocks(G));
  TimedScope ts("k_openness", as_stream(stream));
  const OutDev od = make_out(out);
  hipStream_t s = as_stream(stream);
  switch (T.n_dist) {
#define FSG_OPEN_ND(N)                                                                       \
  case N:                                                                                    \
    hipLaunchKernelGGL(k_openness<N>, grd, blk, 0, s, W, reg->r0, reg->c0, reg->h, reg->w, T, \
                       od, G);                                                               \
    break;
    FSG_OPEN_ND(1) FSG_OPEN_ND(2) FSG_OPEN_ND(3) FSG_OPEN_ND(4) FSG_OPEN_ND(5)
    FSG_OPEN_ND(6) FSG_OPEN_ND(7) FSG_OPEN_ND(8) FSG_OPEN_ND(9) FSG_OPEN_ND(10)
#undef FSG_OPEN_ND
    default:
      hipLaunchKernelGGL(k_openness<0>, grd, blk, 0, s, W, reg->r0, reg->c0, reg->h, reg->w, T,
                         od, G);
  }
  FSG_CHECK_LAUNCH("k_openness");
  return FSG_OK;
}

extern "C" int fsg_openness_points(const fsg_window* win, const int32_t* points, int32_t n,
                                   const fsg_openness_params* p, float* out, void* stream) {
  int rc = validate_window(win, "fsg_openness_points");
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!points || !out)))
    return fail(FSG_E_INVALID, "fsg_openness_points: bad point list");
  OpenTab T;
  if ((rc = openness_table(p, T, "fsg_openness_points"))) return rc;
  if (n == 0) return FSG_OK;
  WinDevR W{win->data, win->ld, win->h, win->w};
  TimedScope ts("k_openness_points", as_stream(stream));
  hipLaunchKernelGGL(k_openness_points, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream),
                     W, (const int2*)points, n, T, out);
  FSG_CHECK_LAUNCH("k_openness_points");
  return FSG_OK;
}

extern "C" size_t fsg_openness_judge_workspace_size(const fsg_window* win, int32_t tile_h,
                                                    int32_t tile_w) {
  if (!win || win->h <= 0 || win->w <= 0 || tile_h <= 0 || tile_w <= 0) return 0;
  const size_t n_tiles = (size_t)((win->h + tile_h - 1) / tile_h) * ((win->w + tile_w - 1) / tile_w);
  return sizeof(double) * 2 * JUDGE_SPLIT * n_tiles;
}

extern "C" int fsg_openness_judge(const fsg_window* win, const fsg_openness_params* p,
                                  const float* approx, int64_t lda, int32_t tile_h,
                                  int32_t tile_w, double sample_frac, uint64_t seed,
                                  int32_t seg_len, float* tile_mae, void* workspace,
                                  size_t workspace_bytes, void* stream) {
  int rc = validate_window(win, "fsg_openness_judge");
  if (rc) return rc;
  if (win->h == 0 || win->w == 0 || !approx || lda < win->w || !tile_mae || tile_h <= 0 ||
      tile_w <= 0 || !(sample_frac > 0.0) || sample_frac > 1.0)
    return fail(FSG_E_INVALID, "fsg_openness_judge: bad arguments");
  const int n_ty = (win->h + tile_h - 1) / tile_h, n_tx = (win->w + tile_w - 1) / tile_w;
  const int min_tw = std::min(tile_w, win->w - (n_tx - 1) * tile_w);
  if (seg_len < 1 || seg_len > 64 || seg_len > min_tw)
    return fail(FSG_E_INVALID, "fsg_openness_judge: seg_len must be 1..64 and fit every tile");
  const int64_t n_tiles = (int64_t)n_ty * n_tx;
  if (n_tiles * JUDGE_SPLIT > 0x7fffffff)
    return fail(FSG_E_INVALID, "fsg_openness_judge: too many tiles");
  if (!workspace || workspace_bytes < fsg_openness_judge_workspace_size(win, tile_h, tile_w))
    return fail(FSG_E_WORKSPACE, "fsg_openness_judge: workspace too small");
  OpenTab T;
  if ((rc = openness_table(p, T, "fsg_openness_judge"))) return rc;
  WinDevR W{win->data, win->ld, win->h, win->w};
  JudgeGeo G{tile_h, tile_w, n_tx, seg_len, sample_frac, seed};
  hipStream_t s = as_stream(stream);
  TimedScope ts("k_openness_judge", s);
  hipLaunchKernelGGL(k_openness_judge, dim3((unsigned)(n_tiles * JUDGE_SPLIT)), dim3(256), 0, s,
                     W, T, approx, lda, G, (double*)workspace);
  FSG_CHECK_LAUNCH("k_openness_judge");
  hipLaunchKernelGGL(k_openness_judge_finish, dim3((unsigned)((n_tiles + 255) / 256)), dim3(256),
                     0, s, (const double*)workspace, (int)n_tiles, tile_mae);
  FSG_CHECK_LAUNCH("k_openness_judge_finish");
  return FSG_OK;
}

extern "C" size_t fsg_ambient_occlusion_workspace_size(const fsg_window* win,
                                                       const fsg_region* reg) {
  if (!win || !reg) return 0;
  int er0, ec0, eh, ew;
  expanded(win, reg, &er0, &ec0, &eh, &ew);
  if (eh <= 0 || ew <= 0) return 0;
  return sizeof(float) * (size_t)eh * (size_t)ew;
}

extern "C" int fsg_ambient_occlusion(const fsg_window* win, const fsg_region* reg,
                                     const fsg_ao_params* p, const fsg_output* out,
                                     void* workspace, size_t workspace_bytes, void* stream) {
  int rc = validate_window(win, "fsg_ambient_occlusion");
  if (rc) return rc;
  rc = validate_region(win, reg, "fsg_ambient_occlusion");
  if (rc) return rc;
  if (!p || !out) return fail(FSG_E_INVALID, "fsg_ambient_occlusion: params/output NULL");
  if ((rc = validate_output(out, reg, "fsg_ambient_occlusion"))) return rc;
  if (p->n_samples < 0 || (p->n_samples > 0 && !p->samples))
    return fail(FSG_E_INVALID, "fsg_ambient_occlusion: bad sample table");
  if (p->n_samples > AO_MAX)
    return fail(FSG_E_UNSUPPORTED, "fsg_ambient_occlusion: more than 256 samples");
  if (reg->h == 0 || reg->w == 0) return FSG_OK;
  const size_t need = fsg_ambient_occlusion_workspace_size(win, reg);
  if (!workspace || workspace_bytes < need)
    return fail(FSG_E_WORKSPACE, "fsg_ambient_occlusion: workspace too small");
  AoTab T{};
  T.n = p->n_samples;
  T.intensity = (float)p->intensity;
  T.quarter_pi = (float)(M_PI / 4);
  T.gamma = (float)(1.0 / 2.2);
  for (int e = 0; e < p->n_samples; e++) {
    const fsg_ao_sample& s = p->samples[e];
    if (s.skip || (s.dx == 0 && s.dy == 0)) {
      T.off[e] = make_short2(0, 0);
      continue;
    }
    if (std::abs(s.dx) > 32767 || std::abs(s.dy) > 32767)
      return fail(FSG_E_UNSUPPORTED, "fsg_ambient_occlusion: offset beyond int16");
    T.off[e] = make_short2((short)s.dx, (short)s.dy);
    T.inv[e] = 1.0f / (float)s.phys;
    T.df[e] = (float)s.dist_factor;
    T.dmax_y = std::max(T.dmax_y, std::abs(s.dy));
    T.dmax_x = std::max(T.dmax_x, std::abs(s.dx));
  }
  int er0, ec0, eh, ew;
  expanded(win, reg, &er0, &ec0, &eh, &ew);
  WinDevR W{win->data, win->ld, win->h, win->w};
  hipStream_t s = as_stream(stream);
  float* raw = (float*)workspace;
  {
    dim3 blk(64, 4), grd((ew + 63) / 64, (eh + 3) / 4);
    TimedScope ts("k_ao_raw", s);
    hipLaunchKernelGGL(k_ao_raw, grd, blk, 0, s, W, er0, ec0, eh, ew, T, raw);
    FSG_CHECK_LAUNCH("k_ao_raw");
  }
  G9 G;
  double w9[9];
  gaussian_weights(1.0, 4, w9);
  for (int d = 0; d <= 4; d++) G.g[d] = w9[4 + d];
  {
    dim3 blk(64, 4), grd((reg->w + 63) / 64, (reg->h + 3) / 4);
    hipLaunchKernelGGL(k_ao_smooth, grd, blk, 0, s, W, reg->r0, reg->c0, reg->h, reg->w, er0,
                       ec0, eh, ew, (const float*)raw, G, T.gamma, make_out(out));
    FSG_CHECK_LAUNCH("k_ao_smooth");
  }
  return FSG_OK;
}

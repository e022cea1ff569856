// Exact percentiles and moments on the device.
//
// Selection: every value maps to an order-preserving 32-bit key (|v| bits for
// the absolute mode, sign-flipped bits otherwise).  Three histogram passes
// (bits 31..21, 20..10, 9..0) narrow each target rank to one key; all target
// ranks (two per percentile: floor(idx) and floor(idx)+1) share each pass.
// The percentile is CuPy's linear method in FP64:
//   idx = (q/100) * (n - 1); lo = floor(idx); w = idx - lo;
//   p = x[lo] * (1 - w) + x[min(lo + 1, n - 1)] * w
// (cupy percentile -> _quantile_unchecked, as called by
// topousm_fast_stat_func, _normalization.py:27, and
// robust_unsigned_stretch_stat_func, _global_stats.py:198-199).
#include <algorithm>
#include <cmath>

#include "fsg_common.hpp"

namespace fsg {

constexpr int PCT_MAX_Q = 2;
constexpr int PCT_MAX_T = 2 * PCT_MAX_Q;
constexpr int NB1 = 2048, NB2 = 2048, NB3 = 1024;

enum { PCT_ABS = 1, PCT_FINITE = 2 };

struct PctState {
  unsigned long long n_valid;
  double pad_sum, pad_sumsq;  // unused (the moments come from PctWs::part)
  int n_targets;
  int pad;
  unsigned long long rank[PCT_MAX_T];
  unsigned long long rem[PCT_MAX_T];
  unsigned int prefix[PCT_MAX_T];
  unsigned int key[PCT_MAX_T];
  double idx[PCT_MAX_Q];
};

constexpr int PCT_BLOCKS = 2048;  // at most this many workgroups per pass

// (count, mean, sum of squared deviations) of a set of values.
struct Moments {
  double n, mean, m2;
};

// The moments of the union of two sets (Chan, Golub and LeVeque 1979): the
// stable pairwise update, applied in a fixed order so that the result does
// not depend on timing.
__host__ __device__ inline Moments merge_moments(const Moments& a, const Moments& b) {
  const double n = a.n + b.n;
  if (b.n == 0.0) return a;
  if (a.n == 0.0) return b;
  const double d = b.mean - a.mean;
  return Moments{n, a.mean + d * (b.n / n), a.m2 + b.m2 + d * d * (a.n * b.n / n)};
}

struct PctWs {
  PctState st;
  double part[PCT_BLOCKS][3];  // per workgroup of the level-1 pass: n, mean, m2
  unsigned long long h1[NB1];
  unsigned long long h2[PCT_MAX_T][NB2];
  unsigned long long h3[PCT_MAX_T][NB3];
};

__device__ __forceinline__ bool pct_valid(float v, int flags) {
  return (flags & PCT_FINITE) ? isfinite(v) : !isnan(v);
}

__device__ __forceinline__ unsigned int pct_key(float v, int flags) {
  const unsigned int b = __float_as_uint(v);
  if (flags & PCT_ABS) return b & 0x7FFFFFFFu;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float pct_value(unsigned int key, int flags) {
  if (flags & PCT_ABS) return __uint_as_float(key);
  const unsigned int b = (key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key;
  return __uint_as_float(b);
}

__global__ void k_pct_zero(unsigned long long* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// Visit every element of x[0, n) once across the grid: float4 loads, four in
// flight per lane, when x is 16-byte aligned; scalar loads for the rest.
template <class F>
__device__ __forceinline__ void pct_visit(const float* __restrict__ x, int64_t n, F&& f) {
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t S = (int64_t)gridDim.x * 256;
  int64_t head = 0;
  if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
    const float4* __restrict__ x4 = (const float4*)x;
    const int64_t n4 = n >> 2;
    int64_t i = tid;
    for (; i + 3 * S < n4; i += 4 * S) {
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = x4[i + u * S];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        f(v[u].x);
        f(v[u].y);
        f(v[u].z);
        f(v[u].w);
      }
    }
    for (; i < n4; i += S) {
      const float4 v = x4[i];
      f(v.x);
      f(v.y);
      f(v.z);
      f(v.w);
    }
    head = n4 << 2;
  }
  for (int64_t i = head + tid; i < n; i += S) f(x[i]);
}

// Level-1 histogram (bits 31..21), valid count and FP64 moments.  Each
// thread sums its values' offsets from its own first value (shifted sums, no
// cancellation for data far from zero) and turns them into (n, mean, m2);
// waves and the workgroup merge those in a fixed order into part[block], and
// k_pct_finish merges the workgroups in block order: the same bits on every
// run, like the reference's two-pass std to a few ulps.
__global__ void __launch_bounds__(256)
k_pct_hist1(const float* __restrict__ x, int64_t n, int flags, PctWs* __restrict__ ws) {
  // one sub-histogram per wave: the level-1 bins (exponent and two mantissa
  // bits) are few and crowded, and waves then never contend for one address
  __shared__ unsigned int hist[4][NB1];
  __shared__ double red[3][4];
  for (int b = threadIdx.x; b < 4 * NB1; b += 256) hist[b / NB1][b % NB1] = 0;
  __syncthreads();
  unsigned int* mine = hist[threadIdx.x >> 6];
  double shift = 0.0, c = 0.0, s1 = 0.0, s2 = 0.0;
  pct_visit(x, n, [&](float v) {
    if (!pct_valid(v, flags)) return;
    atomicAdd(&mine[pct_key(v, flags) >> 21], 1u);
    if (c == 0.0) shift = (double)v;
    const double d = (double)v - shift;
    c += 1.0;
    s1 += d;
    s2 += d * d;
  });
  Moments m{c, c > 0.0 ? shift + s1 / c : 0.0, c > 0.0 ? fmax(s2 - s1 * (s1 / c), 0.0) : 0.0};
  for (int off = 32; off > 0; off >>= 1) {
    const Moments o{__shfl_down(m.n, off, 64), __shfl_down(m.mean, off, 64),
                    __shfl_down(m.m2, off, 64)};
    m = merge_moments(m, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = m.n;
    red[1][threadIdx.x >> 6] = m.mean;
    red[2][threadIdx.x >> 6] = m.m2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Moments b{red[0][0], red[1][0], red[2][0]};
    for (int w = 1; w < 4; w++) b = merge_moments(b, Moments{red[0][w], red[1][w], red[2][w]});
    ws->part[blockIdx.x][0] = b.n;
    ws->part[blockIdx.x][1] = b.mean;
    ws->part[blockIdx.x][2] = b.m2;
  }
  for (int b = threadIdx.x; b < NB1; b += 256) {
    const unsigned int t = hist[0][b] + hist[1][b] + hist[2][b] + hist[3][b];
    if (t) atomicAdd(&ws->h1[b], (unsigned long long)t);
  }
}

// Level-2/3 histograms restricted to each target's current prefix (targets
// sharing a prefix are histogrammed once, under the first of them).
template <int LEVEL>
__global__ void __launch_bounds__(256)
k_pct_hist(const float* __restrict__ x, int64_t n, int flags, PctWs* __restrict__ ws) {
  constexpr int NB = (LEVEL == 2) ? NB2 : NB3;
  constexpr unsigned int MASK = (LEVEL == 2) ? 0xFFE00000u : 0xFFFFFC00u;
  constexpr int SHIFT = (LEVEL == 2) ? 10 : 0;
  constexpr unsigned int BM = (LEVEL == 2) ? 0x7FFu : 0x3FFu;
  __shared__ unsigned int hist[PCT_MAX_T][NB];
  const int nt = ws->st.n_targets;
  unsigned int pre[PCT_MAX_T];
  bool act[PCT_MAX_T];
  for (int t = 0; t < PCT_MAX_T; t++) {
    pre[t] = (t < nt) ? ws->st.prefix[t] : 0xFFFFFFFFu;
    bool dup = false;
    for (int u = 0; u < t; u++) dup |= (pre[u] == pre[t]);
    act[t] = t < nt && !dup;
  }
  for (int t = 0; t < nt; t++)
    for (int b = threadIdx.x; b < NB; b += 256) hist[t][b] = 0;
  __syncthreads();
  pct_visit(x, n, [&](float v) {
    if (!pct_valid(v, flags)) return;
    const unsigned int k = pct_key(v, flags);
    const unsigned int km = k & MASK;
#pragma unroll
    for (int t = 0; t < PCT_MAX_T; t++)
      if (act[t] && km == pre[t]) atomicAdd(&hist[t][(k >> SHIFT) & BM], 1u);
  });
  __syncthreads();
  for (int t = 0; t < nt; t++)
    for (int b = threadIdx.x; b < NB; b += 256)
      if (hist[t][b]) {
        if (LEVEL == 2)
          atomicAdd(&ws->h2[t][b], (unsigned long long)hist[t][b]);
        else
          atomicAdd(&ws->h3[t][b], (unsigned long long)hist[t][b]);
      }
}

// Find, in a histogram of nb bins, the bin holding 0-based rank `rem`;
// returns the bin and rewrites rem as the rank inside it.  256 threads.
__device__ int find_bin(const unsigned long long* h, int nb, unsigned long long* rem) {
  __shared__ unsigned long long part[256];
  __shared__ int found;
  __shared__ unsigned long long found_rem;
  const int per = nb / 256;
  unsigned long long loc = 0;
  for (int b = 0; b < per; b++) loc += h[threadIdx.x * per + b];
  part[threadIdx.x] = loc;
  if (threadIdx.x == 0) found = -1;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long acc = 0;
    for (int t = 0; t < 256; t++) {
      const unsigned long long v = part[t];
      part[t] = acc;
      acc += v;
    }
  }
  __syncthreads();
  unsigned long long cum = part[threadIdx.x];
  const unsigned long long r = *rem;
  for (int b = 0; b < per; b++) {
    const unsigned long long c = h[threadIdx.x * per + b];
    if (c > 0 && r >= cum && r < cum + c) {
      found = threadIdx.x * per + b;
      found_rem = r - cum;
    }
    cum += c;
  }
  __syncthreads();
  const int f = found < 0 ? 0 : found;
  if (threadIdx.x == 0 && found >= 0) *rem = found_rem;
  __syncthreads();
  return f;
}

// Level-1 selection: ranks from the valid count, then bins.
__global__ void __launch_bounds__(256)
k_pct_select1(PctWs* __restrict__ ws, const double* __restrict__ q, int nq) {
  __shared__ unsigned long long total;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  unsigned long long loc = 0;
  for (int b = threadIdx.x; b < NB1; b += 256) loc += ws->h1[b];
  atomicAdd(&total, loc);
  __syncthreads();
  const unsigned long long n = total;
  if (threadIdx.x == 0) {
    ws->st.n_valid = n;
    ws->st.n_targets = 2 * nq;
    for (int j = 0; j < nq; j++) {
      const double qq = q[j] / 100.0;
      const double idx = n > 0 ? qq * ((double)n - 1.0) : 0.0;
      ws->st.idx[j] = idx;
      unsigned long long lo = n > 0 ? (unsigned long long)floor(idx) : 0ull;
      if (n > 0 && lo > n - 1) lo = n - 1;
      unsigned long long hi = lo + 1;
      if (n > 0 && hi > n - 1) hi = n - 1;
      ws->st.rank[2 * j] = lo;
      ws->st.rank[2 * j + 1] = hi;
      ws->st.rem[2 * j] = lo;
      ws->st.rem[2 * j + 1] = hi;
    }
  }
  __syncthreads();
  if (n == 0) return;
  for (int t = 0; t < 2 * nq; t++) {
    const int b = find_bin(ws->h1, NB1, &ws->st.rem[t]);
    if (threadIdx.x == 0) ws->st.prefix[t] = (unsigned int)b << 21;
    __syncthreads();
  }
}

template <int LEVEL>
__global__ void __launch_bounds__(256) k_pct_select(PctWs* __restrict__ ws) {
  const unsigned long long n = ws->st.n_valid;
  if (n == 0) return;
  const int nt = ws->st.n_targets;
  // snapshot: the histogram pass saw these prefixes
  __shared__ unsigned int pre[PCT_MAX_T];
  if (threadIdx.x < PCT_MAX_T) pre[threadIdx.x] = ws->st.prefix[threadIdx.x];
  __syncthreads();
  for (int t = 0; t < nt; t++) {
    // targets sharing a prefix were histogrammed once, under the first index
    int src = t;
    for (int u = 0; u < t; u++)
      if (pre[u] == pre[t] && src == t) src = u;
    __syncthreads();
    if (LEVEL == 2) {
      const int b = find_bin(ws->h2[src], NB2, &ws->st.rem[t]);
      if (threadIdx.x == 0) ws->st.prefix[t] |= (unsigned int)b << 10;
    } else {
      const int b = find_bin(ws->h3[src], NB3, &ws->st.rem[t]);
      if (threadIdx.x == 0) ws->st.key[t] = ws->st.prefix[t] | (unsigned int)b;
    }
    __syncthreads();
  }
}

__global__ void k_pct_finish(PctWs* __restrict__ ws, int nq, int flags, double* __restrict__ out,
                             int64_t* __restrict__ n_valid, double* __restrict__ moments,
                             int n_blocks) {
  // the workgroups' moments: lane t merges its run of blocks in order, lane 0
  // then the 64 runs in lane order (a fixed order: the same bits every run)
  Moments m{0.0, 0.0, 0.0};
  if (moments) {
    const int per = (n_blocks + 63) / 64;
    const int b0 = threadIdx.x * per, b1 = min(n_blocks, b0 + per);
    for (int b = b0; b < b1; b++)
      m = merge_moments(m, Moments{ws->part[b][0], ws->part[b][1], ws->part[b][2]});
    Moments all = m;
    for (int t = 1; t < 64; t++) {
      const Moments o{__shfl(m.n, t, 64), __shfl(m.mean, t, 64), __shfl(m.m2, t, 64)};
      all = merge_moments(all, o);
    }
    m = all;
  }
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long n = ws->st.n_valid;
  for (int j = 0; j < nq; j++) {
    if (n == 0) {
      out[j] = __builtin_nan("");
      continue;
    }
    const double idx = ws->st.idx[j];
    const double lo = floor(idx);
    const double w = idx - lo;
    const double a = (double)pct_value(ws->st.key[2 * j], flags);
    const double b = (double)pct_value(ws->st.key[2 * j + 1], flags);
    out[j] = a * (1.0 - w) + b * w;
  }
  if (n_valid) *n_valid = (int64_t)n;
  if (moments) {
    if (n == 0) {
      moments[0] = moments[1] = __builtin_nan("");
    } else {
      moments[0] = m.mean;
      moments[1] = sqrt(m.m2 / m.n);
    }
  }
}

__global__ void k_finalize(int kind, const double* __restrict__ pct, const double* __restrict__ mom,
                           const int64_t* __restrict__ n_valid, double* __restrict__ stats) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t n = n_valid ? *n_valid : 1;
  if (kind == 0) {
    double s = 1.0;
    if (n > 0) {
      const double p = pct[0];
      if (p > 1e-9) {
        s = p;
      } else {
        const double sd = mom ? mom[1] : 0.0;
        s = (sd > 1e-9) ? sd : 1.0;
      }
    }
    stats[0] = s;
    stats[1] = 0.0;
  } else {
    if (n <= 0) {
      stats[0] = 0.0;
      stats[1] = 0.0;
      return;
    }
    const double lo = pct[0], hi = pct[1];
    const double sc = hi - lo;
    stats[0] = lo;
    // the pre-pass discards stats whose scale is not > 1e-9
    // (_norm_stats.py:288-289), leaving the stretch a no-op
    stats[1] = (sc > 1e-9) ? sc : 0.0;
  }
}

}  // namespace fsg

using namespace fsg;

extern "C" size_t fsg_percentile_workspace_size(int64_t n) {
  (void)n;
  return sizeof(PctWs);
}

extern "C" int fsg_percentile(const float* values, int64_t n, int32_t flags, const double* q,
                              int32_t nq, double* out, int64_t* n_valid, double* moments,
                              void* workspace, size_t workspace_bytes, void* stream) {
  if (n < 0 || (n > 0 && !values)) return fail(FSG_E_INVALID, "fsg_percentile: bad values");
  if (nq < 1 || nq > PCT_MAX_Q || !q || !out)
    return fail(FSG_E_INVALID, "fsg_percentile: 1..2 device quantiles and an output required");
  if (!workspace || workspace_bytes < sizeof(PctWs))
    return fail(FSG_E_WORKSPACE, "fsg_percentile: workspace too small");
  hipStream_t s = as_stream(stream);
  PctWs* ws = (PctWs*)workspace;
  const size_t words = sizeof(PctWs) / sizeof(unsigned long long);
  hipLaunchKernelGGL(k_pct_zero, dim3(256), dim3(256), 0, s, (unsigned long long*)ws, words);
  FSG_CHECK_LAUNCH("k_pct_zero");
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(PCT_BLOCKS, (n + 255) / 256));
  hipLaunchKernelGGL(k_pct_hist1, dim3(blocks), dim3(256), 0, s, values, n, flags, ws);
  FSG_CHECK_LAUNCH("k_pct_hist1");
  hipLaunchKernelGGL(k_pct_select1, dim3(1), dim3(256), 0, s, ws, q, nq);
  FSG_CHECK_LAUNCH("k_pct_select1");
  hipLaunchKernelGGL(k_pct_hist<2>, dim3(blocks), dim3(256), 0, s, values, n, flags, ws);
  FSG_CHECK_LAUNCH("k_pct_hist2");
  hipLaunchKernelGGL(k_pct_select<2>, dim3(1), dim3(256), 0, s, ws);
  FSG_CHECK_LAUNCH("k_pct_select2");
  hipLaunchKernelGGL(k_pct_hist<3>, dim3(blocks), dim3(256), 0, s, values, n, flags, ws);
  FSG_CHECK_LAUNCH("k_pct_hist3");
  hipLaunchKernelGGL(k_pct_select<3>, dim3(1), dim3(256), 0, s, ws);
  FSG_CHECK_LAUNCH("k_pct_select3");
  hipLaunchKernelGGL(k_pct_finish, dim3(1), dim3(64), 0, s, ws, nq, flags, out, n_valid, moments,
                     blocks);
  FSG_CHECK_LAUNCH("k_pct_finish");
  return FSG_OK;
}

extern "C" int fsg_finalize_stats(int32_t kind, const double* pct, const double* moments,
                                  const int64_t* n_valid, double* stats, void* stream) {
  if ((kind != 0 && kind != 1) || !pct || !stats)
    return fail(FSG_E_INVALID, "fsg_finalize_stats: bad arguments");
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64), 0, as_stream(stream), kind, pct, moments,
                     n_valid, stats);
  FSG_CHECK_LAUNCH("k_finalize");
  return FSG_OK;
}

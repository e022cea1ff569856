// Library plumbing: error reporting, validation, numpy-compatible host math,
// NoData preparation, region copies and block-mean decimation.
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "fsg_common.hpp"

namespace fsg {

static thread_local std::string g_last_error;
static thread_local int64_t g_launches = 0;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return FSG_OK;
  return fail(FSG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void count_launch(int n) { g_launches += n; }

// ------------------------------------------------------------ kernel timing
struct EvPair {
  hipEvent_t a, b;
};
static std::mutex g_tmu;
static std::atomic<bool> g_timing{false};
static std::map<std::string, std::vector<EvPair>> g_events;
static std::map<std::string, hipEvent_t> g_open;

bool timing_on() { return g_timing.load(std::memory_order_relaxed); }

void timing_start(const char* name, hipStream_t s) {
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  hipEventRecord(e, s);
  std::lock_guard<std::mutex> lk(g_tmu);
  g_open[name] = e;
}

void timing_stop(const char* name, hipStream_t s) {
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  hipEventRecord(e, s);
  std::lock_guard<std::mutex> lk(g_tmu);
  auto it = g_open.find(name);
  if (it == g_open.end()) {
    hipEventDestroy(e);
    return;
  }
  g_events[name].push_back(EvPair{it->second, e});
  g_open.erase(it);
}

int validate_window(const fsg_window* w, const char* who) {
  if (!w) return fail(FSG_E_INVALID, std::string(who) + ": window is NULL");
  if (w->h < 0 || w->w < 0 || (w->h > 0 && w->w > 0 && w->data == nullptr))
    return fail(FSG_E_INVALID, std::string(who) + ": bad window");
  if (w->ld < w->w) return fail(FSG_E_INVALID, std::string(who) + ": window ld < w");
  return FSG_OK;
}

int validate_region(const fsg_window* w, const fsg_region* r, const char* who) {
  if (!r) return fail(FSG_E_INVALID, std::string(who) + ": region is NULL");
  if (r->r0 < 0 || r->c0 < 0 || r->h < 0 || r->w < 0 || (int64_t)r->r0 + r->h > w->h ||
      (int64_t)r->c0 + r->w > w->w)
    return fail(FSG_E_INVALID, std::string(who) + ": region outside window");
  return FSG_OK;
}

int validate_output(const fsg_output* o, const fsg_region* r, const char* who) {
  if (!o) return FSG_OK;
  if (r->h > 0 && r->w > 0 && o->data == nullptr)
    return fail(FSG_E_INVALID, std::string(who) + ": output data is NULL");
  if (o->ld < r->w) return fail(FSG_E_INVALID, std::string(who) + ": output ld < region w");
  if (o->dtype < FSG_OUT_F32 || o->dtype > FSG_OUT_U8)
    return fail(FSG_E_INVALID, std::string(who) + ": bad output dtype");
  if (o->post_kind < FSG_POST_NONE || o->post_kind > FSG_POST_STRETCH)
    return fail(FSG_E_INVALID, std::string(who) + ": bad post_kind");
  return FSG_OK;
}

double np_pairwise_sum(const double* a, int64_t n) {
  if (n < 8) {
    double res = 0.;
    for (int64_t i = 0; i < n; i++) res += a[i];
    return res;
  } else if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; j++) r[j] = a[j];
    int64_t i;
    for (i = 8; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; j++) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += a[i];
    return res;
  } else {
    int64_t n2 = n / 2;
    n2 -= n2 % 8;
    return np_pairwise_sum(a, n2) + np_pairwise_sum(a + n2, n - n2);
  }
}

void gaussian_weights(double sigma, int radius, double* w) {
  const double sigma2 = sigma * sigma;
  const int n = 2 * radius + 1;
  std::vector<double> phi(n);
  for (int i = 0; i < n; i++) {
    const double x = (double)(i - radius);
    phi[i] = std::exp(-0.5 / sigma2 * (x * x));
  }
  const double s = np_pairwise_sum(phi.data(), n);
  for (int i = 0; i < n; i++) w[i] = phi[i] / s;
}

}  // namespace fsg

using namespace fsg;

// ------------------------------------------------------------------ kernels

// Launches that map rows to blockIdx.y cap gridDim.y here and stride over the
// rest (gridDim.y is limited to 65535).
constexpr int kMaxGridRows = 32768;

// Four columns per lane: one 16-byte load per row (dword-aligned, any row
// stride), a store only for the cells that match the sentinel.
__global__ void k_nodata_to_nan(float* __restrict__ ras, int64_t ld, int h, int w,
                                int has_nodata, double nodata) {
  const int c = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= w || !has_nodata) return;
  for (int r = blockIdx.y; r < h; r += gridDim.y) {   // grid-stride over rows
    float* p = ras + (int64_t)r * ld + c;
    float v[4];
    if (c + 3 < w) {
      float4 q;
      __builtin_memcpy(&q, p, sizeof(q));
      v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++) v[k] = c + k < w ? p[k] : __builtin_nanf("");
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (!isnan(v[k]) && fabs((double)v[k] - nodata) <= 1e-6) p[k] = __builtin_nanf("");
  }
}

// Counts the NaN cells of each window: blockIdx.z = window, blockIdx.y = row
// group, x = columns (no per-element division).
struct WinChunk {
  int4 w[32];  // r0, c0, h, w
};

__global__ void __launch_bounds__(256)
k_count_nan(const float* __restrict__ ras, int64_t ld, WinChunk wins, int n_windows,
            unsigned long long* __restrict__ counts) {
  const int wi = blockIdx.z;
  if (wi >= n_windows) return;
  const int4 wd = wins.w[wi];
  unsigned long long local = 0;
  // 16-byte loads (dword-aligned: windows at any column, any row stride)
  const int w4 = wd.w & ~3;
  for (int r = blockIdx.y; r < wd.z; r += gridDim.y) {
    const float* row = ras + (int64_t)(wd.x + r) * ld + wd.y;
    for (int c = 4 * (blockIdx.x * 256 + threadIdx.x); c < w4; c += 4 * gridDim.x * 256) {
      float4 q;
      __builtin_memcpy(&q, row + c, sizeof(q));
      local += (isnan(q.x) ? 1ull : 0ull) + (isnan(q.y) ? 1ull : 0ull) +
               (isnan(q.z) ? 1ull : 0ull) + (isnan(q.w) ? 1ull : 0ull);
    }
    for (int c = w4 + blockIdx.x * 256 + threadIdx.x; c < wd.w; c += gridDim.x * 256)
      local += isnan(row[c]) ? 1ull : 0ull;
  }
  for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(&counts[wi], local);
}

__global__ void k_copy_region(const float* __restrict__ src, int64_t ld, int r0, int c0,
                              int h, int w, float* __restrict__ dst) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= w) return;
  for (int r = blockIdx.y; r < h; r += gridDim.y)
    dst[(int64_t)r * w + c] = src[(int64_t)(r0 + r) * ld + c0 + c];
}

__global__ void k_encode(const float* __restrict__ src, int64_t ld, int r0, int c0, int h, int w,
                         OutDev out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= w) return;
  const PostK pp = load_post(out);
  for (int r = blockIdx.y; r < h; r += gridDim.y)
    store_out(out, r, c, apply_post(pp, src[(int64_t)(r0 + r) * ld + c0 + c]));
}

// NaN-aware f x f block mean: sum in f64 (exact for cell sizes used here),
// rounded once to f32, then f32 total / f32 count (_nan_utils.py:638-643).
__global__ void k_block_mean(const float* __restrict__ src, int64_t ld, int h, int w, int f,
                             float* __restrict__ out, int64_t ldo, int oh, int ow) {
  const int oc = blockIdx.x * blockDim.x + threadIdx.x;
  if (oc >= ow) return;
  for (int orow = blockIdx.y; orow < oh; orow += gridDim.y) {   // grid-stride over rows
    const int r0 = orow * f, c0 = oc * f;
    const int r1 = min(r0 + f, h), c1 = min(c0 + f, w);
    double tot = 0.0;
    int cnt = 0;
    for (int r = r0; r < r1; r++) {
      const float* row = src + (int64_t)r * ld;
      for (int c = c0; c < c1; c++) {
        const float v = row[c];
        if (isfinite(v)) {
          tot += (double)v;
          cnt++;
        }
      }
    }
    out[(int64_t)orow * ldo + oc] = cnt > 0 ? (float)tot / (float)cnt : __builtin_nanf("");
  }
}

// Power-of-two factor: lanes sum their column over the f rows, then groups of
// f lanes reduce by shuffles (coalesced reads).
__global__ void __launch_bounds__(256)
k_block_mean_p2(const float* __restrict__ src, int64_t ld, int h, int w, int f,
                float* __restrict__ out, int64_t ldo, int oh, int ow) {
  // grid-stride over output rows (block-uniform, so the shuffles see every lane)
  for (int orow = blockIdx.y; orow < oh; orow += gridDim.y) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int r0 = orow * f, r1 = min(r0 + f, h);
  double tot = 0.0;
  int cnt = 0;
  if (c < w) {
    for (int r = r0; r < r1; r++) {
      const float v = src[(int64_t)r * ld + c];
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
    if (oc < ow && orow < oh)
      out[(int64_t)orow * ldo + oc] = cnt > 0 ? (float)tot / (float)cnt : __builtin_nanf("");
  }
  }
}

// Power-of-two factor >= 4 on 16-byte aligned rows: each lane loads a float4
// (four columns of one cell) from each of the f rows -- f independent 16-byte
// loads in flight -- and groups of f/4 lanes reduce by shuffles.
__global__ void __launch_bounds__(256)
k_block_mean_p2v(const float* __restrict__ src, int64_t ld, int h, int w, int f,
                 float* __restrict__ out, int64_t ldo, int oh, int ow) {
  // grid-stride over output rows (block-uniform, so the shuffles see every lane)
  for (int orow = blockIdx.y; orow < oh; orow += gridDim.y) {
  const int c = (blockIdx.x * 256 + threadIdx.x) * 4;
  const int r0 = orow * f, r1 = min(r0 + f, h);
  double tot = 0.0;
  int cnt = 0;
  if (c < w) {
    for (int r = r0; r < r1; r++) {
      const float* p = src + (int64_t)r * ld + c;
      float v[4];
      if (c + 3 < w) {
        const float4 q = *(const float4*)p;
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
      } else {
        for (int k = 0; k < 4; k++) v[k] = c + k < w ? p[k] : __builtin_nanf("");
      }
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (isfinite(v[k])) {
          tot += (double)v[k];
          cnt++;
        }
    }
  }
  for (int off = 1; off < f / 4; off <<= 1) {
    tot += __shfl_down(tot, off, 64);
    cnt += __shfl_down(cnt, off, 64);
  }
  if ((threadIdx.x & (f / 4 - 1)) == 0) {
    const int oc = c / f;
    if (oc < ow && orow < oh)
      out[(int64_t)orow * ldo + oc] = cnt > 0 ? (float)tot / (float)cnt : __builtin_nanf("");
  }
  }
}

// One read of a raster for its NaN count and NaN-aware cell means at up to
// three power-of-two factors 4..16 (fsg_raster_scan).  A block covers 1024
// source columns (four per lane) of F source rows, F the largest factor:
// each lane keeps float64 sums and counts of its four columns per 4-row
// group, then a cell of factor f adds f/4 groups and f/4 lanes.  Cell sums
// are exact in float64, so every cell equals k_block_mean_p2v's (and the
// topousm decimation's) bit for bit.
constexpr int SCAN_MAX_OUT = 3;
struct ScanOut {
  float* p;
  int64_t ld;
  int f, oh, ow;
};
struct ScanArgs {
  const float* x;
  int64_t ld;
  int h, w;
  int n_out;
  ScanOut o[SCAN_MAX_OUT];
  unsigned long long* nan_count;
  int vec;
};

template <int F>
__global__ void __launch_bounds__(256) k_raster_scan(ScanArgs A) {
  constexpr int G = F / 4;  // 4-row groups per band
  const int c = (blockIdx.x * 256 + threadIdx.x) * 4;
  unsigned long long nn = 0;
  // bands of F rows, grid-stride over gridDim.y (block-uniform)
  for (int rb = blockIdx.y * F; rb < A.h; rb += gridDim.y * F) {
    double sg[G];
    int cg[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      sg[g] = 0.0;
      cg[g] = 0;
    }
    if (c < A.w) {
      // every load of the band first (F float4 per lane in flight), then the sums
      float v[F][4];
#pragma unroll
      for (int e = 0; e < F; e++) {
        const int r = min(rb + e, A.h - 1);
        const float* p = A.x + (int64_t)r * A.ld + c;
        if (A.vec && c + 3 < A.w) {
          const float4 q = *(const float4*)p;
          v[e][0] = q.x; v[e][1] = q.y; v[e][2] = q.z; v[e][3] = q.w;
        } else {
#pragma unroll
          for (int k = 0; k < 4; k++) v[e][k] = c + k < A.w ? p[k] : 0.0f;
        }
      }
#pragma unroll
      for (int e = 0; e < F; e++) {
        if (rb + e < A.h) {
#pragma unroll
          for (int k = 0; k < 4; k++) {
            if (c + k < A.w) {
              const float x = v[e][k];
              if (isnan(x)) nn++;
              if (isfinite(x)) {
                sg[e / 4] += (double)x;
                cg[e / 4]++;
              }
            }
          }
        }
      }
    }
    for (int k = 0; k < A.n_out; k++) {
      const ScanOut& O = A.o[k];
      const int gpc = O.f / 4;  // groups (and lanes) per cell
      for (int q = 0; q < G / gpc; q++) {
        double tot = 0.0;
        int cnt = 0;
#pragma unroll
        for (int g = 0; g < G; g++)
          if (g / gpc == q) {
            tot += sg[g];
            cnt += cg[g];
          }
        for (int off = 1; off < gpc; off <<= 1) {
          tot += __shfl_down(tot, off, 64);
          cnt += __shfl_down(cnt, off, 64);
        }
        const int orow = (rb + q * O.f) / O.f;
        const int oc = c / O.f;
        if ((threadIdx.x & (gpc - 1)) == 0 && oc < O.ow && orow < O.oh && c < A.w)
          O.p[(int64_t)orow * O.ld + oc] = cnt > 0 ? (float)tot / (float)cnt : __builtin_nanf("");
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) nn += __shfl_down(nn, off, 64);
  if ((threadIdx.x & 63) == 0 && nn) atomicAdd(A.nan_count, nn);
}

// -------------------------------------------------------------------- API

extern "C" int fsg_abi_version(void) { return FSG_ABI_VERSION; }

extern "C" const char* fsg_last_error(void) { return g_last_error.c_str(); }

extern "C" int64_t fsg_launch_count(void) { return g_launches; }

extern "C" int fsg_timing_enable(int32_t on) {
  g_timing.store(on != 0);
  return FSG_OK;
}

// Sum of the recorded launch durations of `name`, their count and, when
// union_ms is given, the length of the union of their [start, end] intervals
// (launches on several streams may overlap; times relative to the first start).
static int timing_collect(const char* name, double* total_ms, int64_t* count,
                          double* union_ms) {
  std::vector<EvPair> evs;
  {
    std::lock_guard<std::mutex> lk(g_tmu);
    auto it = g_events.find(name);
    if (it != g_events.end()) {
      evs.swap(it->second);
      g_events.erase(it);
    }
  }
  double tot = 0.0;
  std::vector<std::pair<double, double>> iv;
  hipError_t err = hipSuccess;
  for (auto& p : evs) {
    float ms = 0.f, t0 = 0.f, t1 = 0.f;
    hipError_t e = hipEventSynchronize(p.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.a, p.b);
    if (e == hipSuccess && union_ms) {
      e = hipEventElapsedTime(&t0, evs[0].a, p.a);
      if (e == hipSuccess) e = hipEventElapsedTime(&t1, evs[0].a, p.b);
      iv.emplace_back((double)t0, (double)t1);
    }
    if (e != hipSuccess && err == hipSuccess) err = e;
    tot += ms;
  }
  for (auto& p : evs) {
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  if (err != hipSuccess) return check_hip(err, "fsg_timing_read");
  *total_ms = tot;
  *count = (int64_t)evs.size();
  if (union_ms) {
    std::sort(iv.begin(), iv.end());
    double u = 0.0, a = 0.0, b = -1.0;
    for (auto& x : iv) {
      if (x.first > b) {
        if (b > a) u += b - a;
        a = x.first;
        b = x.second;
      } else if (x.second > b) {
        b = x.second;
      }
    }
    if (b > a) u += b - a;
    *union_ms = u;
  }
  return FSG_OK;
}

extern "C" int fsg_timing_read(const char* name, double* total_ms, int64_t* count) {
  if (!name || !total_ms || !count) return fail(FSG_E_INVALID, "fsg_timing_read: NULL argument");
  return timing_collect(name, total_ms, count, nullptr);
}

extern "C" int fsg_timing_read_union(const char* name, double* total_ms, int64_t* count,
                                     double* union_ms) {
  if (!name || !total_ms || !count || !union_ms)
    return fail(FSG_E_INVALID, "fsg_timing_read_union: NULL argument");
  return timing_collect(name, total_ms, count, union_ms);
}

extern "C" int fsg_nodata_to_nan(float* raster, int64_t ld, int32_t h, int32_t w,
                                 int32_t has_nodata, double nodata,
                                 const int32_t* windows, int32_t n_windows,
                                 int64_t* counts, void* stream) {
  if (h < 0 || w < 0 || ld < w || (h > 0 && w > 0 && !raster))
    return fail(FSG_E_INVALID, "fsg_nodata_to_nan: bad raster");
  if (n_windows < 0 || (n_windows > 0 && (!windows || !counts)))
    return fail(FSG_E_INVALID, "fsg_nodata_to_nan: bad windows");
  hipStream_t s = as_stream(stream);
  if (has_nodata && h > 0 && w > 0) {
    dim3 blk(256), grd((w + 1023) / 1024, std::min(h, kMaxGridRows));
    if (std::isnan(nodata)) has_nodata = 0;  // NaN sentinel: isnan only
    if (has_nodata) {
      hipLaunchKernelGGL(k_nodata_to_nan, grd, blk, 0, s, raster, ld, h, w, has_nodata, nodata);
      FSG_CHECK_LAUNCH("k_nodata_to_nan");
    }
  }
  for (int base = 0; base < n_windows; base += 32) {
    WinChunk chunk{};
    const int n = std::min(32, n_windows - base);
    for (int i = 0; i < n; i++) {
      const int32_t* q = windows + 4 * (base + i);
      if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[3] < 0 || q[0] + q[2] > h || q[1] + q[3] > w)
        return fail(FSG_E_INVALID, "fsg_nodata_to_nan: window outside raster");
      chunk.w[i] = make_int4(q[0], q[1], q[2], q[3]);
    }
    int maxw = 1, maxh = 1;
    for (int i = 0; i < n; i++) {
      maxw = std::max(maxw, chunk.w[i].w);
      maxh = std::max(maxh, chunk.w[i].z);
    }
    dim3 blk(256), grd(std::min(4, (maxw + 1023) / 1024), std::min(512, maxh), n);
    hipLaunchKernelGGL(k_count_nan, grd, blk, 0, s, raster, ld, chunk, n,
                       (unsigned long long*)(counts + base));
    FSG_CHECK_LAUNCH("k_count_nan");
  }
  return FSG_OK;
}

extern "C" int fsg_copy_region(const fsg_window* win, const fsg_region* reg, float* dst,
                               void* stream) {
  int rc = validate_window(win, "fsg_copy_region");
  if (rc) return rc;
  rc = validate_region(win, reg, "fsg_copy_region");
  if (rc) return rc;
  if (reg->h == 0 || reg->w == 0) return FSG_OK;
  if (!dst) return fail(FSG_E_INVALID, "fsg_copy_region: dst is NULL");
  dim3 blk(256), grd((reg->w + 255) / 256, std::min(reg->h, kMaxGridRows));
  hipLaunchKernelGGL(k_copy_region, grd, blk, 0, as_stream(stream), win->data, win->ld,
                     reg->r0, reg->c0, reg->h, reg->w, dst);
  FSG_CHECK_LAUNCH("k_copy_region");
  return FSG_OK;
}

extern "C" int fsg_encode(const fsg_window* src, const fsg_region* reg, const fsg_output* out,
                          void* stream) {
  int rc = validate_window(src, "fsg_encode");
  if (rc) return rc;
  rc = validate_region(src, reg, "fsg_encode");
  if (rc) return rc;
  if (!out) return fail(FSG_E_INVALID, "fsg_encode: output NULL");
  if ((rc = validate_output(out, reg, "fsg_encode"))) return rc;
  if (reg->h == 0 || reg->w == 0) return FSG_OK;
  dim3 blk(256), grd((reg->w + 255) / 256, std::min(reg->h, kMaxGridRows));
  hipLaunchKernelGGL(k_encode, grd, blk, 0, as_stream(stream), src->data, src->ld, reg->r0,
                     reg->c0, reg->h, reg->w, make_out(out));
  FSG_CHECK_LAUNCH("k_encode");
  return FSG_OK;
}

extern "C" int fsg_raster_scan(const fsg_window* raster, int32_t n_out,
                               const fsg_scan_out* outs, uint64_t* nan_count, void* stream) {
  int rc = validate_window(raster, "fsg_raster_scan");
  if (rc) return rc;
  if (raster->h == 0 || raster->w == 0) return fail(FSG_E_INVALID, "fsg_raster_scan: empty");
  if (!nan_count) return fail(FSG_E_INVALID, "fsg_raster_scan: nan_count NULL");
  if (n_out < 0 || n_out > SCAN_MAX_OUT || (n_out > 0 && !outs))
    return fail(FSG_E_INVALID, "fsg_raster_scan: 0..3 outputs");
  ScanArgs A{};
  A.x = raster->data;
  A.ld = raster->ld;
  A.h = raster->h;
  A.w = raster->w;
  A.n_out = n_out;
  A.nan_count = (unsigned long long*)nan_count;
  A.vec = (raster->ld % 4 == 0 && (reinterpret_cast<uintptr_t>(raster->data) & 15) == 0) ? 1 : 0;
  int F = 4;
  for (int k = 0; k < n_out; k++) {
    const int f = outs[k].factor;
    if (f != 4 && f != 8 && f != 16)
      return fail(FSG_E_UNSUPPORTED, "fsg_raster_scan: factors are 4, 8 or 16");
    const int oh = (raster->h + f - 1) / f, ow = (raster->w + f - 1) / f;
    if (!outs[k].out || outs[k].ld_out < ow)
      return fail(FSG_E_INVALID, "fsg_raster_scan: bad output");
    A.o[k] = ScanOut{outs[k].out, outs[k].ld_out, f, oh, ow};
    F = std::max(F, f);
  }
  hipStream_t s = as_stream(stream);
  dim3 blk(256), grd((raster->w + 1023) / 1024, std::min((raster->h + F - 1) / F, 32768));
  if (F == 4)
    hipLaunchKernelGGL(k_raster_scan<4>, grd, blk, 0, s, A);
  else if (F == 8)
    hipLaunchKernelGGL(k_raster_scan<8>, grd, blk, 0, s, A);
  else
    hipLaunchKernelGGL(k_raster_scan<16>, grd, blk, 0, s, A);
  FSG_CHECK_LAUNCH("k_raster_scan");
  return FSG_OK;
}

// Area-weighted NaN-aware average of a source raster onto an oh x ow grid
// (GDAL's Resampling.average for any ratio).  Output cell (i, j) covers source
// rows [i*ry, (i+1)*ry) and columns [j*rx, (j+1)*rx), ry = h/oh, rx = w/ow;
// each source pixel weighs the length of its overlap on each axis.  The sums
// run in float64 (row by row: the row's weighted sum, then times the row
// weight); the result is (float)(sum(w*v) / sum(w)) over finite pixels, NaN
// when none.  For an integer ratio every weight is 1.0, so a cell is the
// float64 mean of its finite pixels rounded once.
__device__ __forceinline__ void area_span(int i, double ratio, int n, int& k0, int& k1,
                                          double& a0, double& a1) {
  a0 = (double)i * ratio;
  a1 = (double)(i + 1) * ratio;
  if (a1 > (double)n) a1 = (double)n;
  k0 = (int)(a0 + 1e-8);
  k1 = (int)ceil(a1 - 1e-8);
  if (k1 > n) k1 = n;
  if (k1 <= k0) k1 = k0 + 1;
}

__device__ __forceinline__ double area_weight(int k, double a0, double a1) {
  const double lo = a0 > (double)k ? a0 : (double)k;
  const double hi = a1 < (double)(k + 1) ? a1 : (double)(k + 1);
  return hi > lo ? hi - lo : 0.0;
}

__global__ void __launch_bounds__(256)
k_area_average(const float* __restrict__ src, int64_t ld, int h, int w, float* __restrict__ out,
               int64_t ldo, int oh, int ow, double ry, double rx) {
  const int oc = blockIdx.x * 64 + (threadIdx.x & 63);
  if (oc >= ow) return;
  int kx0, kx1;
  double x0, x1;
  area_span(oc, rx, w, kx0, kx1, x0, x1);
  for (int orow = blockIdx.y * 4 + (threadIdx.x >> 6); orow < oh; orow += gridDim.y * 4) {
    int ky0, ky1;
    double y0, y1;
    area_span(orow, ry, h, ky0, ky1, y0, y1);
    double tot = 0.0, wsum = 0.0;
    for (int k = ky0; k < ky1; k++) {
      const double wy = area_weight(k, y0, y1);
      const float* row = src + (int64_t)k * ld;
      double rv = 0.0, rw = 0.0;
      for (int c = kx0; c < kx1; c++) {
        const float v = row[c];
        if (isfinite(v)) {
          const double wx = area_weight(c, x0, x1);
          rv += wx * (double)v;
          rw += wx;
        }
      }
      tot += wy * rv;
      wsum += wy * rw;
    }
    out[(int64_t)orow * ldo + oc] =
        wsum > 0.0 ? (float)(tot / wsum) : __builtin_nanf("");
  }
}

extern "C" int fsg_block_mean(const fsg_window* src, int32_t factor, float* out,
                              int64_t ld_out, void* stream) {
  int rc = validate_window(src, "fsg_block_mean");
  if (rc) return rc;
  if (factor < 1) return fail(FSG_E_INVALID, "fsg_block_mean: factor < 1");
  const int oh = std::max(1, (src->h + factor - 1) / factor);
  const int ow = std::max(1, (src->w + factor - 1) / factor);
  if (src->h == 0 || src->w == 0) return fail(FSG_E_INVALID, "fsg_block_mean: empty source");
  if (!out || ld_out < ow) return fail(FSG_E_INVALID, "fsg_block_mean: bad output");
  if (factor >= 4 && factor <= 256 && (factor & (factor - 1)) == 0 && src->ld % 4 == 0 &&
      (reinterpret_cast<uintptr_t>(src->data) & 15) == 0) {
    dim3 blk(256), grd((src->w + 1023) / 1024, std::min(oh, kMaxGridRows));
    hipLaunchKernelGGL(k_block_mean_p2v, grd, blk, 0, as_stream(stream), src->data, src->ld,
                       src->h, src->w, factor, out, ld_out, oh, ow);
    FSG_CHECK_LAUNCH("k_block_mean_p2v");
    return FSG_OK;
  }
  if (factor <= 64 && (factor & (factor - 1)) == 0) {
    dim3 blk(256), grd((src->w + 255) / 256, std::min(oh, kMaxGridRows));
    hipLaunchKernelGGL(k_block_mean_p2, grd, blk, 0, as_stream(stream), src->data, src->ld,
                       src->h, src->w, factor, out, ld_out, oh, ow);
    FSG_CHECK_LAUNCH("k_block_mean_p2");
    return FSG_OK;
  }
  dim3 blk(128), grd((ow + 127) / 128, std::min(oh, kMaxGridRows));
  hipLaunchKernelGGL(k_block_mean, grd, blk, 0, as_stream(stream), src->data, src->ld, src->h,
                     src->w, factor, out, ld_out, oh, ow);
  FSG_CHECK_LAUNCH("k_block_mean");
  return FSG_OK;
}


extern "C" int fsg_overview_average(const fsg_window* src, int32_t oh, int32_t ow, float* out,
                                    int64_t ld_out, void* stream) {
  int rc = validate_window(src, "fsg_overview_average");
  if (rc) return rc;
  if (src->h == 0 || src->w == 0) return fail(FSG_E_INVALID, "fsg_overview_average: empty source");
  if (oh < 1 || ow < 1 || oh > src->h || ow > src->w)
    return fail(FSG_E_INVALID, "fsg_overview_average: need 1 <= oh <= h and 1 <= ow <= w");
  if (!out || ld_out < ow) return fail(FSG_E_INVALID, "fsg_overview_average: bad output");
  const double ry = (double)src->h / (double)oh, rx = (double)src->w / (double)ow;
  dim3 blk(256), grd((ow + 63) / 64, std::min((oh + 3) / 4, 16384));
  hipLaunchKernelGGL(k_area_average, grd, blk, 0, as_stream(stream), src->data, src->ld, src->h,
                     src->w, out, ld_out, oh, ow, ry, rx);
  FSG_CHECK_LAUNCH("k_area_average");
  return FSG_OK;
}
